"""The product's event-driven transport loop (AdrRun) against the oracle's
restatement of the reference loop, with the oracle's step as the device:
same stop step, same ladder, same picked snapshots, same info.  Also the
campaign helpers on the host."""
import numpy as np
import pytest

from oracle import prtlb_oracle as O
from prtlb.adr import AdrRun, TransportSolver
from prtlb import campaign as C
import prtlb


class _NumpyDevice:
    """Stands in for TransportSolver: the oracle's float32 step on the host."""

    def __init__(self, mat, vel, pe, da, ns, chem, da_abio):
        self.S = O.adr_setup(mat, vel, pe, da, ns, chem, da_abio)
        S = self.S
        self.C = O.adr_initial(S)
        self.dt, self.ns, self.nd, self.shape, self.R = S.dt, S.ns, S.nd, S.m.shape, S.R
        self.ch, self.umax, self.dmax, self.biofilm = S.ch, S.umax, S.dmax, S.biofilm
        self.pe_grid, self.l_ref = S.pe_grid, S.l_ref
        self.surface_voxels = float(S.surf.sum())

    def step_n(self, k):
        for _ in range(k):
            self.C = O.adr_step(self.S, self.C)

    def pore_means(self):
        return np.array([float(self.C[j][self.S.mf].mean()) for j in range(self.ns)])

    def snapshot(self):
        class _Snap:
            def __init__(s, a):
                s.a = a

            def reshape(s, shp):
                return s

            def cpu(s):
                return s

            def numpy(s):
                return s.a
        return _Snap(np.where(self.S.mf, self.C, 0.0).astype(self.R).reshape((self.ns,) + self.shape))

    def outlet_concentration(self, vel=None):
        return {"mean": np.zeros(self.ns)}

    # ---- deferred checks: the sums "arrive" only after `lag` polls
    lag = 1
    steps_done = 0

    @property
    def n_pore(self):
        return int(self.S.mf.sum())

    _means_of = TransportSolver._means_of

    def check_async(self, keep_state=True):
        class _Host:
            def __init__(s, a):
                s.a = a

            def numpy(s):
                return s.a

        class _Event:
            def __init__(s, lag):
                s.left = lag

            def query(s):
                s.left -= 1
                return s.left < 0

            def synchronize(s):
                s.left = -1

        sums = np.array([np.add.reduce(self.C[j][self.S.mf]) for j in range(self.ns)], np.float32)
        return _Host(sums), _Event(self.lag), (self.C.copy() if keep_state else None)

    def restore(self, state):
        self.C = state.copy()


@pytest.mark.parametrize("kw", [
    dict(pe=100.0, da=1.0, n_t=5, n_species=1),                   # auto stop
    dict(pe=30.0, da=1.0, n_t=4, n_species=4, steps=300),
    dict(pe=30.0, da=1.0, n_t=11, n_species=1, steps=200),
    dict(pe=20.0, da=0.5, n_t=3, n_species=2, max_steps=700),     # cap
    dict(pe=10.0, da=0.5, n_t=2, n_species=2, steps=0)])
def test_adrrun_equals_reference_loop(kw):
    g = np.full((40, 6), O.PORE, np.uint8)
    v = np.zeros((3, 40, 6), np.float32)
    v[0] = 1.0
    kw = dict(kw)
    n_t, ns = kw.pop("n_t"), kw.pop("n_species")
    steps, max_steps = kw.pop("steps", None), kw.pop("max_steps", 100000)
    info_o = {}
    out_o, tn_o = O.solve_adr(g, v, kw["pe"], kw["da"], n_t, n_species=ns, steps=steps,
                              max_steps=max_steps, info=info_o)
    T = _NumpyDevice(g, v, kw["pe"], kw["da"], ns, None, 0.0)
    run = AdrRun(T, n_t, steps=steps, max_steps=max_steps)
    while not run.finished:
        tgt = run.next_target()
        T.step_n(tgt - run.n)
        run.arrive(tgt)
    info = {}
    out, tn = run.result(info, kw["da"], 0.0)
    assert np.array_equal(out, out_o) and np.array_equal(tn, tn_o)
    for k in ("converged", "steps", "ladder", "repeated_snapshots", "dt"):
        assert info[k] == info_o[k], k


@pytest.mark.parametrize("lag", [0, 1, 2, 3])
@pytest.mark.parametrize("kw", [
    dict(pe=100.0, da=1.0, n_t=5, n_species=1),                   # auto stop
    dict(pe=40.0, da=2.0, n_t=6, n_species=4),                    # auto stop, 4 chemicals
    dict(pe=30.0, da=1.0, n_t=4, n_species=4, steps=300),
    dict(pe=20.0, da=0.5, n_t=3, n_species=2, max_steps=700)])    # cap
def test_adrrun_deferred_checks_equal_reference_loop(kw, lag):
    """Checks decided late (their sums arrive `lag` polls after the device
    passed them, at most two left undecided, as solve_adr drives it): the
    run that settled at an earlier check is put back to that step, so out,
    t_norm and info equal the reference loop's."""
    g = np.full((40, 6), O.PORE, np.uint8)
    v = np.zeros((3, 40, 6), np.float32)
    v[0] = 1.0
    kw = dict(kw)
    n_t, ns = kw.pop("n_t"), kw.pop("n_species")
    steps, max_steps = kw.pop("steps", None), kw.pop("max_steps", 100000)
    info_o = {}
    out_o, tn_o = O.solve_adr(g, v, kw["pe"], kw["da"], n_t, n_species=ns, steps=steps,
                              max_steps=max_steps, info=info_o)
    T = _NumpyDevice(g, v, kw["pe"], kw["da"], ns, None, 0.0)
    T.lag = lag
    run = AdrRun(T, n_t, steps=steps, max_steps=max_steps)
    late = 0
    while not run.finished:
        tgt = run.next_target()
        T.step_n(tgt - run.n)
        T.steps_done = tgt
        run.arrive(tgt, defer=True)
        if not run.finished:
            late = max(late, len(run.pending))
            run.resolve(block=len(run.pending) > 2)
    info = {}
    out, tn = run.result(info, kw["da"], 0.0)
    assert np.array_equal(out, out_o) and np.array_equal(tn, tn_o)
    for k in ("converged", "steps", "ladder", "repeated_snapshots", "dt"):
        assert info[k] == info_o[k], k
    assert T.steps_done == info_o["steps"]
    if lag:
        assert late >= 1


def test_campaign_helpers():
    a = C.campaign_specs(n_geom=5, seed=3)
    b = C.campaign_specs(n_geom=5, seed=3)
    assert a == b and all(48 <= s["shape"][0] <= 160 for s in a)
    assert all(0.1 <= s["pe"] <= 100 and 0.01 <= s["da"] <= 10 for s in a)
    g = np.full((10, 6, 6), prtlb.PORE, np.uint8)
    assert C.percolates(g)
    g[5] = prtlb.SOLID
    assert not C.percolates(g)
    v = np.zeros((3, 4, 4), np.float32)
    v[0] = 0.05
    f = C.lowered_force(v, 2e-6)
    assert f == pytest.approx(1e-6 * (0.02 / np.sqrt(3.0)) / 0.05)
    v[0] = 1e-6
    assert C.lowered_force(v, 2e-6) == 2e-6          # never raised
    assert C.pressure_for_peclet(1e-5, 0.01, 20.0, 64, 0.5) == pytest.approx(1e-5 * 20 / (0.01 * 64 / 0.5))


def test_force_probe_and_peclet_scaling_by_linearity():
    """The builders' probe rule and Peclet rescale (3D/tools/build_dataset_3d.py:
    685-726, 3D/tools/settings_and_units.py:1349-1413), pinned by a slit, where
    the flow is exactly linear in the body force (unidirectional flow has no
    inertia): re-solving at campaign.lowered_force puts the peak speed at
    Mach 0.02, and re-solving at pressure_for_peclet gives the wanted Peclet,
    both to the solver's convergence.  Flows from the oracle's C iteration."""
    from oracle import stokes_c as SC
    h = 12
    g = np.full((3, 2, h + 2), O.PORE, np.uint8)
    g[:, :, 0] = g[:, :, -1] = O.SOLID

    def flow(force):
        return SC.stokes(g, nit=400000, force=force, tau_lb=0.8, tol=1e-13, return_f64=True)[1][0]

    probe = flow(C.PROBE_FORCE)
    f = C.lowered_force(probe, 1e-3)                  # a seed high enough to be lowered
    assert f < 1e-3
    v = flow(f)
    mach = float(np.sqrt((v ** 2).sum(0)).max()) * np.sqrt(3.0)
    assert abs(mach / 0.02 - 1) < 1e-8
    assert C.lowered_force(probe, 1e-9) == 1e-9       # never raised
    # Peclet: the axial pore mean of the seed flow, then the rescaled force
    d_lat = 0.1 * 1e-3
    u0 = C.axial_mean(v, g)
    assert u0 == pytest.approx(float(v[0][g == O.PORE].mean()), rel=1e-15)
    for pe in (0.3, 10.0, 250.0):
        fp = C.pressure_for_peclet(f, u0, pe, g.shape[0], d_lat)
        u1 = C.axial_mean(flow(fp), g)
        assert abs(C.peclet_from_flow(u1, g.shape[0], d_lat) / pe - 1) < 1e-8


def test_transport_queue_order():
    """The server's transport admission order: first come first served, or
    fewest pore voxels first with requests past max_wait ahead of all."""
    from types import SimpleNamespace as NS
    from prtlb.serve import transport_queue
    reqs = [NS(key=k, n_pore=n, t_ready=t) for k, n, t in
            (("a", 500, 1.0), ("b", 100, 2.0), ("c", 300, 3.0), ("d", 100, 4.0), ("e", 900, 5.0))]
    assert [r.key for r in transport_queue(reqs, "arrival", 10.0, 60.0)] == list("abcde")
    assert [r.key for r in transport_queue(reqs, "smallest", 10.0, 60.0)] == list("bdcae")
    # a and b have waited past max_wait: they go first, in ready order
    assert [r.key for r in transport_queue(reqs, "smallest", 10.0, 7.5)] == list("abdce")
