"""Stokes on the GPU against the reference's golden vectors, the oracle and
closed forms.  FP64 fields within 1e-12 relative (max |dv| / max |v|)."""
import contextlib
import io

import numpy as np
import pytest
import torch

from conftest import golden
from oracle import prtlb_oracle as O
import prtlb

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def solver(g):
    return prtlb.stokes_d2q9 if np.asarray(g).ndim == 2 else prtlb.stokes_d3q19


@pytest.mark.parametrize("name", ["stokes2d_fixed", "stokes2d_conv",
                                  "stokes3d_fixed", "stokes3d_conv"])
def test_golden_fp64(name):
    d = golden(name)
    mt = d["meta"]
    print("case", name, d["g"].shape, flush=True)
    info = {}
    v32, v64 = solver(d["g"])(d["g"], nit=mt["nit"], force=mt["force"], tau_lb=mt["tau"],
                              tol=mt["tol"], info=info, return_f64=True)
    assert rel(v64, d["v64"]) < 1e-12
    ulp = np.spacing(np.abs(d["v32"]).astype(np.float32))
    assert (np.abs(v32.astype(np.float64) - d["v32"]) <= ulp).all()
    assert info["flow_iters"] == mt["info"]["flow_iters"]
    assert info["flow_converged"] == mt["info"]["flow_converged"]


@pytest.mark.parametrize("shape,seed", [((40, 24), 0), ((20, 16, 12), 1), ((17, 9, 13), 2)])
def test_vs_oracle_fixed_steps(shape, seed):
    if len(shape) == 2:
        g = prtlb.disk_pack(shape, 3.0, 0.6, seed=seed)
    else:
        g = prtlb.sphere_pack(shape, (2.0, 3.5), 0.55, seed=seed)
    print("case", shape, seed, flush=True)
    nit = 300
    _, (v64o, _) = O.stokes(g, nit=nit, force=1e-5, tau_lb=0.8, tol=0.0, return_f64=True)
    v32, v64 = solver(g)(g, nit=nit, force=1e-5, tau_lb=0.8, tol=0.0, return_f64=True)
    assert rel(v64, v64o) < 1e-12
    # fluid-only update gives the same field bit for bit
    _, v64s = solver(g)(g, nit=nit, force=1e-5, tau_lb=0.8, tol=0.0, return_f64=True,
                        skip_solid=True)
    assert np.array_equal(v64s, v64)


@pytest.mark.parametrize("shape", [(6, 5, 1), (1, 7, 9), (9, 1, 8), (3, 40), (2, 2, 70)])
@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_thin_grids_vs_oracle(shape, storage):
    """Degenerate extents: one-node axes (each node is its own periodic
    neighbour) and rows shorter or longer than a cache line of populations,
    which exercise the row padding of the population layout."""
    rng = np.random.default_rng(sum(shape))
    g = np.where(rng.random(shape) < 0.25, prtlb.SOLID, prtlb.PORE).astype(np.uint8)
    g[0] = prtlb.PORE
    print("case", shape, storage, flush=True)
    _, (v64o, _) = O.stokes(g, nit=60, force=1e-5, tol=0.0, return_f64=True)
    _, v64 = solver(g)(g, nit=60, force=1e-5, tol=0.0, return_f64=True, storage=storage)
    tol = 1e-12 if storage == "fp64" else 1e-4
    assert rel(v64, v64o) < tol


def test_reduced_precision_errors():
    """FP32 and FP16-shifted storage against the FP64 solve of a converged
    disk pack: velocity field and permeability (FP16 stores with dithered
    rounding, which keeps the permeability inside the campaign's 1e-4)."""
    g = prtlb.disk_pack((64, 48), 4.0, 0.6, seed=5)
    info64 = {}
    ref64 = prtlb.stokes_d2q9(g, nit=20000, force=2e-6, tol=1e-10, info=info64, return_f64=True)[1]
    errs, perm = {}, {}
    for st in ("fp32", "fp16"):
        info = {}
        v = prtlb.stokes_d2q9(g, nit=20000, force=2e-6, tol=1e-10, info=info, storage=st)
        errs[st] = rel(v, ref64)
        perm[st] = abs(info["permeability"] / info64["permeability"] - 1)
        print(st, "iterations", info["flow_iters"], "FP64", info64["flow_iters"], flush=True)
    print("storage errors vs FP64: velocity", errs, "permeability", perm, flush=True)
    assert errs["fp32"] < 1e-4 and perm["fp32"] < 1e-5
    assert errs["fp16"] < 5e-2 and perm["fp16"] < 3e-4


def _slit2d(H, nx=12):
    g = np.full((nx, H + 2), prtlb.PORE, np.uint8)
    g[:, 0] = g[:, -1] = prtlb.SOLID
    return g


def _slit3d(H, nx=8, nz=8):
    g = np.full((nx, H + 2, nz), prtlb.PORE, np.uint8)
    g[:, 0, :] = g[:, -1, :] = prtlb.SOLID
    return g


def parabola(H, force, tau):
    nu = (tau - 0.5) / 3.0
    y = np.arange(1, H + 1) - 0.5
    return (force / (2 * nu)) * y * (H - y)


def test_poiseuille_2d_3d():
    v = prtlb.stokes_d2q9(_slit2d(20), nit=40000, force=2e-6, tau_lb=0.8)
    ux, ex = v[0][6, 1:-1], parabola(20, 2e-6, 0.8)
    assert abs(ux.max() / ex.max() - 1) < 0.01
    assert np.abs(ux - ex).max() / ex.max() < 0.015
    # a least-squares parabola through the profile
    # (3D/tools/test_flow_solvers.py:245-262): its curvature -g/(2 nu) gives
    # the viscosity (tau - 1/2)/3, its roots the no-slip walls half a voxel
    # beyond the last fluid nodes, and nothing is left over after the fit
    y = np.arange(1, 21) - 0.5
    c = np.polyfit(y, ux, 2)
    assert abs(-2e-6 / (2 * c[0]) / 0.1 - 1) < 0.01, c
    lo, hi = np.sort(np.roots(c))
    assert abs(((0.5 - lo) + (hi - 19.5)) / 2 - 0.5) < 0.1, (lo, hi)
    assert np.abs(ux - np.polyval(c, y)).max() / ux.max() < 0.01
    v3 = prtlb.stokes_d3q19(_slit3d(16), nit=40000, force=2e-6, tau_lb=0.8)
    ux3, ex3 = v3[0][4, 1:-1, 4], parabola(16, 2e-6, 0.8)
    assert abs(ux3.max() / ex3.max() - 1) < 0.015
    assert np.abs(ux3 - ex3).max() / ex3.max() < 0.02


@pytest.mark.parametrize("tau", [0.55, 0.8, 0.9330127, 1.2, 2.0])
def test_wall_stays_at_midlink(tau):
    for H in (4, 8, 20):
        v = prtlb.stokes_d2q9(_slit2d(H), nit=20000, force=2e-6, tau_lb=tau)
        ux, ex = v[0][6, 1:-1], parabola(H, 2e-6, tau)
        assert np.abs(ux - ex).max() / ex.max() < 0.005, (tau, H)


def test_d2q9_equals_d3q19_and_slit_permeability():
    for H in (4, 8):
        i2, i3 = {}, {}
        u2 = prtlb.stokes_d2q9(_slit2d(H), nit=20000, force=2e-6, info=i2)[0].max()
        u3 = prtlb.stokes_d3q19(_slit3d(H), nit=20000, force=2e-6, info=i3)[0].max()
        assert abs(u2 - u3) / u2 < 1e-3
        # Darcy permeability of the slit: closed form of the node-sampled parabola
        k = O.slit_permeability(H)
        i = {}
        prtlb.stokes_d2q9(_slit2d(H), nit=200000, force=2e-6, tol=1e-12, info=i)
        assert abs(i["permeability"] / k - 1) < 1e-6, (i["permeability"], k)


def test_mass_conserved_full_size():
    # Sum of all populations is invariant: collision conserves mass, the Guo
    # term sums to zero, solid reversal and streaming permute.
    g = prtlb.sphere_pack((128, 128, 128), (4.0, 8.0), 0.35, seed=0)
    s = prtlb.StokesSolver(g, force=2e-6)
    m0 = float(s.populations().sum(dtype=torch.float64))
    s.run(200)
    m1 = float(s.populations().sum(dtype=torch.float64))
    assert abs(m1 - m0) / m0 < 1e-12
    v32, v64, sux = s.velocity(f64=True)
    solid = s.solid.reshape(s.shape).bool()
    assert float(v64[:, solid].abs().max()) == 0.0
    assert np.isfinite(s.permeability(sux))


def _duct3d(nx=16, H=14):
    g = np.full((nx, H + 2, H + 2), prtlb.PORE, np.uint8)
    g[:, 0, :] = g[:, -1, :] = prtlb.SOLID
    g[:, :, 0] = g[:, :, -1] = prtlb.SOLID
    return g


def test_porous_run_and_default_iteration_counts():
    """3D/tools/test_flow_solvers.py:284-304 and :372-386: a porous 2D run
    has exactly zero velocity in solid and stays below Mach 0.05, a walled
    3D duct has none in its walls, and the default iteration counts (1200 in
    2D, 400 in 3D) are within 5% of the mean speed at eight times as many.
    A seeded, inlet-connected disk pack of porosity 0.7 on 64 x 32 stands in
    for the reference's random blob of the same size and porosity."""
    g, _ = prtlb.keep_inlet_connected(prtlb.disk_pack((64, 32), 3.0, 0.7, seed=3))
    v = prtlb.stokes_d2q9(g, nit=3000)
    assert float(np.abs(v[:, g != prtlb.PORE]).max()) == 0.0
    assert float(np.sqrt((v[:2] ** 2).sum(0)).max()) / np.sqrt(1.0 / 3.0) < 0.05
    d = _duct3d()
    v3 = prtlb.stokes_d3q19(d, nit=2000)
    assert float(np.abs(v3[:, d != prtlb.PORE]).max()) == 0.0
    for solver, geom, default in ((prtlb.stokes_d2q9, g, 1200), (prtlb.stokes_d3q19, d, 400)):
        a = solver(geom, nit=default)
        b = solver(geom, nit=default * 8)
        m = geom == prtlb.PORE
        sa = np.sqrt((a[:2] ** 2).sum(0))[m].mean()
        sb = np.sqrt((b[:2] ** 2).sum(0))[m].mean()
        print("default", default, "mean speed", sa, sb, flush=True)
        assert abs(sa - sb) / max(sb, 1e-30) < 0.05


def test_plane_flux_constant_and_solid_zero():
    g = np.full((16, 16, 16), prtlb.PORE, np.uint8)
    g[:, 0, :] = g[:, -1, :] = prtlb.SOLID
    g[:, :, 0] = g[:, :, -1] = prtlb.SOLID
    v32, v64 = prtlb.stokes_d3q19(g, nit=400, force=2e-6, return_f64=True, as_tensor=True)
    flux = prtlb.plane_flux(v64).cpu().numpy()
    assert (flux.max() - flux.min()) / abs(flux.mean()) < 0.01
    assert float(v32[:, torch.from_numpy(g != prtlb.PORE).cuda()].abs().max()) == 0.0


def test_mach_warning_and_refusals():
    g = np.full((16, 16, 16), prtlb.PORE, np.uint8)
    g[:, 0, :] = g[:, -1, :] = prtlb.SOLID
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        prtlb.stokes_d3q19(g, nit=200, force=3e-3, tau_lb=0.8)
    assert "Mach" in buf.getvalue()
    with pytest.raises(SystemExit):
        prtlb.stokes_d2q9(np.zeros((8, 8), np.uint8), nit=200)
    with pytest.raises(ValueError):
        prtlb.stokes_d2q9(np.full((8, 8), 2, np.uint8), tau_lb=0.5)
    with pytest.raises(ValueError):
        prtlb.stokes_d3q19(np.full((8, 8), 2, np.uint8))


def test_graph_replays_match_eager():
    # every replay of a captured tick gives the eager result, with eager work
    # and synchronisations between replays
    g = prtlb.disk_pack((128, 128), 6.0, 0.6, seed=0)
    eager = prtlb.StokesSolver(g, force=2e-6)
    graphed = prtlb.StokesSolver(g, force=2e-6)
    for k in range(4):
        eager.run(50, reduce=True)
        ue = eager.mean_abs_ux()
        graphed.tick_graph(50)
        ug = graphed.mean_abs_ux_ready()
        torch.cuda.synchronize()
        assert ug == ue, k
        assert torch.equal(eager.populations(), graphed.populations()), k
        _ = torch.ones(1000, device="cuda").sum().item()     # unrelated eager work
    i1, i2 = {}, {}
    v1 = prtlb.stokes_d2q9(g, nit=5000, tol=1e-8, info=i1, graph=True)
    v2 = prtlb.stokes_d2q9(g, nit=5000, tol=1e-8, info=i2, graph=False)
    assert np.array_equal(v1, v2) and i1["flow_iters"] == i2["flow_iters"]


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16"])
def test_launch_shapes_give_identical_bits(storage):
    """Every launch shape the autotuner may pick (block shape, nodes per
    thread, nontemporal loads, XCD-contiguous block order) computes the same
    populations and the same |u_x| partial sums as the default shape."""
    from prtlb import _lib
    from prtlb.autotune import CANDIDATES
    from prtlb.stokes import StokesSolver
    g = prtlb.sphere_pack((40, 36, 130), (2.0, 5.0), 0.4, seed=4)
    ref = StokesSolver(g, force=2e-5, storage=storage)
    ref.run(57, reduce=True)
    f_ref = ref.f[ref.latest].cpu().numpy()
    m_ref = ref.mean_abs_ux()
    for bz, by, npt, ntl, xcd in CANDIDATES:
        if (1 << bz) * (1 << by) > 512:
            continue
        S = StokesSolver(g, force=2e-5, storage=storage)
        S.tune = _lib.lbm_tune(bz, by, npt, bool(ntl), bool(xcd))
        S.run(57, reduce=True)
        assert np.array_equal(S.f[S.latest].cpu().numpy().view(np.uint8), f_ref.view(np.uint8)), \
            (bz, by, npt, ntl, xcd)
        assert S.mean_abs_ux() == pytest.approx(m_ref, rel=1e-12)


def test_solver_handles_match_stateless_calls():
    """The handle API (create / init / step_n / reduce / velocity / destroy)
    gives the bits of the StokesSolver path over the same buffers, and the
    transport handle those of TransportSolver."""
    import ctypes
    import torch
    from prtlb import _lib
    from prtlb.device import ptr, stream_ptr
    from prtlb.stokes import StokesSolver
    g = prtlb.sphere_pack((30, 22, 26), (2.0, 4.0), 0.45, seed=13)
    S = StokesSolver(g, force=3e-5, storage="fp64")
    S.run(120, reduce=True)
    m_ref = S.mean_abs_ux()
    v32_ref, v64_ref, sux_ref = S.velocity(f64=True)
    L = _lib.lib()
    fa = torch.empty_like(S.f[0])
    fb = torch.empty_like(S.f[0])
    part = torch.zeros(L.prt_lbm_partials_count(19, *S.dims), dtype=torch.float64, device=S.dev)
    scal = torch.zeros(1, dtype=torch.float64, device=S.dev)
    h = ctypes.c_void_p()
    _lib.call("prt_lbm_solver_create", 19, 0, *S.dims, ptr(S.solid), S.n_fluid, ptr(fa), ptr(fb),
              ptr(part), ptr(scal), 0.8, 3.0 / 16.0, 3e-5, 1024.0, 0, stream_ptr(S.dev), ctypes.byref(h))
    try:
        _lib.call("prt_lbm_solver_init", h)
        _lib.call("prt_lbm_solver_step_n", h, 70)
        _lib.call("prt_lbm_solver_step_n", h, 50)
        mean = ctypes.c_double()
        _lib.call("prt_lbm_solver_reduce", h, 0, ctypes.byref(mean))
        assert mean.value == m_ref
        v64 = torch.empty((3, S.n), dtype=torch.float64, device=S.dev)
        _lib.call("prt_lbm_solver_velocity", h, ptr(v64), None)
        assert np.array_equal(v64.cpu().numpy().reshape(v64_ref.shape), v64_ref.cpu().numpy())
        sux = ctypes.c_double()
        _lib.call("prt_lbm_solver_reduce", h, 1, ctypes.byref(sux))
        assert sux.value == pytest.approx(sux_ref, rel=1e-12)
        # 120 steps: the newest state is back in the first buffer, as in S
        assert L.prt_lbm_solver_newest(h) == (fa if S.latest == 0 else fb).data_ptr()
    finally:
        _lib.call("prt_lbm_solver_destroy", h)
    # transport handle against TransportSolver
    v = prtlb.stokes_d3q19(g, nit=300, force=3e-5, tol=0.0)
    T = prtlb.TransportSolver(g, v, 10.0, 1.0, 1, chem=prtlb.chemistry.monod(0.5))
    c_a = T.c[0].clone()
    c_b = T.c[1].clone()
    T.use_graph = False
    T.step_n(37)
    ha = ctypes.c_void_p()
    _lib.call("prt_adr_solver_create", ctypes.byref(T.params), *T.dims, ptr(T.table),
              T.n_pore, ptr(c_a), ptr(c_b), stream_ptr(T.dev), ctypes.byref(ha))
    try:
        _lib.call("prt_adr_solver_step_n", ha, 20)
        _lib.call("prt_adr_solver_step_n", ha, 17)
        snap = torch.empty_like(c_a)
        _lib.call("prt_adr_solver_snapshot", ha, ptr(snap))
        torch.cuda.synchronize()
        assert np.array_equal(snap.cpu().numpy(), T.current().cpu().numpy())
    finally:
        _lib.call("prt_adr_solver_destroy", ha)
