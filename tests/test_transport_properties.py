"""Properties the transport scheme must have, checked on the oracle (CPU)
and on the GPU path with the same inputs: a uniform field stays uniform
under any flow (the div-u correction), pure diffusion in a closed pore space
conserves mass, no new extrema appear (MUSCL-minmod + clip), a well-mixed
first-order reaction follows its discrete closed form exactly, and an
irreversible C -> S sorption conserves C + S.  A dropped or sign-flipped
term in the face fluxes, the div-u term, the reaction cap or the product
coupling fails one of them."""
import numpy as np
import pytest

from oracle import prtlb_oracle as O

BACKENDS = [pytest.param("oracle", id="oracle"),
            pytest.param("gpu", id="gpu", marks=pytest.mark.gpu)]

NAN = float("nan")


def _chem(backend, **kw):
    if backend == "oracle":
        return O.Chemistry(**kw)
    import prtlb
    return prtlb.Chemistry(**kw)


def _flow(g):
    """A short Stokes solve (FP64 oracle, float32 out): not converged, so not
    divergence free -- the harder case for the div-u correction."""
    return O.stokes(g, nit=300, force=1e-4, tol=0.0)


def _evolve(backend, g, v, pe, da, ns, chem, steps, c0=None, da_abio=0.0):
    """`steps` transport steps from the chemistry's initial state (or c0,
    shaped (ns, n)); returns (C (ns, n) float32, dt, pore mask)."""
    if backend == "oracle":
        S = O.adr_setup(g, v, pe, da, ns, chem, da_abio)
        C = O.adr_initial(S) if c0 is None else np.array(c0, np.float32)
        for _ in range(steps):
            C = O.adr_step(S, C)
        return C, S.dt, S.mf.reshape(g.shape)
    import torch
    import prtlb
    T = prtlb.TransportSolver(g, v, pe, da, ns, chem, da_abio)
    if c0 is not None:
        T.load_state(torch.from_numpy(np.ascontiguousarray(c0, np.float32)))
    T.step_n(steps)
    return T.current().cpu().numpy(), T.dt, T.m


@pytest.fixture(scope="module")
def rock():
    rng = np.random.default_rng(5)
    g = np.where(rng.random((36, 22)) < 0.25, O.SOLID, O.PORE).astype(np.uint8)
    g = O.keep_inlet_connected(g)[0]
    return g, _flow(g)


@pytest.mark.parametrize("backend", BACKENDS)
def test_uniform_field_stays_uniform(backend, rock):
    g, v = rock
    ch = _chem(backend, names=["C"], inlet=[1.0], c_init=[1.0], d_rel=[1.0], immobile=[False],
               biotic_on=False)
    C, dt, m = _evolve(backend, g, v, 20.0, 0.0, 1, ch, 300)
    c = C[0].reshape(g.shape)
    assert np.abs(c[m] - 1.0).max() < 1e-5
    assert (c[~m] == 0).all()


@pytest.mark.parametrize("backend", BACKENDS)
def test_closed_box_diffusion_conserves_mass(backend, rock):
    g, _ = rock
    v = np.zeros((3,) + g.shape, np.float32)
    ch = _chem(backend, names=["C"], inlet=[NAN], c_init=[0.0], d_rel=[1.0], immobile=[False],
               biotic_on=False)
    rng = np.random.default_rng(1)
    m = (g == O.PORE).ravel()
    c0 = np.where(m, rng.random(g.size), 0.0).astype(np.float32)[None]
    C, dt, _ = _evolve(backend, g, v, 5.0, 0.0, 1, ch, 400, c0=c0)
    before = float(c0.astype(np.float64).sum())
    after = float(C.astype(np.float64).sum())
    assert abs(after - before) / before < 2e-6
    assert C.max() <= c0.max() and C[0][m].min() >= c0[0][m].min()   # diffusion only smooths
    assert np.abs(C - c0).max() > 1e-3                                 # and it did move


@pytest.mark.parametrize("backend", BACKENDS)
def test_no_new_extrema(backend, rock):
    g, v = rock
    ch = _chem(backend, names=["C"], inlet=[1.0], c_init=[0.0], d_rel=[1.0], immobile=[False],
               biotic_on=False)
    C, dt, m = _evolve(backend, g, v, 200.0, 0.0, 1, ch, 500)
    c = C[0].reshape(g.shape)[m]
    assert c.min() >= 0.0 and c.max() <= 1.0 + 1e-6
    assert c.max() > 0.5                                   # the front entered


@pytest.mark.parametrize("backend", BACKENDS)
def test_well_mixed_first_order_discrete_closed_form(backend, rock):
    g, _ = rock
    v = np.zeros((3,) + g.shape, np.float32)
    ch = _chem(backend, names=["C"], inlet=[NAN], c_init=[1.0], d_rel=[1.0], immobile=[False],
               biotic_on=False, abiotic_index=0, abiotic_surface_only=False)
    steps = 250
    C, dt, m = _evolve(backend, g, v, 10.0, 0.0, 1, ch, steps, da_abio=3.0)
    # the float32 recurrence of one well-mixed cell: R = min(ka c, c/2/dt),
    # c <- max(c + dt (0 - R), 0), ka = Da / nx
    f = np.float32
    ka = f(3.0 / g.shape[0])
    dt = f(dt)
    c = f(1.0)
    for _ in range(steps):
        r = min(f(ka * c), f(f(f(0.5) * c) / dt))
        c = max(f(c + f(dt * f(f(0.0) - r))), f(0.0))
    got = C[0].reshape(g.shape)[m]
    assert (got == c).all()
    assert abs(float(c) - np.exp(-float(ka) * float(dt) * steps)) < 1e-2   # and the ODE


@pytest.mark.parametrize("backend", BACKENDS)
def test_irreversible_sorption_conserves_total(backend, rock):
    g, _ = rock
    v = np.zeros((3,) + g.shape, np.float32)
    ch = _chem(backend, names=["C", "S"], inlet=[NAN, NAN], c_init=[1.0, 0.0], d_rel=[1.0, 0.0],
               immobile=[False, True], biotic_on=False, abiotic_index=0, abiotic_product_index=1,
               abiotic_surface_only=False)
    C, dt, m = _evolve(backend, g, v, 10.0, 0.0, 2, ch, 300, da_abio=2.0)
    tot = (C[0] + C[1]).reshape(g.shape)[m]
    assert np.abs(tot - 1.0).max() < 1e-6
    assert C[1].reshape(g.shape)[m].min() > 0.05                    # S really grew


@pytest.mark.gpu
def test_gpu_equals_oracle_on_the_property_cases(rock):
    """The GPU field equals the oracle's bit for bit on these cases too."""
    g, v = rock
    for da_abio, kw in ((0.0, dict(names=["C"], inlet=[1.0], c_init=[0.0], d_rel=[1.0],
                                   immobile=[False], biotic_on=False)),
                        (1.0, dict(names=["C", "S"], inlet=[1.0, NAN], c_init=[0.0, 0.0],
                                   d_rel=[1.0, 0.0], immobile=[False, True], biotic_on=False,
                                   abiotic_index=0, abiotic_product_index=1,
                                   abiotic_surface_only=True))):
        ns = len(kw["names"])
        a, _, _ = _evolve("oracle", g, v, 50.0, 0.0, ns, _chem("oracle", **kw), 200,
                          da_abio=da_abio)
        b, _, _ = _evolve("gpu", g, v, 50.0, 0.0, ns, _chem("gpu", **kw), 200, da_abio=da_abio)
        assert np.array_equal(a, b)
