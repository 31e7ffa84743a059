"""The oracle's own pins that need no GPU.

* The C restatement of the FP64 Stokes iteration (oracle/stokes_c.c) equals
  the NumPy oracle bit for bit, stops at the same iteration, and meets the
  reference's golden vectors as the NumPy oracle does.
* The C restatement of the float32 transport step (oracle/adr_c.c) equals
  the NumPy oracle's adr_step bit for bit over every kinetics kind, in 2D
  and 3D, snapshots and final state (float32 words compared).
* The oracle's Darcy permeability (a quantity the reference does not
  compute) equals the closed form of a slit, k = (2H^3 + H) / (24 (H + 2)),
  for H = 4..20: the node-sampled Poiseuille parabola, which TRT with the
  magic parameter 3/16 reproduces exactly with the wall at mid-link
  (3D/tools/test_flow_solvers.py:77-134).  A mean over pore nodes instead of
  the box, a missing viscosity or a wrong force scaling fails it by far.
* The oracle's outlet concentration equals plug flow's exp(-Da) within the
  reference self-test's 0.05 (3D/tools/prtlb_3d.py:1868-1884) on a channel
  with solid walls, where a mean over the whole outlet slice (walls
  included) or over the inlet slice would miss by more than that.
"""
import numpy as np
import pytest

from conftest import golden
from oracle import prtlb_oracle as O
from oracle import stokes_c as SC


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _disks(shape, seed, phi=0.6, r=3.0):
    rng = np.random.default_rng(seed)
    g = np.full(shape, O.PORE, np.uint8)
    X, Y = np.indices(shape)
    while (g == O.PORE).mean() > phi:
        cx, cy = rng.random(2) * shape
        dx = np.minimum(np.abs(X - cx), shape[0] - np.abs(X - cx))
        dy = np.minimum(np.abs(Y - cy), shape[1] - np.abs(Y - cy))
        g[dx * dx + dy * dy <= r * r] = O.SOLID
    return g


def _spheres(shape, seed, phi=0.55, r=2.5):
    rng = np.random.default_rng(seed)
    g = np.full(shape, O.PORE, np.uint8)
    I = np.indices(shape)
    while (g == O.PORE).mean() > phi:
        c = rng.random(3) * shape
        d2 = sum(np.minimum(np.abs(I[a] - c[a]), shape[a] - np.abs(I[a] - c[a])) ** 2 for a in range(3))
        g[d2 <= r * r] = O.SOLID
    return g


@pytest.mark.parametrize("g", [_disks((40, 24), 0), _disks((31, 17), 1),
                               _spheres((20, 16, 12), 1), _spheres((9, 13, 7), 2)],
                         ids=["2d-40x24", "2d-31x17", "3d-20x16x12", "3d-9x13x7"])
def test_c_equals_numpy_oracle_fixed_steps(g):
    v32a, (v64a, fa) = O.stokes(g, nit=120, force=1e-5, tol=0.0, return_f64=True)
    v32b, (v64b, fb) = SC.stokes(g, nit=120, force=1e-5, tol=0.0, return_f64=True)
    assert np.array_equal(v64a, v64b)
    assert np.array_equal(fa, fb)
    assert np.array_equal(v32a, v32b)


def test_c_equals_numpy_oracle_converged():
    g = _spheres((16, 12, 10), 3)
    ia, ib = {}, {}
    a = O.stokes(g, force=1e-5, tol=1e-8, info=ia)
    b = SC.stokes(g, force=1e-5, tol=1e-8, info=ib)
    assert ia["flow_iters"] == ib["flow_iters"] and ia["flow_converged"] and ib["flow_converged"]
    assert rel(b, a) < 1e-6
    # the change is a difference of two means summed in different orders
    assert abs(ia["flow_change"] - ib["flow_change"]) <= 1e-13


def test_c_resume_from_populations():
    g = _disks((24, 20), 4)
    _, (v_all, f_all) = SC.stokes(g, nit=300, force=1e-5, tol=0.0, return_f64=True)
    _, (_, f_mid) = SC.stokes(g, nit=170, force=1e-5, tol=0.0, return_f64=True)
    _, (v_res, f_res) = SC.stokes(g, nit=130, force=1e-5, tol=0.0, return_f64=True, f_init=f_mid)
    assert np.array_equal(f_res, f_all) and np.array_equal(v_res, v_all)


@pytest.mark.parametrize("name", ["stokes2d_fixed", "stokes2d_conv", "stokes3d_fixed", "stokes3d_conv"])
def test_c_golden(name):
    d = golden(name)
    mt = d["meta"]
    info = {}
    v32, (v64, _) = SC.stokes(d["g"], nit=mt["nit"], force=mt["force"], tau_lb=mt["tau"],
                              tol=mt["tol"], info=info, return_f64=True)
    assert rel(v64, d["v64"]) < 1e-12
    ulp = np.spacing(np.abs(d["v32"]).astype(np.float32))
    assert (np.abs(v32.astype(np.float64) - d["v32"]) <= ulp).all()
    assert info["flow_iters"] == mt["info"]["flow_iters"]
    assert info["flow_converged"] == mt["info"]["flow_converged"]


@pytest.mark.parametrize("h", [4, 5, 6, 8, 10, 12, 16, 20])
def test_permeability_slit_closed_form(h):
    g = np.full((2, h + 2), O.PORE, np.uint8)
    g[:, 0] = O.SOLID
    g[:, -1] = O.SOLID
    info = {}
    _, (v64, _) = SC.stokes(g, nit=200000, force=1e-6, tau_lb=0.8, tol=1e-14, info=info,
                            return_f64=True)
    assert info["flow_converged"]
    k = O.permeability(v64, g, 0.8, 1e-6)
    assert abs(k / O.slit_permeability(h) - 1) < 1e-9
    # the same field with the walls along z of a 3D box gives the same k
    if h in (4, 8):
        g3 = np.repeat(g[:, None, :], 2, axis=1)
        _, (w64, _) = SC.stokes(g3, nit=200000, force=1e-6, tau_lb=0.8, tol=1e-14,
                                return_f64=True)
        assert abs(O.permeability(w64, g3, 0.8, 1e-6) / O.slit_permeability(h) - 1) < 1e-9


@pytest.mark.parametrize("da", [0.5, 1.0, 2.0])
def test_outlet_concentration_plug_flow(da):
    g = np.full((80, 10), O.PORE, np.uint8)
    g[:, 0] = O.SOLID
    g[:, -1] = O.SOLID
    v = np.zeros((3, 80, 10), np.float32)
    v[0, :, 1:-1] = 1.0
    ch = O.Chemistry(["C"], [1.0], [0.0], [1.0], [False], biotic_on=True)
    out, _, C = O.solve_adr(g, v, pe=4000.0, da=da, n_t=3, n_species=1, chem=ch,
                            return_state=True)
    mean, flux = O.outlet_concentration(C, g, v)
    assert abs(mean[0] - np.exp(-da)) < 0.05
    # uniform velocity over the open rows: the flux-weighted mean is the mean
    assert abs(flux[0] - mean[0]) < 1e-12


# ---- the C restatement of the float32 transport step (oracle/adr_c.c)
from oracle import adr_c as AC  # noqa: E402


def _adr_cases():
    g3 = _spheres((14, 12, 10), 3, phi=0.6, r=2.0)
    g2 = _disks((36, 22), 4, phi=0.65, r=2.5)
    gbf = O.coat_with_biofilm(_spheres((12, 11, 9), 5, phi=0.65, r=2.0), 3, 1)[0]
    nan = np.nan
    monod = O.Chemistry(["C"], [1.0], [0.0], [1.0], [False], ks_donor=0.5)
    full = O.Chemistry(["Ac", "A", "P", "Bio"], [1.0, 1.0, 0.0, nan], [0.0, 0.0, 0.0, 0.1],
                       [1.0, 0.8, 1.0, 0.0], [False, False, False, True], ks_donor=0.3,
                       ks_acceptor=0.15, alpha=0.4, yield_=0.04, decay_lattice=0.006,
                       biomass_coupled=True)
    surface = O.Chemistry(["C"], [1.0], [0.0], [1.0], [False], biotic_on=False,
                          abiotic_surface_only=True, abiotic_index=0)
    sorb = O.Chemistry(["C", "S"], [1.0, nan], [0.0, 0.0], [1.0, 0.0], [False, True],
                       biotic_on=False, abiotic_index=0, abiotic_product_index=1)
    sorb_surf = O.Chemistry(["C", "S"], [1.0, nan], [0.0, 0.0], [1.0, 0.0], [False, True],
                            biotic_on=False, abiotic_index=0, abiotic_surface_only=True)
    dual = O.default_chemistry(2)
    dual.ks_donor = 0.2
    biofilm = O.Chemistry(["Ac", "A", "P", "Bio"], [1.0, 1.0, 0.0, nan], [0.0, 0.0, 0.0, 0.1],
                          [1.0, 1.0, 1.0, 0.0], [False, False, False, True], ks_donor=0.4,
                          d_biofilm_rel=[0.3, 0.3, 0.3, 0.0], biofilm_code=3, seed_in_biofilm=True,
                          decay_lattice=0.006, abiotic_product_index=0)
    return [
        ("3d-monod", g3, monod, dict(pe=10.0, da=1.0, n_species=1)),
        ("3d-dual-monod-coupled-abiotic", g3, full, dict(pe=3.0, da=2.0, n_species=4, da_abio=0.5)),
        ("2d-surface-first-order", g2, surface, dict(pe=40.0, da=0.0, n_species=1, da_abio=1.0)),
        ("3d-irreversible-sorption", g3, sorb, dict(pe=0.5, da=0.0, n_species=2, da_abio=2.0)),
        ("2d-reversible-sorption", g2, sorb_surf,
         dict(pe=5.0, da=0.0, n_species=2, sorption=dict(da_sorb=1.5, da_desorb=0.7, c=0, s=1))),
        ("2d-dual-monod", g2, dual, dict(pe=20.0, da=3.0, n_species=2)),
        ("3d-biofilm", gbf, biofilm, dict(pe=8.0, da=1.5, n_species=4, da_abio=0.3)),
    ]


@pytest.mark.parametrize("case", _adr_cases(), ids=lambda c: c[0])
def test_c_transport_step_equals_numpy_oracle(case):
    """oracle/adr_c.c against prtlb_oracle.adr_step, every bit of every
    snapshot (float32 words compared, so -0 against +0 counts too)."""
    name, g, chem, kw = case
    fn = O.stokes_d2q9 if g.ndim == 2 else O.stokes_d3q19
    v = fn(g, nit=150, force=2e-5, tol=0.0)
    a_info, b_info = {}, {}
    a = O.solve_adr(g, v, n_t=4, steps=240, chem=chem, info=a_info, return_state=True, **kw)
    b = O.solve_adr(g, v, n_t=4, steps=240, chem=chem, info=b_info, return_state=True,
                    stepper=AC.stepper, **kw)
    for x, y in zip(a, b):
        assert x.dtype == y.dtype == np.float32
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name
    assert a_info == b_info
    assert float(a[2].max()) > 0.0, name       # something moved in
