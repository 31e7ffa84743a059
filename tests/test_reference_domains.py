"""Released 2D domains of the reference (the 148 x 64 pore maps of its
Monod, irreversible-sorption and reversible-sorption cases, and two of the
3000 training domains, porosity 0.49 and 0.74; copied as data into
tests/golden/reference_domains.npz by make_domains.py there): real, not
generated, pore spaces through the whole pipeline -- D2Q9 flow to 1e-8,
then transport to its automatic stop -- on the GPU against the oracle."""
import os

import numpy as np
import pytest
from scipy import ndimage

import prtlb
from conftest import oracle_chem
from oracle import adr_c as AC
from oracle import prtlb_oracle as O
from prtlb import campaign as C

DOMAINS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "reference_domains.npz")
# case: (kinetics, Pe, Da, reversible-sorption rates), Pe and Da taken from
# the corners of the reference's parameter grid for the case
# (2D_scripts/prt2d_model.py:72-109: Monod Pe 1..10, Da 0.05..0.5;
# irreversible Pe 1..10, Da_A 74..740; reversible Da_A 0.5..2, Da_D 0.02..0.2);
# the two training domains with first-order surface kinetics at the low
# corner and Monod at the high one
CASES = {
    "monod": ("monod", 10.0, 0.5, None),
    "irreversible_sorption": ("sorption_irreversible", 10.0, 740.0, None),
    "reversible_sorption": ("sorption_reversible", 10.0, 2.0, dict(da_sorb=2.0, da_desorb=0.2)),
    "dat_1": ("first_order", 1.0, 0.05, None),
    "dat_3000": ("monod", 10.0, 0.5, None),
}


def _grid(case, m, released_order=True):
    """A released flat array as the (148, 64) grid, x (the flow direction)
    along axis 0: the case domains (.npz) are (64, 148) in C order and
    transposed, the training domains (.dat) (148, 64) in C order
    (3D/tools/build_transfer_set_2d_to_3d.py:96-104); released_order=False
    gives the other order."""
    npz = not case.startswith("dat_")
    return m.reshape(64, 148).T if npz == released_order else m.reshape(148, 64)


def reference_domain(case):
    """Simulator material codes of a released domain: Jung's codes 1 (pore)
    and 2 (interface) both PORE, 0 SOLID (3D/tools/
    build_transfer_set_2d_to_3d.py:84-92)."""
    with np.load(DOMAINS, allow_pickle=False) as z:
        m = z[case]
    return np.where(_grid(case, m) > 0, prtlb.PORE, prtlb.SOLID).astype(np.uint8)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def within_ulp(a, b):
    ulp = np.spacing(np.abs(np.asarray(b, np.float32)))
    return bool((np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)) <= ulp).all())


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_domain_reading(case):
    """The reading agrees with what the reference measured on these files
    (3D/tools/build_geometry_3d.py:15-33, build_transfer_set_2d_to_3d.py:
    84-104): the first and last x-slices all pore (inlet and outlet
    buffers); the case domains' porosity just under 0.5 with the interface
    counted as pore (0.493 on the Monod domain); one face-connected pore
    cluster, where the other index order breaks the pore space into
    several (the reference picks the order with fewer clusters); the inlet
    connectivity rule (G1) then removes nothing."""
    g = reference_domain(case)
    assert g.shape == (148, 64)
    p = g == prtlb.PORE
    assert p[0].all() and p[-1].all()
    if not case.startswith("dat_"):
        assert 0.49 < p.mean() < 0.50
    if case == "monod":
        assert round(float(p.mean()), 3) == 0.493
    assert ndimage.label(p)[1] == 1
    with np.load(DOMAINS, allow_pickle=False) as z:
        other = _grid(case, z[case], released_order=False) > 0
    assert ndimage.label(other)[1] > 5
    g2, n = prtlb.keep_inlet_connected(g)
    assert n == 0 and np.array_equal(g2, g)


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_domain_pipeline_against_oracle(case):
    """Flow: D2Q9 FP64, tau 0.8, magic 3/16, to 1e-8 -- the stop iteration
    equal, the FP64 field to 1e-12, the float32 field within an ulp,
    permeability to 1e-10.  Transport on that flow: the case's kinetics,
    automatic stop (tol 2e-3, 100 000-step cap; the Monod cases settle,
    the others reach the cap), five snapshots -- the stop step,
    snapshots and t_norm bit for bit with the oracle (its steps computed
    in C, oracle/adr_c.c), the mobile chemical's outlet concentration to
    1e-10; and the reference's physics checks find nothing (the sorbed
    chemical under its own rules)."""
    kin, pe, da, sorption = CASES[case]
    g = reference_domain(case)
    print("domain", case, g.shape, "porosity %.4f" % (g == prtlb.PORE).mean(), flush=True)
    info, oinfo = {}, {}
    v32, v64 = prtlb.stokes_d2q9(g, tau_lb=0.8, magic=3.0 / 16.0, tol=1e-8, info=info,
                                 return_f64=True)
    v32o, (v64o, _) = O.stokes(g, tau_lb=0.8, magic=3.0 / 16.0, tol=1e-8, info=oinfo,
                               return_f64=True)
    print("flow", info["flow_iters"], oinfo["flow_iters"], flush=True)
    assert info["flow_iters"] == oinfo["flow_iters"] and info["flow_converged"]
    assert rel(v64, v64o) < 1e-12
    assert within_ulp(v32, v32o)
    k_o = O.permeability(v64o, g, 0.8, 2e-6)
    assert abs(info["permeability"] / k_o - 1) < 1e-10, (info["permeability"], k_o)

    akw = C._adr_kw(dict(da=da), kin)
    if sorption is not None:
        akw["sorption"] = sorption
    ch, ns, da_t = akw.pop("chem"), akw.pop("n_species"), akw.pop("da")
    ainfo, oainfo = {}, {}
    out, tn = prtlb.solve_adr(g, v32, pe, da_t, 5, n_species=ns, chem=ch, info=ainfo,
                              max_steps=100000, **akw)
    outo, tno, Co = O.solve_adr(g, v32, pe, da_t, 5, n_species=ns, chem=oracle_chem(ch),
                                info=oainfo, max_steps=100000, return_state=True,
                                stepper=AC.stepper, **akw)
    print("transport", ainfo["steps"], ainfo["converged"], oainfo["steps"], oainfo["converged"],
          flush=True)
    assert ainfo["steps"] == oainfo["steps"]
    assert ainfo["converged"] == oainfo["converged"] == (kin == "monod")
    assert np.array_equal(out, outo) and np.array_equal(tn, tno)
    oc, _ = O.outlet_concentration(Co, g)
    assert oc[0] > 0
    assert abs(float(ainfo["outlet_concentration"][0]) / oc[0] - 1) < 1e-10
    view, vch = C.physics_view(out, ch, kin)
    assert prtlb.check_physics_device(view, g, chem=vch, da=da_t,
                                      da_abio=akw.get("da_abio")) == []
    assert C.sorbed_complaints(out, g, kin) == []
