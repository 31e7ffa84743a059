"""The oracle against vectors the reference itself wrote (tests/golden).

Stokes: the FP64 pre-cast field within 1e-12 relative, the float32 field
within one ulp, the convergence stop iteration exactly.  Transport: float32
outputs bit for bit (NumPy 2 casting), t_norm and info exactly.
"""
import numpy as np
import pytest

from conftest import golden
from oracle import prtlb_oracle as O


def _nan(v):
    if isinstance(v, list):
        return [float("nan") if x == "nan" else x for x in v]
    return float("nan") if v == "nan" else v


def chem_from(kw):
    if kw is None:
        return None
    return O.Chemistry(**{k: _nan(v) for k, v in kw.items()})


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("name", ["stokes2d_fixed", "stokes2d_conv",
                                  "stokes3d_fixed", "stokes3d_conv"])
def test_stokes_golden(name):
    d = golden(name)
    mt = d["meta"]
    info = {}
    v32, (v64, f) = O.stokes(d["g"], nit=mt["nit"], force=mt["force"],
                             tau_lb=mt["tau"], tol=mt["tol"], info=info,
                             return_f64=True)
    assert rel(v64, d["v64"]) < 1e-12
    # one float32 ulp at most
    ulp = np.spacing(np.abs(d["v32"]).astype(np.float32))
    assert (np.abs(v32.astype(np.float64) - d["v32"]) <= ulp).all()
    assert info["flow_iters"] == mt["info"]["flow_iters"]
    assert info["flow_converged"] == mt["info"]["flow_converged"]
    if np.isfinite(mt["info"]["flow_change"]):
        assert abs(info["flow_change"] - mt["info"]["flow_change"]) <= \
            1e-9 * abs(mt["info"]["flow_change"]) + 1e-15


ADR = ["adr2d_channel4", "adr2d_auto", "adr2d_disks_abio",
       "adr2d_disks_default4", "adr2d_sorption_irrev", "adr2d_coupled",
       "adr2d_biofilm", "adr3d_monod", "adr3d_default3_abio"]


@pytest.mark.parametrize("name", ADR)
def test_adr_golden_bit_exact(name):
    d = golden(name)
    mt = d["meta"]
    info = {}
    out, tn = O.solve_adr(d["mat"], d["vel"], info=info,
                          chem=chem_from(mt["chem"]), **mt["kw"])
    assert out.dtype == np.float32 and out.shape == d["out"].shape
    assert np.array_equal(out, d["out"])          # bit for bit (+0 == -0)
    assert np.array_equal(tn, d["t_norm"])
    for k, v in mt["info"].items():
        if k == "chemistry":
            continue
        assert info[k] == pytest.approx(v, rel=0, abs=0), k


def test_geometry_golden():
    d = golden("geometry")
    mt = d["meta"]
    gk, n = O.keep_inlet_connected(d["g"])
    assert np.array_equal(gk, d["kept"]) and n == mt["removed"]
    g1, n1 = O.coat_with_biofilm(gk, 3, 1)
    g2, n2 = O.coat_with_biofilm(gk, 4, 2)
    assert np.array_equal(g1, d["coat1"]) and n1 == mt["n1"]
    assert np.array_equal(g2, d["coat2"]) and n2 == mt["n2"]
    gk2, n2d = O.keep_inlet_connected(d["g2d"])
    assert np.array_equal(gk2, d["kept2d"]) and n2d == mt["removed2d"]


KEYWORDS = {"nan": ["not finite"], "good": [], "over_cap": ["above"],
            "negative": ["negative"], "zero": ["exactly zero"],
            "wrap": ["wrapping"]}


@pytest.mark.parametrize("case", sorted(KEYWORDS))
def test_check_physics_golden(case):
    d = golden("check_physics")
    ref = d["meta"]["complaints"][case]
    got = O.check_physics(d["out_" + case], d["mat_" + case])
    assert len(got) == len(ref)
    for kw in KEYWORDS[case]:
        assert any(kw in s for s in ref) and any(kw in s for s in got)


def test_oracle_geodesic_pinned_by_bfs_and_closed_forms():
    """The relaxation geodesic (3D/tools/build_practice_dataset.py:53-72)
    equals breadth-first distances once it has converged, is x along a
    straight channel, and moves one voxel per pass (Jacobi, not in-place)."""
    rng = np.random.default_rng(3)
    for shape, p in [((12, 9, 7), 0.3), ((20, 6, 5), 0.45), ((15, 15, 1), 0.35)]:
        g = np.where(rng.random(shape) < p, 0, 2).astype(np.uint8)
        d, sw = O.geodesic(g)
        assert sw < 4 * shape[0]
        assert np.array_equal(d, O.bfs_distance(g))
    g = np.full((10, 3, 3), 2, np.uint8)
    d, _ = O.geodesic(g)
    assert np.array_equal(d, np.broadcast_to(np.arange(10, dtype=np.float32)[:, None, None], g.shape))
    d1, sw1 = O.geodesic(g, max_sweeps=1)
    assert sw1 == 1 and d1[1, 0, 0] == 1.0 and d1[2, 0, 0] == np.float32(1e9)
    g[:, 1, 1] = 0                                 # solid core: distance 0 there
    d, _ = O.geodesic(g)
    assert (d[:, 1, 1] == 0).all()
