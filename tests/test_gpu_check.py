"""GPU geodesic, distance transform and check_physics passes against the
oracle, SciPy and the host check_physics."""
import numpy as np
import pytest

from conftest import golden
from oracle import prtlb_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import prtlb  # noqa: E402
from prtlb import chemistry  # noqa: E402


@pytest.mark.parametrize("shape,phi,seed", [((40, 24, 20), 0.35, 0), ((33, 17, 29), 0.25, 3),
                                            ((64, 48), 0.5, 1)])
def test_geodesic_matches_oracle(shape, phi, seed):
    g = (prtlb.sphere_pack(shape, (2.0, 5.0), phi, seed=seed, connect=False) if len(shape) == 3
         else prtlb.disk_pack(shape, 4.0, phi, seed=seed, connect=False))
    g3 = g if g.ndim == 3 else g[:, :, None]
    ref, sw_ref = O.geodesic(g3)
    d, sw = prtlb.geodesic_device(g)
    assert np.array_equal(d.cpu().numpy().reshape(g3.shape), ref)
    assert sw >= sw_ref


def test_geodesic_sweep_cap_matches_oracle():
    g = prtlb.sphere_pack((48, 20, 20), (2.0, 5.0), 0.4, seed=7, connect=False)
    for cap in (1, 3, 8, 11):
        ref, _ = O.geodesic(g, max_sweeps=cap)
        d, sw = prtlb.geodesic_device(g, max_sweeps=cap)
        assert sw <= cap
        assert np.array_equal(d.cpu().numpy(), ref), cap


@pytest.mark.parametrize("shape,phi,seed", [((40, 24, 20), 0.35, 0), ((31, 45, 17), 0.6, 2),
                                            ((70, 50), 0.55, 4)])
def test_edt_matches_scipy(shape, phi, seed):
    from scipy import ndimage
    g = (prtlb.sphere_pack(shape, (2.0, 6.0), phi, seed=seed, connect=False) if len(shape) == 3
         else prtlb.disk_pack(shape, 5.0, phi, seed=seed, connect=False))
    ref = np.zeros(g.shape, np.float32)
    ref[...] = ndimage.distance_transform_edt(g == prtlb.PORE)
    got = prtlb.edt_device(g).cpu().numpy()
    assert np.array_equal(got, ref)


def test_edt_refuses_all_pore():
    with pytest.raises(ValueError):
        prtlb.edt_device(np.full((8, 8, 8), prtlb.PORE, np.uint8))


def _transport_case():
    g = prtlb.disk_pack((48, 32), 3.0, 0.7, seed=2)
    assert (g[-1] == prtlb.PORE).any()
    v = prtlb.stokes_d2q9(g, nit=600, tol=0)
    ch = chemistry.sorption_irreversible(True)
    out, tn = prtlb.solve_adr(g, v, 10.0, 0.0, 5, n_species=2, steps=400, chem=ch, da_abio=1.0)
    return g, out, ch


def _same(out, g, **kw):
    host = prtlb.check_physics(out, g, **kw)
    dev = prtlb.check_physics_device(out, g, **kw)
    assert dev == host
    return host


@pytest.mark.parametrize("case", ["nan", "good", "over_cap", "negative", "zero", "wrap"])
def test_check_physics_device_matches_host_on_golden(case):
    d = golden("check_physics")
    got = _same(d["out_" + case], d["mat_" + case])
    assert len(got) == len(d["meta"]["complaints"][case])


def test_check_physics_device_each_rule():
    g, out, ch = _transport_case()
    # the sorbed chemical has no feed, so the reference's cap rule flags it
    # (cap 0); every planted fault below adds to that baseline
    base = _same(out, g, chem=ch, da_abio=1.0)
    assert len(base) == 1 and "above the 0.000" in base[0]
    m = g == prtlb.PORE
    bad = out.copy()
    bad[2, 0, 5, 5] = np.nan
    bad[3, 1, 7, 7] = np.inf
    assert len(_same(bad, g, chem=ch)) == 1
    hot = out.copy()
    hot[-1, 0][m] *= 3.0                          # above the feed
    hot[1, 1][m] = -0.5                           # negative
    assert len(_same(hot, g, chem=ch)) >= 3
    lost = out.copy()
    lost[-1, 0] = 0.0                             # donor lost at the end
    assert len(_same(lost, g, chem=ch, da_abio=1.0)) >= 2
    never = out.copy()
    never[:, 0] = 0.0                             # the reactant never there
    assert len(_same(never, g, chem=ch, da_abio=1.0)) >= 3
    chk = out.copy()
    xs = np.arange(g.shape[0])[:, None]
    chk[-1, 0] = np.where(m, 0.5 + 0.4 * (-1.0) ** xs, 0.0)   # checkerboard along x
    assert any("checkerboard" in s for s in _same(chk, g, chem=ch))
    early = out.copy()
    early[:, 0, g.shape[0] // 2] = 0.0
    early[:, 0, -1] = 0.5                          # outlet before mid-plane
    assert any("wrapping" in s for s in _same(early, g))
    cp = out.copy()
    cp[:, 1] = 0.37 * cp[:, 0]                     # a rescaled copy
    assert any("rescaled copy" in s for s in _same(cp, g, chem=ch))


def test_check_physics_device_3d_and_tensor_input():
    g = prtlb.sphere_pack((24, 16, 16), (2.0, 4.0), 0.45, seed=1)
    v = prtlb.stokes_d3q19(g, nit=300, tol=0)
    out, tn = prtlb.solve_adr(g, v, 5.0, 1.0, 4, n_species=1, steps=300,
                              chem=chemistry.monod(0.5))
    host = prtlb.check_physics(out, g, da=1.0)
    dev = prtlb.check_physics_device(torch.from_numpy(out).cuda(), g, da=1.0)
    assert dev == host
