"""The harmonic pressure prior and its gradient on the GPU against the
oracle (scipy spsolve of the same system): the float32 field within 2e-6
(CG to a relative residual of 1e-12 against a direct solve), the dropped
count exactly, and the gradient bit for bit when both start from the same
field.  Besides generated pore maps, the cases include the reference's
released Monod domain, read as its own feature comparison reads it.
The flow descriptors must match bit for bit."""
import os

import numpy as np
import pytest
import torch

import prtlb
from oracle import features_oracle as F
from prtlb import features

pytestmark = pytest.mark.gpu


def _released_domain():
    """The reference's released 2D Monod domain (tests/golden/
    reference_domains.npz), read as the reference's feature comparison
    reads a released domain (3D/model/test_reference_parity.py:138-150):
    a (64, 148) pore map with its top and bottom rows blocked, transposed
    so that flow runs along axis 0."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "reference_domains.npz")
    with np.load(path, allow_pickle=False) as z:
        m = z["monod"].reshape(64, 148) > 0
    m[0] = m[-1] = False
    return np.ascontiguousarray(m.T)


def _cases():
    rng = np.random.default_rng(0)
    p2 = rng.random((64, 40)) > 0.3
    p2[0] = p2[-1] = True
    p3 = prtlb.sphere_pack((40, 32, 28), (3.0, 6.0), 0.45, seed=2, connect=False) == prtlb.PORE
    p3[0] = p3[-1] = True
    slab = np.ones((21, 9), bool)
    return [("random_2d", p2), ("spheres_3d", p3), ("slab_2d", slab),
            ("released_monod_2d", _released_domain())]


@pytest.mark.parametrize("name,pore", _cases(), ids=[c[0] for c in _cases()])
def test_pressure_matches_direct_solve(name, pore):
    want, wd = F.harmonic_pressure(pore)
    got, gd = features.harmonic_pressure(pore)
    assert gd == wd
    assert got.dtype == np.float32 and got.shape == pore.shape
    assert np.abs(got - want).max() < 2e-6, name
    assert (got[~pore] == 0).all()


@pytest.mark.parametrize("name,pore", _cases(), ids=[c[0] for c in _cases()])
def test_gradient_bit_exact_from_the_same_field(name, pore):
    P, _, _ = features.harmonic_pressure_device(pore)
    for normalise in (False, True):
        got = features.harmonic_gradient_device(pore, normalise, P=P).cpu().numpy()
        want = F.harmonic_gradient(pore, normalise, P=P.cpu().numpy())
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (name, normalise)


def test_gradient_end_to_end_and_no_pore():
    pore = _cases()[1][1]
    got = features.harmonic_gradient(pore)
    want = F.harmonic_gradient(pore)
    assert np.abs(got - want).max() < 1e-4
    assert np.abs(got).max() <= 1.0
    # nothing reaches the inlet: zero field, every pore voxel dropped
    blocked = pore.copy()
    blocked[0] = False
    blocked[1] = False
    P, dropped = features.harmonic_pressure(blocked)
    assert (P == 0).all() and dropped == int(blocked.sum())
    assert (features.harmonic_gradient(blocked) == 0).all()


def test_device_tensor_input():
    pore = _cases()[0][1]
    t = torch.from_numpy(pore.astype(np.uint8)).cuda()
    a, _, iters = features.harmonic_pressure_device(t)
    b, _ = features.harmonic_pressure(pore)
    assert np.array_equal(a.cpu().numpy(), b) and iters > 0


def _desc_cases():
    rng = np.random.default_rng(5)
    p2 = rng.random((60, 36)) > 0.35
    p3 = prtlb.sphere_pack((36, 30, 26), (3.0, 6.0), 0.45, seed=4) == prtlb.PORE
    chan = np.zeros((40, 9), bool)
    chan[:, 1:8] = True
    return [("random_2d", p2), ("spheres_3d", p3), ("channel_2d", chan),
            ("released_monod_2d", _released_domain())]


@pytest.mark.parametrize("name,pore", _desc_cases(), ids=[c[0] for c in _desc_cases()])
def test_descriptors_bit_exact(name, pore):
    for fn, args in ((F.uprm_map, ()), (F.local_thickness, ())):
        want = fn(pore, *args)
        got = (features.uprm_map(pore) if fn is F.uprm_map
               else features.mis_map(pore, buf=0))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (name, fn.__name__)
    for buf in (3, 10):
        want = F.mis_map(pore, buf=buf)
        got = features.mis_map(pore, buf=buf)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (name, "mis", buf)
    want, wr = F.dw2_map(pore)
    got, gr = features.dw2_map(pore)
    assert gr == wr and np.array_equal(got.view(np.uint32), want.view(np.uint32)), (name, "dw2")
    want, wr = F.dw2_map(pore, 2.0, 30.0)
    got, gr = features.dw2_map(pore, 2.0, 30.0)
    assert gr == wr and np.array_equal(got.view(np.uint32), want.view(np.uint32)), (name, "dw2 range")


def test_all_features_dict():
    pore = _desc_cases()[1][1]
    f = features.all_features(pore, buf=4)
    assert set(f) == {"mis", "uprm", "dw2", "dw2_range"}
    assert np.array_equal(f["mis"], F.mis_map(pore, buf=4))
    assert np.array_equal(f["uprm"], F.uprm_map(pore))


def test_extruded_2d_volume_reproduces_the_2d_maps_plane_for_plane():
    """A 3D volume built by repeating a 2D mask along the last axis has, in
    every plane, the 2D answer (no wall along that axis): the check the
    reference's features were held to for their 3D generalisation."""
    rng = np.random.default_rng(11)
    p2 = rng.random((48, 30)) > 0.35
    p3 = np.repeat(p2[:, :, None], 5, axis=2)
    u2, u3 = features.uprm_map(p2), features.uprm_map(p3)
    m2, m3 = features.mis_map(p2, buf=6), features.mis_map(p3, buf=6)
    d2, r2 = features.dw2_map(p2)
    d3, r3 = features.dw2_map(p3)
    P2, _ = features.harmonic_pressure(p2)
    P3, _ = features.harmonic_pressure(p3)
    assert r2 == r3
    for k in range(5):
        assert np.array_equal(u3[:, :, k], u2)
        assert np.array_equal(m3[:, :, k], m2)
        assert np.array_equal(d3[:, :, k], d2)
        assert np.abs(P3[:, :, k] - P2).max() < 2e-6
