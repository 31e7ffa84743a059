"""Pins of the harmonic-pressure oracle (oracle/features_oracle.py) by what
the mathematics fixes, so that the GPU comparison checks against a known
answer: the exact linear ramp of an open slab and of a walled channel (the
reference's own self-test case, 3D/tools/harmonic_pressure.py:210-218), the
discrete harmonic property and the maximum principle on a random pack,
zero on pockets the inlet does not reach, and mirror symmetry.  A wrong
degree (counting walls), a wrong Dirichlet face or a wrong sign fails one
of them."""
import numpy as np
import pytest

from oracle import features_oracle as F


def _pack2(shape, frac, seed):
    rng = np.random.default_rng(seed)
    return rng.random(shape) > frac


@pytest.mark.parametrize("shape", [(21, 9), (17, 6, 5)])
def test_open_slab_is_the_linear_ramp(shape):
    pore = np.ones(shape, bool)
    P, dropped = F.harmonic_pressure(pore)
    ramp = np.linspace(1.0, 0.0, shape[0]).reshape((-1,) + (1,) * (len(shape) - 1))
    assert dropped == 0
    assert np.abs(P - ramp).max() < 1e-6
    g = F.harmonic_gradient(pore, normalise=False, P=P)
    assert np.abs(g[0] + 1.0 / (shape[0] - 1)).max() < 1e-6
    assert np.abs(g[1:]).max() == 0.0
    gn = F.harmonic_gradient(pore, P=P)
    assert np.abs(gn[0] + 1.0).max() < 1e-5 and np.abs(gn).max() <= 1.0


def test_walled_channel_is_the_ramp_and_walls_carry_no_flux():
    pore = np.zeros((30, 12), bool)
    pore[:, 3:9] = True
    P, _ = F.harmonic_pressure(pore)
    ramp = np.linspace(1.0, 0.0, 30)[:, None]
    assert np.abs(P[:, 3:9] - ramp).max() < 1e-6
    assert (P[:, :3] == 0).all() and (P[:, 9:] == 0).all()


def test_random_pack_is_discretely_harmonic_and_bounded():
    pore = _pack2((40, 24), 0.35, 3)
    pore[0] = pore[-1] = True
    P, dropped = F.harmonic_pressure(pore)
    live = F.inlet_connected(pore)
    assert dropped == int(pore.sum() - live.sum()) and dropped > 0
    assert (P[pore & ~live] == 0).all()
    assert P.min() >= 0.0 and P.max() <= 1.0 + 1e-7
    Pd = P.astype(np.float64)
    worst = 0.0
    for x, y in np.argwhere(live):
        if x in (0, pore.shape[0] - 1):
            continue
        s, d = 0.0, 0
        for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            a, b = x + dx, y + dy
            if 0 <= a < pore.shape[0] and 0 <= b < pore.shape[1] and live[a, b]:
                s += Pd[a, b]
                d += 1
        if d:
            worst = max(worst, abs(s - d * Pd[x, y]))
    assert worst < 1e-5
    assert (P[0][live[0]] == 1.0).all() and (P[-1][live[-1]] == 0.0).all()


def test_mirror_symmetry():
    half = _pack2((24, 8, 6), 0.3, 9)
    pore = np.concatenate([half, half[:, ::-1]], axis=1)
    pore[0] = pore[-1] = True
    P, _ = F.harmonic_pressure(pore)
    assert np.abs(P - P[:, ::-1]).max() < 1e-6


def test_gradient_rules_at_walls():
    """One-sided differences next to a wall, zero in a one-voxel slot."""
    pore = np.zeros((6, 5), bool)
    pore[:, 1:4] = True          # a channel three voxels wide
    P = np.tile(np.array([0.0, 0.5, 0.75, 0.5, 0.0], np.float32), (6, 1))
    g = F.harmonic_gradient(pore, normalise=False, P=P)
    # y: central in the middle column, one-sided on the two wall columns
    assert g[1][2, 2] == np.float32((0.5 - 0.5) * 0.5)
    assert g[1][2, 1] == np.float32(0.75 - 0.5)
    assert g[1][2, 3] == np.float32(0.5 - 0.75)
    slot = np.zeros((6, 5), bool)
    slot[:, 2] = True
    gs = F.harmonic_gradient(slot, normalise=False, P=P)
    assert (gs[1] == 0).all()


# ---------------------------------------------------------------- descriptors
def _widest_path_closure(pore, d):
    """Max-min path values from the inlet by repeated relaxation to a fixed
    point (a different algorithm from the oracle's priority queue)."""
    best = np.where(pore & (np.arange(pore.shape[0])[:, None] == 0), d, 0).astype(np.float64)
    while True:
        new = best.copy()
        for ax in range(pore.ndim):
            for s in (-1, 1):
                nb = np.full(pore.shape, 0.0)
                src = [slice(None)] * pore.ndim
                dst = [slice(None)] * pore.ndim
                if s == 1:
                    src[ax], dst[ax] = slice(0, -1), slice(1, None)
                else:
                    src[ax], dst[ax] = slice(1, None), slice(0, -1)
                okn = np.zeros(pore.shape, bool)
                nb[tuple(dst)] = best[tuple(src)]
                okn[tuple(dst)] = pore[tuple(src)]
                cand = np.where(okn & pore, np.minimum(nb, d), 0.0)
                new = np.maximum(new, cand)
        if np.array_equal(new, best):
            return best
        best = new


def test_uprm_is_the_widest_path():
    rng = np.random.default_rng(4)
    pore = rng.random((30, 22)) > 0.3
    u = F.uprm_map(pore)
    want = _widest_path_closure(pore, F.edt(pore).astype(np.float64))
    assert np.array_equal(u.astype(np.float64), want)
    assert (u <= F.edt(pore)).all() and (u[0][pore[0]] == F.edt(pore)[0][pore[0]]).all()


def test_local_thickness_of_a_channel_is_its_half_width():
    pore = np.zeros((24, 9), bool)
    pore[:, 1:8] = True                 # seven voxels wide: centre row 4 lies 4 from the walls
    lt = F.local_thickness(pore)
    assert (lt[:, 1:8] == 4.0).all() and (lt[~pore] == 0).all()
    # a wide slot and a narrow throat in series: the throat keeps its own radius
    pore2 = np.zeros((40, 13), bool)
    pore2[:, 1:12] = True
    pore2[18:22, 1:5] = pore2[18:22, 8:12] = False
    lt2 = F.local_thickness(pore2)
    assert lt2[20, 6] < lt2[5, 6]
    assert lt2.max() == lt2[5, 6]


def test_mis_paints_the_buffers_and_dw2_scales_to_unit():
    pore = np.zeros((40, 9), bool)
    pore[:, 1:8] = True
    m = F.mis_map(pore, buf=5)
    assert (m[:, 1:8] == 4.0).all()      # the buffers take the interior value
    dw2, (lo, hi) = F.dw2_map(pore)
    assert lo == 1.0 and hi == 16.0
    assert dw2[pore].min() == 0.0 and dw2[pore].max() == 1.0
    assert (dw2[~pore] == 0).all()
