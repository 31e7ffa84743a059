"""Material codes, geometry helpers and the seeded pore-geometry generators.

SOLID / WALL / PORE and the three helpers keep the reference's names and
behaviour (3D/tools/prtlb_3d.py:478, :530-614).  The disk and sphere packs
are this build's own seeded generators: the reference's dataset generators
threshold Gaussian fields, so overlapping disks/spheres with periodic
minimum-image distances are defined here, once, for every workload.
"""
import numpy as np

SOLID, WALL, PORE = 0, 1, 2


def keep_inlet_connected(g):
    """Pore voxels not face-connected to the x = 0 pore become SOLID.

    Returns (g, n_removed), 3D/tools/prtlb_3d.py:530-564.  Face connectivity,
    no wrap (scipy.ndimage.label's default cross structure)."""
    from scipy import ndimage
    g = np.asarray(g)
    pore = g == PORE
    if not pore.any():
        return g, 0
    lab, _ = ndimage.label(pore)
    seeds = np.unique(lab[0][pore[0]])
    seeds = seeds[seeds != 0]
    if seeds.size == 0:
        return g, int(pore.sum())
    drop = pore & ~np.isin(lab, seeds)
    if not drop.any():
        return g, 0
    g = g.copy()
    g[drop] = SOLID
    return g, int(drop.sum())


def coat_with_biofilm(g, code, layers=1):
    """Mark pore voxels within `layers` face-dilations of non-pore as `code`.

    Returns (g, n_marked), 3D/tools/prtlb_3d.py:567-593."""
    from scipy import ndimage
    g = np.asarray(g).copy()
    if code is None or layers <= 0:
        return g, 0
    grown = g != PORE
    for _ in range(int(layers)):
        grown = ndimage.binary_dilation(grown)
    lining = grown & (g == PORE)
    g[lining] = code
    return g, int(lining.sum())


def assert_finite_distance(d, g, name="geodesic"):
    """A complaint string if d is not finite (or >= 1e6) on some pore voxel,
    else None.  3D/tools/prtlb_3d.py:596-614."""
    d = np.asarray(d)
    m = np.asarray(g) == PORE
    if not m.any():
        return None
    v = d[m]
    if np.isfinite(v).all() and float(v.max()) < 1e6:
        return None
    n = int((~np.isfinite(v)).sum() + (v >= 1e6).sum())
    return ("%s is not finite on %d of %d pore voxels (max %.4g). Those voxels "
            "cannot be reached from the inlet; run keep_inlet_connected first."
            % (name, n, int(m.sum()), float(v.max())))


def _grain_pack(shape, radius, porosity, seed, max_grains=10 ** 7):
    """Drop grains (disks in 2D, spheres in 3D) with centres U[0, n) per axis
    and radius `radius` (or U(lo, hi) if a pair) until the pore fraction is at
    most `porosity`.  Distances are periodic minimum-image, so the pack tiles.
    Each grain is stamped through its bounding box only."""
    shape = tuple(int(s) for s in shape)
    rng = np.random.default_rng(seed)
    solid = np.zeros(shape, bool)
    total = solid.size
    n_solid = 0
    nd = len(shape)
    for _ in range(max_grains):
        if (total - n_solid) <= porosity * total:
            break
        c = rng.uniform(0.0, 1.0, size=nd) * np.array(shape, np.float64)
        r = float(radius) if np.isscalar(radius) else float(rng.uniform(*radius))
        d2 = None
        idx = []
        for a, n in enumerate(shape):
            k = np.arange(int(np.floor(c[a] - r)), int(np.ceil(c[a] + r)) + 1)
            i = np.unique(np.mod(k, n))
            d = np.abs(i - c[a])
            d = np.minimum(d, n - d)
            dd = (d * d).reshape([-1 if b == a else 1 for b in range(nd)])
            d2 = dd if d2 is None else d2 + dd
            idx.append(i)
        box = np.ix_(*idx)
        hit = d2 <= r * r
        before = solid[box]
        n_solid += int((hit & ~before).sum())
        solid[box] = before | hit
    return np.where(solid, SOLID, PORE).astype(np.uint8)


def disk_pack(shape=(128, 128), radius=6.0, porosity=0.6, seed=0, connect=True):
    """Random overlapping solid disks (2D), the 128x128 r=6 phi~0.6 workload."""
    g = _grain_pack(shape, radius, porosity, seed)
    return keep_inlet_connected(g)[0] if connect else g


def sphere_pack(shape=(128, 128, 128), radius=(4.0, 8.0), porosity=0.35, seed=0,
                connect=True):
    """Random overlapping solid spheres (3D), radii U(lo, hi)."""
    g = _grain_pack(shape, radius, porosity, seed)
    return keep_inlet_connected(g)[0] if connect else g
