"""Material codes, geometry helpers and the seeded pore-geometry generators.

SOLID / WALL / PORE and the three helpers keep the reference's names and
behaviour (3D/tools/prtlb_3d.py:478, :530-614).  The disk and sphere packs
are this package's own seeded generators: the reference's dataset generators
threshold Gaussian fields, so overlapping disks/spheres with periodic
minimum-image distances are defined here, once, for every workload.
"""
import ctypes

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


def _candidates(shape, radius, seed, k):
    """The first k grains the seeded generator draws, in order."""
    rng = np.random.default_rng(seed)
    nd = len(shape)
    cs = np.empty((k, nd))
    rs = np.empty(k)
    ext = np.array(shape, np.float64)
    for i in range(k):
        cs[i] = rng.uniform(0.0, 1.0, size=nd) * ext
        rs[i] = float(radius) if np.isscalar(radius) else float(rng.uniform(*radius))
    return cs, rs


def _cell_lists(shape, cs, rs, cell):
    """Per cell of edge `cell`, the grains whose periodic bounding box meets
    it, in ascending grain order (CSR: start offsets, items)."""
    nd = len(shape)
    ncell = [(n + cell - 1) // cell for n in shape]
    pairs_cell, pairs_k = [], []
    for k in range(len(rs)):
        axes = []
        for a, n in enumerate(shape):
            v = np.arange(int(np.floor(cs[k, a] - rs[k])), int(np.ceil(cs[k, a] + rs[k])) + 1)
            axes.append(np.unique(np.mod(v, n) // cell))
        grid = np.ix_(*axes)
        flat = np.zeros([len(x) for x in axes], np.int64)
        mul = 1
        for a in range(nd - 1, -1, -1):
            flat = flat + grid[a] * mul
            mul *= ncell[a]
        flat = flat.ravel()
        pairs_cell.append(flat)
        pairs_k.append(np.full(flat.size, k, np.int32))
    cells = np.concatenate(pairs_cell)
    ks = np.concatenate(pairs_k)
    order = np.lexsort((ks, cells))
    cells, ks = cells[order], ks[order]
    start = np.zeros(int(np.prod(ncell)) + 1, np.int32)
    np.add.at(start, cells + 1, 1)
    return np.cumsum(start).astype(np.int32), ks


def grain_pack_device(shape, radius, porosity, seed, device=None, cell=16):
    """The seeded grain pack of `_grain_pack`, built on the GPU; returns the
    material grid as a uint8 device tensor, bit-identical to the host one."""
    import torch
    from ._lib import call
    from .device import device as _dev
    from .device import ptr, stream_ptr
    dev = _dev(device)
    shape = tuple(int(s) for s in shape)
    nd = len(shape)
    N = int(np.prod(shape))
    dims = (shape[0], shape[1], 1 if nd == 2 else shape[2])
    mean_r = float(radius) if np.isscalar(radius) else 0.5 * (radius[0] + radius[1])
    vol = np.pi * mean_r ** 2 if nd == 2 else 4.0 / 3.0 * np.pi * mean_r ** 3
    k = int(1.5 * max(-np.log(max(porosity, 1e-6)), 0.05) * N / vol) + 64
    s = stream_ptr(dev)
    while True:
        cs, rs = _candidates(shape, radius, seed, k)
        start, items = _cell_lists(shape, cs, rs, cell)
        t_cs = torch.from_numpy(cs).to(dev)
        t_rs = torch.from_numpy(rs).to(dev)
        t_start = torch.from_numpy(start).to(dev)
        t_items = torch.from_numpy(items).to(dev)
        first = torch.empty(N, dtype=torch.int32, device=dev)
        call("prt_grain_first_cover", *dims, ptr(t_cs), ptr(t_rs), k, ptr(t_start), ptr(t_items),
             cell, ptr(first), s)
        hist = torch.zeros(k + 1, dtype=torch.int32, device=dev)
        call("prt_grain_histogram", ptr(first), N, ptr(hist), s)
        h = hist.cpu().numpy().astype(np.int64)
        solid_after = np.concatenate([[0], np.cumsum(h[:k])])     # solid voxels after j grains
        ok = np.flatnonzero((N - solid_after) <= porosity * N)
        if ok.size:
            kstar = int(ok[0])
            break
        k *= 2
    g = torch.empty(N, dtype=torch.uint8, device=dev)
    call("prt_grain_mask", ptr(first), N, kstar, SOLID, PORE, ptr(g), s)
    return g.reshape(shape)


def keep_inlet_connected_device(g, max_sweeps=1 << 20):
    """keep_inlet_connected on a uint8 device tensor, in place; returns
    (g, n_removed)."""
    import torch
    from ._lib import call
    from .device import ptr, stream_ptr
    shape = tuple(g.shape)
    dims = (shape[0], shape[1], 1 if len(shape) == 2 else shape[2])
    reach = torch.empty(g.numel(), dtype=torch.uint8, device=g.device)
    scratch = torch.zeros(2, dtype=torch.int32, device=g.device)
    removed = ctypes.c_int()
    call("prt_keep_inlet_connected", ptr(g), *dims, PORE, SOLID, ptr(reach), ptr(scratch),
         int(max_sweeps), ctypes.byref(removed), stream_ptr(g.device))
    return g, int(removed.value)


def coat_with_biofilm_device(g, code, layers=1):
    """coat_with_biofilm on a uint8 device tensor, in place; (g, n_marked)."""
    import torch
    from ._lib import call
    from .device import ptr, stream_ptr
    if code is None or layers <= 0:
        return g, 0
    shape = tuple(g.shape)
    dims = (shape[0], shape[1], 1 if len(shape) == 2 else shape[2])
    wa = torch.empty(g.numel(), dtype=torch.uint8, device=g.device)
    wb = torch.empty_like(wa)
    scratch = torch.zeros(1, dtype=torch.int32, device=g.device)
    marked = ctypes.c_int()
    call("prt_coat_with_biofilm", ptr(g), *dims, PORE, int(code), int(layers), ptr(wa), ptr(wb),
         ptr(scratch), ctypes.byref(marked), stream_ptr(g.device))
    return g, int(marked.value)


def _pack(shape, radius, porosity, seed, connect, device):
    if device is None:
        g = _grain_pack(shape, radius, porosity, seed)
        return keep_inlet_connected(g)[0] if connect else g
    g = grain_pack_device(shape, radius, porosity, seed, device=device)
    if connect:
        keep_inlet_connected_device(g)
    return g.cpu().numpy()


def disk_pack(shape=(128, 128), radius=6.0, porosity=0.6, seed=0, connect=True, device=None):
    """Random overlapping solid disks (2D), the 128x128 r=6 phi~0.6 workload.
    device: build it on that GPU (bit-identical to the host build)."""
    return _pack(shape, radius, porosity, seed, connect, device)


def sphere_pack(shape=(128, 128, 128), radius=(4.0, 8.0), porosity=0.35, seed=0,
                connect=True, device=None):
    """Random overlapping solid spheres (3D), radii U(lo, hi).
    device: build it on that GPU (bit-identical to the host build)."""
    return _pack(shape, radius, porosity, seed, connect, device)


def _device_grid(g, device=None):
    import torch
    from .device import device as _dev
    if isinstance(g, torch.Tensor):
        t = g.to(torch.uint8).contiguous()
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(g), np.uint8)).to(_dev(device))
    shape = tuple(t.shape)
    if len(shape) not in (2, 3):
        raise ValueError("grid must be 2D or 3D")
    dims = (shape[0], shape[1], 1 if len(shape) == 2 else shape[2])
    return t, shape, dims


def geodesic_device(g, max_sweeps=None, device=None):
    """Geodesic distance from the inlet through the pore on the GPU: the
    relaxation the reference's dataset builders store as geom/gdf
    (3D/tools/build_practice_dataset.py:53-72), float32, bit-identical.
    Returns (distance tensor shaped like g, sweeps run)."""
    import torch
    from ._lib import call
    from .device import ptr, stream_ptr
    t, shape, dims = _device_grid(g, device)
    n = t.numel()
    if max_sweeps is None:
        max_sweeps = 4 * shape[0]
    da = torch.empty(n, dtype=torch.float32, device=t.device)
    db = torch.empty_like(da)
    out = torch.empty_like(da)
    flag = torch.zeros(1, dtype=torch.int32, device=t.device)
    sw = ctypes.c_int()
    call("prt_geodesic", ptr(t), *dims, PORE, ptr(da), ptr(db), ptr(flag), int(max_sweeps), ptr(out),
         ctypes.byref(sw), stream_ptr(t.device))
    return out.reshape(shape), int(sw.value)


def edt_device(g, device=None):
    """Euclidean distance of each pore voxel to the nearest non-pore voxel on
    the GPU, float32: scipy.ndimage.distance_transform_edt(g == PORE) as the
    reference's dataset builders store it (3D/tools/build_dataset_3d.py:659),
    bit-identical.  The grid must hold a non-pore voxel."""
    import torch
    from ._lib import call
    from .device import ptr, stream_ptr
    t, shape, dims = _device_grid(g, device)
    if not bool((t != PORE).any().item()):
        raise ValueError("edt_device: the grid has no non-pore voxel to measure from")
    n = t.numel()
    wa = torch.empty(n, dtype=torch.int32, device=t.device)
    wb = torch.empty_like(wa)
    out = torch.empty(n, dtype=torch.float32, device=t.device)
    call("prt_edt", ptr(t), *dims, PORE, ptr(wa), ptr(wb), ptr(out), stream_ptr(t.device))
    return out.reshape(shape)
