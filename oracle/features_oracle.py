"""NumPy/SciPy restatement of the harmonic pressure prior and its gradient
(reference: shahram444/PRT-DeepONet-GUI, 3D/tools/harmonic_pressure.py).

TEST INFRASTRUCTURE ONLY: the checker for prtlb.features.  Product code
never imports it.

The linear system is assembled as the reference assembles it -- one row per
inlet-connected pore voxel, a Dirichlet row on x = 0 (P = 1) and x = nx - 1
(P = 0), otherwise the sum over connected face neighbours of (P_nb - P) --
but vectorised over the voxels instead of a Python loop, and solved with the
same library routine (scipy.sparse.linalg.spsolve).  The gradient keeps the
reference's float32 expressions.
"""
import numpy as np
from scipy import ndimage
from scipy.sparse import csr_matrix
from scipy.sparse.linalg import spsolve


def inlet_connected(pore):
    """Pore voxels face-connected to a pore voxel on x = 0
    (3D/tools/harmonic_pressure.py:76-88: label with the cross structure,
    keep the labels present on the inlet face)."""
    pore = np.asarray(pore, bool)
    lab, _ = ndimage.label(pore, ndimage.generate_binary_structure(pore.ndim, 1))
    keep = np.setdiff1d(np.unique(lab[0]), [0])
    return np.isin(lab, keep) & pore


def harmonic_pressure(pore):
    """(float32 field, dropped): 3D/tools/harmonic_pressure.py:91-146."""
    pore = np.asarray(pore, bool)
    live = inlet_connected(pore)
    dropped = int(pore.sum() - live.sum())
    n = int(live.sum())
    if n == 0:
        return np.zeros(pore.shape, np.float32), dropped
    idx = -np.ones(pore.shape, np.int64)
    idx[live] = np.arange(n)
    cells = np.argwhere(live)
    nx = pore.shape[0]
    first = cells[:, 0] == 0
    last = (cells[:, 0] == nx - 1) & ~first
    inner = ~first & ~last
    rows, cols, vals = [], [], []
    b = np.zeros(n)
    k = np.arange(n)
    # Dirichlet rows: P = 1 on the inlet, 0 on the outlet
    rows += [k[first], k[last]]
    cols += [k[first], k[last]]
    vals += [np.ones(first.sum()), np.ones(last.sum())]
    b[first] = 1.0
    deg = np.zeros(n, np.int64)
    for ax in range(pore.ndim):
        for d in (-1, 1):
            nb = cells.copy()
            nb[:, ax] += d
            ok = inner & (nb[:, ax] >= 0) & (nb[:, ax] < pore.shape[ax])
            j = np.full(n, -1)
            j[ok] = idx[tuple(nb[ok].T)]
            ok &= j >= 0
            rows.append(k[ok])
            cols.append(j[ok])
            vals.append(np.ones(ok.sum()))
            deg += ok
    rows.append(k[inner])
    cols.append(k[inner])
    vals.append(np.where(deg[inner] > 0, -deg[inner].astype(float), 1.0))
    A = csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    p = spsolve(A, b)
    field = np.zeros(pore.shape, np.float32)
    field[live] = np.nan_to_num(p, nan=0.0, posinf=0.0, neginf=0.0)
    return field, dropped


def harmonic_gradient(pore, normalise=True, P=None):
    """3D/tools/harmonic_pressure.py:148-197; P: a given pressure field
    (float32), else solved here."""
    pore = np.asarray(pore, bool)
    if P is None:
        P = harmonic_pressure(pore)[0]
    P = np.asarray(P, np.float32)
    nd = pore.ndim
    out = np.zeros((nd,) + pore.shape, np.float32)
    for ax in range(nd):
        if pore.shape[ax] < 2:
            continue
        # neighbour values and validity along this axis, no wrap
        Pp = np.zeros_like(P)
        Pm = np.zeros_like(P)
        okp = np.zeros_like(pore)
        okm = np.zeros_like(pore)
        lo = [slice(None)] * nd
        hi = [slice(None)] * nd
        lo[ax], hi[ax] = slice(0, -1), slice(1, None)
        lo, hi = tuple(lo), tuple(hi)
        Pp[lo], okp[lo] = P[hi], pore[hi]
        Pm[hi], okm[hi] = P[lo], pore[lo]
        both = okp & okm & pore
        onlyp = okp & ~okm & pore
        onlym = okm & ~okp & pore
        g = np.zeros(pore.shape, np.float32)
        g[both] = (Pp[both] - Pm[both]) * 0.5
        g[onlyp] = Pp[onlyp] - P[onlyp]
        g[onlym] = P[onlym] - Pm[onlym]
        out[ax] = g
    out[:, ~pore] = 0.0
    if normalise:
        m = np.abs(out[:, pore]).max() if pore.any() else 0.0
        if m > 0:
            out /= m
    return out


# ---------------------------------------------------------------------------
# flow descriptors (reference: 3D/tools/flow_features.py)
# ---------------------------------------------------------------------------
def edt(pore):
    """distance_transform_edt of the pore mask, float32 (flow_features.py
    uses it for mis, uprm and dw2)."""
    return ndimage.distance_transform_edt(np.asarray(pore, bool)).astype(np.float32)


def disk(radius, ndim):
    """The ball structuring element: offsets in [-ceil r, ceil r] with
    |offset|^2 <= radius * radius, evaluated in the type radius has
    (flow_features.py:119-124)."""
    r = int(np.ceil(radius))
    grid = np.ogrid[tuple(slice(-r, r + 1) for _ in range(ndim))]
    d2 = sum(g * g for g in grid)
    return d2 <= radius * radius


def local_thickness(pore, r_step=0.5):
    """flow_features.py:128-150: radii from the largest distance down in
    steps of r_step; every pore voxel takes the first (largest) radius whose
    ball, centred on a voxel at least that far from a wall, covers it."""
    pore = np.asarray(pore, bool)
    d = edt(pore)
    rmax = float(d.max()) if pore.any() else 0.0
    out = np.zeros(pore.shape, np.float32)
    if rmax <= 0:
        return out
    for r in np.arange(rmax, 0, -r_step):
        centres = pore & (d >= r - 1e-6)
        if not centres.any():
            continue
        painted = ndimage.binary_dilation(centres, structure=disk(r, pore.ndim)) & pore
        new = painted & (out == 0)
        out[new] = r
    out[~pore] = 0.0
    return out


def paint_into_buffers(pore, interior, buf):
    """flow_features.py:153-177: interior values grown into the inlet and
    outlet buffers, largest first."""
    nx = pore.shape[0]
    full = np.zeros(pore.shape, np.float32)
    core = (slice(buf, nx - buf),) + (slice(None),) * (pore.ndim - 1)
    full[core] = interior
    seed = np.zeros_like(full)
    seed[core] = interior
    ends = np.zeros(pore.shape, bool)
    ends[:buf] = True
    ends[nx - buf:] = True
    for r in np.unique(seed[seed > 0])[::-1]:
        reached = ndimage.binary_dilation(seed >= r - 1e-6, structure=disk(r, pore.ndim))
        for lo, hi in ((0, buf), (nx - buf, nx)):
            region = np.zeros(pore.shape, bool)
            region[lo:hi] = True
            grow = reached & region & (full < r)
            full[grow] = r
    return full, ends


def mis_map(pore, buf=10, r_step=0.5):
    """flow_features.py:180-201."""
    pore = np.asarray(pore, bool)
    nx = pore.shape[0]
    if buf <= 0 or 2 * buf >= nx:
        return local_thickness(pore, r_step)
    core = (slice(buf, nx - buf),) + (slice(None),) * (pore.ndim - 1)
    inner = local_thickness(pore[core], r_step)
    full, ends = paint_into_buffers(pore, inner, buf)
    vmin = float(inner[inner > 0].min()) if (inner > 0).any() else 0.0
    full[pore & ends & (full <= 0)] = vmin
    full[~pore] = 0.0
    return full.astype(np.float32)


def uprm_map(pore):
    """flow_features.py:205-257: the widest path from the inlet, by a
    max-first priority queue over face neighbours; a value passing a voxel
    is capped at that voxel's wall distance."""
    import heapq
    pore = np.asarray(pore, bool)
    d = edt(pore)
    best = np.zeros(pore.shape, np.float32)
    done = np.zeros(pore.shape, bool)
    heap = []
    for c in np.argwhere(pore[0]):
        i = (0,) + tuple(int(v) for v in c)
        best[i] = d[i]
        heapq.heappush(heap, (-float(best[i]), i))
    nd = pore.ndim
    steps = [tuple(s if k == ax else 0 for k in range(nd)) for ax in range(nd) for s in (-1, 1)]
    while heap:
        v, i = heapq.heappop(heap)
        if done[i]:
            continue
        done[i] = True
        for s in steps:
            j = tuple(i[k] + s[k] for k in range(nd))
            if any(j[k] < 0 or j[k] >= pore.shape[k] for k in range(nd)):
                continue
            if not pore[j] or done[j]:
                continue
            cand = min(-v, float(d[j]))
            if cand > best[j]:
                best[j] = cand
                heapq.heappush(heap, (-cand, j))
    best[~pore] = 0.0
    return best


def dw2_map(pore, dw2_min=None, dw2_max=None):
    """flow_features.py:260-280: squared wall distance scaled to [0, 1] over
    the pore; returns (map, (lo, hi))."""
    pore = np.asarray(pore, bool)
    f2 = (edt(pore) * edt(pore)).astype(np.float32)
    vals = f2[pore]
    lo = float(dw2_min) if dw2_min is not None else (float(vals.min()) if vals.size else 0.0)
    hi = float(dw2_max) if dw2_max is not None else (float(vals.max()) if vals.size else 1.0)
    if hi <= lo:
        hi = lo + 1.0
    out = np.zeros_like(f2)
    out[pore] = np.clip((f2[pore] - lo) / (hi - lo), 0.0, 1.0)
    return out.astype(np.float32), (lo, hi)
