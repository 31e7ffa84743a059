"""Per-rock flow features on the GPU: the harmonic pressure prior and its
gradient (3D/tools/harmonic_pressure.py) and the MIS, UPRM and dw2 pore
descriptors (3D/tools/flow_features.py), the inputs the reference's velocity
operator is trained on (3D/model/train_velocity.py:194,238,
3D/model/predict_velocity.py:115,142).

Same names, arguments and return values as the reference modules:

    harmonic_pressure(pore, report=False) -> (float32 field, dropped)
    harmonic_gradient(pore, normalise=True) -> float32 (ndim,) + pore.shape
    mis_map(pore, buf=10, r_step=0.5), uprm_map(pore) -> float32 maps
    dw2_map(pore, dw2_min=None, dw2_max=None) -> (float32 map, (lo, hi))
    all_features(pore, buf=10, dw2_range=None, r_step=0.5) -> dict

with NumPy arrays in and out; the *_device forms keep the result on the GPU.
Flow runs along axis 0.  The reference solves the sparse system directly
(scipy spsolve); here it is FP64 conjugate gradients to a relative residual
of 1e-12, so the float32 field agrees with the direct solve to rounding.
The descriptors are bit-identical to the reference's maps.
"""
import ctypes

import numpy as np
import torch

from ._lib import PrtError, call, lib
from .device import device as _device
from .device import ptr, stream_ptr

__all__ = ["harmonic_pressure", "harmonic_gradient", "harmonic_pressure_device",
           "harmonic_gradient_device", "mis_map", "uprm_map", "dw2_map", "all_features",
           "mis_map_device", "uprm_map_device", "dw2_map_device", "all_features_device"]

DEFAULT_BUFFER = 10      # 3D/tools/flow_features.py:86


def _mask(pore, device=None):
    """(uint8 0/1 device mask, shape, kernel dims) of a 2D or 3D pore mask."""
    if isinstance(pore, torch.Tensor):
        t = (pore != 0).to(torch.uint8).contiguous()
    else:
        a = np.asarray(pore, bool)
        t = torch.from_numpy(np.ascontiguousarray(a).astype(np.uint8)).to(_device(device))
    shape = tuple(t.shape)
    if len(shape) not in (2, 3):
        raise ValueError("pore must be 2D or 3D, got %dD" % len(shape))
    dims = (shape[0], 1, shape[1]) if len(shape) == 2 else shape
    return t, shape, dims


def harmonic_pressure_device(pore, tol=1e-12, max_iter=None, device=None):
    """The harmonic pressure as a float32 device tensor shaped like pore,
    with the count of pore voxels the inlet does not reach and the CG
    iterations: (field, dropped, iterations)."""
    t, shape, dims = _mask(pore, device)
    n = t.numel()
    wb = int(lib().prt_harmonic_work_size(*dims))
    work = torch.empty(wb, dtype=torch.uint8, device=t.device)
    out = torch.empty(n, dtype=torch.float32, device=t.device)
    if max_iter is None:
        max_iter = max(20000, 40 * max(shape))
    dropped, iters = ctypes.c_int(), ctypes.c_int()
    call("prt_harmonic_pressure", ptr(t), *dims, ptr(work), wb, float(tol), int(max_iter), ptr(out),
         ctypes.byref(dropped), ctypes.byref(iters), stream_ptr(t.device))
    return out.reshape(shape), int(dropped.value), int(iters.value)


def harmonic_pressure(pore, report=False):
    """Laplace's equation on the pore space, flow along axis 0
    (3D/tools/harmonic_pressure.py:91-146): (float32 field, dropped)."""
    field, dropped, _ = harmonic_pressure_device(pore)
    if report:
        n = int(np.asarray(pore, bool).sum()) - dropped
        if n == 0:
            print("  no pore voxel connects to the inlet; pressure is zero everywhere")
        else:
            print("  solved %d cells, %d pore voxels dropped as unreachable" % (n, dropped))
    return field.cpu().numpy(), dropped


def harmonic_gradient_device(pore, normalise=True, P=None, device=None):
    """The gradient components (ndim,) + pore.shape as a float32 device
    tensor; P: a precomputed pressure field (device tensor), else solved."""
    t, shape, dims = _mask(pore, device)
    if P is None:
        P = harmonic_pressure_device(t)[0]
    P = P.to(device=t.device, dtype=torch.float32).contiguous()
    if tuple(P.shape) != shape:
        raise PrtError("harmonic_gradient: P has shape %s, pore %s" % (tuple(P.shape), shape))
    nd = len(shape)
    n = t.numel()
    out = torch.empty((nd, n), dtype=torch.float32, device=t.device)
    scratch = torch.empty((n + 255) // 256 + 1, dtype=torch.float32, device=t.device)
    call("prt_harmonic_gradient", ptr(t), *dims, nd, ptr(P), int(bool(normalise)), ptr(out),
         ptr(scratch), stream_ptr(t.device))
    return out.reshape((nd,) + shape)


def harmonic_gradient(pore, normalise=True):
    """3D/tools/harmonic_pressure.py:148-197, NumPy in and out."""
    return harmonic_gradient_device(pore, normalise).cpu().numpy()


def _edt(t, shape, dims):
    """Wall distance of every pore voxel (mask t, 1 = pore), float32."""
    if not bool((t == 0).any().item()):
        raise ValueError("the pore mask has no wall voxel to measure distances from")
    n = t.numel()
    wa = torch.empty(n, dtype=torch.int32, device=t.device)
    wb = torch.empty_like(wa)
    out = torch.empty(n, dtype=torch.float32, device=t.device)
    call("prt_edt", ptr(t), *dims, 1, ptr(wa), ptr(wb), ptr(out), stream_ptr(t.device))
    return out


def uprm_map_device(pore, device=None):
    """UPRM (3D/tools/flow_features.py:205-257) as a float32 device tensor:
    the largest sphere the inlet can deliver to each voxel."""
    t, shape, dims = _mask(pore, device)
    e = _edt(t, shape, dims)
    best = torch.empty_like(e)
    flag = torch.zeros(1, dtype=torch.int32, device=t.device)
    sweeps = ctypes.c_int()
    call("prt_uprm", ptr(t), *dims, ptr(e), ptr(best), ptr(flag), 1 << 30, ctypes.byref(sweeps),
         stream_ptr(t.device))
    return best.reshape(shape)


def dw2_map_device(pore, dw2_min=None, dw2_max=None, device=None):
    """dw2 (3D/tools/flow_features.py:260-280): (float32 device map, (lo, hi))."""
    t, shape, dims = _mask(pore, device)
    e = _edt(t, shape, dims)
    f2 = (e * e)[t.reshape(-1) != 0]
    lo = float(dw2_min) if dw2_min is not None else (float(f2.min().item()) if f2.numel() else 0.0)
    hi = float(dw2_max) if dw2_max is not None else (float(f2.max().item()) if f2.numel() else 1.0)
    if hi <= lo:
        hi = lo + 1.0
    out = torch.empty_like(e)
    # the reference subtracts and divides in float32 (NumPy casts the Python
    # floats lo and hi - lo to the array's type)
    call("prt_dw2", ptr(t), t.numel(), ptr(e), ctypes.c_float(np.float32(lo)),
         ctypes.c_float(np.float32(hi - lo)), ptr(out), stream_ptr(t.device))
    return out.reshape(shape), (lo, hi)


def _local_thickness(t, shape, dims, r_step):
    """Local thickness of mask t (flow_features.py:128-150), float32 device."""
    e = _edt(t, shape, dims)
    rmax = float(e.max().item()) if bool(t.any().item()) else 0.0
    out = torch.zeros_like(e)
    if rmax <= 0:
        return out
    rc = torch.empty(t.numel(), dtype=torch.float64, device=t.device)
    call("prt_local_thickness", ptr(t), *dims, ptr(e), rmax, float(r_step), ptr(rc), ptr(out),
         stream_ptr(t.device))
    return out


def mis_map_device(pore, buf=DEFAULT_BUFFER, r_step=0.5, device=None):
    """MIS (3D/tools/flow_features.py:180-201) as a float32 device tensor: the
    local thickness of the interior, painted into the inlet and outlet
    buffers of `buf` voxels."""
    t, shape, dims = _mask(pore, device)
    nx = shape[0]
    if buf <= 0 or 2 * buf >= nx:
        return _local_thickness(t, shape, dims, r_step).reshape(shape)
    core = t[buf:nx - buf].contiguous()
    cshape = tuple(core.shape)
    cdims = (cshape[0], 1, cshape[1]) if len(cshape) == 2 else cshape
    inner = _local_thickness(core, cshape, cdims, r_step)
    full = torch.zeros(shape, dtype=torch.float32, device=t.device)
    full[buf:nx - buf] = inner.reshape(cshape)
    pos = inner[inner > 0]
    vmin = float(pos.min().item()) if pos.numel() else 0.0
    vmax = float(inner.max().item()) if inner.numel() else 0.0
    call("prt_paint_buffers", ptr(t), *dims, int(buf), ctypes.c_float(vmax), ctypes.c_float(vmin),
         ptr(full), stream_ptr(t.device))
    return full


def all_features_device(pore, buf=DEFAULT_BUFFER, dw2_range=None, r_step=0.5, device=None):
    """MIS, UPRM and dw2 in one call (3D/tools/flow_features.py:297-313),
    device tensors in a dict with 'mis', 'uprm', 'dw2' and 'dw2_range'."""
    lo_hi = dw2_range if dw2_range is not None else (None, None)
    dw2, used = dw2_map_device(pore, lo_hi[0], lo_hi[1], device=device)
    return {"mis": mis_map_device(pore, buf=buf, r_step=r_step, device=device),
            "uprm": uprm_map_device(pore, device=device), "dw2": dw2, "dw2_range": used}


def mis_map(pore, buf=DEFAULT_BUFFER, r_step=0.5):
    return mis_map_device(pore, buf, r_step).cpu().numpy()


def uprm_map(pore, connectivity=1):
    if connectivity != 1:
        raise ValueError("uprm_map only supports face connectivity, to match the solver")
    return uprm_map_device(pore).cpu().numpy()


def dw2_map(pore, dw2_min=None, dw2_max=None):
    m, used = dw2_map_device(pore, dw2_min, dw2_max)
    return m.cpu().numpy(), used


def all_features(pore, buf=DEFAULT_BUFFER, dw2_range=None, r_step=0.5):
    f = all_features_device(pore, buf, dw2_range, r_step)
    return {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in f.items()}
