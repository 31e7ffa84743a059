"""The C restatement of the oracle's FP64 Stokes solve.  TEST INFRASTRUCTURE ONLY.

`stokes` here takes the same arguments and returns the same values as
`prtlb_oracle.stokes` (3D/tools/prtlb_2d.py:348-486,
3D/tools/prtlb_3d.py:348-456), iterating in oracle/stokes_c.c instead of
vectorised NumPy, so that a 128^3 D3Q19 box runs in seconds instead of
minutes.  The convergence rule, the final moments, the float32 cast and the
refusals are the NumPy oracle's own (FlowConvergence, checked_flow).

Pinned in tests/test_oracle_c.py against the NumPy oracle (fields within
1e-13, the same stop iteration) and the reference's golden vectors.
"""
import ctypes
import os
import subprocess

import numpy as np

from . import prtlb_oracle as _O

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_stokes.so")
SRC = os.path.join(HERE, "stokes_c.c")
_lib = None


def build(verbose=False):
    """Compile oracle/stokes_c.c with gcc (FP64, no contraction, OpenMP)."""
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c99", "-ffp-contract=off", "-fno-fast-math",
           "-fopenmp", "-Wall", "-o", LIB + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        L = ctypes.CDLL(LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        u8 = ctypes.POINTER(ctypes.c_uint8)
        L.oracle_stokes_init.argtypes = [ctypes.c_int] * 4 + [dp]
        L.oracle_stokes_run.argtypes = ([ctypes.c_int] * 4 + [u8, dp, dp, ctypes.c_int]
                                        + [ctypes.c_double] * 3 + [dp])
        L.oracle_stokes_velocity.argtypes = [ctypes.c_int] * 4 + [u8, dp, ctypes.c_double, dp, dp]
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def stokes(g, nit=30000, force=2e-6, tau_lb=0.8, magic=3.0 / 16.0, tol=1e-6,
           info=None, return_f64=False, every=50, f_init=None):
    """prtlb_oracle.stokes, iterated in C.  With return_f64 also returns
    (v64, f): the FP64 pre-cast field (3, *grid) and the pre-collision
    populations (Q, *grid).  f_init: start from these pre-collision
    populations (Q, *grid) instead of the rest state w."""
    g = np.asarray(g)
    nd = g.ndim
    if nd not in (2, 3):
        raise ValueError("grid must be 2D or 3D")
    q = 9 if nd == 2 else 19
    X, Y, Z = (g.shape[0], 1, g.shape[1]) if nd == 2 else g.shape
    solid = np.ascontiguousarray(g != _O.PORE).astype(np.uint8).ravel()
    fluid = solid == 0
    n_fluid = int(fluid.sum())
    om_p = 1.0 / float(tau_lb)
    om_m = _O.trt_minus(om_p, magic)
    gf = float(force)
    L = lib()
    f = np.empty(q * g.size)
    tmp = np.empty_like(f)
    if f_init is None:
        L.oracle_stokes_init(q, X, Y, Z, _p(f))
    else:
        f_init = np.asarray(f_init, np.float64)
        if f_init.shape != (q,) + g.shape:
            raise ValueError("f_init must have shape %s" % (((q,) + g.shape),))
        f[:] = f_init.ravel()
    conv = _O.FlowConvergence(tol=tol, max_it=nit, every=every, name="%dD flow" % nd)
    s = ctypes.c_double(0.0)
    it = 0
    while it < nit:
        chunk = min(conv.every - (it % conv.every), nit - it) if conv.tol > 0 else nit - it
        rc = L.oracle_stokes_run(q, X, Y, Z, _p(solid, ctypes.c_uint8), _p(f), _p(tmp), chunk,
                                 om_p, om_m, gf, ctypes.byref(s))
        if rc:
            raise RuntimeError("oracle_stokes_run failed")
        it += chunk
        um = s.value / n_fluid if n_fluid else 0.0
        if conv.done(it - 1, um):
            break
        conv.iters = it
    conv.report(info)
    v3 = np.empty(3 * g.size)
    L.oracle_stokes_velocity(q, X, Y, Z, _p(solid, ctypes.c_uint8), _p(f), gf, _p(v3), None)
    v3 = v3.reshape(3, g.size)
    v = np.zeros((3, g.size))
    if nd == 2:
        v[0], v[1] = v3[0], v3[2]   # the 2D grid ran as (nx, 1, ny)
    else:
        v[:] = v3
    v = v.reshape((3,) + g.shape)
    out = _O.checked_flow(v.astype(np.float32), solid.reshape(g.shape).astype(bool), tau_lb, force)
    if return_f64:
        return out, (v, f.reshape((q,) + g.shape))
    return out
