"""The C restatement of the oracle's float32 transport step.  TEST
INFRASTRUCTURE ONLY.

`stepper(S)` returns a function step(S, C) -> C that does what
`prtlb_oracle.adr_step` does (3D/tools/prtlb_3d.py:1392-1413: pin the fed
inlet in place, transport, reactions, Euler, clip), computed in
oracle/adr_c.c from the oracle's own set-up S (`prtlb_oracle.adr_setup`).
Handed to `prtlb_oracle.solve_adr(..., stepper=...)`, it leaves the snapshot
ladder, the stop rule and `info` to the NumPy oracle and only runs the steps
in C, so a 128^3 transport run of 20 000 steps takes seconds instead of
hours.

Pinned in tests/test_oracle_c.py: bit for bit against `adr_step` over every
kinetics kind (Monod, dual Monod with coupled biomass, first order on the
surface, abiotic product, irreversible and reversible sorption, biofilm) in
2D and 3D.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_adr.so")
SRC = os.path.join(HERE, "adr_c.c")
_lib = None


class AdrRx(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "biotic_on", "has_ksd", "has_ksa", "coupled", "alpha_pos", "abio_j", "abio_on",
        "abio_surf", "abio_mk", "sorption", "sorb_c", "sorb_s")] + [(n, ctypes.c_float) for n in (
            "k", "ka", "kd", "dt", "ks_d", "ks_a", "c3_init", "alpha_dt", "neg_alpha", "c2",
            "yield_", "k_sorb", "k_des")]


def build(verbose=False):
    """Compile oracle/adr_c.c with gcc (float32 as written: no contraction,
    no fast math; OpenMP over voxels)."""
    cmd = ["gcc", "-O3", "-fPIC", "-shared", "-std=c99", "-ffp-contract=off", "-fno-fast-math",
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
        P = ctypes.c_void_p
        L.oracle_adr_step.restype = None
        L.oracle_adr_step.argtypes = ([ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, P, P, P, ctypes.c_int64, P, P,
                                       ctypes.POINTER(AdrRx), P, P])
        _lib = L
    return _lib


def _rx(S):
    """The reaction constants as the oracle's float32 expressions form them
    (prtlb_oracle.adr_reaction; Python floats enter as float32, NEP 50)."""
    ch, ns, f32 = S.ch, S.ns, np.float32
    rx = AdrRx()
    rx.biotic_on = int(bool(getattr(ch, "biotic_on", True)))
    rx.has_ksd = int(ch.ks_donor > 0)
    rx.has_ksa = int(not (ch.ks_acceptor <= 0))
    rx.coupled = int(bool(ch.biomass_coupled and ns > 3 and ch.c_init[3] > 0))
    rx.alpha_pos = int(ch.alpha > 0)
    j = getattr(ch, "abiotic_index", 2)
    rx.abio_j = int(j)
    rx.abio_on = int(not (S.ka <= 0 or ns <= j))
    rx.abio_surf = int(bool(ch.abiotic_surface_only))
    mk = getattr(ch, "abiotic_product_index", None)
    rx.abio_mk = int(mk) if (mk is not None and mk < ns and mk != j) else -1
    rx.sorption = int(bool(S.sorption))
    rx.sorb_c = int(S.sorption.get("c", 0)) if S.sorption else 0
    rx.sorb_s = int(S.sorption.get("s", 1)) if S.sorption else 0
    rx.k, rx.ka, rx.kd, rx.dt = S.k, S.ka, S.kd, S.dt
    rx.ks_d, rx.ks_a = f32(ch.ks_donor), f32(ch.ks_acceptor)
    rx.c3_init = f32(ch.c_init[3]) if ns > 3 else f32(1.0)
    rx.alpha_dt = f32(float(ch.alpha) * S.dt)
    rx.neg_alpha = f32(-ch.alpha)
    rx.c2 = f32(1.0 - (ch.yield_ if ns > 3 else 0.0))
    rx.yield_ = f32(ch.yield_)
    rx.k_sorb, rx.k_des = S.k_sorb, S.k_des
    return rx


def stepper(S):
    """step(S, C) -> new C for the oracle set-up S (float32 only).

    The oracle's set-up arrays are taken over the pore voxels into one
    record per voxel, with each face select resolved into the voxel it
    reads (np.where(o_lo, C[m1], C) reads C[o_lo ? m1 : n];
    np.where(f_lo, C[m2], Cm1) reads C[f_lo ? m2 : that]).  Two output buffers take turns; their solid voxels
    stay zero, as np.where(mf, ..., 0.0) makes them, and the oracle keeps
    copies of what it records, so a buffer is reused two steps later."""
    if np.dtype(S.R) != np.float32:
        raise ValueError("the C transport step is float32 only")
    L = lib()
    N, nd, ns = S.mf.size, S.nd, S.ns
    if N >= 2 ** 31:
        raise ValueError("grid too large for the 32-bit voxel indices")
    if ns > 8:
        raise ValueError("the C transport step holds at most 8 chemicals per voxel")
    mov = np.asarray(S.moving, np.int32)
    pore = np.flatnonzero(S.mf)
    n = np.arange(N)
    nm = len(mov)
    per_ax = 6 + 2 * nm
    W = 1 + nd * per_ax + 3
    rec = np.zeros((pore.size, W), np.uint32)
    f32 = rec.view(np.float32)

    def put_f(col, a):
        f32[:, col] = np.asarray(a, np.float32)[pore]

    rec[:, 0] = pore.astype(np.uint32)
    for ax, fa in enumerate(S.faces):
        b = 1 + ax * per_ax
        im1 = np.where(fa["o_lo"], fa["m1"], n)
        ip1 = np.where(fa["o_hi"], fa["p1"], n)
        rec[:, b + 0] = np.where(fa["f_lo"], fa["m2"], im1)[pore]
        rec[:, b + 1] = im1[pore]
        rec[:, b + 2] = ip1[pore]
        rec[:, b + 3] = np.where(fa["f_hi"], fa["p2"], ip1)[pore]
        put_f(b + 4, fa["u_lo"])
        put_f(b + 5, fa["u_hi"])
        for q in range(nm):
            put_f(b + 6 + 2 * q, fa["d_lo"][q])
            put_f(b + 7 + 2 * q, fa["d_hi"][q])
    t = 1 + nd * per_ax
    put_f(t, S.divu)
    put_f(t + 1, S.surf)
    put_f(t + 2, S.surf_sorb if S.surf_sorb is not None else np.zeros(N))
    keep = dict(
        mov=mov, rec=np.ascontiguousarray(rec),
        inlet=np.ascontiguousarray(S.inlet, np.int64),
        fed=np.ascontiguousarray(S.fed, np.float32),
        pin=np.array([int(j in S.moving and S.fed[j] == S.fed[j]) for j in range(ns)], np.int32))
    rx = _rx(S)
    order = ("mov", "rec")
    bufs = [np.zeros((ns, N), np.float32) for _ in range(2)]
    turn = [0]

    def step(S_, C):
        C = np.ascontiguousarray(C, np.float32)
        assert C.shape == (ns, N)
        out = bufs[turn[0]]
        if out is C:
            turn[0] ^= 1
            out = bufs[turn[0]]
        turn[0] ^= 1
        L.oracle_adr_step(N, pore.size, ns, nd, len(mov), *[keep[k].ctypes.data for k in order],
                          keep["inlet"].ctypes.data, keep["inlet"].size, keep["fed"].ctypes.data,
                          keep["pin"].ctypes.data, ctypes.byref(rx), C.ctypes.data, out.ctypes.data)
        return out

    step.keep = keep
    return step
