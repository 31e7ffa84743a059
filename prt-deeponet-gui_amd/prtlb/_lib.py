"""ctypes binding of libprtlb_hip.so (declarations in include/prtlb.h).

The library is required: there is no CPU fallback anywhere in `prtlb`.  If it
is missing or no HIP device is visible, every solver raises.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libprtlb_hip.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "prtlb.h")

PRT_STORE = {"fp64": 0, "fp32": 1, "fp16": 2}
PRT_LBM_SKIP_SOLID = 1


PRT_LBM_NT_LOADS = 1 << 18
PRT_LBM_SPARSE = 1 << 1
PRT_LBM_XCD = 1 << 19
PRT_LBM_SLAB = 1 << 20


def lbm_tune(bz_log2, by_log2, npt, nt_loads=False, xcd=False):
    """PRT_LBM_TUNE (| PRT_LBM_NT_LOADS | PRT_LBM_XCD): launch bits for
    prt_lbm_step / _run."""
    return ((bz_log2 << 8) | (by_log2 << 12) | ((npt - 1) << 16) | (PRT_LBM_NT_LOADS if nt_loads else 0)
            | (PRT_LBM_XCD if xcd else 0))


def batch_shape(t_log2, e_log2, xcd):
    """PRT_BATCH_SHAPE: launch shape of a batched step (2^t_log2 threads per
    block, 2^e_log2 list entries per thread, XCD-contiguous block order)."""
    return t_log2 | (e_log2 << 4) | (int(bool(xcd)) << 8)


BATCH_SHAPE_LBM_DEFAULT = batch_shape(8, 0, 1)
BATCH_SHAPE_ADR_DEFAULT = batch_shape(8, 0, 0)
# every shape the batched steps take (prt_batch_shape_tile is non-zero)
BATCH_SHAPES = [batch_shape(t, e, x) for x in (0, 1)
                for t, e in ((8, 0), (7, 1), (6, 2), (8, 1), (7, 2))]


def batch_shape_name(code):
    return "%dx%d%s" % (1 << (code & 15), 1 << ((code >> 4) & 15), "x" if code >> 8 else "")


class PrtError(RuntimeError):
    """A call into libprtlb_hip.so failed."""


class AdrParams(ctypes.Structure):
    """Mirror of prt_adr_params (include/prtlb.h)."""
    _fields_ = [
        ("nd", ctypes.c_int32), ("ns", ctypes.c_int32),
        ("biotic_on", ctypes.c_int32), ("coupled", ctypes.c_int32),
        ("abio_index", ctypes.c_int32), ("abio_target", ctypes.c_int32),
        ("abio_product", ctypes.c_int32), ("surface_only", ctypes.c_int32),
        ("sorption", ctypes.c_int32), ("sorb_c", ctypes.c_int32),
        ("sorb_s", ctypes.c_int32), ("write_solid", ctypes.c_int32),
        ("dt", ctypes.c_double), ("k", ctypes.c_double), ("ka", ctypes.c_double),
        ("kd", ctypes.c_double), ("ks_donor", ctypes.c_double),
        ("ks_acceptor", ctypes.c_double), ("alpha_dt", ctypes.c_double),
        ("neg_alpha", ctypes.c_double), ("alpha_pos", ctypes.c_double),
        ("c2", ctypes.c_double), ("yield_", ctypes.c_double),
        ("c3_init", ctypes.c_double), ("k_sorb", ctypes.c_double),
        ("k_desorb", ctypes.c_double),
        ("mobile", ctypes.c_int32 * 8), ("pinned", ctypes.c_int32 * 8),
        ("fed", ctypes.c_double * 8), ("dface", (ctypes.c_double * 4) * 8),
        ("immobile_list", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class LbmSegment(ctypes.Structure):
    """Mirror of prt_lbm_segment."""
    _fields_ = [("solid", ctypes.c_void_p), ("f_a", ctypes.c_void_p), ("f_b", ctypes.c_void_p),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("phase", ctypes.c_int32), ("omega_p", ctypes.c_double),
                ("omega_m", ctypes.c_double), ("force", ctypes.c_double),
                ("fp16_scale", ctypes.c_double), ("link", ctypes.c_void_p),
                ("n_fluid", ctypes.c_int64), ("link_bits", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class AdrSegment(ctypes.Structure):
    """Mirror of prt_adr_segment."""
    _fields_ = [("params", ctypes.POINTER(AdrParams)), ("nx", ctypes.c_int32),
                ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("phase", ctypes.c_int32),
                ("pore_table", ctypes.c_void_p), ("n_pore", ctypes.c_int64),
                ("c_a", ctypes.c_void_p), ("c_b", ctypes.c_void_p)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double

_SIGS = {
    "prt_last_error": (ctypes.c_char_p, []),
    "prt_version": (ctypes.c_char_p, []),
    "prt_adr_params_size": (ctypes.c_size_t, []),
    "prt_lbm_fill_rest": (_I, [_I, _I, _I, _I, _I, _P, _P]),
    "prt_lbm_pop_elems": (_I64, [_I, _I, _I, _I, _I]),
    "prt_lbm_lattice": (_I, [_I, _P, _P, _P]),
    "prt_lbm_step": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _D, _D, _D, _D, _I, _P, _P]),
    "prt_lbm_partials_count": (_I64, [_I, _I, _I, _I]),
    "prt_lbm_run": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _I, _D, _D, _D, _D, _I, _P, _P]),
    "prt_sum_f64": (_I, [_P, _I64, _P, _P]),
    "prt_lbm_static_partials_max": (_I64, []),
    "prt_lbm_run_sum_static": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _I, _D, _D, _D, _D, _I, _P, _P]),
    "prt_lbm_velocity": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _I, _D, _D, _P, _P, _P, _P]),
    "prt_plane_flux": (_I, [_P, _I, _I, _I, _P, _P]),
    "prt_adr_step": (_I, [ctypes.POINTER(AdrParams), _I, _I, _I, _P, _P, _P, _P, _P]),
    "prt_adr_step_f64": (_I, [ctypes.POINTER(AdrParams), _I, _I, _I, _P, _P, _P, _P, _P]),
    "prt_adr_run": (_I, [ctypes.POINTER(AdrParams), _I, _I, _I, _P, _P, _P, _P, _I, _P]),
    "prt_lbm_link_stride": (_I64, [_I64]),
    "prt_lbm_launch_blocks": (_I64, [_I, _I, _I, _I, _I]),
    "prt_trt_minus": (_D, [_D, _D]),
    "prt_lbm_solver_create": (_I, [_I, _I, _I, _I, _I, _P, ctypes.c_longlong, _P, _P, _P, _P, _D, _D, _D,
                                   _D, _I, _P, _P]),
    "prt_lbm_solver_init": (_I, [_P]),
    "prt_lbm_solver_step_n": (_I, [_P, _I]),
    "prt_lbm_solver_reduce": (_I, [_P, _I, _P]),
    "prt_lbm_solver_velocity": (_I, [_P, _P, _P]),
    "prt_lbm_solver_newest": (_P, [_P]),
    "prt_lbm_solver_destroy": (_I, [_P]),
    "prt_adr_solver_create": (_I, [ctypes.POINTER(AdrParams), _I, _I, _I, _P, ctypes.c_longlong,
                                   _P, _P, _P, _P]),
    "prt_adr_solver_step_n": (_I, [_P, _I]),
    "prt_adr_solver_newest": (_P, [_P]),
    "prt_adr_solver_snapshot": (_I, [_P, _P]),
    "prt_adr_solver_destroy": (_I, [_P]),
    "prt_lbm_link_table": (_I, [_I, _I, _I, _I, _P, _I64, _P, _P, _P]),
    "prt_lbm_sparse_fill_rest": (_I, [_I, _I, _I64, _P, _P]),
    "prt_lbm_sparse_partials_count": (_I, [_I64]),
    "prt_lbm_link_pack16": (_I, [_I, _I64, _P, _P, _P, _P]),
    "prt_lbm_sparse_run": (_I, [_I, _I, _I64, _P, _I, _P, _P, _I, _D, _D, _D, _D, _P, _P]),
    "prt_lbm_sparse_velocity": (_I, [_I, _I, _I64, _I64, _P, _P, _I, _P, _P, _D, _D, _P, _P, _P,
                                     _P]),
    "prt_geodesic": (_I, [_P, _I, _I, _I, _I, _P, _P, _P, _I, _P, _P, _P]),
    "prt_edt": (_I, [_P, _I, _I, _I, _I, _P, _P, _P, _P]),
    "prt_harmonic_work_size": (_I64, [_I, _I, _I]),
    "prt_harmonic_pressure": (_I, [_P, _I, _I, _I, _P, _I64, _D, _I, _P, _P, _P, _P]),
    "prt_harmonic_gradient": (_I, [_P, _I, _I, _I, _I, _P, _I, _P, _P, _P]),
    "prt_uprm": (_I, [_P, _I, _I, _I, _P, _P, _P, _I, _P, _P]),
    "prt_dw2": (_I, [_P, _I64, _P, ctypes.c_float, ctypes.c_float, _P, _P]),
    "prt_local_thickness": (_I, [_P, _I, _I, _I, _P, _D, _D, _P, _P, _P]),
    "prt_paint_buffers": (_I, [_P, _I, _I, _I, _I, ctypes.c_float, ctypes.c_float, _P, _P]),
    "prt_check_blocks": (_I, [_I64]),
    "prt_check_nonfinite": (_I, [_P, _I64, _I, _P, _P]),
    "prt_check_pore_range": (_I, [_P, _I, _I, _I64, _P, _I, _P, _P, _P]),
    "prt_check_plane_max": (_I, [_P, _I, _I, _I, _I64, _I, _P, _P]),
    "prt_check_checker": (_I, [_P, _P, _I, _I64, _I, _P, _P]),
    "prt_check_moments": (_I, [_P, _I, _I, _P, _I64, _I, _P, _P, _P]),
    "prt_check_copy_residual": (_I, [_P, _I, _I, _I, ctypes.c_float, _P, _I64, _I, _P, _P]),
    "prt_adr_table_bytes": (_I64, [_I, _I, _I64]),
    "prt_adr_pore_table": (_I, [_I, _I, _I, _I, _P, _P, _P, _I64, _P, _P]),
    "prt_adr_pore_table_f64": (_I, [_I, _I, _I, _I, _P, _P, _P, _I64, _P, _P]),
    "prt_adr_run_list": (_I, [ctypes.POINTER(AdrParams), _I, _I, _I, _P, _I64, _P, _P, _I, _P]),
    "prt_adr_step_list_f64": (_I, [ctypes.POINTER(AdrParams), _I, _I, _I, _P, _I64, _P, _P, _P]),
    "prt_lbm_batch_desc_size": (_I64, []),
    "prt_adr_batch_desc_size": (_I64, []),
    "prt_lbm_batch_plan": (_I, [_I, _I, _P, _I, _P, _P, _I64, ctypes.POINTER(_I64)]),
    "prt_lbm_pop_stride": (_I64, [_I64, _I]),
    "prt_stream_copy": (_I, [_P, _P, _I64, _I, _P]),
    "prt_adr_batch_plan": (_I, [_P, _I, _P, _P, _I64, ctypes.POINTER(_I64),
                                ctypes.POINTER(ctypes.c_int)]),
    "prt_lbm_batch_step": (_I, [_I, _I, _P, _P, _I64, _I, _I, _P, _P]),
    "prt_adr_batch_step": (_I, [_P, _P, _I64, _I, _I, _P, _P]),
    "prt_seg_sums": (_I, [_P, _P, _I, _I, _P, _P, _P]),
    "prt_batch_shape_tile": (_I, [_I]),
    "prt_lbm_halo_elems": (_I64, [_I, _I, _I, _I, _I]),
    "prt_lbm_halo_pack": (_I, [_I, _I, _I, _I, _I, _P, _I, _I, _P, _P]),
    "prt_lbm_halo_unpack": (_I, [_I, _I, _I, _I, _I, _P, _I, _I, _P, _P]),
    "prt_lbm_batch_plan_shaped": (_I, [_I, _I, _P, _I, _I, _P, _P, _I64, ctypes.POINTER(_I64),
                                       ctypes.POINTER(_I64)]),
    "prt_adr_batch_plan_shaped": (_I, [_P, _I, _I, _P, _P, _I64, ctypes.POINTER(_I64),
                                       ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_int)]),
    "prt_lbm_batch_step_shaped": (_I, [_I, _I, _P, _P, _I64, _I, _I, _I, _P, _P]),
    "prt_adr_batch_step_shaped": (_I, [_P, _P, _I64, _I, _I, _I, _P, _P]),
    "prt_grain_first_cover": (_I, [_I, _I, _I, _P, _P, _I, _P, _P, _I, _P, _P]),
    "prt_grain_histogram": (_I, [_P, _I64, _P, _P]),
    "prt_grain_mask": (_I, [_P, _I64, _I, _I, _I, _P, _P]),
    "prt_keep_inlet_connected": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _I,
                                      ctypes.POINTER(ctypes.c_int), _P]),
    "prt_coat_with_biofilm": (_I, [_P, _I, _I, _I, _I, _I, _I, _P, _P, _P,
                                   ctypes.POINTER(ctypes.c_int), _P]),
    "prt_pairwise_plan": (_I, [_I64, _P, _I64, ctypes.POINTER(_I64),
                               ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "prt_pairwise_sum_f32": (_I, [_P, _I64, _I, _P, _I64, _P, _I64, _I64, _I64, _P, _P, _P]),
}

_lib = None


def lib():
    """The loaded library (raises PrtError if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PrtError("libprtlb_hip.so is not built (%s); run "
                           "`python -m prtlb.build` in prt-deeponet-gui_amd"
                           % LIB_PATH)
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def call(name, *args):
    """Call a status-returning entry point; raise PrtError on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().prt_last_error().decode(errors="replace")
        raise PrtError("%s failed (%d): %s" % (name, rc, msg))
    return rc


def header_symbols(path=HEADER):
    """Function names declared in include/prtlb.h."""
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(prt_[a-z0-9_]+)\s*\(", src)))


def pairwise_plan(n):
    """NumPy's float32 pairwise-sum tree for n values (host int32 array)."""
    L, I, H = _I64(), _I64(), _I64()
    call("prt_pairwise_plan", n, None, 0, ctypes.byref(L), ctypes.byref(I), ctypes.byref(H))
    size = 2 * L.value + 2 * I.value + H.value + 1 + (n + 8191) // 8192
    plan = np.zeros(size, np.int32)
    call("prt_pairwise_plan", n, plan.ctypes.data, size, ctypes.byref(L),
         ctypes.byref(I), ctypes.byref(H))
    return plan, L.value, I.value, H.value
