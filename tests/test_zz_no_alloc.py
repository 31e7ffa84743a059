"""The allocation-free convergence tick, prt_lbm_run_sum_static, through
the C ABI: the same populations and the same sum, bit for bit, as
prt_lbm_run with a partials buffer followed by prt_sum_f64, on two calls in
a row (the second on a smaller grid, so the shared device array holds the
first call's partials), and a clear refusal beyond the array's size."""
import ctypes

import numpy as np
import pytest
import torch

import prtlb
from prtlb import _lib
from prtlb.device import ptr, stream_ptr

pytestmark = pytest.mark.gpu


def _pair(g, storage):
    """Two solvers on the same grid: one for each path."""
    return [prtlb.StokesSolver(g, force=1e-5, tau_lb=0.8, storage=storage) for _ in range(2)]


def _tick_existing(S, steps, flags):
    """prt_lbm_run with a partials buffer, then prt_sum_f64 over the whole
    buffer (the way StokesSolver sums it)."""
    n = int(_lib.lib().prt_lbm_partials_count(S.q, *S.dims))
    part = torch.zeros(n, dtype=torch.float64, device=S.dev)
    out = torch.zeros(1, dtype=torch.float64, device=S.dev)
    a, b = S.f[S.latest], S.f[1 - S.latest]
    _lib.call("prt_lbm_run", S.q, S.st, *S.dims, ptr(S.solid), ptr(a), ptr(b), steps, S.om_p, S.om_m,
              S.force, S.fp16_scale, flags, ptr(part), stream_ptr(S.dev))
    _lib.call("prt_sum_f64", ptr(part), n, ptr(out), stream_ptr(S.dev))
    if steps % 2:
        S.latest = 1 - S.latest
    return out


def _tick_static(S, steps, flags):
    out = torch.full((1,), -1.0, dtype=torch.float64, device=S.dev)
    a, b = S.f[S.latest], S.f[1 - S.latest]
    _lib.call("prt_lbm_run_sum_static", S.q, S.st, *S.dims, ptr(S.solid), ptr(a), ptr(b), steps,
              S.om_p, S.om_m, S.force, S.fp16_scale, flags, ptr(out), stream_ptr(S.dev))
    if steps % 2:
        S.latest = 1 - S.latest
    return out


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_static_tick_equals_run_and_sum(storage):
    big = prtlb.sphere_pack((48, 40, 36), (3.0, 5.0), 0.45, seed=3)
    small = prtlb.disk_pack((30, 22), 3.0, 0.6, seed=4)
    for g, steps, flags in ((big, 51, 0), (small, 40, 0), (big, 7, _lib.lbm_tune(6, 0, 1))):
        print("case", g.shape, storage, steps, flags, flush=True)
        A, B = _pair(g, storage)
        for rep in range(2):
            sa = _tick_existing(A, steps, flags)
            sb = _tick_static(B, steps, flags)
            torch.cuda.synchronize()
            va, vb = float(sa.item()), float(sb.item())
            assert va > 0 and np.float64(va).tobytes() == np.float64(vb).tobytes(), (va, vb)
            for x, y in zip(A.f, B.f):
                assert torch.equal(x, y)


def test_static_tick_refuses_beyond_its_array():
    L = _lib.lib()
    cap = int(L.prt_lbm_static_partials_max())
    flags = _lib.lbm_tune(6, 0, 1)
    assert L.prt_lbm_launch_blocks(19, 400, 400, 400, flags) > cap
    dev = torch.device("cuda:0")
    tiny = torch.zeros(16, dtype=torch.float64, device=dev)
    solid = torch.zeros(16, dtype=torch.uint8, device=dev)
    rc = L.prt_lbm_run_sum_static(19, 0, 400, 400, 400, ptr(solid), ptr(tiny), ptr(tiny[8:]), 5, 1.25,
                                  0.9, 1e-6, 0.0, flags, ptr(tiny), stream_ptr(dev))
    assert rc == -3                                   # PRT_E_RANGE, before any launch
    assert b"static scratch" in L.prt_last_error()
    assert float(tiny.abs().sum()) == 0.0
