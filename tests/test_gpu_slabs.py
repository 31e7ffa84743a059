"""A box split along x into slabs (one GPU's share each of a multi-GPU
solve), stepped with halo exchanges on one device: populations and velocity
equal the undivided box's bit for bit."""
import numpy as np
import pytest
import torch

import prtlb
from prtlb import _lib
from prtlb.device import ptr, stream_ptr
from prtlb.slab import SlabSet, split_planes
from prtlb.stokes import StokesSolver

pytestmark = pytest.mark.gpu


def _box_populations(S):
    X, Y, Z = S.rows
    return S.f[S.latest].reshape(X, Y, S.q, -1)


@pytest.mark.parametrize("q,storage,skip,parts", [
    (19, "fp64", False, 4), (19, "fp64", True, 3), (19, "fp32", True, 4), (19, "fp32", False, 5),
    (19, "fp16", True, 4), (9, "fp64", False, 4), (9, "fp32", True, 3)])
def test_slabs_equal_the_box(q, storage, skip, parts):
    g = (prtlb.sphere_pack((40, 18, 22), (2.0, 4.0), 0.5, seed=7) if q == 19
         else prtlb.disk_pack((61, 37), 3.0, 0.6, seed=8))
    steps = 73
    S = StokesSolver(g, force=3e-5, storage=storage, skip_solid=skip)
    S.run(steps)
    T = SlabSet(g, parts, force=3e-5, storage=storage, skip_solid=skip)
    T.run(steps)
    box = _box_populations(S)
    Z = S.rows[2]
    assert torch.equal(T.populations()[..., :Z].contiguous().view(torch.uint8),
                       box[..., :Z].contiguous().view(torch.uint8))
    v_box = S.velocity(f64=True)[1]
    assert torch.equal(T.velocity(), v_box)


def test_split_planes_and_halo_messages():
    assert split_planes(10, 3) == [(0, 4), (4, 7), (7, 10)]
    with pytest.raises(ValueError):
        split_planes(3, 4)
    L = _lib.lib()
    # 5 D3Q19 / 3 D2Q9 populations of one plane, row runs padded to 128 bytes
    assert L.prt_lbm_halo_elems(19, 0, 6, 8, 20) == 5 * 8 * 32
    assert L.prt_lbm_halo_elems(9, 1, 6, 20, 1) == 3 * 1 * 32
    # pack then unpack into another plane copies exactly those runs
    dims = (5, 6, 7)
    elems = L.prt_lbm_pop_elems(19, 0, *dims)
    dev = torch.device("cuda:0")
    f = torch.arange(elems, dtype=torch.float64, device=dev)
    n = L.prt_lbm_halo_elems(19, 0, *dims)
    buf = torch.empty(n, dtype=torch.float64, device=dev)
    s = stream_ptr(dev)
    _lib.call("prt_lbm_halo_pack", 19, 0, *dims, ptr(f), 3, 1, ptr(buf), s)
    g = f.clone()
    _lib.call("prt_lbm_halo_unpack", 19, 0, *dims, ptr(g), 0, 1, ptr(buf), s)
    c = np.zeros((19, 3), np.int32)
    w = np.zeros(19)
    o = np.zeros(19, np.int32)
    _lib.call("prt_lbm_lattice", 19, c.ctypes.data, w.ctypes.data, o.ctypes.data)
    plus = [i for i in range(19) if c[i, 0] == 1]
    assert len(plus) == 5
    F, G = f.reshape(5, 6, 19, -1), g.reshape(5, 6, 19, -1)
    want = F.clone()
    want[0, :, plus] = F[3, :, plus]
    assert torch.equal(G, want)
    assert torch.equal(buf.reshape(5, 6, -1), F[3, :, plus].permute(1, 0, 2))
    for bad in ((5, 1), (-1, 1), (0, 0), (0, 2)):
        with pytest.raises(_lib.PrtError):
            _lib.call("prt_lbm_halo_pack", 19, 0, *dims, ptr(f), bad[0], bad[1], ptr(buf), s)
    with pytest.raises(_lib.PrtError):                 # a slab needs an own plane
        _lib.call("prt_lbm_step", 19, 0, 2, 6, 7, ptr(torch.zeros(84, dtype=torch.uint8, device=dev)),
                  ptr(f), ptr(g), 1.25, 1.6, 1e-6, 1.0, _lib.PRT_LBM_SLAB, None, s)


@pytest.mark.parametrize("storage,skip", [("fp64", False), ("fp32", True)])
def test_slab_solve_matches_the_box_solve(storage, skip):
    """The converging solve on slabs (what the drop-in runs for a grid
    beyond the 32-bit element index) stops at the box's iteration with the
    box's field and permeability."""
    g = prtlb.sphere_pack((40, 20, 20), (3.0, 5.0), 0.6, seed=3)
    S = StokesSolver(g, force=2e-6, storage=storage, skip_solid=skip)
    ib, isl = {}, {}
    S.solve(nit=6000, tol=1e-4, info=ib)
    vb, _, sb = S.velocity()
    T = SlabSet(g, 3, force=2e-6, storage=storage, skip_solid=skip)
    T.solve(nit=6000, tol=1e-4, info=isl)
    vs, ss = T.velocity32()
    print("iterations", ib["flow_iters"], isl["flow_iters"], flush=True)
    assert ib["flow_iters"] == isl["flow_iters"]
    assert torch.equal(vb, vs)
    assert T.permeability(ss) == pytest.approx(S.permeability(sb), rel=1e-12)
