"""Launch shapes of the batched steps (PRT_BATCH_SHAPE): every shape the
tuner may pick gives the same bits as the default shape -- fields and the
per-member convergence sums -- and tuning leaves every member's state as it
found it."""

import numpy as np
import pytest
import torch

import prtlb
from prtlb import _lib, autotune
from prtlb.batch import AdrBatch, LbmBatch
from prtlb.device import ptr, stream_ptr

pytestmark = pytest.mark.gpu


def _lbm_geoms(q):
    if q == 19:
        # 2100-2300 fluid nodes each: several tiles of either size, ragged ends
        return ([prtlb.sphere_pack((20, 16, 12), (2.0, 3.5), 0.55, seed=s) for s in range(3)]
                + [prtlb.sphere_pack((33, 17, 9), (2.0, 3.0), 0.6, seed=9)])
    return [prtlb.disk_pack((40, 30), 3.0, 0.6, seed=1), prtlb.disk_pack((64, 20), 3.0, 0.6, seed=2),
            prtlb.disk_pack((90, 70), 4.0, 0.55, seed=3)]


def _run_lbm(q, storage, sparse, shape, geoms, ticks=3):
    lb = LbmBatch(q, storage=storage, sparse=sparse, skip_solid=True)
    lb.autotune = False
    lb.shape = shape
    for i, g in enumerate(geoms):
        lb.add(i, g, force=1e-5 * (i + 1), nit=1000, tol=0.0)
    sums = []
    for _ in range(ticks):
        lb.tick()
        sums.append(lb.sums.cpu().numpy().copy())
    fields = [lb.newest(lb.members[i]).cpu().numpy().copy() for i in range(len(geoms))]
    return fields, sums


@pytest.mark.parametrize("q,storage,sparse", [(19, "fp32", True), (19, "fp64", True),
                                              (19, "fp16", True), (9, "fp32", True),
                                              (19, "fp32", False)])
def test_lbm_batch_shapes_same_bits(q, storage, sparse):
    geoms = _lbm_geoms(q)
    ref_f, ref_s = _run_lbm(q, storage, sparse, _lib.BATCH_SHAPE_LBM_DEFAULT, geoms)
    lb = LbmBatch(q, storage=storage, sparse=sparse)
    shapes = autotune.batch_candidates(lb)
    assert len(shapes) == (10 if sparse else 2)
    for shape in shapes:
        print("shape", _lib.batch_shape_name(shape), flush=True)
        f, s = _run_lbm(q, storage, sparse, shape, geoms)
        for a, b in zip(ref_f, f):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), _lib.batch_shape_name(shape)
        for a, b in zip(ref_s, s):
            assert np.array_equal(a, b), _lib.batch_shape_name(shape)


def _adr_members(kind):
    g1 = prtlb.disk_pack((48, 24), 3.0, 0.65, seed=3)
    g2 = prtlb.sphere_pack((24, 16, 12), (2.0, 3.0), 0.55, seed=4)
    g3 = prtlb.sphere_pack((30, 20, 18), (2.0, 4.0), 0.5, seed=6)
    v1 = prtlb.stokes_d2q9(g1, nit=1500, force=1e-5, tol=1e-6)
    v2 = prtlb.stokes_d3q19(g2, nit=800, force=1e-5, tol=1e-6)
    v3 = prtlb.stokes_d3q19(g3, nit=800, force=1e-5, tol=1e-6)
    sorb = prtlb.chemistry.sorption_irreversible()
    if kind == "two":
        return [("a", g1, v1, 20.0, 1.0, 4, dict(n_species=2, steps=300, chem=sorb, da_abio=2.0)),
                ("b", g2, v2, 5.0, 0.3, 3, dict(n_species=1, steps=260)),
                ("c", g3, v3, 12.0, 0.0, 3, dict(n_species=2, steps=280, chem=sorb, da_abio=0.7))]
    return [("a", g1, v1, 20.0, 1.0, 4, dict(n_species=4, steps=300)),
            ("b", g3, v3, 8.0, 0.5, 3, dict(n_species=3, steps=250))]


def _run_adr(members, shape=None, tune=False):
    ab = AdrBatch()
    ab.autotune = tune
    if shape is not None:
        ab.shape = shape
    for key, g, v, pe, da, nt, kw in members:
        ab.add(key, g, v, pe, da, nt, **kw)
    got = {}
    while len(ab):
        for m in ab.advance():
            info = {}
            got[m.key] = ab.finish(m, info) + (info,)
    return got, ab


@pytest.mark.parametrize("kind", ["two", "four"])
def test_adr_batch_shapes_same_bits(kind):
    members = _adr_members(kind)
    ref, _ = _run_adr(members, _lib.BATCH_SHAPE_ADR_DEFAULT)
    for shape in _lib.BATCH_SHAPES:
        print("shape", _lib.batch_shape_name(shape), flush=True)
        got, ab = _run_adr(members, shape)
        assert ab.planned_shape == shape
        for key in ref:
            assert np.array_equal(ref[key][0], got[key][0]), (key, _lib.batch_shape_name(shape))
            assert np.array_equal(ref[key][1], got[key][1])
            assert ref[key][2]["steps"] == got[key][2]["steps"]


def test_adr_batch_partials_per_tile():
    """The transport step's FP64 partial sums (C ABI only): one slot per
    256 entries of each member and chemical, equal to the sum of the new
    field over that tile's pore voxels."""
    members = _adr_members("two")
    ab = AdrBatch()
    ab.autotune = False
    for key, g, v, pe, da, nt, kw in members:
        ab.add(key, g, v, pe, da, nt, **kw)
    ab._plan()
    mem = list(ab.members.values())
    w = 2
    ntiles = sum((m.T.n_pore + 255) // 256 for m in mem)
    part = torch.full((ntiles * w,), np.nan, dtype=torch.float64, device=ab.dev)
    s = stream_ptr(ab.dev)
    _lib.call("prt_adr_batch_step_shaped", ptr(ab.desc), ptr(ab.blk_seg), ab.n_blocks,
              ab.max_species, ab.parity, _lib.BATCH_SHAPE_ADR_DEFAULT, ptr(part), s)
    with pytest.raises(_lib.PrtError):
        _lib.call("prt_adr_batch_step_shaped", ptr(ab.desc), ptr(ab.blk_seg), ab.n_blocks,
                  ab.max_species, ab.parity, _lib.batch_shape(7, 1, 0), ptr(part), s)
    p = part.cpu().numpy().reshape(ntiles, w)
    t0 = 0
    for m in mem:
        T = m.T
        newest = T.c[1 - ((ab.parity ^ m.phase) & 1)]          # the buffer just written
        field = newest.reshape(T.ns, -1).double()
        for j in range(w):
            if j < T.ns:
                listed = bool(T.params.immobile_list) and j in T._immobile
                row = field[j, :T.n_pore] if listed else field[j][T.pore_idx.long()]
                want = torch.nn.functional.pad(row, (0, -T.n_pore % 256)).reshape(-1, 256).sum(1)
                want = want.cpu().numpy()
            else:
                want = np.zeros((T.n_pore + 255) // 256)
            got = p[t0:t0 + len(want), j]
            assert np.allclose(got, want, rtol=1e-12, atol=1e-300), (m.key, j)
        t0 += (T.n_pore + 255) // 256
    assert t0 == ntiles


def test_tune_batch_restores_state():
    geoms = _lbm_geoms(19)
    ref_f, ref_s = _run_lbm(19, "fp32", True, _lib.BATCH_SHAPE_LBM_DEFAULT, geoms, ticks=4)
    lb = LbmBatch(19, storage="fp32")
    lb.autotune = False
    for i, g in enumerate(geoms):
        lb.add(i, g, force=1e-5 * (i + 1), nit=1000, tol=0.0)
    lb.tick()
    lb.tick()
    before = [t.clone() for m in lb.members.values() for t in m.f]
    changed = autotune.tune_batch(lb, force=True)
    entry = autotune._CACHE[autotune._batch_key(lb)]
    print("batch tune", entry["name"], entry["table"], flush=True)
    assert len(entry["table"]) == 10 and lb.shape in _lib.BATCH_SHAPES
    after = [t for m in lb.members.values() for t in m.f]
    for a, b in zip(before, after):
        assert torch.equal(a.view(torch.uint8), b.view(torch.uint8))
    assert changed == (lb.shape != lb.planned_shape)
    if changed:
        lb._plan()
    sums = []
    for _ in range(2):
        lb.tick()
        sums.append(lb.sums.cpu().numpy())
    for i in range(len(geoms)):
        assert np.array_equal(lb.newest(lb.members[i]).cpu().numpy().view(np.uint8),
                              ref_f[i].view(np.uint8))
    for a, b in zip(ref_s[2:], sums):
        assert np.array_equal(a, b)


def test_adr_batch_tuned_equals_untuned():
    """A transport batch that tunes its shape on its first plan (the
    default) returns what an untuned one does."""
    members = _adr_members("two")
    ref, _ = _run_adr(members, tune=False)
    for k in [k for k in autotune._CACHE if k.startswith("batch_adr")]:
        del autotune._CACHE[k]
    got, ab = _run_adr(members, tune=True)
    keys = [k for k in autotune._CACHE if k.startswith("batch_adr")]
    assert keys, "no transport batch was tuned"
    print("tuned", {k: autotune._CACHE[k]["name"] for k in keys}, flush=True)
    for key in ref:
        assert np.array_equal(ref[key][0], got[key][0]) and np.array_equal(ref[key][1], got[key][1])
