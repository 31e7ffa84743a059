"""Batched solves: a member of a batch gives the same bits as the same solve
alone, and the continuous-batching server returns, for every request, what
that request gives when served alone."""
import numpy as np
import pytest

import prtlb
from prtlb import campaign as C
from prtlb.batch import AdrBatch, LbmBatch
from prtlb.serve import Request, Server

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sparse", [True, False])
def test_lbm_batch_equals_single(sparse):
    geoms = [prtlb.sphere_pack((20, 16, 12), (2.0, 3.5), 0.55, seed=s) for s in range(3)]
    geoms.append(prtlb.sphere_pack((33, 17, 9), (2.0, 3.0), 0.6, seed=9))
    lb = LbmBatch(19, storage="fp32", sparse=sparse)
    for i, g in enumerate(geoms):
        lb.add(i, g, force=1e-5 * (i + 1), nit=300, tol=0.0)
    for _ in range(6):
        lb.tick()
    for i, g in enumerate(geoms):
        m = lb.members[i]
        v_b = lb.finish(m).cpu().numpy()
        v_s = prtlb.stokes_d3q19(g, nit=300, force=1e-5 * (i + 1), tol=0.0, storage="fp32",
                                 skip_solid=True)
        assert np.array_equal(v_b, v_s), i


def test_lbm_batch_join_late_and_converge():
    g1 = prtlb.disk_pack((40, 30), 3.0, 0.6, seed=1)
    g2 = prtlb.disk_pack((64, 20), 3.0, 0.6, seed=2)
    lb = LbmBatch(9, storage="fp64", skip_solid=False)
    lb.add("a", g1, force=2e-6, tol=1e-6, nit=20000)
    for _ in range(3):
        lb.tick()
    lb.add("b", g2, force=2e-6, tol=1e-6, nit=20000)      # joins 150 iterations later
    got = {}
    while len(lb):
        for m in lb.tick():
            info = {}
            got[m.key] = (lb.finish(m, info).cpu().numpy(), info)
    for key, g in (("a", g1), ("b", g2)):
        info = {}
        v = prtlb.stokes_d2q9(g, nit=20000, force=2e-6, tol=1e-6, info=info)
        assert np.array_equal(got[key][0], v)
        assert got[key][1]["flow_iters"] == info["flow_iters"]


@pytest.mark.parametrize("q,storage", [(19, "fp64"), (19, "fp32"), (19, "fp16"), (9, "fp64"),
                                       (9, "fp32")])
def test_fluid_list_step_equals_dense_fluid_only(q, storage):
    """The fluid-list kernels against the dense fluid-only kernels: the
    compact populations equal the dense ones at every fluid node, and the
    velocity fields are equal, bit for bit."""
    import torch
    from prtlb import _lib
    from prtlb.device import ptr, stream_ptr
    from prtlb.stokes import StokesSolver
    g = (prtlb.sphere_pack((26, 18, 14), (2.0, 4.0), 0.5, seed=11) if q == 19
         else prtlb.disk_pack((50, 34), 3.0, 0.6, seed=12))
    dims = (g.shape[0], g.shape[1], 1 if g.ndim == 2 else g.shape[2])
    force, steps = 3e-5, 137
    S = StokesSolver(g, force=force, storage=storage, skip_solid=True)
    S.run(steps)
    dev = S.dev
    st = _lib.PRT_STORE[storage]
    fidx = torch.from_numpy(np.flatnonzero(g.reshape(-1) == prtlb.PORE).astype(np.int32)).to(dev)
    nf = fidx.numel()
    lk = _lib.lib().prt_lbm_link_stride(nf)
    link = torch.empty((q - 1) * lk, dtype=torch.int32, device=dev)
    work = torch.empty(g.size, dtype=torch.int32, device=dev)
    s = stream_ptr(dev)
    _lib.call("prt_lbm_link_table", q, *dims, ptr(fidx), nf, ptr(work), ptr(link), s)
    ld = _lib.lib().prt_lbm_pop_stride(nf, st)
    fa = torch.empty((q, ld), dtype=S.f[0].dtype, device=dev)
    fb = torch.empty_like(fa)
    for b in (fa, fb):
        _lib.call("prt_lbm_sparse_fill_rest", q, st, nf, ptr(b), s)
    _lib.call("prt_lbm_sparse_run", q, st, nf, ptr(link), 32, ptr(fa), ptr(fb), steps, S.om_p, S.om_m,
              force, S.fp16_scale, None, s)
    newest, other = (fb, fa) if steps % 2 else (fa, fb)
    dense_new = S.populations().reshape(q, -1)
    v_dense = S.velocity(f64=True)[1].cpu().numpy()
    v64 = torch.zeros((3, g.size), dtype=torch.float64, device=dev)
    part = torch.zeros(_lib.lib().prt_lbm_sparse_partials_count(nf), dtype=torch.float64, device=dev)
    _lib.call("prt_lbm_sparse_velocity", q, st, g.size, nf, ptr(fidx), ptr(link), 32, ptr(newest),
              ptr(other), force, S.fp16_scale, ptr(v64), None, ptr(part), s)
    assert np.array_equal(v64.cpu().numpy().reshape(v_dense.shape), v_dense)
    d = np.ascontiguousarray(dense_new.cpu().numpy()[:, fidx.cpu().numpy()])
    # tiled list layout: (tiles, q, 256) -> (q, entries)
    c = newest.reshape(-1, q, 256).permute(1, 0, 2).reshape(q, -1)[:, :nf]
    c = np.ascontiguousarray(c.cpu().numpy())
    assert np.array_equal(d.view(np.uint8), c.view(np.uint8))

    # the 16-bit link table: decodes to the 32-bit one and gives the same bits
    l16 = torch.empty((q - 1) * lk, dtype=torch.int16, device=dev)
    over = torch.ones(1, dtype=torch.int32, device=dev)
    _lib.call("prt_lbm_link_pack16", q, nf, ptr(link), ptr(l16), ptr(over), s)
    assert int(over.item()) == 0
    d16 = l16.cpu().numpy().reshape(q - 1, lk)[:, :nf].astype(np.int64)
    k = np.arange(nf)
    dec = np.where(d16 == -32768, -1, (k + d16) % nf)
    assert np.array_equal(dec, link.cpu().numpy().reshape(q - 1, lk)[:, :nf])
    ga, gb = torch.empty_like(fa), torch.empty_like(fa)
    for b in (ga, gb):
        _lib.call("prt_lbm_sparse_fill_rest", q, st, nf, ptr(b), s)
    _lib.call("prt_lbm_sparse_run", q, st, nf, ptr(l16), 16, ptr(ga), ptr(gb), steps, S.om_p,
              S.om_m, force, S.fp16_scale, None, s)
    new16, other16 = (gb, ga) if steps % 2 else (ga, gb)
    assert torch.equal(new16.view(torch.uint8), newest.view(torch.uint8))
    v16 = torch.zeros_like(v64)
    _lib.call("prt_lbm_sparse_velocity", q, st, g.size, nf, ptr(fidx), ptr(l16), 16, ptr(new16),
              ptr(other16), force, S.fp16_scale, ptr(v16), None, ptr(part), s)
    assert torch.equal(v16, v64)


def test_link_pack16_reports_overflow():
    """A list with more fluid nodes per x-plane than a 16-bit step reaches
    keeps the 32-bit table: the pack reports it and LbmBatch falls back."""
    import torch
    from prtlb import _lib
    from prtlb.batch import LbmBatch
    from prtlb.device import ptr, stream_ptr
    g = np.full((3, 200, 200), prtlb.PORE, np.uint8)     # 40 000 fluid nodes per plane
    g[:, 0, :] = prtlb.SOLID
    dev = torch.device("cuda:0")
    fidx = torch.from_numpy(np.flatnonzero(g.reshape(-1) == prtlb.PORE).astype(np.int32)).to(dev)
    nf = fidx.numel()
    lk = _lib.lib().prt_lbm_link_stride(nf)
    link = torch.empty(18 * lk, dtype=torch.int32, device=dev)
    work = torch.empty(g.size, dtype=torch.int32, device=dev)
    s = stream_ptr(dev)
    _lib.call("prt_lbm_link_table", 19, *g.shape, ptr(fidx), nf, ptr(work), ptr(link), s)
    l16 = torch.empty(18 * lk, dtype=torch.int16, device=dev)
    over = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.call("prt_lbm_link_pack16", 19, nf, ptr(link), ptr(l16), ptr(over), s)
    assert int(over.item()) == 1
    b = LbmBatch(19, storage="fp32")
    m = b.add(0, g, force=1e-6)
    assert m.link_bits == 32 and m.link.dtype == torch.int32
    b2 = LbmBatch(19, storage="fp32")
    m2 = b2.add(0, prtlb.sphere_pack((24, 20, 16), (2.0, 4.0), 0.5, seed=5), force=1e-6)
    assert m2.link_bits == 16 and m2.link.dtype == torch.int16


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_fluid_list_batch_join_late_and_converge(storage):
    g1 = prtlb.sphere_pack((24, 18, 14), (2.0, 4.0), 0.5, seed=21)
    g2 = prtlb.sphere_pack((30, 12, 16), (2.0, 3.0), 0.55, seed=22)
    lb = LbmBatch(19, storage=storage)
    assert lb.sparse
    lb.add("a", g1, force=2e-6, tol=1e-6, nit=20000)
    for _ in range(3):
        lb.tick()
    lb.add("b", g2, force=2e-6, tol=1e-6, nit=20000)
    got = {}
    while len(lb):
        for m in lb.tick():
            info = {}
            got[m.key] = (lb.finish(m, info).cpu().numpy(), info)
    for key, g in (("a", g1), ("b", g2)):
        info = {}
        v = prtlb.stokes_d3q19(g, nit=20000, force=2e-6, tol=1e-6, info=info, storage=storage,
                               skip_solid=True)
        assert np.array_equal(got[key][0], v)
        assert got[key][1]["flow_iters"] == info["flow_iters"]
        assert got[key][1]["permeability"] == pytest.approx(info["permeability"], rel=1e-12)


def test_adr_batch_equals_single():
    g1 = prtlb.disk_pack((48, 24), 3.0, 0.65, seed=3)
    g2 = prtlb.sphere_pack((24, 16, 12), (2.0, 3.0), 0.55, seed=4)
    v1 = prtlb.stokes_d2q9(g1, nit=2000, force=1e-5, tol=1e-6)
    v2 = prtlb.stokes_d3q19(g2, nit=1000, force=1e-5, tol=1e-6)
    ab = AdrBatch()
    ab.add("a", g1, v1, 20.0, 1.0, 5, n_species=2)                    # auto stop
    ab.add("b", g2, v2, 5.0, 0.0, 4, n_species=2, steps=700,
           chem=prtlb.chemistry.sorption_irreversible(), da_abio=2.0)
    got = {}
    while len(ab):
        for m in ab.advance():
            info = {}
            got[m.key] = ab.finish(m, info) + (info,)
    ia, ib = {}, {}
    oa, ta = prtlb.solve_adr(g1, v1, 20.0, 1.0, 5, n_species=2, info=ia)
    ob, tb = prtlb.solve_adr(g2, v2, 5.0, 0.0, 4, n_species=2, steps=700, info=ib,
                             chem=prtlb.chemistry.sorption_irreversible(), da_abio=2.0)
    assert np.array_equal(got["a"][0], oa) and np.array_equal(got["a"][1], ta)
    assert got["a"][2]["steps"] == ia["steps"]
    assert np.array_equal(got["b"][0], ob) and np.array_equal(got["b"][1], tb)


def test_adr_batch_deferred_checks_restore_the_stop_step():
    """A member whose convergence checks are decided several checks after
    the device passed them is put back to the step it settled at: the same
    snapshots, t_norm and stop step as solve_adr."""
    g1 = np.full((40, 8), prtlb.PORE, np.uint8)
    g1[:, 0] = g1[:, -1] = prtlb.SOLID
    v1 = np.zeros((3, 40, 8), np.float32)
    v1[0] = 1.0
    g2 = prtlb.disk_pack((40, 20), 3.0, 0.6, seed=5)
    v2 = prtlb.stokes_d2q9(g2, nit=2000, force=1e-5, tol=1e-6)
    ab = AdrBatch()
    ab.eager = False
    ab.check_lag = 3
    ab.add("a", g1, v1, 30.0, 1.0, 3, n_species=4)                          # settles
    ab.add("b", g2, v2, 8.0, 0.5, 4, n_species=1, steps=40000)
    got = {}
    while len(ab):
        for m in ab.advance():
            info = {}
            got[m.key] = ab.finish(m, info) + (info,)
    ia, ib = {}, {}
    oa, ta = prtlb.solve_adr(g1, v1, 30.0, 1.0, 3, n_species=4, info=ia)
    ob, tb = prtlb.solve_adr(g2, v2, 8.0, 0.5, 4, n_species=1, steps=40000, info=ib)
    print("stop step", ia["steps"], flush=True)
    assert ia["converged"] and ia["steps"] < 40000
    assert np.array_equal(got["a"][0], oa) and np.array_equal(got["a"][1], ta)
    assert got["a"][2]["steps"] == ia["steps"] and got["a"][2]["converged"]
    assert np.array_equal(got["b"][0], ob) and np.array_equal(got["b"][1], tb)


def test_server_result_same_alone_and_in_mix():
    specs = [dict(shape=(96, 96), radius=6.0, porosity=0.6, geom_seed=1),
             dict(shape=(40, 40, 40), radius=(3.0, 6.0), porosity=0.4, geom_seed=2),
             dict(shape=(32, 32, 32), radius=(3.0, 6.0), porosity=0.45, geom_seed=3),
             dict(shape=(128, 128), radius=6.0, porosity=0.6, geom_seed=4)]
    pes, das = [3.0, 10.0, 0.5, 40.0], [0.1, 1.0, 5.0, 0.3]
    geoms = [C.make_geometry(s) for s in specs]

    def reqs():
        return [Request(i, geoms[i], pes[i], das[i], arrival=0.0) for i in range(4)]

    mix = Server(nit=30000, adr_steps=None, max_steps=3000).run(reqs(), realtime=False)
    for r in mix:
        alone = Server(nit=30000, adr_steps=None, max_steps=3000).run(
            [Request(r.key, geoms[r.key], pes[r.key], das[r.key])], realtime=False)[0]
        assert np.array_equal(r.result["velocity"], alone.result["velocity"])
        assert np.array_equal(r.result["out"], alone.result["out"])
        assert r.result["info"]["steps"] == alone.result["info"]["steps"]


def test_campaign_batched_matches_sequential_fp32():
    specs = C.campaign_specs(n_geom=3, edge=(32, 40), radius=(2.0, 4.0), seed=5)
    geoms = {s["key"]: C.make_geometry(s) for s in specs}
    rb, _ = C.run_batched(specs, geoms, storage="fp32", adr_steps=200)
    rs, _ = C.run_sequential(specs, geoms, storage="fp32", adr_steps=200)
    for k in rb:
        assert rb[k]["flow_iters"] == rs[k]["flow_iters"]
        assert rb[k]["permeability"] == pytest.approx(rs[k]["permeability"], rel=1e-12)
        assert np.array_equal(rb[k]["out_last"], rs[k]["out_last"])


def test_dataset_layout_and_content():
    """The campaign dataset: the reference's group/dataset names, dtypes and
    shapes; each stored run equals the same transport solve alone."""
    from prtlb import dataset as D
    ds = D.build_dataset(n_geom=2, shape=(24, 20, 20), n_sets=2, n_times=4, radius=(2.0, 4.0),
                         adr_steps=300, seed=3)
    G, S = 2, 4
    assert ds["geom/material"].shape == (G, 24, 20, 20) and ds["geom/material"].dtype == np.uint8
    assert ds["geom/gdf"].dtype == np.float32 and ds["geom/edt"].dtype == np.float32
    assert ds["samples/conc"].shape == (S, 4, 2, 24, 20, 20)
    assert ds["samples/conc"].dtype == np.float16 and ds["samples/velocity"].dtype == np.float16
    assert ds["samples/conc.attrs.conc_scale"].shape == (2,)
    assert list(ds["samples/geom_index"]) == [0, 0, 1, 1]
    from oracle import prtlb_oracle as O
    for i in range(G):
        ref, _ = O.geodesic(ds["geom/material"][i])
        assert np.array_equal(ds["geom/gdf"][i], ref)
    k = 3
    i, pe, da = ds["_draws"][k]
    g = ds["geom/material"][i]
    v = ds["_flows"][i]
    assert np.array_equal(ds["samples/velocity"][k], v.astype(np.float16))
    kw = C._adr_kw(dict(da=da), "sorption_irreversible")
    out, tn = prtlb.solve_adr(g, v, pe, kw.pop("da"), 4, steps=300, **kw)
    assert np.array_equal(ds["samples/conc"][k], out.astype(np.float16))
    assert np.array_equal(ds["samples/t_norm"][k], tn)
    import os, tempfile
    path = os.path.join(tempfile.mkdtemp(), "ds.npz")
    D.write_dataset(path, ds)
    z = np.load(path, allow_pickle=False)
    assert "samples/conc" in z.files and z["samples/settled"].dtype == bool


def test_dataset_2d_layout_and_content():
    """A 2D campaign (disk packs, D2Q9) in the 2D builder's file layout
    (3D/tools/build_dataset_2d.py:896-937): grids with a trailing axis of one
    voxel, dimension 2, z velocity zero; each stored run equals the same
    transport solve alone."""
    from prtlb import dataset as D
    ds = D.build_dataset(n_geom=2, shape=(40, 32), n_sets=2, n_times=3, radius=(2.0, 4.0),
                         porosity=(0.4, 0.6), adr_steps=300, seed=5)
    G, S = 2, 4
    assert int(ds[".attrs.dimension"]) == 2
    assert list(ds[".attrs.shape"]) == [40, 32, 1]
    assert ds["geom/material"].shape == (G, 40, 32, 1)
    assert ds["geom/gdf"].shape == (G, 40, 32, 1) and ds["geom/edt"].shape == (G, 40, 32, 1)
    assert ds["samples/conc"].shape == (S, 3, 2, 40, 32, 1)
    assert ds["samples/velocity"].shape == (S, 3, 40, 32, 1)
    assert not ds["samples/velocity"][:, 2].any()
    from oracle import prtlb_oracle as O
    for i in range(G):
        ref, _ = O.geodesic(ds["geom/material"][i, ..., 0])
        assert np.array_equal(ds["geom/gdf"][i, ..., 0], ref)
    k = 2
    i, pe, da = ds["_draws"][k]
    g = ds["geom/material"][i, ..., 0]
    v = ds["_flows"][i]
    kw = C._adr_kw(dict(da=da), "sorption_irreversible")
    out, tn = prtlb.solve_adr(g, v, pe, kw.pop("da"), 3, steps=300, **kw)
    assert np.array_equal(ds["samples/conc"][k][..., 0], out.astype(np.float16))
    assert np.array_equal(ds["samples/t_norm"][k], tn)


def test_dataset_stops_on_physics_complaint(monkeypatch, capsys):
    """A planted complaint (the host's check returns one for run 1) stops
    the campaign before a dataset exists, with the reference's PHYSICS line
    (3D/tools/build_dataset_3d.py:843-860)."""
    from prtlb import dataset as D
    real = D.check_physics_device
    calls = []

    def planted(out, mat, **kw):
        calls.append(1)
        bad = real(out, mat, **kw)
        return bad + (["C: planted complaint"] if len(calls) == 2 else [])

    monkeypatch.setattr(D, "check_physics_device", planted)
    with pytest.raises(SystemExit, match="1 physics complaint"):
        D.build_dataset(n_geom=1, shape=(20, 16, 16), n_sets=2, n_times=3, radius=(2.0, 4.0),
                        adr_steps=100, seed=4)
    assert "PHYSICS run 1" in capsys.readouterr().out
    calls.clear()
    ds = D.build_dataset(n_geom=1, shape=(20, 16, 16), n_sets=2, n_times=3, radius=(2.0, 4.0),
                         adr_steps=100, seed=4, keep_complaints=True)
    assert list(ds["_complaints"]) == [1]


def test_adr_batch_with_an_all_solid_member():
    """A member without pore voxels next to a normal one, and alone: each
    result equals its solve_adr."""
    g_empty = np.full((12, 10), prtlb.SOLID, np.uint8)
    v_empty = np.zeros((3, 12, 10), np.float32)
    g = prtlb.disk_pack((40, 20), 3.0, 0.6, seed=6)
    v = prtlb.stokes_d2q9(g, nit=1500, force=1e-5, tol=1e-6)
    for members in (("empty", "pack"), ("empty",)):
        ab = AdrBatch()
        if "empty" in members:
            ab.add("empty", g_empty, v_empty, 5.0, 1.0, 3, n_species=1, steps=300)
        if "pack" in members:
            ab.add("pack", g, v, 5.0, 1.0, 3, n_species=1, steps=300)
        got = {}
        while len(ab):
            for m in ab.advance():
                got[m.key] = ab.finish(m, {})
        oe, te = prtlb.solve_adr(g_empty, v_empty, 5.0, 1.0, 3, n_species=1, steps=300)
        assert np.array_equal(got["empty"][0], oe) and np.array_equal(got["empty"][1], te)
        assert float(np.abs(oe).max()) == 0.0
        if "pack" in members:
            op, tp = prtlb.solve_adr(g, v, 5.0, 1.0, 3, n_species=1, steps=300)
            assert np.array_equal(got["pack"][0], op) and np.array_equal(got["pack"][1], tp)


def test_server_admission_under_a_binding_budget():
    """An HBM budget smaller than the mix's footprint: requests wait, are
    admitted as others leave (never holding more than the budget), and
    each result equals the same request served alone."""
    specs = [dict(shape=(40, 40, 40), radius=(3.0, 6.0), porosity=0.4, geom_seed=11),
             dict(shape=(32, 32, 32), radius=(3.0, 6.0), porosity=0.45, geom_seed=12),
             dict(shape=(96, 96), radius=6.0, porosity=0.6, geom_seed=13),
             dict(shape=(36, 36, 36), radius=(3.0, 6.0), porosity=0.35, geom_seed=14)]
    geoms = [C.make_geometry(s) for s in specs]
    pes, das = [4.0, 12.0, 1.0, 30.0], [0.3, 2.0, 0.05, 1.0]

    def reqs():
        return [Request(i, geoms[i], pes[i], das[i], arrival=0.0) for i in range(len(geoms))]

    probe = Server()
    foot = [probe._footprint(r) for r in reqs()]
    budget = 1.05 * max(foot)
    assert sum(foot) > 2 * budget                     # the budget binds
    with Server(hbm_budget=budget, nit=30000, adr_steps=None, max_steps=2000) as S:
        mix = S.run(reqs(), realtime=False)
        assert S.peak_reserved <= budget and S.in_use() == 0
    by_key = {r.key: r for r in mix}
    waited = [r for r in mix if any(o.round_done is not None and o.round_done <= r.round_admit
                                    for o in mix if o is not r)]
    assert waited, [(r.key, r.round_admit, r.round_done) for r in mix]
    for k in range(len(geoms)):
        alone = Server(nit=30000, adr_steps=None, max_steps=2000).run(
            [Request(k, geoms[k], pes[k], das[k])], realtime=False)[0]
        r = by_key[k]
        assert np.array_equal(r.result["velocity"], alone.result["velocity"])
        assert np.array_equal(r.result["out"], alone.result["out"])
        assert r.result["info"]["steps"] == alone.result["info"]["steps"]


def test_campaign_mix_edges_48_to_160_batch_equals_single():
    """A config-4-shaped mix: eight rocks spanning the campaign's edges
    48..160 (radii 3..8, porosity 0.25..0.5, Pe and Da log-uniform),
    irreversible sorption, 300 transport steps.  Batched (one fluid-list
    launch for all flows, one pore-list launch for all transport) equals the
    same solves one at a time: flow iterations and transport fields bit for
    bit, permeability to 1e-12 (its FP64 sum runs over other blocks)."""
    specs = C.campaign_specs(n_geom=8, seed=11)
    for sp, e in zip(specs, (48, 64, 80, 96, 112, 128, 144, 160)):
        sp["shape"] = (e, e, e)
    geoms = {s["key"]: C.make_geometry(s) for s in specs}
    print("edges", [s["shape"][0] for s in specs], flush=True)
    rb, _ = C.run_batched(specs, geoms, storage="fp32", tol=1e-4, adr_steps=300)
    rs, _ = C.run_sequential(specs, geoms, storage="fp32", tol=1e-4, adr_steps=300)
    for k in rb:
        assert rb[k]["flow_iters"] == rs[k]["flow_iters"], k
        assert rb[k]["permeability"] == pytest.approx(rs[k]["permeability"], rel=1e-12), k
        assert np.array_equal(rb[k]["out_last"], rs[k]["out_last"]), k
        assert rb[k]["adr_steps"] == rs[k]["adr_steps"] == 300
