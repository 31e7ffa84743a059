"""The batched-campaign and serving workloads at their full sizes: the
64-rock campaign (edges 48..160) and requests drawn from the serving mix
(256^2..1024^2 disk packs, 64^3..256^3 sphere packs).  Shortened runs check
each member against the same solve run one at a time / served alone, and the
largest members' transport fields against the oracle (its steps computed in
C, oracle/adr_c.c) bit for bit; three members run their whole transport
(settled, or to the 100 000-step cap) against the oracle, step count
included."""
import numpy as np
import pytest

import prtlb
from conftest import oracle_chem
from oracle import adr_c as AC
from oracle import prtlb_oracle as O
from prtlb import campaign as C
from prtlb import serve
from prtlb.batch import AdrBatch
from prtlb.serve import Request, Server

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_campaign_64_rocks_batch_equals_single():
    """The 64 rocks of the campaign study (the same seeded generator and
    parameters), flow to 1e-4 and 100 irreversible-sorption steps: one
    fluid-list batch for every flow and one pore-list batch for every
    transport run give each rock's flow iterations and transport field bit
    for bit as its solves one at a time, permeability to 1e-12."""
    specs = C.campaign_specs(n_geom=64, seed=0)
    edges = [s["shape"][0] for s in specs]
    assert len(specs) == 64 and min(edges) < 64 and max(edges) > 150
    geoms = {s["key"]: C.make_geometry(s) for s in specs}
    print("box nodes %d, pore nodes %d" % (sum(g.size for g in geoms.values()),
                                          sum(int((g == 2).sum()) for g in geoms.values())),
          flush=True)
    rb, tb = C.run_batched(specs, geoms, storage="fp32", tol=1e-4, adr_steps=100)
    print("batched", tb, flush=True)
    rs, ts = C.run_sequential(specs, geoms, storage="fp32", tol=1e-4, adr_steps=100)
    print("one at a time", ts, flush=True)
    assert set(rb) == set(rs) == {s["key"] for s in specs}
    for k in rb:
        assert rb[k]["flow_iters"] == rs[k]["flow_iters"], k
        assert rb[k]["permeability"] == pytest.approx(rs[k]["permeability"], rel=1e-12), k
        assert np.array_equal(rb[k]["out_last"], rs[k]["out_last"]), k
        assert rb[k]["adr_steps"] == rs[k]["adr_steps"] == 100


@pytest.mark.timeout(900)
def test_serving_mix_at_full_sizes_alone_equals_mix():
    """Six requests of the serving study's own draw (its seeded Poisson
    mix): the largest 2D and 3D ones, the smallest 3D one and three more,
    reversible sorption.  Served together (600 flow iterations each, 300
    transport steps) every result equals the request served alone."""
    pool = serve.poisson_requests(32, seed=0)
    d2 = sorted((r for r in pool if r.g.ndim == 2), key=lambda r: r.g.size)
    d3 = sorted((r for r in pool if r.g.ndim == 3), key=lambda r: r.g.size)
    pick = [d2[-1], d3[-1], d3[0], d2[0], d3[len(d3) // 2], d3[len(d3) // 3]]
    print("sizes", [r.g.shape for r in pick], flush=True)
    assert max(r.g.shape[0] for r in pick if r.g.ndim == 2) > 900
    assert max(r.g.shape[0] for r in pick if r.g.ndim == 3) > 230

    def reqs(rs):
        return [Request(r.key, r.g, r.pe, r.da, arrival=0.0) for r in rs]

    kw = dict(nit=600, tol=0.0, adr_steps=None, max_steps=300)
    with Server(**kw) as S:
        mix = S.run(reqs(pick), realtime=False)
    by_key = {r.key: r for r in mix}
    # the largest 2D and 3D requests' transport against the oracle, on the
    # velocity the server handed its transport stage
    for r in pick[:2]:
        m = by_key[r.key]
        akw = C._adr_kw(dict(da=r.da), r.kinetics)
        info = {}
        outo, tno = O.solve_adr(r.g, m.result["velocity"], r.pe, akw["da"], 5, n_species=2,
                                max_steps=300, chem=oracle_chem(akw["chem"]), info=info,
                                sorption=akw["sorption"], stepper=AC.stepper)
        print("oracle", r.g.shape, info["steps"], flush=True)
        assert np.array_equal(m.result["out"], outo), r.g.shape
        assert np.array_equal(m.result["t_norm"], tno), r.g.shape
    for r in pick:
        with Server(**kw) as S:
            alone = S.run(reqs([r]), realtime=False)[0]
        m = by_key[r.key]
        assert np.array_equal(m.result["velocity"], alone.result["velocity"]), r.g.shape
        assert np.array_equal(m.result["out"], alone.result["out"]), r.g.shape
        assert m.result["info"]["steps"] == alone.result["info"]["steps"] == 300


@pytest.mark.timeout(900)
def test_campaign_transport_batch_against_oracle():
    """Rocks of the 64-rock campaign draw -- the three largest (edge near
    160), the median and the smallest -- with their own Pe and Da,
    irreversible sorption, 1000 steps, all in one transport batch: every
    member's snapshots equal the oracle's bit for bit."""
    specs = C.campaign_specs(n_geom=64, seed=0)
    by_size = sorted(specs, key=lambda s: s["shape"][0])
    pick = by_size[-3:] + [by_size[len(by_size) // 2], by_size[0]]
    print("edges", [s["shape"][0] for s in pick], flush=True)
    assert max(s["shape"][0] for s in pick) > 150 and min(s["shape"][0] for s in pick) < 56
    geoms = {s["key"]: C.make_geometry(s) for s in pick}
    vels = {k: prtlb.stokes_d3q19(g, nit=400, force=2e-6, tol=0.0, storage="fp32")
            for k, g in geoms.items()}
    ab = AdrBatch()
    for sp in pick:
        kw = C._adr_kw(sp, "sorption_irreversible")
        ab.add(sp["key"], geoms[sp["key"]], vels[sp["key"]], sp["pe"], kw.pop("da"), 4,
               steps=1000, **kw)
    got = {}
    while len(ab):
        for m in ab.advance():
            got[m.key] = ab.finish(m, {})
    for sp in pick:
        k = sp["key"]
        kw = C._adr_kw(sp, "sorption_irreversible")
        outo, tno = O.solve_adr(geoms[k], vels[k], sp["pe"], kw["da"], 4, n_species=2, steps=1000,
                                chem=oracle_chem(kw["chem"]), da_abio=kw["da_abio"],
                                stepper=AC.stepper)
        print("rock", k, geoms[k].shape, flush=True)
        assert np.array_equal(got[k][0], outo), k
        assert np.array_equal(got[k][1], tno), k


@pytest.mark.timeout(900)
def test_served_request_settles_at_the_oracles_step():
    """A serving-mix request that settles before the 100 000-step cap
    (request 29 of the 32-request draw: a 116^3 sphere pack, Pe 59.9, Da
    1.76, reversible sorption, 763 636 pore voxels): served, its transport
    stops at the step the oracle's own stop rule picks (NumPy-order pore
    means every check), with every snapshot and t_norm bit for bit."""
    pool = serve.poisson_requests(32, seed=0)
    r = next(q for q in pool if q.key == 29)
    assert r.g.shape == (116, 116, 116)
    with Server() as S:
        got = S.run([Request(r.key, r.g, r.pe, r.da, arrival=0.0)], realtime=False)[0]
    info = got.result["info"]
    print("served: %d steps, converged %s" % (info["steps"], info["converged"]), flush=True)
    assert info["converged"] and info["steps"] < 100000
    akw = C._adr_kw(dict(da=r.da), r.kinetics)
    oinfo = {}
    outo, tno = O.solve_adr(r.g, got.result["velocity"], r.pe, akw["da"], 5, n_species=2,
                            max_steps=100000, chem=oracle_chem(akw["chem"]), info=oinfo,
                            sorption=akw["sorption"], stepper=AC.stepper)
    print("oracle: %d steps, converged %s" % (oinfo["steps"], oinfo["converged"]), flush=True)
    assert oinfo["steps"] == info["steps"] and oinfo["converged"]
    assert np.array_equal(got.result["out"], outo)
    assert np.array_equal(got.result["t_norm"], tno)


@pytest.mark.timeout(900)
def test_campaign_smallest_rocks_full_length_against_oracle():
    """The two smallest rocks of the 64-rock draw at the campaign's full
    length: irreversible sorption in automatic mode, which never settles,
    so both run to the 100 000-step cap in one transport batch; snapshots,
    t_norm and the step count equal the oracle's bit for bit."""
    specs = sorted(C.campaign_specs(n_geom=64, seed=0), key=lambda s: s["shape"][0])[:2]
    geoms = {s["key"]: C.make_geometry(s) for s in specs}
    vels = {k: prtlb.stokes_d3q19(g, nit=400, force=2e-6, tol=0.0, storage="fp32")
            for k, g in geoms.items()}
    ab = AdrBatch()
    for sp in specs:
        kw = C._adr_kw(sp, "sorption_irreversible")
        ab.add(sp["key"], geoms[sp["key"]], vels[sp["key"]], sp["pe"], kw.pop("da"), 5,
               max_steps=100000, **kw)
    got, infos = {}, {}
    while len(ab):
        for m in ab.advance():
            infos[m.key] = {}
            got[m.key] = ab.finish(m, infos[m.key])
    for sp in specs:
        k = sp["key"]
        kw = C._adr_kw(sp, "sorption_irreversible")
        oinfo = {}
        outo, tno = O.solve_adr(geoms[k], vels[k], sp["pe"], kw["da"], 5, n_species=2,
                                max_steps=100000, chem=oracle_chem(kw["chem"]),
                                da_abio=kw["da_abio"], info=oinfo, stepper=AC.stepper)
        print("rock", k, geoms[k].shape, infos[k]["steps"], oinfo["steps"], flush=True)
        assert infos[k]["steps"] == oinfo["steps"] == 100000
        assert not infos[k]["converged"] and not oinfo["converged"]
        assert np.array_equal(got[k][0], outo), k
        assert np.array_equal(got[k][1], tno), k


@pytest.mark.timeout(900)
def test_served_2d_request_full_length_against_oracle():
    """The smallest 2D request of the serving draw (a 284^2 disk pack,
    reversible sorption) served alone through its whole transport, settled
    or to the 100 000-step cap: the step count, snapshots and t_norm equal
    the oracle's bit for bit."""
    pool = serve.poisson_requests(32, seed=0)
    r = min((q for q in pool if q.g.ndim == 2), key=lambda q: q.g.size)
    with Server() as S:
        got = S.run([Request(r.key, r.g, r.pe, r.da, arrival=0.0)], realtime=False)[0]
    info = got.result["info"]
    akw = C._adr_kw(dict(da=r.da), r.kinetics)
    oinfo = {}
    outo, tno = O.solve_adr(r.g, got.result["velocity"], r.pe, akw["da"], 5, n_species=2,
                            max_steps=100000, chem=oracle_chem(akw["chem"]), info=oinfo,
                            sorption=akw["sorption"], stepper=AC.stepper)
    print("request", r.key, r.g.shape, "served", info["steps"], info["converged"], "oracle",
          oinfo["steps"], oinfo["converged"], flush=True)
    assert info["steps"] == oinfo["steps"] and info["converged"] == oinfo["converged"]
    assert np.array_equal(got.result["out"], outo)
    assert np.array_equal(got.result["t_norm"], tno)
