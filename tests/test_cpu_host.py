"""CPU-only checks of the product's host side: the C ABI loads and exports
what include/prtlb.h declares, the NumPy-order sum plan, the host set-up of
the transport step against the oracle's, the snapshot bookkeeping, the
geometry helpers and check_physics against the reference's golden vectors.
No kernel is launched here."""
import ctypes

import numpy as np
import pytest

from conftest import golden
from oracle import prtlb_oracle as O
from prtlb import _lib
from prtlb.adr import adr_host_setup, ladder, pick_snapshots
import prtlb


def test_library_exports_header():
    L = _lib.lib()
    syms = _lib.header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert L.prt_adr_params_size() == ctypes.sizeof(_lib.AdrParams)
    assert b"gfx950" in L.prt_version()


def _eval_plan(a, plan, L, I, H):
    """Evaluate the plan on the CPU with float32 arithmetic: trees per
    8192-element chunk, then the chunk sums left to right from zero."""
    a = np.asarray(a, np.float32)
    ls, ll = plan[:L], plan[L:2 * L]
    lf, rt = plan[2 * L:2 * L + I], plan[2 * L + I:2 * L + 2 * I]
    v = np.zeros(L + I, np.float32)
    for k in range(L):
        seg = a[ls[k]:ls[k] + ll[k]]
        if len(seg) < 8:
            r = np.float32(0)
            for x in seg:
                r = np.float32(r + x)
        else:
            acc = seg[:8].copy()
            i = 8
            while i < len(seg) - len(seg) % 8:
                acc = (acc + seg[i:i + 8]).astype(np.float32)
                i += 8
            r = np.float32(np.float32(np.float32(acc[0] + acc[1]) + np.float32(acc[2] + acc[3]))
                           + np.float32(np.float32(acc[4] + acc[5]) + np.float32(acc[6] + acc[7])))
            for x in seg[i:]:
                r = np.float32(r + x)
        v[k] = r
    for k in range(I):
        v[L + k] = np.float32(v[lf[k]] + v[rt[k]])
    roots = plan[2 * L + 2 * I + H + 1:]
    assert len(roots) == (len(a) + 8191) // 8192
    acc = np.float32(0)
    for r in roots:
        acc = np.float32(acc + v[r])
    return acc


@pytest.mark.parametrize("n", [1, 7, 8, 100, 128, 129, 1000, 4097, 33333, 100000,
                               131072, 250001, 1000003])
def test_pairwise_plan_matches_numpy(n):
    rng = np.random.default_rng(n)
    a = (rng.random(n) * rng.choice([1e-3, 1.0, 1e3], n)).astype(np.float32)
    plan, L, I, H = _lib.pairwise_plan(n)
    assert _eval_plan(a, plan, L, I, H) == np.add.reduce(a)
    # levels are ordered children-first
    lo = plan[2 * L + 2 * I:2 * L + 2 * I + H + 1]
    assert lo[0] == 0 and lo[-1] == I


ADR = ["adr2d_channel4", "adr2d_disks_abio", "adr2d_disks_default4",
       "adr2d_sorption_irrev", "adr2d_coupled", "adr2d_biofilm", "adr3d_monod",
       "adr3d_default3_abio"]


def _chem(kw, mod):
    if kw is None:
        return None
    fix = {k: ([float("nan") if x == "nan" else x for x in v] if isinstance(v, list)
               else (float("nan") if v == "nan" else v)) for k, v in kw.items()}
    return mod.Chemistry(**fix)


@pytest.mark.parametrize("name", ADR)
def test_host_setup_matches_oracle(name):
    d = golden(name)
    kw = dict(d["meta"]["kw"])
    args = (d["mat"], d["vel"], kw["pe"], kw["da"], kw["n_species"])
    T = adr_host_setup(*args, chem=_chem(d["meta"]["chem"], prtlb),
                       da_abio=kw.get("da_abio", 0.0))
    S = O.adr_setup(*args, chem=_chem(d["meta"]["chem"], O), da_abio=kw.get("da_abio", 0.0))
    assert T.dt == S.dt and T.dt.dtype == np.float32
    assert (T.D, T.k, T.ka, T.kd) == (S.D, S.k, S.ka, S.kd)
    assert np.array_equal(T.u_host, np.stack([S.u[a].ravel() for a in range(S.nd)]))
    assert np.array_equal(T.c_host, O.adr_initial(S))
    # face diffusivities: the kernel's 2x2 table reproduces every face value
    p = T.params
    bf = S.biofilm.ravel().astype(int)
    for ax, fa in enumerate(S.faces):
        for r, j in enumerate(S.moving if len(S.moving) < S.ns else range(S.ns)):
            tab = np.array(p.dface[j], np.float32)
            lo = np.where(fa["o_lo"], tab[2 * bf + bf[fa["m1"]]], 0).astype(np.float32)
            assert np.array_equal(lo, fa["d_lo"][r])
    assert T.surface_voxels == pytest.approx(float(S.surf.sum()), abs=0)


@pytest.mark.parametrize("limit,n_t", [(1, 3), (2, 2), (200, 11), (300, 4), (5000, 6), (20000, 9)])
def test_ladder_and_pick_match_reference_loop(limit, n_t):
    # the reference's loop (3D/tools/prtlb_3d.py:1488-1546) with a fixed count
    keep, nxt = [0], 1
    for n in range(1, limit + 1):
        if n >= nxt:
            keep.append(n)
            nxt = max(n + 1, int(np.ceil(n * 1.12)))
    if keep[-1] != limit:
        keep.append(limit)
    assert ladder(limit, 1.12) == keep
    ks = np.asarray(keep)
    want = O.wanted_steps(limit, n_t)
    used, picked = set(), []
    for w in want:
        i = int(np.abs(ks - w).argmin())
        if i in used:
            i = O.nearest_unused(i, used, len(ks))
        used.add(i)
        picked.append(i)
    assert pick_snapshots(keep, limit, n_t) == picked


def test_geometry_helpers_match_reference():
    d = golden("geometry")
    mt = d["meta"]
    gk, n = prtlb.keep_inlet_connected(d["g"])
    assert np.array_equal(gk, d["kept"]) and n == mt["removed"]
    g1, n1 = prtlb.coat_with_biofilm(gk, 3, 1)
    g2, n2 = prtlb.coat_with_biofilm(gk, 4, 2)
    assert np.array_equal(g1, d["coat1"]) and n1 == mt["n1"]
    assert np.array_equal(g2, d["coat2"]) and n2 == mt["n2"]
    assert prtlb.assert_finite_distance(np.zeros((3, 3)), np.full((3, 3), 2)) is None
    bad = np.zeros((3, 3)); bad[1, 1] = np.inf
    assert "not finite" in prtlb.assert_finite_distance(bad, np.full((3, 3), 2))


def test_generators_hit_porosity_and_are_seeded():
    g = prtlb.disk_pack((128, 128), 6.0, 0.6, seed=0, connect=False)
    phi = float((g == prtlb.PORE).mean())
    assert 0.55 < phi <= 0.60
    assert np.array_equal(g, prtlb.disk_pack((128, 128), 6.0, 0.6, seed=0, connect=False))
    assert not np.array_equal(g, prtlb.disk_pack((128, 128), 6.0, 0.6, seed=1, connect=False))
    s = prtlb.sphere_pack((48, 40, 32), (4.0, 8.0), 0.35, seed=3, connect=False)
    phi3 = float((s == prtlb.PORE).mean())
    assert 0.28 < phi3 <= 0.35
    # periodic minimum image: the pack tiles, so a grain crossing x = 0
    # shows on both faces somewhere
    assert (s[0] == 0).any() and (s[-1] == 0).any()


KEYS = {"nan": ["not finite"], "good": [], "over_cap": ["above"],
        "negative": ["negative"], "zero": ["exactly zero"], "wrap": ["wrapping"]}


@pytest.mark.parametrize("case", sorted(KEYS))
def test_check_physics_matches_reference(case):
    d = golden("check_physics")
    ref = d["meta"]["complaints"][case]
    got = prtlb.check_physics(d["out_" + case], d["mat_" + case])
    assert len(got) == len(ref)
    for kw in KEYS[case]:
        assert any(kw in s for s in got)


@pytest.mark.parametrize("q,nd", [(9, 2), (19, 3)])
def test_lattice_tables(q, nd):
    """The kernels' lattice: the identities 3D/tools/test_flow_solvers.py:51-73
    checks (sum w = 1, sum w c = 0, sum w c c = I/3, opposites) to 1e-12, and
    the reference's velocity order and weights exactly."""
    c = np.zeros(3 * q, np.int32)
    w = np.zeros(q, np.float64)
    opp = np.zeros(q, np.int32)
    _lib.call("prt_lbm_lattice", q, c.ctypes.data, w.ctypes.data, opp.ctypes.data)
    c = c.reshape(q, 3)
    assert not c[:, nd:].any()
    c = c[:, :nd].astype(np.float64)
    assert abs(w.sum() - 1.0) < 1e-12
    assert np.abs(w @ c).max() < 1e-12
    assert np.abs(np.einsum("i,ia,ib->ab", w, c, c) - np.eye(nd) / 3.0).max() < 1e-12
    assert (c[opp] == -c).all() and (opp[opp] == np.arange(q)).all()
    e, w_ref, opp_ref = O.lattice(nd)
    assert np.array_equal(c, e) and np.array_equal(w, w_ref) and np.array_equal(opp, opp_ref)
    with pytest.raises(_lib.PrtError):
        _lib.call("prt_lbm_lattice", 15, c.ctypes.data, w.ctypes.data, opp.ctypes.data)


def test_grid_peclet_warning():
    # 3D/tools/prtlb_3d.py:2320-2334: a Peclet number the grid cannot resolve
    # is warned about at set-up, an ordinary one is not
    import contextlib
    import io
    g = np.full((80, 8), prtlb.PORE, np.uint8)
    v = np.zeros((3, 80, 8), np.float32)
    v[0] = 1.0
    ch = prtlb.Chemistry(names=["Ac"], inlet=[1.0], c_init=[0.0], d_rel=[1.0], immobile=[False],
                         biotic_on=False)
    for pe, warned in ((5000.0, True), (50.0, False)):
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            adr_host_setup(g, v, pe, 0.0, 1, chem=ch)
        assert ("grid Peclet" in buf.getvalue()) == warned, pe


def test_host_errors():
    with pytest.raises(ValueError):
        prtlb.trt_minus(1.0 / 0.5)
    g = np.full((10, 4), 2, np.uint8)
    v = np.zeros((3, 10, 4), np.float32)
    with pytest.raises(ValueError):
        adr_host_setup(g, v, 10.0, 0.0, 2, da_abio=1.0)      # abiotic index 2, 2 chemicals
    with pytest.raises(ValueError):
        adr_host_setup(np.zeros((4,), np.uint8), np.zeros((3, 4), np.float32), 1.0, 0.0, 1)
    with pytest.raises(ValueError):
        adr_host_setup(g, v, 10.0, 0.0, 3, chem=prtlb.default_chemistry(2))


def test_device_required():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(prtlb.PrtError):
        prtlb.stokes_d2q9(np.full((8, 6), 2, np.uint8), nit=2)


def test_fastdiv_identity():
    """The multiply-high division the transport kernels use to turn a flat
    index into (x, y, z): with l = ceil(log2 d) and
    m = floor(2^32 (2^l - d) / d) + 1, (mulhi(n, m) + n) >> l == n // d for
    every 0 <= n < 2^31 (Granlund & Montgomery 1994).  Checked on edge
    numerators and random ones for the divisors grids produce."""
    rng = np.random.default_rng(0)
    divisors = list(range(1, 600)) + [1 << k for k in range(1, 21)] + \
        [(1 << k) - 1 for k in range(2, 21)] + [160 * 160, 256 * 256, 384 * 384, 1024 * 1024, 65535]
    edge = np.array([0, 1, 2, (1 << 31) - 1, (1 << 31) - 2, (1 << 30), (1 << 30) - 1], np.uint64)
    for d in divisors:
        l = 0
        while (1 << l) < d:
            l += 1
        m = ((((1 << l) - d) << 32) // d + 1) & 0xFFFFFFFF
        assert m < (1 << 32)
        n = np.concatenate([edge, rng.integers(0, 1 << 31, 2000).astype(np.uint64),
                            (np.arange(1, 50, dtype=np.uint64) * d) - 1,
                            np.arange(1, 50, dtype=np.uint64) * d])
        n = n[n < (1 << 31)]
        q = (((n * np.uint64(m)) >> np.uint64(32)) + n) >> np.uint64(l)
        assert np.array_equal(q, n // np.uint64(d)), d


def test_handle_api_host_side():
    """The handle API's host-only parts: omega- agrees with the host rule,
    tau <= 1/2 is refused, and handles are created and destroyed without a
    device (no device call happens before a step)."""
    import ctypes
    L = _lib.lib()
    for tau in (0.55, 0.8, 1.0, 2.0):
        assert L.prt_trt_minus(1.0 / tau, 3.0 / 16.0) == pytest.approx(prtlb.trt_minus(1.0 / tau), rel=1e-15)
    assert np.isnan(L.prt_trt_minus(1.0 / 0.5, 3.0 / 16.0))
    buf = (ctypes.c_double * 8)()
    p = ctypes.addressof(buf)
    h = ctypes.c_void_p()
    _lib.call("prt_lbm_solver_create", 19, 0, 4, 4, 4, p, 10, p + 8, p + 16, p + 24, p + 32,
              0.8, 3.0 / 16.0, 2e-6, 1024.0, 0, None, ctypes.byref(h))
    assert h.value
    assert L.prt_lbm_solver_newest(h) == p + 8
    _lib.call("prt_lbm_solver_destroy", h)
    with pytest.raises(_lib.PrtError):
        _lib.call("prt_lbm_solver_create", 19, 0, 4, 4, 4, p, 10, p + 8, p + 8, p + 24, p + 32,
                  0.8, 3.0 / 16.0, 2e-6, 1024.0, 0, None, ctypes.byref(h))     # aliased buffers
    with pytest.raises(_lib.PrtError):
        _lib.call("prt_lbm_solver_create", 19, 0, 4, 4, 4, p, 10, p + 8, p + 16, p + 24, p + 32,
                  0.5, 3.0 / 16.0, 2e-6, 1024.0, 0, None, ctypes.byref(h))     # tau = 1/2
    assert L.prt_lbm_launch_blocks(19, 8, 8, 130, 0) > 0


def test_write_dataset_refuses_complaints(tmp_path):
    """A dataset whose runs drew check_physics complaints is not written
    (3D/tools/build_dataset_2d.py:889-894); a clean one is."""
    from prtlb import dataset as D
    shape = (6, 5, 4)
    mat = np.full((1,) + shape, prtlb.PORE, np.uint8)
    conc = np.zeros((2, 3, 1) + shape, np.float32)
    ds = D.write_layout(mat, np.zeros((1,) + shape, np.float32), np.zeros((1,) + shape, np.float32),
                        conc, np.zeros((2, 3) + shape, np.float32), np.zeros(2, np.int32),
                        np.zeros((2, 3), np.float32), np.zeros((2, 3), np.float32),
                        np.ones(2, bool), np.zeros(2, np.float32), species=["C"])
    ds["_complaints"] = {1: ["C: the field went non-finite"]}
    path = tmp_path / "bad.npz"
    with pytest.raises(SystemExit, match="NOT been written"):
        D.write_dataset(str(path), ds)
    assert not path.exists()
    ds["_complaints"] = {}
    D.write_dataset(str(tmp_path / "good.npz"), ds)
    z = np.load(str(tmp_path / "good.npz"), allow_pickle=False)
    assert z["geom/material"].shape == (1,) + shape


def test_chemistry_matches_the_oracle():
    """The package's Chemistry holds what the oracle's restatement of the
    reference's holds: built-in chemicals, truncation, decay, refusals."""
    attrs = ("names", "inlet", "c_init", "d_rel", "d_biofilm_rel", "immobile", "ks_donor",
             "ks_acceptor", "alpha", "yield_", "da_decay", "decay_lattice", "biotic_on",
             "biomass_coupled", "abiotic_index", "abiotic_product_index", "abiotic_surface_only",
             "biofilm_code", "seed_in_biofilm", "l_ref_voxels", "source")

    def same(a, b):
        for k in attrs:
            x, y = getattr(a, k), getattr(b, k)
            if isinstance(x, np.ndarray):
                assert x.dtype == y.dtype and np.array_equal(x, y, equal_nan=x.dtype.kind == "f"), k
            else:
                assert x == y, k

    for ns in (1, 2, 3, 4, 7):
        same(prtlb.default_chemistry(ns), O.default_chemistry(ns))
        same(prtlb.default_chemistry(4).truncate(ns), O.default_chemistry(4).truncate(ns))
    kw = dict(ks_donor=0.3, da_decay=0.5, biofilm_code=3, l_ref_voxels=12, abiotic_product_index=1)
    a = prtlb.Chemistry(["x", "y"], [1, np.nan], [0, 0.2], [1, 0.5], [0, 1], **kw)
    b = O.Chemistry(["x", "y"], [1, np.nan], [0, 0.2], [1, 0.5], [0, 1], **kw)
    same(a, b)
    assert a.decay_for(40) == b.decay_for(40) == 0.5 / 40
    assert a.truncate(1).inlet.shape == (1,) and a.inlet.shape == (2,)
    with pytest.raises(ValueError):
        prtlb.Chemistry(["x", "y"], [1.0], [0, 0], [1, 1], [0, 0])


def test_sorption_physics_view():
    """Irreversible sorption keeps the sorbed S at index 1, the acceptor's
    slot, where check_physics would cap it at zero.  S is immobile and keeps
    what the passing solute leaves, so it may exceed the feed: through
    campaign.physics_view the mobile solute alone meets the donor's rules
    (a solute above its feed is still named), and sorbed_complaints holds S
    to what does bound it (finite, not negative)."""
    from prtlb import campaign as C
    ch = prtlb.chemistry.sorption_irreversible()
    mat = np.full((30, 6), prtlb.PORE, np.uint8)
    x = np.linspace(0.0, 1.0, 30, dtype=np.float32)[:, None] * np.ones((1, 6), np.float32)
    out = np.zeros((3, 2, 30, 6), np.float32)
    for t, f in enumerate((0.2, 0.6, 1.0)):
        front = np.clip((f - x) * 3.0, 0.0, 1.0)
        out[t, 0] = front * np.exp(-x)
        out[t, 1] = 0.6 * front * (1 - np.exp(-x)) + 0.05 * t
    raw = prtlb.check_physics(out, mat, chem=ch, da_abio=1.0)
    assert any("S reaches" in s for s in raw)          # the acceptor's cap of zero
    view, vch = C.physics_view(out, ch, "sorption_irreversible")
    assert view.shape == (3, 1, 30, 6) and list(vch.names) == ["C"]
    assert prtlb.check_physics(view, mat, chem=vch, da_abio=1.0) == []
    assert C.sorbed_complaints(out, mat, "sorption_irreversible") == []
    out[2, 1, 5, 3] = 40.0                             # S well above the feed: sound
    view, vch = C.physics_view(out, ch, "sorption_irreversible")
    assert prtlb.check_physics(view, mat, chem=vch, da_abio=1.0) == []
    assert C.sorbed_complaints(out, mat, "sorption_irreversible") == []
    out[2, 1, 5, 3] = -0.5
    assert C.sorbed_complaints(out, mat, "sorption_irreversible") == [
        "S reaches -0.5, which is negative"]
    out[2, 1, 5, 3] = np.nan
    bad = C.sorbed_complaints(out, mat, "sorption_reversible")
    assert len(bad) == 1 and "S is not finite: 1 NaN" in bad[0]
    assert C.sorbed_complaints(out, mat, "monod") == []
    out[2, 1, 5, 3] = 0.0
    out[2, 0, 7, 2] = 1.5                              # the solute above its feed
    view, vch = C.physics_view(out, ch, "sorption_irreversible")
    bad = prtlb.check_physics(view, mat, chem=vch, da_abio=1.0)
    assert len(bad) == 1 and "C reaches 1.500, above the 1.000" in bad[0]


def test_batch_shapes_and_plans_host_side():
    """The batched steps' launch shapes: the shape codes the tuner sweeps
    are the ones the library takes, and a shaped plan lays out blocks of the
    shape's tile and partial-sum slots per 256 entries (plans only fill host
    descriptors; the member buffers are never touched here)."""
    L = _lib.lib()
    tiles = {_lib.batch_shape(t, e, x): (1 << t) * (1 << e)
             for x in (0, 1) for t in range(5, 10) for e in range(4)}
    for code, tile in tiles.items():
        want = tile if (code in _lib.BATCH_SHAPES) else 0
        assert L.prt_batch_shape_tile(code) == want, _lib.batch_shape_name(code)
    assert len(_lib.BATCH_SHAPES) == 10
    assert {L.prt_batch_shape_tile(c) for c in _lib.BATCH_SHAPES} == {256, 512}
    assert L.prt_batch_shape_tile(3 << 8 | 8) == 0

    sizes = [1, 255, 256, 257, 700, 1024, 3000]
    segs = (_lib.LbmSegment * len(sizes))()
    base = 1 << 20
    for i, n in enumerate(sizes):
        a = base + 0x10000 * i
        segs[i] = _lib.LbmSegment(None, a, a + 0x8000, n, 1, 1, 0, 1.25, 1.6, 1e-6, 1024.0,
                                  a + 0x4000, n, 16, 0)
    dsz = L.prt_lbm_batch_desc_size()
    for code in _lib.BATCH_SHAPES:
        tile = L.prt_batch_shape_tile(code)
        desc = np.zeros(dsz * len(sizes), np.uint8)
        nb, nt = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("prt_lbm_batch_plan_shaped", 19, 1, segs, len(sizes), code, desc.ctypes.data,
                  None, 0, ctypes.byref(nb), ctypes.byref(nt))
        nblk = [(n + tile - 1) // tile for n in sizes]
        assert nb.value == sum(nblk)
        assert nt.value == sum((n + 255) // 256 for n in sizes)
        blk = np.zeros(nb.value, np.int32)
        _lib.call("prt_lbm_batch_plan_shaped", 19, 1, segs, len(sizes), code, desc.ctypes.data,
                  blk.ctypes.data, nb.value, ctypes.byref(nb), ctypes.byref(nt))
        assert np.array_equal(blk, np.repeat(np.arange(len(sizes)), nblk))
        with pytest.raises(_lib.PrtError):     # the block table too short
            _lib.call("prt_lbm_batch_plan_shaped", 19, 1, segs, len(sizes), code,
                      desc.ctypes.data, blk.ctypes.data, nb.value - 1, ctypes.byref(nb),
                      ctypes.byref(nt))
    # the unshaped plan is the default shape's
    nb0 = ctypes.c_int64()
    _lib.call("prt_lbm_batch_plan", 19, 1, segs, len(sizes), desc.ctypes.data, None, 0,
              ctypes.byref(nb0))
    assert nb0.value == sum((n + 255) // 256 for n in sizes)
    with pytest.raises(_lib.PrtError):
        _lib.call("prt_lbm_batch_plan_shaped", 19, 1, segs, len(sizes), 7, desc.ctypes.data,
                  None, 0, ctypes.byref(nb), ctypes.byref(nt))
    # dense members: 256 threads with one node each, either block order
    dense = (_lib.LbmSegment * 1)()
    dense[0] = _lib.LbmSegment(base, base + 0x100, base + 0x200, 4, 4, 4, 0, 1.25, 1.6, 1e-6, 1.0,
                               None, 0, 0, 0)
    for code in _lib.BATCH_SHAPES:
        ok = (code & 0xFF) == 8
        if ok:
            _lib.call("prt_lbm_batch_plan_shaped", 19, 1, dense, 1, code, desc.ctypes.data, None,
                      0, ctypes.byref(nb), ctypes.byref(nt))
        else:
            with pytest.raises(_lib.PrtError):
                _lib.call("prt_lbm_batch_plan_shaped", 19, 1, dense, 1, code, desc.ctypes.data,
                          None, 0, ctypes.byref(nb), ctypes.byref(nt))


def test_slab_count_for_grids_beyond_the_32_bit_index():
    """Grids whose populations exceed the kernels' 32-bit element index are
    solved as x-slabs of one device: the fewest slabs that each fit."""
    from prtlb.slab import parts_needed, slab_dims, split_planes
    L = _lib.lib()
    assert parts_needed((256, 256, 256), "fp64") == 1
    assert parts_needed((2048, 2048), "fp32") == 1
    for shape, st in (((1024, 1024, 1024), "fp32"), ((768, 768, 768), "fp64"),
                      ((1100, 900, 1000), "fp16"), ((30000, 30000), "fp64")):
        code = _lib.PRT_STORE[st]
        dims = (shape[0], shape[1], 1 if len(shape) == 2 else shape[2])
        q = 9 if len(shape) == 2 else 19
        assert L.prt_lbm_pop_elems(q, code, *dims) < 0          # the box is refused
        p = parts_needed(shape, st)
        widest = max(x1 - x0 for x0, x1 in split_planes(shape[0], p))
        assert L.prt_lbm_pop_elems(q, code, *slab_dims(shape, widest)) > 0
        fewer = -(-shape[0] // (p - 1))
        assert L.prt_lbm_pop_elems(q, code, *slab_dims(shape, fewer)) < 0


def test_host_campaign_estimate_units():
    """bench.cpu_campaign_estimate: node iterations over the whole-host
    D3Q19 rate plus pore-voxel steps over the per-core transport rate times
    the processes, and the GPU's batched rate over the host's."""
    import bench
    campaign = {"n_solves": 64,
                "work": {"stokes_node_iterations": 3_600_000_000,
                         "transport_pore_voxel_steps": 7_200_000_000},
                "stages": {"total": {"batched_solves_per_hour": 3600.0}}}
    cpu = {"all_cores_procs": 10, "d3q19_64cube_mlups_all_cores": 10.0,
           "adr_3d_64cube_sorption_pore_voxel_steps_per_s_1core": 2_000_000.0}
    est = bench.cpu_campaign_estimate(campaign, cpu)
    assert est["stokes_s"] == 360.0          # 3.6e9 / 10e6
    assert est["transport_s"] == 360.0       # 7.2e9 / (2e6 * 10)
    assert est["total_h"] == 0.2
    assert est["solves_per_hour"] == 320.0   # 64 solves in 720 s
    assert est["gpu_batched_over_host"] == 11.2


def _adr_params_ok():
    p = _lib.AdrParams()
    p.nd, p.ns, p.dt = 3, 1, 0.1
    p.abio_index = p.abio_target = p.abio_product = -1
    p.mobile[0] = 1
    return p


@pytest.mark.parametrize("case", [
    "q", "grid", "d2q9_nz", "lbm_range", "static_range", "table_nd", "adr_ns", "adr_dt", "adr_dface",
    "adr_alias", "adr_steps", "adr_species_index"])
def test_abi_refuses_bad_arguments_before_any_device_call(case):
    """The C ABI's argument checks run on the host before any HIP call:
    each bad argument returns its status (PRT_E_ARG -1, PRT_E_RANGE -3) and
    a message, on a machine without a GPU."""
    import ctypes
    L = _lib.lib()
    buf = (ctypes.c_double * 16)()
    a, b, s = ctypes.addressof(buf), ctypes.addressof(buf) + 64, ctypes.addressof(buf) + 96
    lbm = lambda q, nx, ny, nz, fn="prt_lbm_run": getattr(L, fn)(
        q, 0, nx, ny, nz, s, a, b, 1, 1.25, 0.9, 1e-6, 1024.0, 0, None, None)
    p = _adr_params_ok()
    want, text = -1, None
    if case == "q":
        rc, text = lbm(7, 4, 4, 4), "q must be 9"
    elif case == "grid":
        rc, text = lbm(19, 4, 4, 0), "grid dimensions"
    elif case == "d2q9_nz":
        rc, text = lbm(9, 4, 4, 2), "D2Q9"
    elif case == "lbm_range":
        rc, want, text = lbm(19, 2000, 2000, 2000), -3, "32-bit"
    elif case == "static_range":
        rc, want = lbm(19, 2000, 2000, 2000, "prt_lbm_run_sum_static"), -3
    elif case == "table_nd":
        rc, text = L.prt_adr_pore_table(4, 4, 4, 4, s, a, b, 1, a, None), "nd must be 2 or 3"
    else:
        cin, cout, steps = a, b, 1
        if case == "adr_ns":
            p.ns, text = 9, "1..8 chemicals"
        elif case == "adr_dt":
            p.dt, text = 0.0, "dt must be positive"
        elif case == "adr_dface":
            p.dface[0][1], text = -1.0, "non-negative"
        elif case == "adr_alias":
            cout, text = a, "aliased"
        elif case == "adr_steps":
            steps, text = -1, "negative step count"
        elif case == "adr_species_index":
            p.abio_index, text = 3, "species index"
        rc = L.prt_adr_run_list(ctypes.byref(p), 4, 4, 4, s, 4, cin, cout, steps, None)
    assert rc == want, (case, rc)
    msg = L.prt_last_error().decode()
    assert msg and (text is None or text in msg), (case, msg)
