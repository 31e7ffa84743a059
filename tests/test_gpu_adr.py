"""Transport on the GPU: bit for bit against the reference's golden vectors
and the oracle (float32), FP64 against the oracle's FP64 variant, and the
closed forms (plug flow, discrete sorption relaxation)."""
import numpy as np
import pytest

from conftest import golden
from oracle import prtlb_oracle as O
import prtlb

pytestmark = pytest.mark.gpu


def _nan(v):
    if isinstance(v, list):
        return [float("nan") if x == "nan" else x for x in v]
    return float("nan") if v == "nan" else v


def chem_from(kw, mod):
    return None if kw is None else mod.Chemistry(**{k: _nan(v) for k, v in kw.items()})


ADR = ["adr2d_channel4", "adr2d_auto", "adr2d_disks_abio", "adr2d_disks_default4",
       "adr2d_sorption_irrev", "adr2d_coupled", "adr2d_biofilm", "adr3d_monod",
       "adr3d_default3_abio"]


@pytest.mark.parametrize("name", ADR)
def test_golden_bit_exact(name):
    d = golden(name)
    mt = d["meta"]
    print("case", name, flush=True)
    info = {}
    out, tn = prtlb.solve_adr(d["mat"], d["vel"], info=info,
                              chem=chem_from(mt["chem"], prtlb), **mt["kw"])
    assert out.dtype == np.float32
    assert np.array_equal(out, d["out"])
    assert np.array_equal(tn, d["t_norm"])
    for k, v in mt["info"].items():
        if k != "chemistry":
            assert info[k] == v, k


def _flow3d(shape, seed):
    g = prtlb.sphere_pack(shape, (2.0, 4.0), 0.5, seed=seed)
    v = prtlb.stokes_d3q19(g, nit=800, force=1e-5, tol=1e-6)
    return g, v


@pytest.mark.parametrize("ns,da,da_abio,ks", [(1, 1.0, 0.0, 0.5), (4, 2.0, 0.7, 0.0), (3, 0.3, 1.0, 0.1)])
def test_vs_oracle_3d(ns, da, da_abio, ks):
    g, v = _flow3d((30, 22, 18), 7)
    print("case 3d", ns, da, da_abio, ks, flush=True)
    ch = prtlb.default_chemistry(ns) if ns > 1 else prtlb.chemistry.monod(ks)
    if ns > 1:
        ch.ks_donor = ks
    cho = O.default_chemistry(ns) if ns > 1 else O.Chemistry(["C"], [1.0], [0.0], [1.0], [False],
                                                               ks_donor=ks)
    if ns > 1:
        cho.ks_donor = ks
    kw = dict(pe=10.0, da=da, n_t=5, n_species=ns, steps=400, da_abio=da_abio)
    out, tn = prtlb.solve_adr(g, v, chem=ch, **kw)
    outo, tno = O.solve_adr(g, v, chem=cho, **kw)
    assert np.array_equal(out, outo) and np.array_equal(tn, tno)


@pytest.mark.parametrize("shape", [(1, 7, 9), (6, 5, 1), (2, 2, 70), (3, 40), (40, 1)])
def test_thin_grids_vs_oracle(shape):
    """Degenerate extents -- a one-plane grid (the inlet pin then reaches
    every voxel), one-voxel y or z, short rows -- bit for bit against the
    oracle with every chemistry kind (4 species: mobile and immobile rows)."""
    rng = np.random.default_rng(sum(shape) + 1)
    g = np.where(rng.random(shape) < 0.2, prtlb.SOLID, prtlb.PORE).astype(np.uint8)
    g[0] = prtlb.PORE
    rngv = np.random.default_rng(7)
    v = np.zeros((3,) + shape, np.float32)
    v[0] = np.where(g == prtlb.PORE, 0.5 + rngv.random(shape), 0.0)
    print("case thin", shape, flush=True)
    kw = dict(pe=10.0, da=1.0, n_t=3, n_species=4, steps=150, da_abio=0.5)
    out, tn = prtlb.solve_adr(g, v, **kw)
    outo, tno = O.solve_adr(g, v, **kw)
    assert np.array_equal(out, outo) and np.array_equal(tn, tno)


def test_auto_stop_matches_oracle():
    # the stop step rests on float32 pore means in NumPy's summation order
    g = prtlb.disk_pack((48, 24), 3.0, 0.7, seed=3)
    v = prtlb.stokes_d2q9(g, nit=3000, force=1e-5, tol=1e-6)
    ch = prtlb.chemistry.first_order(False)
    cho = O.Chemistry(["C"], [1.0], [0.0], [1.0], [False], biotic_on=False, abiotic_index=0)
    i1, i2 = {}, {}
    out, tn = prtlb.solve_adr(g, v, pe=20.0, da=0.0, n_t=6, n_species=1, chem=ch,
                              da_abio=1.0, info=i1)
    outo, tno = O.solve_adr(g, v, pe=20.0, da=0.0, n_t=6, n_species=1, chem=cho,
                            da_abio=1.0, info=i2)
    assert i1["converged"] and i2["converged"]
    assert i1["steps"] == i2["steps"]
    assert np.array_equal(out, outo) and np.array_equal(tn, tno)
    # a run that reaches the step cap stops at the same place as well
    i1, i2 = {}, {}
    out, tn = prtlb.solve_adr(g, v, pe=20.0, da=1.0, n_t=4, n_species=2, info=i1,
                              max_steps=3000)
    outo, tno = O.solve_adr(g, v, pe=20.0, da=1.0, n_t=4, n_species=2, info=i2,
                            max_steps=3000)
    assert i1["steps"] == i2["steps"] and np.array_equal(out, outo)


def test_fp64_variant_vs_oracle():
    g, v = _flow3d((20, 14, 12), 9)
    kw = dict(pe=8.0, da=1.0, n_t=3, n_species=4, steps=150, da_abio=0.5)
    out, _ = prtlb.solve_adr(g, v, precision="fp64", **kw)
    outo, _ = O.solve_adr(g, v, real=np.float64, **kw)
    assert out.dtype == np.float64
    scale = max(np.abs(outo).max(), 1e-300)
    assert np.abs(out - outo).max() / scale < 1e-12


def test_reversible_sorption_vs_oracle_and_closed_form():
    g = prtlb.disk_pack((40, 20), 3.0, 0.65, seed=11)
    v = prtlb.stokes_d2q9(g, nit=2000, force=1e-5, tol=1e-6)
    sorb = dict(da_sorb=3.0, da_desorb=1.0, c=0, s=1)
    kw = dict(pe=5.0, da=0.0, n_t=4, n_species=2, steps=300, sorption=sorb)
    out, tn = prtlb.solve_adr(g, v, chem=prtlb.chemistry.sorption_reversible(True), **kw)
    cho = O.Chemistry(["C", "S"], [1.0, float("nan")], [0.0, 0.0], [1.0, 0.0], [False, True],
                      biotic_on=False, abiotic_index=0, abiotic_surface_only=True)
    outo, tno = O.solve_adr(g, v, chem=cho, **kw)
    assert np.array_equal(out, outo)
    # well-mixed closed cell: uniform C0 = 1, S0 = 0, no feed, no flow; the
    # explicit step relaxes the gap to C_inf by (1 - (ka + kd) dt) exactly
    box = np.full((12, 10), prtlb.PORE, np.uint8)
    still = np.zeros((3, 12, 10), np.float32)
    ch = prtlb.Chemistry(["C", "S"], [float("nan"), float("nan")], [1.0, 0.0], [1.0, 0.0],
                         [False, True], biotic_on=False, abiotic_index=0)
    info = {}
    out, tn, T = prtlb.solve_adr(box, still, pe=1.0, da=0.0, n_t=2, n_species=2, steps=200,
                                 chem=ch, sorption=dict(da_sorb=2.0, da_desorb=1.0),
                                 info=info, return_state=True)
    ka, kd, dt = float(T.params.k_sorb), float(T.params.k_desorb), info["dt"]
    cinf = kd / (ka + kd)
    want = cinf + (1.0 - cinf) * (1.0 - (ka + kd) * dt) ** 200
    c_end = out[-1, 0]
    assert np.abs(c_end - want).max() < 5e-5
    assert np.abs(out[-1, 0] + out[-1, 1] - 1.0).max() < 5e-5


@pytest.mark.parametrize("da", [0.5, 1.0, 2.0])
def test_plug_flow_outlet(da):
    g = np.full((80, 8), prtlb.PORE, np.uint8)
    v = np.zeros((3, 80, 8), np.float32)
    v[0] = 1.0
    info = {}
    ch = prtlb.Chemistry(["C"], [1.0], [0.0], [1.0], [False], biotic_on=True)
    out, _ = prtlb.solve_adr(g, v, pe=4000.0, da=da, n_t=3, n_species=1, chem=ch, info=info)
    assert abs(float(out[-1, 0, -1].mean()) - np.exp(-da)) < 0.05
    assert abs(float(info["outlet_concentration"][0]) - np.exp(-da)) < 0.05
    assert out.max() <= 1.0 + 1e-4


def test_sealed_half_and_larger_parity():
    g = np.full((60, 8), prtlb.PORE, np.uint8)
    g[30:] = prtlb.SOLID
    v = np.zeros((3, 60, 8), np.float32)
    v[0, :30] = 1.0
    out, _ = prtlb.solve_adr(g, v, pe=50.0, da=0.1, n_t=5, n_species=1)
    assert float(out[:, 0, 35:].max()) == 0.0
    # a 64^3 sphere pack, 100 steps, bit for bit against the oracle
    gs = prtlb.sphere_pack((64, 64, 64), (4.0, 8.0), 0.35, seed=4)
    vs = prtlb.stokes_d3q19(gs, nit=600, force=2e-6, tol=1e-5)
    kw = dict(pe=10.0, da=1.0, n_t=3, n_species=1, steps=100)
    out, _ = prtlb.solve_adr(gs, vs, chem=prtlb.chemistry.monod(0.5), **kw)
    outo, _ = O.solve_adr(gs, vs, chem=O.Chemistry(["C"], [1.0], [0.0], [1.0], [False],
                                                   ks_donor=0.5), **kw)
    assert np.array_equal(out, outo)


def test_graph_replays_match_eager():
    g = prtlb.disk_pack((128, 128), 6.0, 0.6, seed=1)
    v = prtlb.stokes_d2q9(g, nit=3000, force=1e-5, tol=1e-6)
    T1 = prtlb.TransportSolver(g, v, 10.0, 0.0, 1, chem=prtlb.chemistry.first_order(True),
                               da_abio=1.0)
    T2 = prtlb.TransportSolver(g, v, 10.0, 0.0, 1, chem=prtlb.chemistry.first_order(True),
                               da_abio=1.0)
    T1.use_graph, T2.use_graph = True, False
    for k in range(5):
        T1.step_n(T1.GRAPH_STEPS)
        T2.step_n(T2.GRAPH_STEPS)
        import torch
        torch.cuda.synchronize()
        assert torch.equal(T1.current(), T2.current()), k
    out1, _ = prtlb.solve_adr(g, v, 10.0, 0.0, 4, n_species=1, steps=1000,
                              chem=prtlb.chemistry.first_order(True), da_abio=1.0)
    outo, _ = O.solve_adr(g, v, 10.0, 0.0, 4, n_species=1, steps=1000,
                          chem=O.Chemistry(["C"], [1.0], [0.0], [1.0], [False], biotic_on=False,
                                           abiotic_index=0, abiotic_surface_only=True),
                          da_abio=1.0)
    assert np.array_equal(out1, outo)


@pytest.mark.parametrize("n", [1, 7, 8, 9, 127, 128, 129, 1000, 8192, 8193, 100003, 9_000_017])
def test_device_pairwise_sum_equals_numpy(n):
    """prt_pairwise_sum_f32 against np.add.reduce on float32 values, bit for
    bit, over two rows gathered through an index: short leaves, ragged
    leaf tails, one and many 8192-element buffers, and more than 1024 of
    them (the chain then runs over several staged tiles)."""
    import torch
    from prtlb import _lib
    from prtlb.device import ptr, stream_ptr
    rng = np.random.default_rng(n)
    ld = n + 5
    rows = (rng.random((2, ld)) * rng.choice([1e-3, 1.0, 1e3], (2, ld))).astype(np.float32)
    idx = np.sort(rng.choice(ld, n, replace=False)).astype(np.int32)
    plan, L, I, H = _lib.pairwise_plan(n)
    dev = torch.device("cuda:0")
    c = torch.from_numpy(rows).to(dev)
    ti = torch.from_numpy(idx).to(dev)
    tp = torch.from_numpy(plan).to(dev)
    scratch = torch.empty(2 * (L + I), dtype=torch.float32, device=dev)
    out = torch.empty(2, dtype=torch.float32, device=dev)
    print("case n", n, flush=True)
    _lib.call("prt_pairwise_sum_f32", ptr(c), ld, 2, ptr(ti), n, ptr(tp), L, I, H, ptr(scratch),
              ptr(out), stream_ptr(dev))
    got = out.cpu().numpy()
    for r in range(2):
        assert got[r] == np.add.reduce(rows[r][idx]), (n, r)
