"""The transport scheme on the GPU against answers that exist without it:
the reference's own scheme checks (3D/tools/prtlb_3d.py:1859-2524,
"THE SCHEME ITSELF"), run through prtlb.solve_adr with the same grids,
chemistries, Peclet and Damkohler numbers and the same tolerances."""
import numpy as np
import pytest
from scipy.special import erfcinv

import prtlb
from prtlb import Chemistry, default_chemistry

pytestmark = pytest.mark.gpu

PORE, SOLID = prtlb.PORE, prtlb.SOLID


def _channel(nx, ny, walls=False):
    g = np.full((nx, ny), PORE, np.uint8)
    if walls:
        g[:, 0] = g[:, -1] = SOLID
    v = np.zeros((3, nx, ny), np.float32)
    v[0] = 1.0
    return g, v


def _tracer():
    return Chemistry(names=["Ac"], inlet=[1.0], c_init=[0.0], d_rel=[1.0],
                     immobile=[False], biotic_on=False)


def test_plug_flow_3d_and_maximum_principle():
    # 3D/tools/prtlb_3d.py:1876-1890
    g = np.full((40, 6, 6), PORE, np.uint8)
    v = np.zeros((3, 40, 6, 6), np.float32)
    v[0] = 1.0
    out, _ = prtlb.solve_adr(g, v, pe=4000.0, da=1.0, n_t=3, n_species=1)
    assert abs(float(out[-1, 0, -1].mean()) - np.exp(-1.0)) < 0.06
    assert out.max() <= 1.0 + 1e-4


@pytest.mark.parametrize("pe,tol", [(20.0, 0.10), (50.0, 0.08), (100.0, 0.08)])
def test_front_width_is_the_erfc_width(pe, tol):
    # 3D/tools/prtlb_3d.py:2266-2297: a step fed into an empty channel
    # spreads as erfc((x - ut)/sqrt(4Dt))/2; its 90%-10% width is
    # 2 erfcinv(0.2) sqrt(4Dt), compared without fitting anything
    nx = 64
    g, v = _channel(nx, 8)
    i0 = {}
    prtlb.solve_adr(g, v, pe=pe, da=0.0, n_t=2, n_species=1, steps=1, chem=_tracer(), info=i0)
    nsteps = int(round((nx * 0.5) / i0["dt"]))
    out, _ = prtlb.solve_adr(g, v, pe=pe, da=0.0, n_t=2, n_species=1, steps=nsteps, chem=_tracer())
    prof = out[-1, 0].mean(1).astype(float)
    xs = np.arange(nx) + 0.5
    x90, x10 = np.interp([0.9, 0.1], prof[::-1], xs[::-1])
    w = abs(x10 - x90)
    w_exact = 2 * erfcinv(0.2) * np.sqrt(4 * (nx / pe) * nsteps * i0["dt"])
    print("Pe", pe, "width", w, "exact", w_exact, flush=True)
    assert abs(w - w_exact) / w_exact < tol


def test_monotone_front_stays_monotone():
    # 3D/tools/prtlb_3d.py:2299-2318: the limiter leaves no bump
    g, v = _channel(80, 8)
    out, _ = prtlb.solve_adr(g, v, pe=150.0, da=0.0, n_t=6, n_species=1, steps=400, chem=_tracer())
    rises = [float(np.diff(out[i, 0].mean(1)).max()) for i in range(out.shape[0])]
    assert max(rises) < 1e-6, rises


@pytest.mark.parametrize("da", [20.0, 100.0, 400.0])
def test_donor_plus_product_never_exceeds_the_feed(da):
    # 3D/tools/prtlb_3d.py:2336-2355 ("biotic only", three chemicals, nearly
    # zero-order donor kinetics, no acceptor demand, zero yield)
    cm = default_chemistry(3)
    cm.ks_donor = 1e-3
    cm.alpha = 0.0
    cm.inlet = np.array([1.0, 1.0, 0.0], np.float32)
    cm.c_init = np.array([0.0, 1.0, 0.0], np.float32)
    g = np.full((40, 8), PORE, np.uint8)
    v = np.stack([np.ones((40, 8), np.float32), np.zeros((40, 8), np.float32),
                  np.zeros((40, 8), np.float32)])
    out, _ = prtlb.solve_adr(g, v, pe=30.0, da=da, n_t=3, n_species=3, chem=cm)
    tot = float((out[-1, 0] + out[-1, 2]).max())
    assert tot <= 1.02, tot


def test_zero_acceptor_half_saturation_is_first_order():
    # 3D/tools/prtlb_3d.py:2357-2366: Ks_A = 0 is first order in the
    # acceptor, not 0/0
    ck = default_chemistry(4)
    ck.ks_acceptor = 0.0
    g = np.full((30, 8), PORE, np.uint8)
    v = np.stack([np.ones((30, 8), np.float32)] * 3)
    out, _ = prtlb.solve_adr(g, v, pe=30.0, da=2.0, n_t=3, n_species=4, chem=ck)
    assert np.isfinite(out).all()


def test_stop_waits_for_the_slowest_chemical():
    # 3D/tools/prtlb_3d.py:2399-2421: at the automatic stop the product and
    # the biomass are within 1% of a run sixteen times as long
    g, v = _channel(40, 8, walls=True)
    chem = default_chemistry(4)
    info = {}
    stopped, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=3, n_species=4, chem=chem, info=info)
    long_run, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=3, n_species=4, chem=chem,
                                  steps=info["steps"] * 16)
    pm = g == PORE
    for j in (2, 3):
        f_stop = float(stopped[-1, j][pm].mean())
        f_long = float(long_run[-1, j][pm].mean())
        print("chemical", j, f_stop, f_long, flush=True)
        assert abs(f_stop - f_long) / max(abs(f_long), 1e-9) < 0.01


def test_peclet_uses_the_axial_velocity():
    # 3D/tools/prtlb_3d.py:2423-2436: a transverse component (faster in
    # speed, same axial mean) does not slow the front down
    g = np.full((40, 16), PORE, np.uint8)
    v_ax = np.zeros((3, 40, 16), np.float32)
    v_ax[0] = 1.0
    v_tt = v_ax.copy()
    v_tt[1] = 0.6
    a, _ = prtlb.solve_adr(g, v_ax, pe=30.0, da=0.0, n_t=2, n_species=1, steps=200, chem=_tracer())
    b, _ = prtlb.solve_adr(g, v_tt, pe=30.0, da=0.0, n_t=2, n_species=1, steps=200, chem=_tracer())

    def reach(f):
        return float(np.argmax(f[-1, 0].mean(1) < 0.5))

    assert abs(reach(a) - reach(b)) <= 1, (reach(a), reach(b))


def test_divergent_field_makes_no_solute():
    # 3D/tools/prtlb_3d.py:2438-2454: a 10% compressible field; a chemical
    # that is only transported never rises above its feed
    g = np.full((24, 16), PORE, np.uint8)
    g[:, 0] = g[:, -1] = SOLID
    xs = np.arange(24)[:, None] * np.ones((1, 16))
    v = np.zeros((3, 24, 16), np.float32)
    v[0] = 1.0 + 0.10 * np.sin(2 * np.pi * xs / 24.0)
    ch = Chemistry(names=["Ac", "A"], inlet=[1.0, 1.0], c_init=[0.0, 1.0], d_rel=[1.0, 1.0],
                   immobile=[False, False], biotic_on=False)
    out, _ = prtlb.solve_adr(g, v, pe=2.0, da=0.0, n_t=3, n_species=2, steps=40000, chem=ch)
    pd = g == PORE
    worst = max(float(out[:, 0][:, pd].max()), float(out[:, 1][:, pd].max()))
    assert worst <= 1.0005, worst


def test_abiotic_reaction_converts_donor_to_product():
    # 3D/tools/prtlb_3d.py:2485-2524 ("abiotic only": the donor is the
    # reactant and P the product, one mole for one; "none" is a pure loss)
    g, v = _channel(60, 16, walls=True)
    ma = g == PORE

    def chem(product):
        c = default_chemistry(4)
        c.biotic_on = False
        c.abiotic_index = 0
        c.abiotic_product_index = 2 if product else None
        c.abiotic_surface_only = False
        c.c_init = np.array([0.0, 0.0, 0.0, 0.0], np.float32)
        return c

    prev = -1.0
    for da_a in (0.5, 5.0):
        c = chem(True)
        out, _ = prtlb.solve_adr(g, v, pe=10.0, da=0.0, n_t=3, n_species=3, chem=c, da_abio=da_a)
        ac = float(out[-1, 0][ma].mean())
        p = float(out[-1, 2][ma].mean())
        print("Da_abio", da_a, "Ac", ac, "P", p, flush=True)
        assert p > prev and p > 0.05
        assert abs(ac + p - 1.0) < 0.01
        assert not prtlb.check_physics(out, g, chem=c, da_abio=da_a)
        prev = p
    out, _ = prtlb.solve_adr(g, v, pe=10.0, da=0.0, n_t=3, n_species=3, chem=chem(False), da_abio=5.0)
    assert float(out[-1, 2][ma].max()) == 0.0
    assert float(out[-1, 0][ma].mean()) < 0.5


def test_check_physics_on_a_sound_run_and_planted_faults():
    # 3D/tools/prtlb_3d.py:2368-2397: quiet on a sound four-chemical run,
    # and each planted fault is named
    g, v = _channel(40, 8, walls=True)
    chem = default_chemistry(4)
    op, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4, chem=chem)
    assert not prtlb.check_physics(op, g, chem=chem)
    b = op.copy()
    b[:, 2] = 5.0
    assert any("P reaches" in c for c in prtlb.check_physics(b, g, chem=chem))
    b = op.copy()
    b[:, 1] = -0.5 * (g == PORE)
    assert any("negative" in c for c in prtlb.check_physics(b, g, chem=chem))
    b = op.copy()
    b[:, 0] = 0.0
    assert any("exactly zero" in c for c in prtlb.check_physics(b, g, chem=chem))
    b = op.copy()
    b[:, 3] = 500.0
    assert any("runaway" in c for c in prtlb.check_physics(b, g, chem=chem))
    xs = np.arange(40)[:, None] * np.ones((1, 8))
    b = op.copy()
    b[-1, 0] = np.where(g == PORE, 0.5 + 0.4 * ((-1.0) ** xs), 0.0)
    assert any("checkerboard" in c for c in prtlb.check_physics(b, g, chem=chem))
    zs = np.arange(8)[None, :] * np.ones((40, 1))
    b = op.copy()
    b[-1, 0] = np.where(g == PORE, 0.5 + 0.4 * ((-1.0) ** zs), 0.0)
    assert not any("checkerboard" in c for c in prtlb.check_physics(b, g, chem=chem))
    z = op.copy()
    z[:, 2] = 0.0
    assert any("never happened" in c for c in prtlb.check_physics(z, g, chem=chem, da_abio=3.78))
    assert not any("never happened" in c for c in prtlb.check_physics(z, g, chem=chem, da_abio=0.0))


# ---- the run's bookkeeping and the chemistry switches
# (3D/tools/prtlb_3d.py:1899-2255), with the reference's Settings edits
# written as the Chemistry attributes they set

def test_snapshot_times_and_info():
    # :1899-1915: n_t = 2 spans 0..1; 11 snapshots of a 200-step run are
    # distinct; an explicit step count is recorded as not converged
    g, v = _channel(40, 6)
    _, tn = prtlb.solve_adr(g, v, pe=100.0, da=1.0, n_t=2, n_species=1)
    assert tn[0] == 0.0 and abs(tn[-1] - 1.0) < 1e-6
    info = {}
    _, tn = prtlb.solve_adr(g, v, pe=100.0, da=1.0, n_t=11, n_species=1, steps=200, info=info)
    assert len(set(tn.tolist())) == 11 and info["repeated_snapshots"] == 0
    assert info["converged"] is False


def test_default_chemistry_run():
    # :1917-1994: the biomass survives at the inlet face; check_physics flags
    # an all-NaN field and is quiet on this run; default chemistry equals
    # chem=None bit for bit; the product never exceeds its donor
    g, v = _channel(40, 8)
    out, tb = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=3, n_species=4)
    assert float(out[-1, 3, 0].mean()) > 0.5 * 0.1
    nan_field = np.full((3, 1, 20, 8), np.nan, np.float32)
    assert any("not finite" in s for s in prtlb.check_physics(nan_field, np.full((20, 8), PORE, np.uint8)))
    assert prtlb.check_physics(out, g) == []
    base, tb = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4)
    same, ts = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4, chem=default_chemistry(4))
    assert np.array_equal(base, same) and np.array_equal(tb, ts)
    assert float(base[:, 2].max()) <= 1.0 + 1e-4


def test_per_chemical_diffusivity_is_used():
    # :2000-2022: with no flow and no reaction a chemical given a tenth of
    # the diffusivity spreads less far, the other one exactly as before
    g = np.full((40, 8), PORE, np.uint8)
    still = np.zeros((3, 40, 8), np.float32)
    s = default_chemistry(2)
    slow = default_chemistry(2)
    slow.d_rel = np.array([1.0, 0.1], np.float32)
    slow.d_biofilm_rel = slow.d_rel.copy()
    a1, _ = prtlb.solve_adr(g, still, pe=1.0, da=1e-9, n_t=2, n_species=2, steps=3000, chem=s)
    a2, _ = prtlb.solve_adr(g, still, pe=1.0, da=1e-9, n_t=2, n_species=2, steps=3000, chem=slow)
    assert float(a2[-1, 1].mean()) < 0.6 * float(a1[-1, 1].mean())
    assert abs(float(a1[-1, 0].mean()) - float(a2[-1, 0].mean())) < 1e-6
    assert prtlb.check_physics(a2, g, chem=slow) == []


def test_unfed_acceptor_only_decreases():
    # :2024-2036: an acceptor placed in the sample (no feed) starts at 1 and
    # then only ever decreases
    g, v = _channel(40, 8)
    c = default_chemistry(4)
    c.inlet = np.array([1.0, np.nan, 0.0, np.nan], np.float32)
    c.c_init = np.array([0.0, 1.0, 0.0, 0.1], np.float32)
    pre, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=6, n_species=4, chem=c)
    prof = [float(pre[t, 1][g == PORE].mean()) for t in range(6)]
    assert abs(prof[0] - 1.0) < 1e-5
    assert all(prof[i + 1] <= prof[i] + 1e-6 for i in range(5)), prof


def test_biomass_mobility_kinetics_and_coupling_switches():
    # :2038-2055, :2160-2187: planktonic biomass sits further downstream than
    # sessile; a donor half-saturation constant and coupling the rate to the
    # biomass each change the product
    g, v = _channel(40, 8)
    base, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4)

    def downstream(f):
        return float(f[-1, 3, 20:].mean()) / max(float(f[-1, 3].mean()), 1e-30)

    pla = default_chemistry(4)
    pla.immobile = np.array([False, False, False, False])
    pla.d_rel = np.array([1.0, 1.0, 1.0, 0.2], np.float32)
    pla.d_biofilm_rel = np.array([1.0, 1.0, 1.0, 0.1], np.float32)
    mob, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4, chem=pla)
    assert downstream(mob) > downstream(base)
    mon = default_chemistry(4)
    mon.ks_donor = 1.0e-3
    o_m, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4, chem=mon)
    assert abs(float(o_m[-1, 2].mean()) - float(base[-1, 2].mean())) > 1e-4
    cp = default_chemistry(4)
    cp.biomass_coupled = True
    o_c, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4, chem=cp)
    assert abs(float(o_c[-1, 2].mean()) - float(base[-1, 2].mean())) > 1e-4


def test_abiotic_on_the_product_and_its_switches():
    # :2057-2140 ("both": the abiotic reaction consumes the product)
    g, v = _channel(40, 8)
    base4, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4)
    ab, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4, da_abio=5.0)
    assert float(ab[-1, 2].mean()) < 0.6 * float(base4[-1, 2].mean())
    non, _ = prtlb.solve_adr(g, v, pe=200.0, da=0.0, n_t=3, n_species=4, da_abio=8.0)
    assert abs(float(non[-1, 0, -1].mean()) - 1.0) < 0.02
    off, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4, da_abio=0.0)
    assert np.array_equal(off, base4)
    # surface-only: no surface in an open channel, some along walls
    s2 = default_chemistry(4)
    s2.abiotic_surface_only = True
    i_open = {}
    prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=4, n_species=4, chem=s2, da_abio=5.0, info=i_open)
    assert i_open["surface_voxels"] == 0.0
    gw, vw = _channel(40, 8, walls=True)
    b_w, _ = prtlb.solve_adr(gw, vw, pe=30.0, da=1.0, n_t=4, n_species=4, chem=s2)
    i_wall = {}
    a_w, _ = prtlb.solve_adr(gw, vw, pe=30.0, da=1.0, n_t=4, n_species=4, chem=s2, da_abio=5.0,
                             info=i_wall)
    assert float(a_w[-1, 2].mean()) < float(b_w[-1, 2].mean()) - 1e-4 and i_wall["surface_voxels"] > 0
    with pytest.raises(ValueError, match="cannot happen"):
        prtlb.solve_adr(gw, vw, pe=30.0, da=1.0, n_t=3, n_species=2, chem=default_chemistry(4), da_abio=5.0)


def test_no_microbes_means_no_biomass():
    # :2236-2252: with the biotic reaction off the biomass is exactly zero,
    # the abiotic reaction still consumes the donor, a biotic run has biomass
    g, v = _channel(40, 8)
    c = default_chemistry(4)
    c.biotic_on = False
    c.abiotic_index = 0
    c.abiotic_product_index = 2
    c.c_init = np.array([0.0, 0.0, 0.0, 0.0], np.float32)
    o_ab, _ = prtlb.solve_adr(g, v, pe=30.0, da=0.0, n_t=3, n_species=4, chem=c, da_abio=2.0)
    assert float(np.abs(o_ab[:, 3]).max()) == 0.0
    assert float(o_ab[-1, 0, -1].mean()) < 0.9
    bio, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=3, n_species=4)
    assert float(bio[:, 3].max()) > 0.05
    info = {}
    prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=3, n_species=2, info=info)
    assert info.get("chemistry") is not None


def test_biofilm_lining():
    # :2189-2234: a biofilm lining is not solid, the attached biomass starts
    # only there, a twentieth of the biofilm diffusivity changes the field,
    # and with no biofilm voxel the biofilm settings change nothing
    BF = 4
    gb = np.full((40, 8), PORE, np.uint8)
    gb[:, 0] = gb[:, -1] = BF
    vb = np.zeros((3, 40, 8), np.float32)
    vb[0] = 1.0

    def lined(d_bf):
        c = default_chemistry(4)
        c.biofilm_code = BF
        c.seed_in_biofilm = True
        c.d_biofilm_rel = np.array(d_bf, np.float32)
        return c

    ob, _ = prtlb.solve_adr(gb, vb, pe=30.0, da=1.0, n_t=3, n_species=4, chem=lined([1.0, 1.0, 1.0, 0.0]))
    lining = gb == BF
    assert float(ob[-1, 0][lining].max()) > 0.1
    assert float(ob[0, 3][lining].min()) > 0.0 and float(ob[0, 3][gb == PORE].max()) == 0.0
    slow = lined([0.05, 0.05, 0.05, 0.0])
    os_, _ = prtlb.solve_adr(gb, vb, pe=30.0, da=1.0, n_t=3, n_species=4, chem=slow)
    assert abs(float(os_[-1, 0][lining].mean()) - float(ob[-1, 0][lining].mean())) > 1e-3
    g, v = _channel(40, 8)
    plain, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=3, n_species=4)
    with_bf, _ = prtlb.solve_adr(g, v, pe=30.0, da=1.0, n_t=3, n_species=4, chem=slow.truncate(4))
    assert np.array_equal(plain, with_bf)
