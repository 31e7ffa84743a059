"""Write the golden vectors the oracle is pinned against.

Runs the REFERENCE simulator (shahram444/PRT-DeepONet-GUI, read-only at
/root/reference) on small seeded inputs and stores inputs and outputs as
compressed .npz fixtures beside this script.  Only this script touches the
reference; the fixtures are data.  NumPy version at generation time is
recorded in every file (float32 casting follows NumPy 2 / NEP 50).

    python tests/golden/make_golden.py

The FP64 pre-cast Stokes velocity is captured by handing the reference module
a thin numpy proxy whose ``stack`` records the array it builds (the reference
returns only the float32 cast).
"""
import json
import os
import sys

import numpy as np

REF = "/root/reference/3D/tools"
HERE = os.path.dirname(os.path.abspath(__file__))


class _NpSpy:
    """numpy, except that np.stack remembers the last array it built."""

    def __init__(self):
        self.last = None

    def __getattr__(self, name):
        return getattr(np, name)

    def stack(self, *a, **k):
        self.last = np.stack(*a, **k)
        return self.last


def disks(shape, r, porosity, seed):
    """Seeded periodic overlapping disks/spheres (inputs only)."""
    rng = np.random.default_rng(seed)
    grids = np.indices(shape).astype(np.float64)
    solid = np.zeros(shape, bool)
    while (~solid).mean() > porosity:
        c = rng.uniform(0, shape, size=len(shape))
        rr = r if np.isscalar(r) else rng.uniform(*r)
        d2 = np.zeros(shape)
        for a, n in enumerate(shape):
            d = np.abs(grids[a] - c[a])
            d = np.minimum(d, n - d)
            d2 += d * d
        solid |= d2 <= rr * rr
    return np.where(solid, 0, 2).astype(np.uint8)


def save(name, **arrs):
    meta = arrs.pop("meta", {})
    meta["numpy"] = np.__version__
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        meta=np.array(json.dumps(meta)), **arrs)
    print("wrote", name, flush=True)


def main():
    sys.path.insert(0, REF)
    import prtlb_2d as r2
    import prtlb_3d as r3
    from settings_and_units import Chemistry

    # ---- Stokes, fixed iteration count (tol 0 never stops early) ----------
    spy2 = _NpSpy()
    r2.np = spy2
    g2, _ = r2.keep_inlet_connected(disks((24, 16), 3.0, 0.62, 1))
    info = {}
    v = r2.stokes_d2q9(g2, nit=200, force=2e-5, tau_lb=0.8, tol=0.0, info=info)
    save("stokes2d_fixed", g=g2, v32=v, v64=spy2.last,
         meta=dict(nit=200, force=2e-5, tau=0.8, tol=0.0, info=info))

    info = {}
    v = r2.stokes_d2q9(g2, nit=6000, force=2e-6, tau_lb=0.55, tol=1e-6, info=info)
    save("stokes2d_conv", g=g2, v32=v, v64=spy2.last,
         meta=dict(nit=6000, force=2e-6, tau=0.55, tol=1e-6, info=info))

    spy3 = _NpSpy()
    r3.np = spy3
    g3, _ = r3.keep_inlet_connected(disks((16, 12, 10), (2.0, 3.5), 0.55, 2))
    info = {}
    v = r3.stokes_d3q19(g3, nit=60, force=3e-5, tau_lb=1.2, tol=0.0, info=info)
    save("stokes3d_fixed", g=g3, v32=v, v64=spy3.last,
         meta=dict(nit=60, force=3e-5, tau=1.2, tol=0.0, info=info))

    g3c, _ = r3.keep_inlet_connected(disks((12, 10, 8), (1.5, 3.0), 0.6, 3))
    info = {}
    v3c = r3.stokes_d3q19(g3c, nit=4000, force=2e-6, tau_lb=0.8, tol=1e-6, info=info)
    save("stokes3d_conv", g=g3c, v32=v3c, v64=spy3.last,
         meta=dict(nit=4000, force=2e-6, tau=0.8, tol=1e-6, info=info))
    r2.np = np
    r3.np = np

    # ---- transport, fixed step counts -------------------------------------
    def adr(name, mod, mat, vel, kw, chem_kw=None):
        info = {}
        chem = Chemistry(**chem_kw) if chem_kw is not None else None
        out, tn = mod.solve_adr(mat, vel, info=info, chem=chem, **kw)
        kw2 = dict(kw)
        save(name, mat=mat, vel=vel, out=out, t_norm=tn,
             meta=dict(kw=kw2, chem=_jsonable(chem_kw), info=_jsonable(info)))
        return out

    g = np.full((40, 8), 2, np.uint8)
    v = np.zeros((3, 40, 8), np.float32)
    v[0] = 1.0
    adr("adr2d_channel4", r2, g, v,
        dict(pe=30.0, da=1.0, n_t=4, n_species=4, steps=300))
    g6 = np.full((40, 6), 2, np.uint8)
    v6 = np.zeros((3, 40, 6), np.float32)
    v6[0] = 1.0
    adr("adr2d_auto", r2, g6, v6, dict(pe=100.0, da=1.0, n_t=5, n_species=1))

    gd, _ = r2.keep_inlet_connected(disks((32, 24), 3.0, 0.6, 4))
    vd = r2.stokes_d2q9(gd, nit=1500, force=2e-5, tol=1e-6)
    cfg1 = dict(names=["C"], inlet=[1.0], c_init=[0.0], d_rel=[1.0],
                immobile=[False], biotic_on=False, abiotic_surface_only=True,
                abiotic_index=0)
    adr("adr2d_disks_abio", r2, gd, vd,
        dict(pe=10.0, da=0.0, n_t=5, n_species=1, steps=400, da_abio=1.0), cfg1)
    adr("adr2d_disks_default4", r2, gd, vd,
        dict(pe=20.0, da=2.0, n_t=4, n_species=4, steps=300))
    sorb = dict(names=["C", "S"], inlet=[1.0, float("nan")], c_init=[0.0, 0.0],
                d_rel=[1.0, 0.0], immobile=[False, True], biotic_on=False,
                abiotic_index=0, abiotic_product_index=1,
                abiotic_surface_only=True)
    adr("adr2d_sorption_irrev", r2, gd, vd,
        dict(pe=5.0, da=0.0, n_t=4, n_species=2, steps=300, da_abio=2.0), sorb)
    coupled = dict(names=["Ac", "A", "P", "Bio"], inlet=[1.0, 1.0, 0.0, float("nan")],
                   c_init=[0.0, 0.2, 0.0, 0.1], d_rel=[1.0, 0.7, 1.3, 0.0],
                   immobile=[False, False, False, True], ks_donor=0.1,
                   ks_acceptor=0.15, alpha=0.4, yield_=0.05, decay_lattice=0.01,
                   biomass_coupled=True, abiotic_index=2, da_decay=None)
    adr("adr2d_coupled", r2, gd, vd,
        dict(pe=15.0, da=4.0, n_t=4, n_species=4, steps=300, da_abio=0.5), coupled)

    gb, nb = r2.coat_with_biofilm(gd, 3, 1)
    biof = dict(names=["Ac", "A", "P", "Bio"], inlet=[1.0, 1.0, 0.0, float("nan")],
                c_init=[0.0, 0.0, 0.0, 0.1], d_rel=[1.0, 1.0, 1.0, 0.0],
                d_biofilm_rel=[0.2, 0.3, 0.25, 0.0],
                immobile=[False, False, False, True], biofilm_code=3,
                seed_in_biofilm=True, decay_lattice=0.006)
    adr("adr2d_biofilm", r2, gb, vd,
        dict(pe=12.0, da=1.5, n_t=4, n_species=4, steps=300), biof)

    gs, _ = r3.keep_inlet_connected(disks((16, 12, 10), (2.0, 3.5), 0.55, 5))
    vs = r3.stokes_d3q19(gs, nit=600, force=2e-5, tol=1e-6)
    cfg2 = dict(names=["C"], inlet=[1.0], c_init=[0.0], d_rel=[1.0],
                immobile=[False], ks_donor=0.5, biotic_on=True)
    adr("adr3d_monod", r3, gs, vs,
        dict(pe=10.0, da=1.0, n_t=4, n_species=1, steps=200), cfg2)
    adr("adr3d_default3_abio", r3, gs, vs,
        dict(pe=40.0, da=0.5, n_t=3, n_species=3, steps=150, da_abio=1.5))

    # ---- geometry helpers --------------------------------------------------
    rng = np.random.default_rng(7)
    gg = np.where(rng.random((14, 12, 10)) < 0.45, 0, 2).astype(np.uint8)
    gk, nrem = r3.keep_inlet_connected(gg)
    gc1, n1 = r3.coat_with_biofilm(gk, 3, 1)
    gc2, n2 = r3.coat_with_biofilm(gk, 4, 2)
    g2d = np.where(rng.random((30, 20)) < 0.42, 0, 2).astype(np.uint8)
    gk2, nrem2 = r2.keep_inlet_connected(g2d)
    save("geometry", g=gg, kept=gk, coat1=gc1, coat2=gc2, g2d=g2d, kept2d=gk2,
         meta=dict(removed=nrem, n1=n1, n2=n2, removed2d=nrem2))

    # ---- check_physics -----------------------------------------------------
    cases = {}
    gp = np.full((20, 8), 2, np.uint8)
    cases["nan"] = (np.full((3, 1, 20, 8), np.nan, np.float32), gp, {})
    good = np.load(os.path.join(HERE, "adr2d_channel4.npz"))["out"]
    cases["good"] = (good, np.full((40, 8), 2, np.uint8), {})
    hot = good.copy()
    hot[-1, 0, 5, 3] = 1.5
    cases["over_cap"] = (hot, np.full((40, 8), 2, np.uint8), {})
    neg = good.copy()
    neg[-1, 1, 7, 2] = -0.01
    cases["negative"] = (neg, np.full((40, 8), 2, np.uint8), {})
    zero = good.copy()
    zero[-1, 1] = 0.0
    cases["zero"] = (zero, np.full((40, 8), 2, np.uint8), {})
    wrap = np.zeros((3, 1, 40, 8), np.float32)
    wrap[1, 0, -1] = 0.5
    cases["wrap"] = (wrap, np.full((40, 8), 2, np.uint8), {})
    res = {}
    arrs = {}
    for k, (o, gm, kw) in cases.items():
        res[k] = r2.check_physics(o, gm, **kw)
        arrs["out_" + k] = o
        arrs["mat_" + k] = gm
    save("check_physics", meta=dict(complaints=res), **arrs)


def _jsonable(d):
    if d is None:
        return None
    out = {}
    for k, v in d.items():
        if isinstance(v, (np.floating, np.integer)):
            v = v.item()
        elif isinstance(v, np.ndarray):
            v = v.tolist()
        if isinstance(v, float) and v != v:
            v = "nan"
        if isinstance(v, list):
            v = ["nan" if (isinstance(x, float) and x != x) else x for x in v]
        out[k] = v
    return out


if __name__ == "__main__":
    main()
