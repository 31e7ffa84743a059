"""Campaign driver: what build_dataset_2d/3d.py do around the solver.

Per rock: a probe Stokes solve at force 1e-6, the force lowered (never
raised) so the peak speed sits at Mach 0.02, the real Stokes solve, the
axial pore mean speed; per (rock, Pe, Da): a transport solve.  Restates the
caller logic of 3D/tools/build_dataset_3d.py:685-726 and
3D/tools/build_dataset_2d.py:715-759 (the HDF5 writing is replaced by
returned arrays; h5py is not a dependency here).

Two executions of the same campaign:
  run_sequential  one solve at a time, stokes_d2q9/3q19 + solve_adr
  run_batched     every rock's Stokes in one LbmBatch, every transport in
                  one AdrBatch, members admitted under an HBM budget
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import chemistry as chemmod
from .adr import adr_host_setup, solve_adr
from .batch import AdrBatch, LbmBatch, adr_footprint, lbm_footprint
from .geometry import PORE, disk_pack, sphere_pack
from .stokes import stokes_d2q9, stokes_d3q19

MACH_TARGET = 0.02
PROBE_FORCE = 1e-6


def lowered_force(probe_v, seed_force):
    """The reference's probe rule (3D/tools/build_dataset_3d.py:689-712):
    lower the seed force so the probe's peak speed would sit at Mach 0.02;
    never raise it."""
    peak = float(np.sqrt((np.asarray(probe_v, np.float64) ** 2).sum(0)).max())
    if peak <= 0:
        return float(seed_force)
    want = PROBE_FORCE * (MACH_TARGET / np.sqrt(3.0)) / peak
    return float(want) if want < seed_force else float(seed_force)


def axial_mean(v, g):
    """Axial pore-mean speed, falling back to the mean speed for a closed
    or recirculating field (3D/tools/build_dataset_3d.py:720-724)."""
    v = np.asarray(v)
    m = np.asarray(g) == PORE
    sp = np.sqrt((v[:3] ** 2).sum(0))
    ax = float(v[0][m].mean())
    spd = float(sp[m].mean())
    return ax if ax > 0.1 * spd else spd


def peclet_from_flow(u_lattice, l_ref_voxels, d_lattice):
    """Pe = u L / D in lattice units (3D/tools/settings_and_units.py:1349-1383)."""
    return float(u_lattice) * float(l_ref_voxels) / max(float(d_lattice), 1e-30)


def pressure_for_peclet(delta_p, u_lattice, peclet, l_ref_voxels, d_lattice):
    """The pressure gradient giving a wanted Peclet on this rock: Stokes flow
    is linear, so seed * wanted / seed Peclet (settings_and_units.py:1396-1413)."""
    seed = peclet_from_flow(u_lattice, l_ref_voxels, d_lattice)
    return 0.0 if seed <= 0 else float(delta_p) * float(peclet) / seed


def campaign_specs(n_geom=64, edge=(48, 160), radius=(3.0, 8.0), porosity=(0.25, 0.5),
                   pe=(0.1, 100.0), da=(0.01, 10.0), seed=0, nd=3):
    """Seeded rocks and parameter sets of the batched-campaign workload:
    cubes with edges U{lo..hi}, radii U(lo, hi), porosity U(lo, hi), Pe and Da
    log-uniform."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_geom):
        e = int(rng.integers(edge[0], edge[1] + 1))
        specs.append(dict(
            key=i, shape=(e,) * nd, radius=tuple(radius),
            porosity=float(rng.uniform(*porosity)), geom_seed=int(rng.integers(2 ** 31)),
            pe=float(np.exp(rng.uniform(np.log(pe[0]), np.log(pe[1])))),
            da=float(np.exp(rng.uniform(np.log(da[0]), np.log(da[1]))))))
    return specs


def percolates(g):
    """True when the pore space connects the x = 0 face to the x = nx-1
    face (face connectivity), i.e. a body force along x can drive flow."""
    from scipy import ndimage
    pore = np.asarray(g) == PORE
    lab, _ = ndimage.label(pore)
    a = set(np.unique(lab[0][pore[0]]).tolist()) - {0}
    b = set(np.unique(lab[-1][pore[-1]]).tolist()) - {0}
    return bool(a & b)


def make_geometry(spec, max_tries=50):
    """The spec's pack, redrawn with the next seed until it percolates along
    x (the reference's builders likewise keep only spanning rocks,
    3D/tools/build_dataset_2d.py:285 keep_spanning_cluster / percolates)."""
    seed = spec["geom_seed"]
    for _ in range(max_tries):
        if len(spec["shape"]) == 2:
            g = disk_pack(spec["shape"], spec["radius"], spec["porosity"], seed=seed)
        else:
            g = sphere_pack(spec["shape"], spec["radius"], spec["porosity"], seed=seed)
        if percolates(g):
            return g
        seed += 1
    raise RuntimeError("no percolating pack for %r in %d draws" % (spec, max_tries))


def _adr_kw(spec, kinetics):
    if kinetics == "sorption_irreversible":
        return dict(chem=chemmod.sorption_irreversible(True), n_species=2, da=0.0,
                    da_abio=spec["da"])
    if kinetics == "sorption_reversible":
        return dict(chem=chemmod.sorption_reversible(True), n_species=2, da=0.0,
                    sorption=dict(da_sorb=spec["da"], da_desorb=0.5 * spec["da"]))
    if kinetics == "monod":
        return dict(chem=chemmod.monod(0.5), n_species=1, da=spec["da"])
    return dict(chem=chemmod.first_order(True), n_species=1, da=0.0, da_abio=spec["da"])


def physics_view(out, chem, kinetics):
    """(snapshots, chemistry) to hand check_physics for a run of `kinetics`.

    check_physics gives each index its role (3D/tools/prtlb_3d.py:1684-1731):
    0 the donor, capped at its feed; 1 the acceptor, capped at its own feed;
    2 the product, capped at its feed plus the donor's; 3 the biomass, held
    below a thousand times its start.  None of these bounds the sorbed S of
    the sorption kinetics (build-defined, index 1): an immobile chemical
    that keeps what the passing solute leaves behind, so irreversibly it
    grows with time without bound (at Da 740 it reaches hundreds of times
    the feed) and reversibly it tends to k_sorb / k_des times the solute.
    The view therefore holds the mobile solute alone, under every rule of
    the donor; sorbed_complaints checks S."""
    if not kinetics.startswith("sorption"):
        return out, chem
    return np.asarray(out)[:, :1], chem.truncate(1)


def sorbed_complaints(out, mat, kinetics):
    """The rules that do hold for the sorbed S of a sorption run (index 1
    of the snapshots), in check_physics' wording (3D/tools/prtlb_3d.py:
    1717-1737): finite, and not negative in the pore.  An empty list for
    other kinetics."""
    if not kinetics.startswith("sorption"):
        return []
    s = np.asarray(out)[:, 1][:, np.asarray(mat) == PORE]
    bad = []
    fin = np.isfinite(s)
    if not fin.all():
        bad.append("S is not finite: %d NaN and %d infinite values -- the integration "
                   "diverged" % (int(np.isnan(s).sum()), int(np.isinf(s).sum())))
    elif s.size and float(s.min()) < -1e-6:
        bad.append("S reaches %.3g, which is negative" % float(s.min()))
    return bad


def run_sequential(specs, geoms, storage="fp64", tau=0.8, tol=1e-6, nit=30000,
                   kinetics="sorption_irreversible", n_t=5, adr_steps=None, max_steps=100000,
                   adr_precision="fp32", progress=None):
    """One solve after another (the reference's own schedule, on the GPU).

    Returns (results by key, seconds) with seconds = {"total", "flow",
    "transport"}: the Stokes solves (probe + real, every rock) and the
    transport solves timed as stages, as run_batched does.  progress, if
    given, is called with a short text after every finished solve."""
    res = {}
    t0 = time.perf_counter()
    flows = {}
    for sp in specs:
        g = geoms[sp["key"]]
        fn = stokes_d2q9 if g.ndim == 2 else stokes_d3q19
        pinfo = {}
        probe = fn(g, nit=nit, force=PROBE_FORCE, tau_lb=tau, tol=tol, info=pinfo, storage=storage)
        force = lowered_force(probe, 2e-6)
        info = {}
        flows[sp["key"]] = fn(g, nit=nit, force=force, tau_lb=tau, tol=tol, info=info,
                              storage=storage)
        res[sp["key"]] = dict(permeability=info["permeability"], flow_iters=info["flow_iters"],
                              probe_iters=pinfo["flow_iters"], force=force)
        if progress:
            progress("flow %s done (%d its)" % (sp["key"], info["flow_iters"]))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for sp in specs:
        g = geoms[sp["key"]]
        ainfo = {}
        kw = _adr_kw(sp, kinetics)
        out, tn = solve_adr(g, flows[sp["key"]], sp["pe"], kw.pop("da"), n_t, steps=adr_steps,
                            max_steps=max_steps, info=ainfo, precision=adr_precision, **kw)
        res[sp["key"]].update(adr_steps=ainfo["steps"], converged=ainfo["converged"],
                              outlet=np.asarray(ainfo["outlet_concentration"]).tolist(),
                              out_last=out[-1])
        if progress:
            progress("transport %s done (%d steps)" % (sp["key"], ainfo["steps"]))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return res, dict(total=t2 - t0, flow=t1 - t0, transport=t2 - t1)


# Pore voxels in a campaign's transport batch at once: up to about this
# many, a batch's pore table (32 bytes per voxel in float32) and fields stay
# in the 256 MB Infinity Cache from step to step; a 64-rock campaign's
# transport stage runs 15% faster in such groups than as one batch.
TRANSPORT_CAP = 4e6


def run_batched(specs, geoms, storage="fp32", tau=0.8, tol=1e-6, nit=30000,
                kinetics="sorption_irreversible", n_t=5, adr_steps=None, max_steps=100000,
                hbm_budget=200e9, progress=None, transport_cap=TRANSPORT_CAP, flow_cap=None):
    """Every Stokes solve in one LbmBatch per lattice, every transport solve
    in one AdrBatch; members admitted while they fit the HBM budget and, for
    transport, while the batch holds at most transport_cap pore voxels (None:
    no cap; a run larger than the cap runs alone), in spec order.
    Returns (results by key, seconds) as run_sequential."""
    res = {sp["key"]: {} for sp in specs}
    t0 = time.perf_counter()
    nd = len(specs[0]["shape"])
    q = 9 if nd == 2 else 19
    # ---- Stokes: probe stage then real stage, one batch
    lb = LbmBatch(q, storage=storage)
    t_note = time.perf_counter()
    pending = [(sp["key"], "probe") for sp in specs]
    n_fluid = {sp["key"]: int((geoms[sp["key"]] == PORE).sum()) for sp in specs}
    velocities = {}
    while pending or len(lb):
        while pending:
            key, stage = pending[0]
            g = geoms[key]
            need = lbm_footprint(g.shape, storage, n_fluid[key] if lb.sparse else None)
            if len(lb) and lb.bytes() + need > hbm_budget:
                break
            if flow_cap and len(lb) and (sum(m.nodes for m in lb.members.values())
                                         + n_fluid[key] > flow_cap):
                break
            pending.pop(0)
            force = PROBE_FORCE if stage == "probe" else res[key]["force"]
            m = lb.add((key, stage), g, force=force, tau_lb=tau, tol=tol, nit=nit)
            m.stage = stage
        finished = lb.tick()
        if progress and time.perf_counter() - t_note > 20.0:
            t_note = time.perf_counter()
            progress("flow batch: %d members running, %d waiting" % (len(lb), len(pending)))
        for m in finished:
            key, stage = m.key
            info = {}
            v = lb.finish(m, info).cpu().numpy()
            if stage == "probe":
                res[key]["force"] = lowered_force(v, 2e-6)
                res[key]["probe_iters"] = m.conv.iters
                pending.append((key, "real"))
            else:
                res[key]["permeability"] = info["permeability"]
                res[key]["flow_iters"] = m.conv.iters
                velocities[key] = v
                if progress:
                    progress("flow %s done (%d its)" % (key, m.conv.iters))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # ---- transport: one batch; the members' host set-ups (NumPy over each
    # whole grid) run on worker threads, in order, while members join
    ab = AdrBatch()
    queue = list(specs)
    n_pore = {sp["key"]: int((geoms[sp["key"]] == PORE).sum()) for sp in specs}
    pool = ThreadPoolExecutor(max_workers=8)

    def prep(sp):
        kw = _adr_kw(sp, kinetics)
        return adr_host_setup(geoms[sp["key"]], velocities[sp["key"]], sp["pe"], kw["da"],
                              kw["n_species"], kw["chem"], kw.get("da_abio", 0.0), "fp32",
                              kw.get("sorption"))

    try:
        hosts = {sp["key"]: pool.submit(prep, sp) for sp in queue}
        while queue or len(ab):
            while queue:
                sp = queue[0]
                g = geoms[sp["key"]]
                kw = _adr_kw(sp, kinetics)
                if len(ab) and ab.bytes() + adr_footprint(g.shape, kw["n_species"], g.ndim) > hbm_budget:
                    break
                if transport_cap and len(ab) and (sum(m.T.n_pore for m in ab.members.values())
                                                  + n_pore[sp["key"]] > transport_cap):
                    break
                queue.pop(0)
                ab.add(sp["key"], g, velocities[sp["key"]], sp["pe"], kw.pop("da"), n_t,
                       steps=adr_steps, max_steps=max_steps, host=hosts.pop(sp["key"]).result(), **kw)
            finished = ab.advance()
            if progress and time.perf_counter() - t_note > 20.0:
                t_note = time.perf_counter()
                progress("transport batch: %d members running, %d waiting" % (len(ab), len(queue)))
            for m in finished:
                ainfo = {}
                out, tn = ab.finish(m, ainfo)
                r = res[m.key]
                r.update(adr_steps=ainfo["steps"], converged=ainfo["converged"],
                         outlet=np.asarray(ainfo["outlet_concentration"]).tolist(),
                         out_last=out[-1])
                if progress:
                    progress("transport %s done (%d steps)" % (m.key, ainfo["steps"]))
    finally:
        pool.shutdown(wait=False, cancel_futures=True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return res, dict(total=t2 - t0, flow=t1 - t0, transport=t2 - t1)
