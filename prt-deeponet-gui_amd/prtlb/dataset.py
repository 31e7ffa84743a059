"""A training-data campaign in the reference's dataset layout.

3D/tools/build_dataset_3d.py builds one HDF5 file per campaign: G rocks of
one shape, each with its material grid, geodesic distance from the inlet
(geom/gdf) and pore distance transform (geom/edt); then, per rock, n_sets
transport runs with drawn Pe and Damkohler numbers, stored as float16
snapshots with a per-chemical scale, float16 velocities, the parameters,
snapshot times, the settled flag and the pressure gradient
(:862-902).  build_dataset does the same with every stage on the GPU --
geometry (seeded packs, connectivity), gdf and edt, the probe and real
Stokes solves of all rocks in one fluid-list batch, all transport runs in
one batch, check_physics on the device -- and write_dataset stores the
result under the same group/dataset names.  h5py is not part of this
build, so the container is an .npz archive whose keys are the HDF5 paths
("geom/material", "samples/conc", ...) and whose attributes are stored as
"<path>.attrs.<name>" arrays; numpy.load(..., allow_pickle=False) reads it.

A 2D campaign (shape (nx, ny): disk packs, D2Q9) is stored as
3D/tools/build_dataset_2d.py stores it (:896-937): every grid with a
trailing axis of one voxel (shape attribute (nx, ny, 1)), the dimension
attribute 2, and velocities with their z component zero.
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import campaign as C
from .adr import adr_host_setup
from .batch import AdrBatch, LbmBatch
from .geometry import PORE, disk_pack, edt_device, geodesic_device, sphere_pack
from .physics import check_physics_device

PARAM_NAMES = ("pe", "da", "da_abio")


def _draw_rock(shape, phi, rng, radius, max_tries=50):
    """A percolating seeded sphere pack, or disk pack for a 2D shape
    (3D/tools/build_dataset_3d.py:618-632, build_dataset_2d.py:642-650: draw,
    keep what the inlet reaches, keep only spanning rocks)."""
    pack = disk_pack if len(shape) == 2 else sphere_pack
    for _ in range(max_tries):
        g = pack(shape, radius, phi, seed=int(rng.integers(1, 2 ** 31)), connect=True)
        if C.percolates(g):
            return g
    raise RuntimeError("no percolating pack at porosity %.3f in %d draws" % (phi, max_tries))


def build_dataset(n_geom=4, shape=(48, 48, 48), n_sets=2, n_times=5, porosity=(0.3, 0.5),
                  radius=(3.0, 6.0), pe=(0.1, 100.0), da=(0.01, 10.0),
                  kinetics="sorption_irreversible", seed=0, tau=0.8, tol=1e-6, nit=30000,
                  adr_steps=None, max_steps=100000, storage="fp32", d_ratio=1e-3, progress=None,
                  keep_complaints=False):
    """Run a campaign of n_geom rocks x n_sets transport runs; returns the
    dataset as a dict of arrays keyed like the reference's HDF5 file.

    A run that check_physics complains about is printed as the reference
    prints it, and then the campaign stops with SystemExit instead of
    returning a dataset that looks like a good one
    (3D/tools/build_dataset_3d.py:843-860).  keep_complaints=True returns
    it anyway with the complaints under "_complaints", for inspection;
    write_dataset still refuses it.

    samples/delta_P is the force giving each run's Peclet on its rock
    (pressure_for_peclet) with the sample length as reference length and
    the lattice diffusivity nu_lattice * d_ratio (d_ratio = D / nu of the
    real fluid; 1e-3 for a solute in water, 3D/tools/settings_and_units.py:
    1349-1413)."""
    rng = np.random.default_rng(seed)
    G, S = int(n_geom), int(n_geom) * int(n_sets)
    nd = len(shape)
    if nd not in (2, 3):
        raise ValueError("shape must be (nx, ny) or (nx, ny, nz)")
    t0 = time.perf_counter()
    mats = [_draw_rock(tuple(shape), float(rng.uniform(*porosity)), rng, radius) for _ in range(G)]
    gdf = np.zeros((G,) + tuple(shape), np.float32)
    edt = np.zeros((G,) + tuple(shape), np.float32)
    for i, g in enumerate(mats):
        gdf[i] = geodesic_device(g)[0].cpu().numpy()
        edt[i] = edt_device(g).cpu().numpy()
    if progress:
        progress("dataset geometry %.1f s" % (time.perf_counter() - t0))
    # ---- Stokes: probe then real, every rock in one fluid-list batch
    lb = LbmBatch(19 if nd == 3 else 9, storage=storage)
    force = {}
    vel = {}
    for i, g in enumerate(mats):
        lb.add((i, "probe"), g, force=C.PROBE_FORCE, tau_lb=tau, tol=tol, nit=nit)
    while len(lb):
        for m in lb.tick():
            i, stage = m.key
            v = lb.finish(m, {}).cpu().numpy()
            if stage == "probe":
                force[i] = C.lowered_force(v, 2e-6)
                lb.add((i, "real"), mats[i], force=force[i], tau_lb=tau, tol=tol, nit=nit)
            else:
                vel[i] = v
    if progress:
        progress("dataset flow %.1f s" % (time.perf_counter() - t0))
    # ---- transport: every (rock, set) in one batch
    draws = []
    for i in range(G):
        for j in range(n_sets):
            p = float(np.exp(rng.uniform(np.log(pe[0]), np.log(pe[1]))))
            d = float(np.exp(rng.uniform(np.log(da[0]), np.log(da[1]))))
            draws.append((i, p, d))
    ab = AdrBatch()

    def prep(draw):
        i, p, d = draw
        kw = C._adr_kw(dict(da=d), kinetics)
        return adr_host_setup(mats[i], vel[i], p, kw["da"], kw["n_species"], kw["chem"],
                              kw.get("da_abio", 0.0), "fp32", kw.get("sorption"))

    # the members' host set-ups run on worker threads while earlier ones
    # join; runs join while the batch holds at most C.TRANSPORT_CAP pore
    # voxels (the batch's table and fields then stay in the Infinity Cache)
    n_pore = [int((m == PORE).sum()) for m in mats]
    results = {}
    with ThreadPoolExecutor(max_workers=8) as pool:
        hosts = [pool.submit(prep, dr) for dr in draws]
        k = 0
        while k < len(draws) or len(ab):
            held = sum(m.T.n_pore for m in ab.members.values())
            while k < len(draws) and not (len(ab) and held + n_pore[draws[k][0]] > C.TRANSPORT_CAP):
                i, p, d = draws[k]
                kw = C._adr_kw(dict(da=d), kinetics)
                ab.add(k, mats[i], vel[i], p, kw.pop("da"), n_times, steps=adr_steps,
                       max_steps=max_steps, host=hosts[k].result(), **kw)
                held += n_pore[i]
                k += 1
            for m in ab.advance():
                info = {}
                out, tn = ab.finish(m, info)
                results[m.key] = (out, tn, info)
    if progress:
        progress("dataset transport %.1f s" % (time.perf_counter() - t0))
    kw0 = C._adr_kw(dict(da=1.0), kinetics)
    chem = kw0["chem"]
    ns = kw0["n_species"]
    conc = np.zeros((S, n_times, ns) + tuple(shape), np.float32)
    vels = np.zeros((S, 3) + tuple(shape), np.float32)
    par = np.zeros((S, len(PARAM_NAMES)), np.float32)
    tn_all = np.zeros((S, n_times), np.float32)
    settled = np.zeros(S, bool)
    dP = np.zeros(S, np.float32)
    gi = np.zeros(S, np.int32)
    complaints = {}
    for k, (i, p, d) in enumerate(draws):
        out, tn, info = results[k]
        conc[k], tn_all[k], vels[k], gi[k] = out, tn, vel[i], i
        kwk = C._adr_kw(dict(da=d), kinetics)
        par[k] = (p, kwk["da"], kwk.get("da_abio", 0.0))
        settled[k] = bool(info.get("converged", True))
        u_ax = C.axial_mean(vel[i], mats[i])
        dP[k] = C.pressure_for_peclet(force[i], u_ax, p, shape[0], (tau - 0.5) / 3.0 * d_ratio)
        view, vchem = C.physics_view(out, chem, kinetics)
        bad = check_physics_device(view, mats[i], chem=vchem, da=kwk["da"],
                                   da_abio=kwk.get("da_abio"))
        bad = bad + C.sorbed_complaints(out, mats[i], kinetics)
        for msg in bad:
            print("     PHYSICS run %d (Pe=%.2f Da_bio=%.2f Da_abio=%.2f): %s"
                  % (k, p, kwk["da"], kwk.get("da_abio", 0.0), msg), flush=True)
        if bad:
            complaints[k] = bad
    n_bad = sum(len(v) for v in complaints.values())
    if n_bad and not keep_complaints:
        raise SystemExit("%d physics complaint(s): this file is not usable as training data, "
                         "and writing it would hide that." % n_bad)
    mat = np.stack(mats)
    if nd == 2:
        # the 2D file's grids carry a trailing axis of one voxel
        mat, gdf, edt, conc, vels = (a[..., None] for a in (mat, gdf, edt, conc, vels))
    ds = write_layout(mat, gdf, edt, conc, vels, gi, par, tn_all, settled, dP,
                      species=list(chem.names)[:ns], dimension=nd)
    # not written: the check_physics complaints per run, and the float32
    # flows and draws the runs used
    ds["_complaints"] = complaints
    ds["_flows"] = vel
    ds["_draws"] = draws
    return ds


def write_layout(mat, gdf, edt, conc, vel, gi, par, tn, settled, dP, species, biofilm=None,
                 dimension=None):
    """The reference's file content as a dict: group/dataset paths to arrays,
    attributes as "<path>.attrs.<name>" (3D/tools/build_dataset_3d.py:862-902).
    conc and velocity are stored float16; conc carries the per-chemical
    scale max over samples and snapshots, floored at 1e-6.  dimension
    (default: the grids' rank) is 2 for a 2D campaign stored with a trailing
    axis of one voxel."""
    S, T, Cn = conc.shape[:3]
    shape = conc.shape[3:]
    scale = np.maximum(conc.reshape(S * T * Cn, -1).max(1).reshape(S, T, Cn).max((0, 1)), 1e-6)
    ds = {
        ".attrs.n_samples": np.int64(S),
        ".attrs.shape": np.array(shape, np.int32),
        ".attrs.species": np.array([str(s) for s in species]),
        ".attrs.param_names": np.array(PARAM_NAMES),
        ".attrs.dimension": np.int64(len(shape) if dimension is None else dimension),
        ".attrs.source": np.array("generated_%dd" % (len(shape) if dimension is None else dimension)),
        "geom/gid": np.arange(mat.shape[0], dtype=np.int32),
        "geom/material": np.asarray(mat, np.uint8),
        "geom/gdf": np.asarray(gdf, np.float32),
        "geom/edt": np.asarray(edt, np.float32),
        "samples/geom_index": np.asarray(gi, np.int32),
        "samples/run_id": np.arange(S, dtype=np.int32),
        "samples/params": np.asarray(par, np.float32),
        "samples/t_norm": np.asarray(tn, np.float32),
        "samples/delta_P": np.asarray(dP, np.float32),
        "samples/settled": np.asarray(settled, bool),
        "samples/conc": np.asarray(conc).astype(np.float16),
        "samples/conc.attrs.conc_scale": scale.astype(np.float32),
        "samples/velocity": np.asarray(vel).astype(np.float16),
    }
    if biofilm is not None and np.asarray(biofilm).any():
        ds["geom/biofilm"] = np.asarray(biofilm, np.uint8)
    return ds


def write_dataset(path, ds):
    """Write the layout dict to an .npz archive (no pickled objects).

    Refuses (SystemExit, as 3D/tools/build_dataset_2d.py:889-894) a dataset
    whose runs drew check_physics complaints: the file would look exactly
    like a good one."""
    bad = ds.get("_complaints") or {}
    n_bad = sum(len(v) for v in bad.values())
    if n_bad:
        raise SystemExit("%d physics complaint(s) in %d run(s): this file is not usable as "
                         "training data, so it has NOT been written." % (n_bad, len(bad)))
    np.savez(path, **{k: v for k, v in ds.items() if not k.startswith("_")})
    return path
