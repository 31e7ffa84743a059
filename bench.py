"""Benchmark of the PRT-LB hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one fused D3Q19 TRT iteration (pull-stream + collide + Guo force +
solid reversal) over the whole 256^3 sphere pack (radii U(4,8), porosity
~0.35, seed 0), FP64 populations: the precision-variant workload's FP64
case, where the 70% HBM target is stated.  Populations, mask and partial-sum
buffers are resident in HBM before the timed region; K steps are timed with
HIP events on the stream the kernels run on.  `value` is box-node MLUPS; the
roofline object prices the 304 B/node algorithmic traffic against 8 TB/s.
Extra keys report the FP32 / FP16-shifted storages, the 128^3 drop-in grid,
the fluid lists, the fused transport step, the batched 64-rock campaign
(solves/hour batched FP32 against one-at-a-time FP64, permeability error),
the 32-request serving run (solves/hour, p50/p99 latency, converged share)
and the CPU baseline (this repo's NumPy oracle on the host, one core, on a
bounded sample of the same 256^3 step), with the host's per-core rates and
its solves/hour for the 64-rock campaign estimated from them.  Prints ONE
JSON line.  Studies run as separate invocations (--precision, --campaign N,
--serve N [--serve-baseline], --adr-only, --fluid-list, --shapes, --tail,
--slab, --big N, --geometry, --features).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "prt-deeponet-gui_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_HBM = 8.0e12          # MI355X HBM3E, bytes/s
BYTES_PER_LU = {("fp64", 19): 304, ("fp32", 19): 152, ("fp16", 19): 76,
                ("fp64", 9): 144, ("fp32", 9): 72, ("fp16", 9): 36}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_lbm(g, storage, steps, warmup, skip_solid=False, tune=True):
    import torch
    import prtlb
    from prtlb import autotune
    s = prtlb.StokesSolver(g, force=2e-6, storage=storage, skip_solid=skip_solid)
    shape = autotune.tune(s, steps=5)["tune"] if tune else 0
    s.run(warmup)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    s.run(steps)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    nodes = s.n_fluid if skip_solid else s.n
    mlups = nodes / (ms * 1e-3) / 1e6
    bpl = BYTES_PER_LU[(storage, s.q)]
    gbs = nodes * bpl / (ms * 1e-3) / 1e9
    out = dict(ms_per_step=ms, mlups=mlups, gbs=gbs, frac=gbs * 1e9 / PEAK_HBM,
               nodes=nodes, fluid_fraction=s.n_fluid / s.n, bytes_per_lu=bpl,
               mlups_fluid=s.n_fluid / (ms * 1e-3) / 1e6, launch_shape=shape)
    del s
    torch.cuda.empty_cache()
    return out


def config1_pipeline():
    """The CPU-runnable workload end to end through the drop-in API (NumPy in,
    NumPy out, so PCIe transfers included): 128x128 disk pack (r 6,
    porosity ~0.6), D2Q9 TRT Stokes in FP64 to 1e-8, then 5000 steps of
    transport with a first-order surface reaction (Pe 10, Da 1)."""
    import torch
    import prtlb
    g = prtlb.disk_pack((128, 128), 6.0, 0.6, seed=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = {}
    v = prtlb.stokes_d2q9(g, tol=1e-8, info=info)
    t1 = time.perf_counter()
    ainfo = {}
    out, tn = prtlb.solve_adr(g, v, 10.0, 0.0, 5, n_species=1, steps=5000, info=ainfo,
                              chem=prtlb.chemistry.first_order(True), da_abio=1.0)
    t2 = time.perf_counter()
    return dict(flow_s=round(t1 - t0, 3), flow_iters=info["flow_iters"],
                permeability=info["permeability"], transport_s=round(t2 - t1, 3),
                transport_steps=ainfo["steps"], total_s=round(t2 - t0, 3))


def time_fluid_list(g, storage, steps, warmup, link16=True):
    """One geometry as a fluid-list member (compact populations + link
    table): time per step, fluid-node updates per second, and bandwidth
    counting the populations (2 Q s per fluid node) and the link table
    (2 or 4 bytes per direction) it reads."""
    import torch
    from prtlb.batch import LbmBatch
    b = LbmBatch(19, storage=storage, link16=link16)
    b.add(0, g, force=1e-6)
    b.step(warmup)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    b.step(steps)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    m = b.members[0]
    nf = m.n_fluid
    sb = {"fp64": 8, "fp32": 4, "fp16": 2}[storage]
    pop_b, link_b = 2 * 19 * sb, m.link_bits // 8 * 18
    gbs = nf * (pop_b + link_b) / (ms * 1e-3) / 1e9
    return dict(ms_per_step=ms, mlups_fluid=nf / (ms * 1e-3) / 1e6, fluid_nodes=nf,
                link_bits=m.link_bits, bytes_per_fluid_node=pop_b + link_b, gbs=gbs,
                frac=gbs * 1e9 / PEAK_HBM)


def time_adr(g, steps, warmup, kinetics="monod", pore_list=True, precision="fp32", v=None):
    import torch
    import prtlb
    if v is None:
        v = prtlb.stokes_d3q19(g, nit=1500, force=2e-6, tol=1e-5, storage="fp32")
    if kinetics == "monod":
        T = prtlb.TransportSolver(g, v, 10.0, 1.0, 1, chem=prtlb.chemistry.monod(0.5),
                                  precision=precision)
    else:
        T = prtlb.TransportSolver(g, v, 10.0, 0.0, 2, chem=prtlb.chemistry.sorption_irreversible(),
                                  da_abio=1.0, precision=precision)
    T.use_graph = False
    T.pore_list = pore_list
    T.step_n(warmup)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    T.step_n(steps)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    vps = T.n_pore / (ms * 1e-3)
    cb = 4 if precision == "fp32" else 8
    bpv = 2 * T.ns * cb + 3 * cb + 1     # 21 B per pore voxel-step for ns = 1 (3D, f32)
    return dict(ms_per_step=ms, pore_voxel_steps_per_s=vps, gbs=vps * bpv / 1e9,
                frac=vps * bpv / PEAK_HBM, bytes_per_voxel=bpv, pore_voxels=T.n_pore)


def measured_traffic(key, shape=None):
    """HBM bytes per launch of this workload from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/), for the launch shape the
    tuner picked (entries measured per shape), or None if not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        entry = json.load(open(path))[key]
    except (OSError, ValueError, KeyError):
        return None
    by_shape = entry.get("by_shape", {})
    if shape is not None and str(shape) in by_shape:
        return by_shape[str(shape)]
    if shape is None or entry.get("launch_shape") == shape:
        return entry.get("traffic_bytes")
    return None


def stream_copy(nbytes=2 * 1024 ** 3, reps=10):
    """Measured device-copy bandwidth (read + write bytes / time) of each
    copy variant of the library: the practical HBM ceiling."""
    import torch
    from prtlb import _lib
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    s = torch.cuda.current_stream().cuda_stream
    out = {}
    for var in (0, 1, 2):
        for _ in range(2):
            _lib.call("prt_stream_copy", a.data_ptr(), b.data_ptr(), nbytes, var, s)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            _lib.call("prt_stream_copy", a.data_ptr(), b.data_ptr(), nbytes, var, s)
        e1.record()
        e1.synchronize()
        out["variant%d_GBs" % var] = round(2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)
    out["best_GBs"] = max(out.values())
    return out


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return dict(model=model, logical_cpus=os.cpu_count(), usable_cpus=usable)


def cpu_baseline(E=256, nit=2):
    """The NumPy oracle (a restatement of the reference's vectorised NumPy
    loop, np.roll streaming) on a bounded sample of the headline workload:
    the D3Q19 FP64 step on the same E^3 sphere pack, one core.  The set-up
    and final moments are timed apart (a 1-iteration run) and taken out."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import prtlb
    from oracle import prtlb_oracle as O
    g = prtlb.sphere_pack((E, E, E), (4.0, 8.0), 0.35, seed=0)
    t0 = time.perf_counter()
    O.stokes(g, nit=1, force=2e-6, tol=0.0)
    t1 = time.perf_counter()
    O.stokes(g, nit=1 + nit, force=2e-6, tol=0.0)
    t2 = time.perf_counter()
    per = max((t2 - t1) - (t1 - t0), 1e-9) / nit
    cpu = host_cpu()
    return dict(value=g.size / per / 1e6, unit="MLUPS", cores=1, kind="port",
                sample="NumPy oracle D3Q19 FP64 TRT step, %d^3 sphere pack (the headline workload), "
                       "%d iterations timed as the difference of a %d- and a 1-iteration run "
                       "(%.1f s per iteration)" % (E, nit, 1 + nit, per),
                cpu_model=cpu["model"], host_logical_cpus=cpu["logical_cpus"],
                host_usable_cpus=cpu["usable_cpus"])


def _oracle_stokes_worker(args):
    """One host process of the whole-host figure: the NumPy oracle's D3Q19
    step on the given grid for nit iterations; returns seconds."""
    g, nit = args
    os.environ["OMP_NUM_THREADS"] = "1"
    sys.path.insert(0, ROOT)
    from oracle import prtlb_oracle as O
    t0 = time.perf_counter()
    O.stokes(g, nit=nit, force=2e-6, tol=0.0)
    return time.perf_counter() - t0


def cpu_campaign_estimate(campaign, cpu):
    """Converged solves per hour of the 64-rock campaign on this host, one
    process per core, from the campaign's own work (every Stokes node
    iteration of the probe and real solves over the dense box, every
    transport pore-voxel step) at the host rates measured here: the
    whole-host D3Q19 rate and the one-core sorption transport rate times the
    processes.  An estimate from measured rates, not a timed campaign (that
    would take the host about the time printed)."""
    w = campaign["work"]
    procs = cpu["all_cores_procs"]
    stokes_s = w["stokes_node_iterations"] / (cpu["d3q19_64cube_mlups_all_cores"] * 1e6)
    adr_s = w["transport_pore_voxel_steps"] / (
        cpu["adr_3d_64cube_sorption_pore_voxel_steps_per_s_1core"] * procs)
    total = stokes_s + adr_s
    gpu = campaign["stages"]["total"]
    return {"kind": "estimate from measured host rates", "procs": procs,
            "stokes_s": round(stokes_s, 1), "transport_s": round(adr_s, 1),
            "total_h": round(total / 3600, 2),
            "solves_per_hour": round(campaign["n_solves"] / total * 3600, 2),
            "gpu_batched_over_host": round(gpu["batched_solves_per_hour"] * total
                                           / (campaign["n_solves"] * 3600), 1)}


def cpu_extras(procs=None):
    """The other host figures of the CPU comparison, each a few seconds:
    the NumPy oracle's D2Q9 step (128^2 disk pack, 200 iterations), its
    transport step in 2D (128^2, 200 steps) and 3D (64^3, 200 steps), one
    core each; and the whole-host rate of the D3Q19 step with one process
    per core (64^3, 20 iterations each), the campaign-style use of a host."""
    import multiprocessing as mp
    import numpy as np
    import prtlb
    from oracle import prtlb_oracle as O
    out = {}
    g2 = prtlb.disk_pack((128, 128), 6.0, 0.6, seed=0)
    t0 = time.perf_counter()
    v2 = O.stokes(g2, nit=200, force=2e-6, tol=0.0)
    dt = time.perf_counter() - t0
    out["d2q9_128sq_mlups_1core"] = round(g2.size * 200 / dt / 1e6, 3)
    t0 = time.perf_counter()
    O.solve_adr(g2, v2, 10.0, 0.0, 3, n_species=1, steps=200,
                chem=O.Chemistry(names=["C"], inlet=[1.0], c_init=[0.0], d_rel=[1.0],
                                 immobile=[False], biotic_on=False, abiotic_index=0,
                                 abiotic_surface_only=True), da_abio=1.0)
    dt = time.perf_counter() - t0
    out["adr_2d_128sq_pore_voxel_steps_per_s_1core"] = round(int((g2 == 2).sum()) * 200 / dt, 1)
    g3 = prtlb.sphere_pack((64, 64, 64), (4.0, 8.0), 0.35, seed=0)
    t0 = time.perf_counter()
    v3 = O.stokes(g3, nit=200, force=2e-6, tol=0.0)
    dt = time.perf_counter() - t0
    out["d3q19_64cube_200it_mlups_1core"] = round(g3.size * 200 / dt / 1e6, 3)
    t0 = time.perf_counter()
    O.solve_adr(g3, v3, 10.0, 1.0, 3, n_species=1, steps=200,
                chem=O.Chemistry(names=["C"], inlet=[1.0], c_init=[0.0], d_rel=[1.0],
                                 immobile=[False], ks_donor=0.5, biotic_on=True))
    dt = time.perf_counter() - t0
    out["adr_3d_64cube_pore_voxel_steps_per_s_1core"] = round(int((g3 == 2).sum()) * 200 / dt, 1)
    # the campaign's kinetics: irreversible sorption (a mobile and an
    # immobile chemical, abiotic rate on the surface)
    t0 = time.perf_counter()
    O.solve_adr(g3, v3, 10.0, 0.0, 3, n_species=2, steps=200, da_abio=1.0,
                chem=O.Chemistry(names=["C", "S"], inlet=[1.0, np.nan], c_init=[0.0, 0.0],
                                 d_rel=[1.0, 0.0], immobile=[False, True], biotic_on=False,
                                 abiotic_index=0, abiotic_product_index=1,
                                 abiotic_surface_only=True))
    dt = time.perf_counter() - t0
    out["adr_3d_64cube_sorption_pore_voxel_steps_per_s_1core"] = round(int((g3 == 2).sum()) * 200
                                                                       / dt, 1)
    procs = procs or min(16, host_cpu()["usable_cpus"] or 1)
    ctx = mp.get_context("spawn")            # fresh interpreters: no GPU state inherited
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        secs = pool.map(_oracle_stokes_worker, [(g3, 20)] * procs)
    wall = time.perf_counter() - t0
    out["d3q19_64cube_mlups_all_cores"] = round(procs * g3.size * 20 / max(secs) / 1e6, 2)
    out["all_cores_procs"] = procs
    out["all_cores_wall_s"] = round(wall, 2)
    return out


def precision_study(cases=((256, 0.25), (256, 0.45), (384, 0.25), (384, 0.45)),
                    nit=20000, tol=1e-8, adr_steps=2000):
    """Storage-precision errors against the FP64 GPU solve on seeded sphere
    packs: velocity (max |dv| / max |v|), Darcy permeability (relative) and,
    after adr_steps of Monod transport (Pe 10, Da 1, Ks 0.5), concentration
    (max |dC|) against the FP64 transport on the FP64 flow."""
    import numpy as np
    import torch
    import prtlb
    out = []
    for E, phi in cases:
        g = prtlb.sphere_pack((E, E, E), (4.0, 8.0), phi, seed=0)
        ref = {}
        res = {"grid": E, "porosity_target": phi, "porosity": float((g == prtlb.PORE).mean())}
        for st in ("fp64", "fp32", "fp16"):
            S = prtlb.StokesSolver(g, force=2e-6, storage=st)
            t0 = time.perf_counter()
            conv = S.solve(nit=nit, tol=tol)
            torch.cuda.synchronize()
            v32, v64, sux = S.velocity(f64=True)
            k = S.permeability(sux)
            r = {"iters": conv.iters, "converged": bool(conv.converged),
                 "solve_s": round(time.perf_counter() - t0, 2), "permeability": k}
            v64 = v64.cpu().numpy()
            del S, v32
            torch.cuda.empty_cache()
            if st == "fp64":
                ref = {"v": v64, "k": k}
            else:
                r["vel_err"] = float(np.abs(v64 - ref["v"]).max() / np.abs(ref["v"]).max())
                r["perm_rel_err"] = abs(k - ref["k"]) / abs(ref["k"])
            vel = v64.astype(np.float32)
            prec = "fp64" if st == "fp64" else "fp32"
            T = prtlb.TransportSolver(g, v64 if st == "fp64" else vel, 10.0, 1.0, 1,
                                      chem=prtlb.chemistry.monod(0.5), precision=prec)
            T.step_n(adr_steps)
            c = T.current().double().cpu().numpy()
            del T
            torch.cuda.empty_cache()
            if st == "fp64":
                ref["c"] = c
            else:
                r["conc_err"] = float(np.abs(c - ref["c"]).max())
            res[st] = r
            log("precision", E, phi, st, {kk: vv for kk, vv in r.items()})
        out.append(res)
    return out


def campaign_study(n_geom=64, edge=(48, 160), seed=0, adr_steps=None, max_steps=100000,
                   adr_precision="fp64", batched_only=False, transport_cap=4e6, flow_cap=None,
                   storage="fp32"):
    """The batched campaign: n_geom seeded sphere packs (edges U{48..160},
    radii U(3,8), porosity U(0.25,0.5)), Pe log-U(0.1,100), Da log-U(0.01,10),
    irreversible sorption.  Runs it twice: FP32 batched (one launch per
    iteration for all rocks' fluid nodes, then one for all transport solves'
    pore voxels) and FP64 one solve at a time over the dense box (the
    reference's schedule and layout).  Reports, per stage and in total,
    solves per hour of each schedule and the permeability error of the FP32
    batched result against the FP64 one."""
    import numpy as np
    from prtlb import campaign as C
    specs = C.campaign_specs(n_geom=n_geom, edge=edge, seed=seed)
    t0 = time.perf_counter()
    geoms = {s["key"]: C.make_geometry(s) for s in specs}
    log("campaign geometry %.1f s" % (time.perf_counter() - t0))
    from prtlb import autotune
    rb, tb = C.run_batched(specs, geoms, storage=storage, adr_steps=adr_steps, max_steps=max_steps,
                           progress=log, transport_cap=transport_cap, flow_cap=flow_cap)
    log("campaign batched %s %r" % (storage, tb))
    # the launch shapes the batches' tuner picked (PRTLB_BATCH_TUNE=0: none)
    shapes = {k: {"picked": v["name"], "ms": {n: round(t, 4) for n, t in v["table"].items()}}
              for k, v in autotune._CACHE.items() if k.startswith("batch_")}
    if batched_only:
        n = len(specs)
        return {"n_solves": n, "transport_cap": transport_cap, "flow_cap": flow_cap,
                "batched_fp32_s": {k: round(v, 2) for k, v in tb.items()},
                "batched_solves_per_hour": round(n / tb["total"] * 3600, 1),
                "batch_tuning": os.environ.get("PRTLB_BATCH_TUNE", "1") != "0",
                "batch_shapes": shapes}
    rs, ts = C.run_sequential(specs, geoms, storage="fp64", adr_steps=adr_steps,
                              max_steps=max_steps, adr_precision=adr_precision, progress=log)
    log("campaign sequential fp64 %r" % ts)
    perr = [abs(rb[k]["permeability"] - rs[k]["permeability"]) / abs(rs[k]["permeability"])
            for k in rb]
    n = len(specs)
    nodes = sum(int(np.prod(s["shape"])) for s in specs)
    pore = sum(int((geoms[s["key"]] == 2).sum()) for s in specs)
    stages = {}
    for st in ("flow", "transport", "total"):
        stages[st] = {"batched_fp32_s": round(tb[st], 2), "sequential_fp64_s": round(ts[st], 2),
                      "batched_solves_per_hour": round(n / tb[st] * 3600, 1),
                      "sequential_solves_per_hour": round(n / ts[st] * 3600, 1),
                      "speedup": round(ts[st] / tb[st], 3)}
    # the sequential schedule's work, for the host estimate (cpu_campaign_estimate)
    work = {"stokes_node_iterations": sum(
                (rs[s["key"]]["probe_iters"] + rs[s["key"]]["flow_iters"]) * int(np.prod(s["shape"]))
                for s in specs),
            "transport_pore_voxel_steps": sum(
                rs[s["key"]]["adr_steps"] * int((geoms[s["key"]] == 2).sum()) for s in specs)}
    return {"n_solves": n, "box_nodes": nodes, "pore_nodes": pore, "transport_cap": transport_cap,
            "batched_storage": storage, "work": work,
            "transport_precision_sequential": adr_precision, "adr_max_steps": max_steps,
            "stages": stages,
            "max_perm_rel_err": max(perr), "median_perm_rel_err": float(np.median(perr)),
            "flow_converged_batched": sum(1 for k in rb if rb[k]["flow_iters"] < 30000),
            "flow_converged_sequential": sum(1 for k in rs if rs[k]["flow_iters"] < 30000),
            "transport_settled_batched": sum(1 for k in rb if rb[k]["converged"]),
            "transport_settled_sequential": sum(1 for k in rs if rs[k]["converged"]),
            "adr_steps_mean": float(np.mean([rs[k]["adr_steps"] for k in rs])),
            "batch_shapes": shapes}


def slab_study(E=256, parts=8, steps=50, warmup=5):
    """One GPU's share of an E^3 solve split along x over `parts` GPUs, at
    its real shape: the slab step (E/parts own planes between two halo
    planes, PRT_LBM_SLAB) and its halo pack and unpack kernels, timed with
    HIP events, beside the whole box's step.  The exchange itself would run
    over xGMI between GPUs; its time is estimated from the message size at
    one link's rate."""
    import torch
    import prtlb
    from prtlb.slab import SlabSet
    from prtlb.stokes import StokesSolver
    g = prtlb.sphere_pack((E, E, E), (4.0, 8.0), 0.35, seed=0)
    out = {"grid": E, "parts": parts}

    def timed(fn, n):
        st = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        for _ in range(warmup):
            fn()
        e0.record(st)
        for _ in range(n):
            fn()
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / n

    for storage, pop in (("fp64", 8), ("fp32", 4)):
        T = SlabSet(g, parts, storage=storage)
        T.exchange(0)
        own = T.slabs[0]["m"] * E * E
        ms = timed(lambda: T.step_slab(0, 0, 1), steps)
        sl = T.slabs[0]

        def halo():
            s = prtlb.device.stream_ptr(T.dev)
            f = prtlb.device.ptr(sl["f"][0])
            for plane, d in ((sl["m"], 1), (1, -1)):
                prtlb._lib.call("prt_lbm_halo_pack", T.q, T.st, *sl["dims"], f, plane, d,
                                prtlb.device.ptr(sl["send"][d]), s)
            for plane, d in ((0, 1), (sl["m"] + 1, -1)):
                prtlb._lib.call("prt_lbm_halo_unpack", T.q, T.st, *sl["dims"], f, plane, d,
                                prtlb.device.ptr(sl["send"][d]), s)
        ms_halo = timed(halo, steps)
        msg, slab_grid = sl["send"][1].numel() * pop, [sl["dims"][0], E, E]
        del T, sl
        torch.cuda.empty_cache()
        S = StokesSolver(g, storage=storage)
        ms_box = timed(lambda: S.run(1), steps)
        del S
        torch.cuda.empty_cache()
        gbs = own * 2 * 19 * pop / (ms * 1e-3) / 1e9
        out[storage] = {"slab_grid": slab_grid, "own_nodes": own,
                        "slab_step_ms": round(ms, 4), "slab_GBs": round(gbs, 1),
                        "slab_frac": round(gbs * 1e9 / PEAK_HBM, 4),
                        "halo_pack_unpack_ms": round(ms_halo, 4),
                        "message_bytes": msg,
                        "xgmi_one_message_ms_at_153GBs": round(msg / 153e9 * 1e3, 4),
                        "box_step_ms": round(ms_box, 4),
                        "box_step_over_parts_ms": round(ms_box / parts, 4)}
        log("slab", storage, out[storage])
    return out


def big_study(E=1024, storage="fp32", steps=20):
    """A box beyond the kernels' 32-bit element index on one GPU: E^3 D3Q19
    (seeded sphere pack, porosity 0.35) solved as x-slabs of one device
    (what stokes_d3q19 does for such a grid), halo exchanges included;
    timed with HIP events over `steps` iterations after 2 warm-up ones."""
    import torch
    import prtlb
    from prtlb.slab import SlabSet
    t0 = time.perf_counter()
    g = prtlb.sphere_pack((E, E, E), (4.0, 8.0), 0.35, seed=0)
    t_geo = time.perf_counter() - t0
    T = SlabSet(g, None, force=2e-6, storage=storage)
    T.run(2)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    T.run(steps)
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    pop = {"fp64": 8, "fp32": 4, "fp16": 2}[storage]
    n = E ** 3
    gbs = n * 2 * 19 * pop / (ms * 1e-3) / 1e9
    held = sum(sum(t.numel() * t.element_size() for t in sl["f"]) for sl in T.slabs)
    return {"grid": E, "storage": storage, "nodes": n, "slabs": len(T.slabs),
            "ms_per_step": round(ms, 3), "mlups": round(n / (ms * 1e-3) / 1e6, 1),
            "GBs": round(gbs, 1), "frac": round(gbs * 1e9 / PEAK_HBM, 4),
            "population_GB": round(held / 1e9, 1), "geometry_s": round(t_geo, 1)}


def geometry_study(edges=(128, 256)):
    """The geometry and check passes on the GPU against their host forms on
    the same grids: seeded sphere pack + inlet connectivity, biofilm coating,
    geodesic distance, distance transform, and check_physics over a
    synthetic 10-snapshot, 2-chemical result.  Wall seconds, synchronised."""
    import numpy as np
    import torch
    import prtlb
    from prtlb import geometry as G
    from scipy import ndimage
    from oracle import prtlb_oracle as O

    def wall(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    out = {}
    # warm-up: the first call of each pass loads its kernels
    gw = prtlb.sphere_pack((32, 32, 32), (4.0, 8.0), 0.35, seed=1, device="cuda")
    tw = torch.from_numpy(gw).cuda()
    G.coat_with_biofilm_device(tw.clone(), 1, 1)
    G.geodesic_device(tw)
    G.edt_device(tw)
    prtlb.check_physics_device(torch.rand((2, 2, 32, 32, 32), device="cuda"), gw)
    for E in edges:
        shape = (E, E, E)
        r = {}
        g, r["pack_gpu_s"] = wall(lambda: prtlb.sphere_pack(shape, (4.0, 8.0), 0.35, seed=0,
                                                             device="cuda"))
        _, r["pack_host_s"] = wall(lambda: prtlb.sphere_pack(shape, (4.0, 8.0), 0.35, seed=0))
        t = torch.from_numpy(g).cuda()
        _, r["coat_gpu_s"] = wall(lambda: G.coat_with_biofilm_device(t.clone(), 1, 1))
        _, r["coat_host_s"] = wall(lambda: G.coat_with_biofilm(g, 1, 1))
        (_, sweeps), r["geodesic_gpu_s"] = wall(lambda: G.geodesic_device(t))
        r["geodesic_sweeps"] = sweeps
        _, r["edt_gpu_s"] = wall(lambda: G.edt_device(t))
        _, r["edt_scipy_s"] = wall(lambda: ndimage.distance_transform_edt(g == prtlb.PORE))
        rng = np.random.default_rng(0)
        snaps = torch.from_numpy(rng.random((10, 2) + shape, dtype=np.float32)).cuda()
        _, r["check_physics_gpu_s"] = wall(lambda: prtlb.check_physics_device(snaps, g))
        if E <= 128:
            host = snaps.cpu().numpy()
            _, r["check_physics_host_s"] = wall(lambda: prtlb.check_physics(host, g))
        out["edge_%d" % E] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        del t, snaps
        torch.cuda.empty_cache()
    # the oracle's geodesic relaxation is NumPy passes over the grid: 64^3 only
    g64 = prtlb.sphere_pack((64, 64, 64), (4.0, 8.0), 0.35, seed=0)
    (_, sw64), gpu64 = wall(lambda: G.geodesic_device(g64))
    _, host64 = wall(lambda: O.geodesic(g64))
    out["geodesic_64"] = {"gpu_s": round(gpu64, 4), "oracle_s": round(host64, 3), "sweeps": sw64}
    return out


def features_study(edge=128, cpu_edge=48, buf=10):
    """The per-rock flow features (harmonic pressure and gradient, MIS, UPRM,
    dw2) on a seeded sphere pack: GPU wall time per map (synchronised, mask
    resident), and the same maps from the NumPy/SciPy oracle on a smaller
    pack on one host core."""
    import numpy as np
    import torch
    import prtlb
    from prtlb import features as Fe
    out = {"gpu_edge": edge, "cpu_edge": cpu_edge, "buffer": buf}

    def wall(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    g = prtlb.sphere_pack((edge, edge, edge), (4.0, 8.0), 0.35, seed=0) == prtlb.PORE
    t = torch.from_numpy(g.astype(np.uint8)).cuda()
    Fe.harmonic_pressure_device(t)                          # warm-up
    (P, dropped, iters), s = wall(lambda: Fe.harmonic_pressure_device(t))
    out["harmonic_pressure_s"] = round(s, 4)
    out["harmonic_cg_iterations"] = iters
    out["harmonic_pore_voxels"] = int(g.sum()) - dropped
    out["harmonic_gradient_s"] = round(wall(lambda: Fe.harmonic_gradient_device(t, P=P))[1], 5)
    out["mis_s"] = round(wall(lambda: Fe.mis_map_device(t, buf=buf))[1], 4)
    out["uprm_s"] = round(wall(lambda: Fe.uprm_map_device(t))[1], 4)
    out["dw2_s"] = round(wall(lambda: Fe.dw2_map_device(t))[1], 5)
    from oracle import features_oracle as O
    gc = prtlb.sphere_pack((cpu_edge,) * 3, (4.0, 8.0), 0.35, seed=0) == prtlb.PORE
    for name, fn in (("harmonic_pressure", lambda: O.harmonic_pressure(gc)),
                     ("mis", lambda: O.mis_map(gc, buf=buf)), ("uprm", lambda: O.uprm_map(gc)),
                     ("dw2", lambda: O.dw2_map(gc))):
        t0 = time.perf_counter()
        fn()
        out["cpu_%s_s" % name] = round(time.perf_counter() - t0, 3)
    out["cpu_note"] = "oracle on %d^3, one core (scipy spsolve, binary_dilation, heapq)" % cpu_edge
    return out


def tail_study(counts=(1, 4, 16), edge=48, reps=4):
    """Latency-bound tail batches: per-launch time of the batched Stokes
    tick (fluid lists; 50 launches and the per-member sums) and of the
    batched transport chunk (32 launches), replayed from a captured HIP graph
    against the same launches issued one by one, for 1, 4 and 16 members of
    an edge^3 sphere pack (porosity 0.35).  A kernel trace of the same run
    gives the kernels' own durations, the floor a persistent multi-step
    launch could approach."""
    import numpy as np
    import torch
    import prtlb
    from prtlb.batch import AdrBatch, LbmBatch
    out = {}
    for n in counts:
        geoms = [prtlb.sphere_pack((edge, edge, edge), (3.0, 6.0), 0.35, seed=i) for i in range(n)]
        row = {}
        for graph in (True, False):
            b = LbmBatch(19, storage="fp32")
            b.use_graph = graph
            for i, g in enumerate(geoms):
                b.add(i, g, force=1e-6, nit=10 ** 6)
            for _ in range(2):                      # both launch parities captured
                b._tick_launches()
                b.parity += b.every
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                b._tick_launches()
                b.parity += b.every
            e1.record()
            torch.cuda.synchronize()
            row["stokes_us_per_launch_" + ("graph" if graph else "eager")] = \
                e0.elapsed_time(e1) * 1e3 / (reps * b.every)
            row["fluid_nodes"] = int(sum(m.n_fluid for m in b.members.values()))
            del b
        for graph in (True, False):
            a = AdrBatch()
            a.use_graph = graph
            for i, g in enumerate(geoms):
                v = np.zeros((3,) + g.shape, np.float32)
                v[0] = np.where(g == prtlb.PORE, 1.0, 0.0)
                a.add(i, g, v, 10.0, 1.0, 3, n_species=2,
                      chem=prtlb.chemistry.sorption_irreversible(), da_abio=1.0)
            a._plan()
            steps = a.GRAPH_STEPS
            for _ in range(2):
                a._graph_chunk() if graph else a._launch(a.parity, steps)
                a.parity += steps
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                a._graph_chunk() if graph else a._launch(a.parity, steps)
                a.parity += steps
            e1.record()
            torch.cuda.synchronize()
            row["transport_us_per_launch_" + ("graph" if graph else "eager")] = \
                e0.elapsed_time(e1) * 1e3 / (reps * steps)
            row["pore_voxels"] = int(sum(m.T.n_pore for m in a.members.values()))
            del a
        out["members_%d" % n] = row
        log("tail %d members: %s" % (n, row))
    return out


def serve_study(n_req=48, rate=4.0, seed=0, budget=200e9, max_steps=100000, transport_cap=4e6,
                transport_order="smallest", baseline=False):
    """Serving: Poisson arrivals (rate per s) of solve requests, 30% D2Q9
    (256^2..1024^2 disk packs), 70% D3Q19 (64^3..256^3 sphere packs), Pe and
    Da log-uniform, reversible sorption; HBM budget admission; continuous
    batching.  Reports p50/p99 time from arrival to converged field and
    solves per hour; with baseline, also the same requests solved one at a
    time in FP64 (serve_sequential)."""
    from prtlb import serve
    t0 = time.perf_counter()
    reqs = serve.poisson_requests(n_req, rate=rate, seed=seed)
    log("serve geometry %.1f s" % (time.perf_counter() - t0))
    t0 = time.perf_counter()
    done = serve.Server(hbm_budget=budget, max_steps=max_steps, transport_cap=transport_cap,
                        transport_order=transport_order).run(reqs, realtime=True, progress=log)
    wall = time.perf_counter() - t0
    out = serve.latency_summary(done, wall)
    out.update(rate_per_s=rate, hbm_budget_GB=budget / 1e9, transport_cap=transport_cap,
               transport_order=transport_order,
               n_2d=sum(1 for r in done if r.g.ndim == 2),
               adr_steps_mean=float(sum(r.result["info"]["steps"] for r in done) / len(done)))
    if baseline:
        out["sequential_fp64"] = serve_sequential(done, max_steps=max_steps,
                                                  served_per_hour=out["solves_per_hour"])
    return out


def serve_sequential(reqs, max_steps=100000, served_per_hour=None):
    """The served requests solved again one at a time in FP64 on the GPU, in
    arrival order, with the server's settings (Stokes force, tau 0.8, tol
    1e-6, 30 000 iterations; transport n_t 5, the same step cap): solves per
    hour of that schedule against the served rate, and the permeability error
    of each served FP32 result against its FP64 solve."""
    import numpy as np
    import torch
    import prtlb
    from prtlb import campaign as C
    t0 = time.perf_counter()
    perr, settled, flow_s = [], 0, 0.0
    for r in sorted(reqs, key=lambda q: q.arrival):
        fn = prtlb.stokes_d2q9 if r.g.ndim == 2 else prtlb.stokes_d3q19
        info = {}
        t1 = time.perf_counter()
        v = fn(r.g, nit=30000, force=r.force, tau_lb=0.8, tol=1e-6, info=info, storage="fp64")
        torch.cuda.synchronize()
        flow_s += time.perf_counter() - t1
        k64 = info["permeability"]
        perr.append(abs(r.result["permeability"] - k64) / abs(k64))
        kw = C._adr_kw(dict(da=r.da), r.kinetics)
        ainfo = {}
        prtlb.solve_adr(r.g, v, r.pe, kw.pop("da"), 5, max_steps=max_steps, info=ainfo,
                        precision="fp64", **kw)
        settled += bool(ainfo["converged"])
        log("serve sequential fp64: request %d done (%d steps)" % (r.key, ainfo["steps"]))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_hour = len(reqs) / wall * 3600.0
    return dict(wall_s=wall, flow_s=flow_s, solves_per_hour=per_hour,
                served_over_sequential=(served_per_hour / per_hour) if served_per_hour else None,
                settled=settled, max_perm_rel_err=float(max(perr)),
                median_perm_rel_err=float(np.median(perr)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grid", type=int, default=256, help="edge of the main cube")
    ap.add_argument("--no-studies", action="store_true",
                    help="leave the 64-rock campaign and the 48-request serving run out")
    ap.add_argument("--storage", default="fp64", choices=["fp64", "fp32", "fp16"])
    ap.add_argument("--stream-only", action="store_true",
                    help="only the device-copy ceiling (counter calibration runs)")
    ap.add_argument("--precision", action="store_true",
                    help="storage-precision error study (256^3, 384^3)")
    ap.add_argument("--campaign", type=int, default=0,
                    help="batched-campaign study with this many rocks")
    ap.add_argument("--batched-only", action="store_true",
                    help="campaign study: the batched schedule only")
    ap.add_argument("--campaign-flow-cap", type=float, default=0,
                    help="campaign study: fluid nodes in the Stokes batch at once (0: no cap)")
    ap.add_argument("--campaign-transport-cap", type=float, default=4e6,
                    help="campaign study: pore voxels in the transport batch at once (0: no cap)")
    ap.add_argument("--campaign-storage", default="fp32", choices=["fp32", "fp16"],
                    help="campaign study: population storage of the batched Stokes stage")
    ap.add_argument("--serve", type=int, default=0,
                    help="serving study with this many Poisson requests")
    ap.add_argument("--serve-baseline", action="store_true",
                    help="serving study: also solve the requests one at a time in FP64")
    ap.add_argument("--adr-steps", type=int, default=None)
    ap.add_argument("--transport-cap", type=float, default=4e6,
                    help="serving: pore voxels allowed in the transport batch at once (0: no cap)")
    ap.add_argument("--transport-order", default="smallest", choices=["arrival", "smallest"],
                    help="serving: which waiting request enters the transport batch first")
    ap.add_argument("--adr-max-steps", type=int, default=100000,
                    help="transport step cap of the campaign and serving studies")
    ap.add_argument("--tail", action="store_true",
                    help="small batches: graph replay against launches one by one")
    ap.add_argument("--adr-only", action="store_true",
                    help="only the transport step (128^3 and 256^3)")
    ap.add_argument("--fluid-list", action="store_true",
                    help="only the fluid-list step on the 128^3 pack (FP32 with 16- and 32-bit "
                         "links, FP64)")
    ap.add_argument("--shapes", action="store_true",
                    help="only the launch-shape tuner on the main grid: every candidate in order "
                         "(2 + 5 iterations each), for counter runs that price each shape")
    ap.add_argument("--slab", action="store_true",
                    help="only one GPU's share of the main cube split along x over 8 GPUs: "
                         "the slab step and its halo kernels")
    ap.add_argument("--big", type=int, default=0,
                    help="only an N^3 FP32 D3Q19 box beyond the 32-bit element index, on slabs")
    ap.add_argument("--geometry", action="store_true",
                    help="only the geometry and check passes, GPU against host")
    ap.add_argument("--features", action="store_true",
                    help="only the flow-feature maps (harmonic pressure, MIS, UPRM, dw2)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import prtlb
    assert torch.cuda.is_available(), "bench needs a HIP device"
    if args.stream_only:
        print(json.dumps({"stream_copy": stream_copy()}), flush=True)
        return
    if args.tail:
        print(json.dumps({"tail": tail_study()}), flush=True)
        return
    if args.adr_only:
        res = {}
        for E in (128, 256):
            g = prtlb.sphere_pack((E, E, E), (4.0, 8.0), 0.35, seed=0)
            v = prtlb.stokes_d3q19(g, nit=1500, force=2e-6, tol=1e-5, storage="fp32")
            for kin in ("monod", "sorption"):
                # pore list and dense box in float32, pore list in FP64
                for name, pl, prec in (("list", True, "fp32"), ("dense", False, "fp32"),
                                       ("list_f64", True, "fp64")):
                    res["adr_%d_%s_%s" % (E, kin, name)] = time_adr(
                        g, args.steps, args.warmup, kin, pl, prec, v=v)
        print(json.dumps({"adr": res}), flush=True)
        return
    if args.precision:
        print(json.dumps({"precision": precision_study()}), flush=True)
        return
    if args.shapes:
        from prtlb import autotune
        E = args.grid
        g = prtlb.sphere_pack((E, E, E), (4.0, 8.0), 0.35, seed=0)
        s = prtlb.StokesSolver(g, force=2e-6, storage=args.storage)
        res = autotune.tune(s, steps=5, force=True, passes=1)
        print(json.dumps({"shapes": {"grid": E, "storage": args.storage, "best": res["tune"],
                                     "order": res["order"], "ms": res["table"]}}), flush=True)
        return
    if args.features:
        print(json.dumps({"features": features_study()}), flush=True)
        return
    if args.geometry:
        print(json.dumps({"geometry": geometry_study()}), flush=True)
        return
    if args.big:
        print(json.dumps({"big": big_study(args.big, storage=args.storage if args.storage != "fp64"
                                           else "fp32")}), flush=True)
        return
    if args.slab:
        print(json.dumps({"slab": slab_study(args.grid)}), flush=True)
        return
    if args.fluid_list:
        import numpy as np
        g = prtlb.sphere_pack((128, 128, 128), (4.0, 8.0), 0.35, seed=0)
        res = {}
        for name, st, l16 in (("fp32_link16", "fp32", True), ("fp32_link32", "fp32", False),
                              ("fp64_link16", "fp64", True)):
            res[name] = time_fluid_list(g, st, args.steps, args.warmup, link16=l16)
        # the same FP32 step on a box with no wall and on a looser pack: the
        # reads of the two-step wall reversal grow with the wall area
        res["fp32_link16_no_walls"] = time_fluid_list(
            np.full((128, 128, 128), prtlb.PORE, np.uint8), "fp32", args.steps, args.warmup)
        res["fp32_link16_porosity_0.70"] = time_fluid_list(
            prtlb.sphere_pack((128, 128, 128), (4.0, 8.0), 0.70, seed=0), "fp32", args.steps,
            args.warmup)
        print(json.dumps({"fluid_list": res}), flush=True)
        return
    if args.campaign:
        print(json.dumps({"campaign": campaign_study(args.campaign, adr_steps=args.adr_steps,
                                                     max_steps=args.adr_max_steps,
                                                     batched_only=args.batched_only,
                                                     transport_cap=args.campaign_transport_cap
                                                     or None,
                                                     flow_cap=args.campaign_flow_cap or None,
                                                     storage=args.campaign_storage)}),
              flush=True)
        return
    if args.serve:
        print(json.dumps({"serve": serve_study(args.serve, max_steps=args.adr_max_steps,
                                               transport_cap=args.transport_cap or None,
                                               transport_order=args.transport_order,
                                               baseline=args.serve_baseline)}),
              flush=True)
        return

    E = args.grid
    t_start = time.perf_counter()
    log("geometry %d^3 ..." % E)
    gmain = prtlb.sphere_pack((E, E, E), (4.0, 8.0), 0.35, seed=0)
    log("%s %d^3 ..." % (args.storage, E))
    main_r = time_lbm(gmain, args.storage, args.steps, args.warmup)
    extras = {}
    if not args.no_extras:
        g128 = prtlb.sphere_pack((128, 128, 128), (4.0, 8.0), 0.35, seed=0)
        for st in ("fp64", "fp32", "fp16"):
            log(st, "128^3 ...")
            extras["d3q19_128_" + st] = time_lbm(g128, st, args.steps, args.warmup)
        log("fluid-only fp64 128^3 ...")
        extras["d3q19_128_fp64_fluid_only"] = time_lbm(g128, "fp64", args.steps, args.warmup,
                                                       skip_solid=True)
        for st in ("fp64", "fp32"):
            log("fluid list", st, "128^3 ...")
            extras["d3q19_128_%s_fluid_list" % st] = time_fluid_list(g128, st, args.steps,
                                                                     args.warmup)
        if E == 256:
            for st in ("fp32", "fp16"):
                log(st, "256^3 ...")
                extras["d3q19_256_" + st] = time_lbm(gmain, st, args.steps, args.warmup)
        del gmain
        log("D2Q9 2048^2 ...")
        g2d = prtlb.disk_pack((2048, 2048), 6.0, 0.6, seed=0, device="cuda")
        for st in ("fp64", "fp32"):
            extras["d2q9_2048sq_" + st] = time_lbm(g2d, st, args.steps, args.warmup)
        del g2d
        log("transport 128^3 / 256^3 ...")
        extras["adr_128_monod_f32"] = time_adr(g128, args.steps, args.warmup)
        g256 = prtlb.sphere_pack((256, 256, 256), (4.0, 8.0), 0.35, seed=0)
        extras["adr_256_monod_f32"] = time_adr(g256, max(args.steps // 2, 20), args.warmup)
        del g256
        torch.cuda.empty_cache()
        log("stream copy ...")
        extras["stream_copy"] = stream_copy()
        log("2D pipeline 128^2 ...")
        extras["d2q9_128sq_pipeline"] = config1_pipeline()
        if not args.no_studies:
            log("campaign 64 rocks ...")
            extras["campaign_64"] = campaign_study(64, max_steps=args.adr_max_steps)
            # 32 Poisson requests keep the default run inside a few minutes;
            # `--serve 48` is the larger study
            log("serve 32 requests ...")
            extras["serve_32"] = serve_study(32, max_steps=args.adr_max_steps)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(E)
        if not args.no_extras:
            log("cpu extras ...")
            extras["cpu_extras"] = cpu_extras()
            if "campaign_64" in extras:
                extras["cpu_extras"]["campaign_64_host"] = cpu_campaign_estimate(
                    extras["campaign_64"], extras["cpu_extras"])
    extras["bench_wall_s"] = round(time.perf_counter() - t_start, 1)
    line = {
        "metric": "D3Q19 TRT MLUPS & %HBM peak; batched solves/hr + p99 latency at fixed error",
        "value": round(main_r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp16": "f16 storage, f32 arithmetic"}[args.storage],
        "data": "synthetic: seeded %d^3 overlapping-sphere pack, radii U(4,8), porosity %.3f"
                % (E, main_r["fluid_fraction"]),
        "config": {"workload": "D3Q19 TRT Stokes step, %d^3 sphere pack (radii U(4,8), porosity "
                               "0.35), %s populations, dense box (every node updated)"
                               % (E, args.storage.upper()),
                   "grid": [E, E, E], "lattice": "D3Q19", "storage": args.storage,
                   "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "achieved": round(main_r["gbs"], 1),
                     "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": round(main_r["frac"], 4),
                     "traffic": measured_traffic("d3q19_%s_%d" % (args.storage, E),
                                                 main_r["launch_shape"])},
        "cpu_baseline": cpu,
        "mlups_fluid": round(main_r["mlups_fluid"], 1),
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
