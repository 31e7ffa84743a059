"""A solve-request server with continuous batching.

Requests (a pore mask, Pe, Da, kinetics) arrive at any time.  Each one is
admitted as soon as the HBM budget has room for it and then runs through
two stages -- Stokes (one member of the LbmBatch of its lattice), then
transport (one member of the AdrBatch) -- while its populations and fields
stay on the device.  Every scheduling round advances ALL admitted requests
with one launch per batch, so requests that arrive while others run are
folded into the same launches as soon as memory frees up; a finished
request leaves its batch and frees its memory at once.

Each request's result is bit-identical to the same request served alone
(tests/test_gpu_batch.py::test_server_result_same_alone_and_in_mix and
::test_server_admission_under_a_binding_budget check it): the batched
kernels run the
single-geometry arithmetic, and every per-request reduction covers that
request's nodes only.  The reference has no serving layer; its solver is a
per-process function (3D/tools/prtlb_3d.py:348, :781).
"""
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import campaign as C
from .adr import adr_host_setup
from .batch import AdrBatch, LbmBatch, adr_footprint, lbm_footprint
from .geometry import PORE


class Request:
    def __init__(self, key, g, pe, da, kinetics="sorption_reversible", arrival=0.0,
                 force=2e-6, probe=False):
        self.key, self.g, self.pe, self.da = key, np.asarray(g), float(pe), float(da)
        self.n_pore = int((self.g == PORE).sum())
        self.kinetics, self.arrival, self.force, self.probe = kinetics, float(arrival), float(force), probe
        self.t_admit = self.t_done = None
        self.round_admit = self.round_done = None   # scheduling rounds of admission and result
        self.result = None


def transport_queue(ready, order, now, max_wait):
    """The requests waiting for the transport batch, in the order they may
    enter: "arrival" keeps the order they became ready; "smallest" puts the
    ones that have waited more than max_wait seconds first (in the order
    they became ready), then the rest by pore voxels, fewest first."""
    ready = list(ready)
    if order == "arrival":
        return ready

    def key(r):
        if now - r.t_ready > max_wait:
            return (0, r.t_ready, 0)
        return (1, r.n_pore, r.t_ready)
    return sorted(ready, key=key)


class Server:
    """Continuous-batching PRT-LB server on one GPU."""

    def __init__(self, hbm_budget=200e9, storage="fp32", tau=0.8, tol=1e-6, nit=30000, n_t=5,
                 adr_steps=None, max_steps=100000, device=None, transport_cap=4e6,
                 transport_order="smallest", max_wait=60.0):
        # transport_cap: pore voxels allowed in the transport batch at once
        # (None: no cap; a request larger than the cap runs alone).  The
        # batched step keeps the device busy from about a million pore
        # voxels on, and up to about four million its pore table and fields
        # stay in the 256 MB Infinity Cache from step to step, so this cap
        # raises throughput, and the requests let in finish without waiting
        # for the whole mix.
        # transport_order: which waiting request enters first, "arrival"
        # (first come first served) or "smallest" (fewest pore voxels
        # first: the shortest jobs, since every member runs the same steps
        # and a step costs it its pore voxels); a request that has waited
        # max_wait seconds goes ahead of the smaller ones, so none starves.
        if transport_order not in ("arrival", "smallest"):
            raise ValueError("transport_order must be 'arrival' or 'smallest'")
        self.transport_cap = transport_cap
        self.transport_order = transport_order
        self.max_wait = float(max_wait)
        self.ready = deque()
        self.budget = float(hbm_budget)
        self.storage, self.tau, self.tol, self.nit = storage, tau, tol, nit
        self.n_t, self.adr_steps, self.max_steps = n_t, adr_steps, max_steps
        self.lb = {9: LbmBatch(9, storage, device=device), 19: LbmBatch(19, storage, device=device)}
        self.ab = AdrBatch(device=device)
        # transport host set-ups (NumPy over the whole grid) run on worker
        # threads, off the loop that keeps the device busy; the pool lives
        # for one run() (or until close())
        self.pool = None
        self.waiting = deque()
        self.done = []
        self.reserved = {}
        self.rounds = 0
        self.peak_reserved = 0.0

    def close(self):
        """Stop the set-up workers; queued set-ups are cancelled."""
        if self.pool is not None:
            self.pool.shutdown(wait=False, cancel_futures=True)
            self.pool = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def submit(self, req):
        self.waiting.append(req)

    def _footprint(self, req):
        g = req.g
        kw = C._adr_kw(dict(da=req.da), req.kinetics)
        n_fluid = req.n_pore if self.lb[9].sparse else None
        return max(lbm_footprint(g.shape, self.storage, n_fluid),
                   adr_footprint(g.shape, kw["n_species"], g.ndim))

    def in_use(self):
        return sum(self.reserved.values())

    def _admit(self, now):
        while self.waiting and self.waiting[0].arrival <= now:
            req = self.waiting[0]
            need = self._footprint(req)
            if self.reserved and self.in_use() + need > self.budget:
                break
            self.waiting.popleft()
            self.reserved[req.key] = need
            self.peak_reserved = max(self.peak_reserved, self.in_use())
            req.t_admit = now
            req.round_admit = self.rounds
            q = 9 if req.g.ndim == 2 else 19
            stage = "probe" if req.probe else "real"
            m = self.lb[q].add((req.key, stage), req.g, force=C.PROBE_FORCE if req.probe else req.force,
                               tau_lb=self.tau, tol=self.tol, nit=self.nit)
            m.req = req

    def _round(self, now):
        self.rounds += 1
        self._admit(now)
        for q, lb in self.lb.items():
            if not len(lb):
                continue
            for m in lb.tick():
                req = m.req
                key, stage = m.key
                info = {}
                v = lb.finish(m, info).cpu().numpy()
                if stage == "probe":
                    req.force = C.lowered_force(v, req.force)
                    m2 = lb.add((key, "real"), req.g, force=req.force, tau_lb=self.tau,
                                tol=self.tol, nit=self.nit)
                    m2.req = req
                    continue
                req.permeability = info["permeability"]
                req.velocity = v
                kw = C._adr_kw(dict(da=req.da), req.kinetics)
                req.prep = self.pool.submit(adr_host_setup, req.g, v, req.pe, kw["da"],
                                            kw["n_species"], kw["chem"], kw.get("da_abio", 0.0),
                                            "fp32", kw.get("sorption"))
                req.t_ready = time.perf_counter()
                self.ready.append(req)
        self._admit_transport()
        if len(self.ab):
            for m in self.ab.advance():
                req = m.req
                ainfo = {}
                out, tn = self.ab.finish(m, ainfo)
                req.result = dict(out=out, t_norm=tn, info=ainfo, permeability=req.permeability,
                                  velocity=req.velocity)
                req.t_done = time.perf_counter()
                req.round_done = self.rounds
                self.reserved.pop(req.key, None)
                self.done.append(req)

    def _admit_transport(self):
        if len(self.ready) > 1:
            self.ready = deque(transport_queue(self.ready, self.transport_order,
                                               time.perf_counter(), self.max_wait))
        while self.ready:
            req = self.ready[0]
            n_pore = req.n_pore
            held = sum(m.T.n_pore for m in self.ab.members.values())
            if self.transport_cap and len(self.ab) and held + n_pore > self.transport_cap:
                break
            # a set-up still on its worker thread is waited for only when the
            # device has nothing else to run
            if not req.prep.done() and (len(self.ab) or any(len(b) for b in self.lb.values())):
                break
            self.ready.popleft()
            kw = C._adr_kw(dict(da=req.da), req.kinetics)
            am = self.ab.add(req.key, req.g, req.velocity, req.pe, kw.pop("da"), self.n_t,
                             steps=self.adr_steps, max_steps=self.max_steps, host=req.prep.result(),
                             **kw)
            am.req = req

    def idle(self):
        return not (self.waiting or self.ready or any(len(b) for b in self.lb.values())
                    or len(self.ab))

    def run(self, requests, realtime=True, progress=None):
        """Serve `requests` (arrival times in seconds from now when realtime,
        else all at once) until every one is done.  Returns them, with
        latency = done - arrival.  progress, if given, is called with a short
        status text every 20 s of serving."""
        t0 = time.perf_counter()
        t_note = t0
        for r in sorted(requests, key=lambda r: r.arrival):
            self.submit(r)
        if self.pool is None:
            self.pool = ThreadPoolExecutor(max_workers=4)
        try:
            self._serve(t0, t_note, realtime, progress)
        finally:
            self.close()
        for r in requests:
            r.latency = (r.t_done - t0) - (r.arrival if realtime else 0.0)
        return requests

    def _serve(self, t0, t_note, realtime, progress):
        while not self.idle():
            if progress and time.perf_counter() - t_note > 20.0:
                t_note = time.perf_counter()
                progress("serve %.0f s: %d done, %d in flow, %d ready, %d in transport, %d waiting, "
                         "%.1f GB held"
                         % (t_note - t0, len(self.done), sum(len(b) for b in self.lb.values()),
                            len(self.ready), len(self.ab), len(self.waiting), self.in_use() / 1e9))
            now = (time.perf_counter() - t0) if realtime else float("inf")
            busy = any(len(b) for b in self.lb.values()) or len(self.ab) or self.ready
            if not busy and self.waiting and realtime and self.waiting[0].arrival > now:
                time.sleep(min(0.01, self.waiting[0].arrival - now))
                continue
            self._round(now)


def poisson_requests(n, rate=4.0, frac_2d=0.3, seed=0, size2d=(256, 1024), size3d=(64, 256),
                     pe=(0.1, 100.0), da=(0.01, 10.0), kinetics="sorption_reversible"):
    """The serving workload: Poisson arrivals, 30% D2Q9 disk packs and 70%
    D3Q19 sphere packs of random size, Pe and Da log-uniform."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        two_d = rng.random() < frac_2d
        if two_d:
            e = int(rng.integers(size2d[0], size2d[1] + 1))
            spec = dict(shape=(e, e), radius=6.0, porosity=float(rng.uniform(0.5, 0.7)))
        else:
            e = int(rng.integers(size3d[0], size3d[1] + 1))
            spec = dict(shape=(e, e, e), radius=(3.0, 8.0), porosity=float(rng.uniform(0.25, 0.5)))
        spec["geom_seed"] = int(rng.integers(2 ** 31))
        g = C.make_geometry(spec)
        p = float(np.exp(rng.uniform(np.log(pe[0]), np.log(pe[1]))))
        d = float(np.exp(rng.uniform(np.log(da[0]), np.log(da[1]))))
        reqs.append(Request(i, g, p, d, kinetics=kinetics, arrival=t))
    return reqs


def latency_summary(reqs, wall):
    """Throughput and latency percentiles over all requests, and separately
    over those whose transport settled (the reference's stop rule) and
    those that ran to the step cap."""
    lat = np.array([r.latency for r in reqs])
    conv = np.array([bool(r.result["info"].get("converged")) for r in reqs])

    def pct(a, q):
        return float(np.percentile(a, q)) if a.size else None

    return dict(n=len(reqs), p50_s=pct(lat, 50), p99_s=pct(lat, 99), mean_s=float(lat.mean()),
                solves_per_hour=len(reqs) / wall * 3600.0, wall_s=wall,
                converged=int(conv.sum()), converged_fraction=float(conv.mean()),
                p50_converged_s=pct(lat[conv], 50), p99_converged_s=pct(lat[conv], 99),
                p50_capped_s=pct(lat[~conv], 50), p99_capped_s=pct(lat[~conv], 99))
