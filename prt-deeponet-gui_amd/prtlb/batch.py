"""Many geometries per launch: batched Stokes and transport.

LbmBatch and AdrBatch hold a changing set of members (one geometry each)
and advance all of them with one kernel launch per iteration
(prt_lbm_batch_step / prt_adr_batch_step).  Members join and leave between
launches, so a caller can keep the device full while solves of very
different sizes start and finish -- continuous batching.  A member's result
is bit-identical to the same solve alone: the per-node arithmetic is the
single-geometry kernels' own, and every per-member reduction runs over that
member's nodes only, in a fixed order.

The reference solves one geometry per process (3D/tools/build_dataset_3d.py
:685-811 loops rocks and parameter sets); this module is the GPU-native
replacement for that loop's inner solves.
"""
import ctypes
import os

import numpy as np
import torch

from ._lib import (BATCH_SHAPE_ADR_DEFAULT, BATCH_SHAPE_LBM_DEFAULT, PRT_LBM_SKIP_SOLID,
                   PRT_LBM_SPARSE, PRT_STORE, AdrSegment, LbmSegment, PrtError, call, lib)
from .adr import AdrRun, TransportSolver
from .device import device as _device
from .device import ptr, stream_ptr
from .geometry import PORE
from .stokes import DTYPES, FlowConvergence, checked_flow, trt_minus

_STORE_BYTES = {"fp64": 8, "fp32": 4, "fp16": 2}


def lbm_footprint(shape, storage, n_fluid=None):
    """Device bytes one Stokes member holds: two population sets (z-rows
    padded to whole 128-byte lines) + mask over the box, or, for a fluid
    list of n_fluid nodes, two compact population sets, the link table (its
    32-bit form, built first), the list and the mask."""
    n = int(np.prod(shape))
    q = 9 if len(shape) == 2 else 19
    if n_fluid is None:
        b = _STORE_BYTES[storage]
        line = 128 // b
        zs = (shape[-1] + line - 1) // line * line
        return (n // shape[-1]) * zs * 2 * q * b + n
    return int(n_fluid) * (2 * q * _STORE_BYTES[storage] + 4 * (q - 1) + 4) + n


def adr_footprint(shape, ns, nd):
    """Device bytes one transport member holds (two fields, u, flags) plus
    a generous allowance for its snapshots (kept on the device)."""
    n = int(np.prod(shape))
    return n * (2 * ns * 4 + nd * 4 + 1) + 12 * n * ns * 4


class _Seg:
    pass


class _Tuned:
    """The launch-shape policy both batch kinds share: the shape is tuned
    (prtlb.autotune.tune_batch) on the first plan with members and again
    only when the batch has grown past twice, or shrunk below a quarter of,
    the entries it was tuned at."""

    def _wants_tuning(self):
        if not self.autotune or not self.members:
            return False
        from .autotune import batch_entries
        n = batch_entries(self)
        if n < 1:
            return False
        if self._tuned_n is not None and self._tuned_n / 4 <= n <= 2 * self._tuned_n:
            return False
        self._tuned_n = n
        return True


def _autotune_default():
    """Batches tune their launch shape unless PRTLB_BATCH_TUNE=0."""
    return os.environ.get("PRTLB_BATCH_TUNE", "1") != "0"


def _width(ns):
    return 1 if ns <= 1 else (2 if ns == 2 else (4 if ns <= 4 else 8))


class LbmBatch(_Tuned):
    """Stokes members on one lattice (q = 9 or 19), advanced together.

    Every member checks FlowConvergence on the same launches (every `every`
    iterations of its own), because members join only at those boundaries.
    """

    def __init__(self, q, storage="fp32", skip_solid=True, every=50, device=None,
                 fp16_scale=1024.0, sparse=None, link16=True):
        if q not in (9, 19):
            raise ValueError("q must be 9 or 19")
        self.q, self.storage, self.st = q, storage, PRT_STORE[storage]
        # fluid-list members (compact populations + link table) by default
        # in the fluid-only mode; the dense box otherwise
        self.sparse = bool(skip_solid) if sparse is None else bool(sparse)
        if self.sparse and not skip_solid:
            raise ValueError("a fluid-list batch is fluid-only (skip_solid=True)")
        self.flags = (PRT_LBM_SKIP_SOLID if skip_solid else 0) | (PRT_LBM_SPARSE if self.sparse else 0)
        # fluid lists keep their link table as 16-bit steps where it fits
        self.link16 = bool(link16)
        self.every = int(every)
        self.dev = _device(device)
        self.fp16_scale = float(fp16_scale)
        self.members = {}
        self.parity = 0
        self._dirty = True
        self.n_blocks = 0
        # launch shape (PRT_BATCH_SHAPE); with autotune the tuner picks it
        # for each new membership size class (prtlb.autotune.tune_batch)
        self.shape = self.default_shape = BATCH_SHAPE_LBM_DEFAULT
        self.autotune = _autotune_default()
        self._tuned_n = None

    def __len__(self):
        return len(self.members)

    def bytes(self):
        return sum(m.bytes for m in self.members.values())

    def add(self, key, g, force=2e-6, tau_lb=0.8, magic=3.0 / 16.0, tol=1e-6, nit=30000):
        """Admit one geometry (its rest state); it steps from the next launch."""
        if key in self.members:
            raise KeyError("member %r already present" % (key,))
        g = np.asarray(g)
        nd = 2 if self.q == 9 else 3
        if g.ndim != nd:
            raise ValueError("a q=%d batch takes %dD grids" % (self.q, nd))
        if nit % self.every:
            raise ValueError("nit must be a multiple of %d in a batch" % self.every)
        m = _Seg()
        m.key, m.shape, m.n = key, g.shape, g.size
        m.dims = (g.shape[0], g.shape[1], 1 if nd == 2 else g.shape[2])
        m.solid = torch.from_numpy(np.ascontiguousarray(g != PORE).astype(np.uint8).reshape(-1)).to(self.dev)
        m.n_fluid = int(g.size - int((g != PORE).sum()))
        m.force, m.tau = float(force), float(tau_lb)
        m.om_p = 1.0 / m.tau
        m.om_m = trt_minus(m.om_p, magic)
        s = stream_ptr(self.dev)
        if self.sparse:
            if m.n_fluid < 1:
                raise PrtError("a fluid-list member needs at least one fluid node")
            m.fidx = torch.from_numpy(np.flatnonzero(g.reshape(-1) == PORE).astype(np.int32)).to(self.dev)
            lk = int(lib().prt_lbm_link_stride(m.n_fluid))
            m.link = torch.empty((self.q - 1) * lk, dtype=torch.int32, device=self.dev)
            work = torch.empty(m.n, dtype=torch.int32, device=self.dev)
            call("prt_lbm_link_table", self.q, *m.dims, ptr(m.fidx), m.n_fluid, ptr(work),
                 ptr(m.link), s)
            m.link_bits = 32
            if self.link16:
                # the 16-bit form where every link fits (halves the table)
                l16 = torch.empty((self.q - 1) * lk, dtype=torch.int16, device=self.dev)
                over = torch.empty(1, dtype=torch.int32, device=self.dev)
                call("prt_lbm_link_pack16", self.q, m.n_fluid, ptr(m.link), ptr(l16), ptr(over), s)
                if int(over.item()) == 0:
                    m.link, m.link_bits = l16, 16
            m.nodes = m.n_fluid
            ld = int(lib().prt_lbm_pop_stride(m.n_fluid, self.st))
            m.f = [torch.empty((self.q, ld), dtype=DTYPES[self.storage], device=self.dev)
                   for _ in range(2)]
            for buf in m.f:
                call("prt_lbm_sparse_fill_rest", self.q, self.st, m.n_fluid, ptr(buf), s)
            torch.cuda.current_stream(self.dev).synchronize()   # work is freed on return
        else:
            m.link, m.nodes = None, m.n
            elems = int(lib().prt_lbm_pop_elems(self.q, self.st, *m.dims))
            m.f = [torch.empty(elems, dtype=DTYPES[self.storage], device=self.dev)
                   for _ in range(2)]
            for buf in m.f:
                call("prt_lbm_fill_rest", self.q, self.st, *m.dims, ptr(buf), s)
        m.phase = self.parity & 1            # newest = f[0] at the current parity
        m.iters = 0
        m.nit = int(nit)
        m.conv = FlowConvergence(tol=tol, every=self.every, max_it=nit,
                                 name="%dD flow" % nd)
        m.bytes = lbm_footprint(g.shape, self.storage, m.n_fluid if self.sparse else None)
        self.members[key] = m
        self._dirty = True
        return m

    def remove(self, key):
        self.members.pop(key)
        self._dirty = True

    def newest(self, m):
        return m.f[0] if ((self.parity ^ m.phase) & 1) == 0 else m.f[1]

    def _plan(self):
        # captured graphs hold the previous plan's pointers: drop them
        self._graphs = {}
        mem = list(self.members.values())
        segs = (LbmSegment * len(mem))()
        for i, m in enumerate(mem):
            segs[i] = LbmSegment(ptr(m.solid), ptr(m.f[0]), ptr(m.f[1]), *m.dims, m.phase,
                                 m.om_p, m.om_m, m.force, self.fp16_scale,
                                 ptr(m.link) if m.link is not None else None, m.nodes,
                                 m.link_bits if m.link is not None else 0, 0)
        dsz = lib().prt_lbm_batch_desc_size()
        desc = np.zeros(dsz * len(mem), np.uint8)
        nb, nt = ctypes.c_int64(), ctypes.c_int64()
        call("prt_lbm_batch_plan_shaped", self.q, self.st, segs, len(mem), self.shape,
             desc.ctypes.data, None, 0, ctypes.byref(nb), ctypes.byref(nt))
        blk = np.zeros(nb.value, np.int32)
        call("prt_lbm_batch_plan_shaped", self.q, self.st, segs, len(mem), self.shape,
             desc.ctypes.data, blk.ctypes.data, nb.value, ctypes.byref(nb), ctypes.byref(nt))
        self.order = mem
        self.desc = torch.from_numpy(desc).to(self.dev)
        self.blk_seg = torch.from_numpy(blk).to(self.dev)
        self.n_blocks = nb.value
        # partial sums go per 256-node tile of each member, whatever the shape
        nblk = np.array([(m.nodes + 255) // 256 for m in mem], np.int32)
        blk0 = np.concatenate([[0], np.cumsum(nblk)[:-1]]).astype(np.int32)
        assert int(nblk.sum()) == nt.value
        self.blk0 = torch.from_numpy(blk0).to(self.dev)
        self.nblk = torch.from_numpy(nblk).to(self.dev)
        self.partials = torch.zeros(nt.value, dtype=torch.float64, device=self.dev)
        self.sums = torch.zeros(len(mem), dtype=torch.float64, device=self.dev)
        self._dirty = False
        self.planned_shape = self.shape
        if self._wants_tuning():
            from . import autotune
            if autotune.tune_batch(self):
                self._plan()            # the tuner picked another shape: plan for it

    def _launch(self, parity, steps, reduce_last):
        s = stream_ptr(self.dev)
        for k in range(steps):
            red = reduce_last and k == steps - 1
            call("prt_lbm_batch_step_shaped", self.q, self.st, ptr(self.desc), ptr(self.blk_seg),
                 self.n_blocks, (parity + k) & 1, self.flags, self.shape,
                 ptr(self.partials) if red else None, s)

    def step(self, steps, reduce_last=False):
        """Advance every member `steps` iterations (one launch each)."""
        if not self.members or steps <= 0:
            return
        if self._dirty:
            self._plan()
        self._launch(self.parity, steps, reduce_last)
        self.parity += steps
        for m in self.members.values():
            m.iters += steps

    # ticks replay a captured HIP graph of the `every` launches and the
    # per-member sums while the membership (the plan) is unchanged; every
    # is even, so each tick starts at the same launch parity
    use_graph = True

    def _tick_launches(self):
        if self._dirty:
            self._plan()
        if not (self.use_graph and self.every % 2 == 0):
            self._launch(self.parity, self.every, True)
            call("prt_seg_sums", ptr(self.blk0), ptr(self.nblk), len(self.order), 1,
                 ptr(self.partials), ptr(self.sums), stream_ptr(self.dev))
            return
        key = self.parity & 1
        if key not in self._graphs:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._launch(self.parity, self.every, True)
                call("prt_seg_sums", ptr(self.blk0), ptr(self.nblk), len(self.order), 1,
                     ptr(self.partials), ptr(self.sums), stream_ptr(self.dev))
            self._graphs[key] = g
        self._graphs[key].replay()

    def tick(self):
        """`every` iterations, then the convergence check of every member.
        Returns the members that finished (converged or hit their cap)."""
        if not self.members:
            return []
        self._tick_launches()
        self.parity += self.every
        for m in self.members.values():
            m.iters += self.every
        sums = self.sums.cpu().numpy()
        done = []
        for i, m in enumerate(self.order):
            um = sums[i] / m.n_fluid if m.n_fluid else 0.0
            if m.conv.done(m.iters - 1, um) or m.iters >= m.nit:
                done.append(m)
        return done

    def finish(self, m, info=None):
        """Velocity of a finished member (float32 (3, *grid) device tensor,
        zero in solid) and its info; removes it from the batch."""
        m.conv.report(info)
        s = stream_ptr(self.dev)
        newest = self.newest(m)
        other = m.f[1] if newest is m.f[0] else m.f[0]
        out = torch.zeros(1, dtype=torch.float64, device=self.dev)
        if self.sparse:
            v32 = torch.zeros((3, m.n), dtype=torch.float32, device=self.dev)
            part = torch.zeros(lib().prt_lbm_sparse_partials_count(m.n_fluid), dtype=torch.float64,
                               device=self.dev)
            call("prt_lbm_sparse_velocity", self.q, self.st, m.n, m.n_fluid, ptr(m.fidx),
                 ptr(m.link), m.link_bits, ptr(newest), ptr(other), m.force, self.fp16_scale, None,
                 ptr(v32), ptr(part), s)
        else:
            v32 = torch.empty((3, m.n), dtype=torch.float32, device=self.dev)
            part = torch.zeros(lib().prt_lbm_partials_count(self.q, *m.dims), dtype=torch.float64,
                               device=self.dev)
            call("prt_lbm_velocity", self.q, self.st, *m.dims, ptr(m.solid), ptr(newest),
                 ptr(other), self.flags, m.force, self.fp16_scale, None, ptr(v32), ptr(part), s)
        call("prt_sum_f64", ptr(part), part.numel(), ptr(out), s)
        sux = float(out.item())
        v32 = v32.reshape((3,) + m.shape)
        checked_flow(v32, m.solid.reshape(m.shape).bool(), m.tau, m.force)
        if info is not None:
            nu = (m.tau - 0.5) / 3.0
            info["permeability"] = nu * (sux / m.n) / m.force
            info["porosity"] = m.n_fluid / m.n
            info["storage"] = self.storage
        self.remove(m.key)
        return v32


class AdrBatch(_Tuned):
    """Float32 transport members advanced together, each with its own event
    schedule (snapshot ladder, convergence checks) from AdrRun."""

    def __init__(self, device=None):
        self.dev = _device(device)
        self.members = {}
        self.parity = 0
        self._dirty = True
        # launch shape (PRT_BATCH_SHAPE); with autotune the tuner picks it
        # for each new membership size class (prtlb.autotune.tune_batch)
        self.shape = self.default_shape = BATCH_SHAPE_ADR_DEFAULT
        self.autotune = _autotune_default()
        self._tuned_n = None

    def __len__(self):
        return len(self.members)

    def bytes(self):
        return sum(m.bytes for m in self.members.values())

    def add(self, key, mat, vel, pe, da, n_t, n_species=2, steps=None, tol=2e-3,
            max_steps=100000, ratio=1.12, chem=None, da_abio=0.0, sorption=None, host=None):
        """host: the member's host set-up made beforehand with
        adr.adr_host_setup(mat, vel, pe, da, n_species, chem, da_abio, "fp32",
        sorption) (for instance on worker threads while the device runs);
        mat, vel, pe, da, n_species, chem, da_abio and sorption must be the
        ones it was made with."""
        if key in self.members:
            raise KeyError("member %r already present" % (key,))
        if host is not None:
            if host.precision != "fp32":
                raise ValueError("a batch member's host set-up must be float32")
            T = host
            T._setup_device(self.dev)
        else:
            T = TransportSolver(mat, vel, pe, da, n_species, chem, da_abio, precision="fp32",
                                sorption=sorption, device=self.dev)
        m = _Seg()
        m.key, m.T = key, T
        m.phase = self.parity & 1
        m.run = AdrRun(T, n_t, steps=steps, tol=tol, max_steps=max_steps, ratio=ratio)
        m.args = (da, da_abio, vel)
        m.bytes = adr_footprint(T.shape, T.ns, T.nd)
        self.members[key] = m
        self._dirty = True
        return m

    def remove(self, key):
        self.members.pop(key)
        self._dirty = True

    def _sync_latest(self):
        for m in self.members.values():
            m.T.latest = 0 if ((self.parity ^ m.phase) & 1) == 0 else 1

    # chunks of GRAPH_STEPS (even) launches replay a captured HIP graph while
    # the membership (the plan) is unchanged
    GRAPH_STEPS = 32
    use_graph = True
    # convergence checks are decided without stopping the device: a member's
    # check is read once its means reach the host (eager), and at the latest
    # when it has more than check_lag undecided checks
    eager = True
    check_lag = 2

    def _launch(self, parity, steps):
        if self.n_blocks < 1:          # every member's pore is empty: nothing to update
            return
        s = stream_ptr(self.dev)
        for k in range(steps):
            call("prt_adr_batch_step_shaped", ptr(self.desc), ptr(self.blk_seg), self.n_blocks,
                 self.max_species, (parity + k) & 1, self.shape, None, s)

    def _graph_chunk(self):
        graphs = getattr(self, "_graphs", None)
        if graphs is None or graphs[0] != self._plan_id:
            graphs = self._graphs = (self._plan_id, {})
        key = self.parity & 1
        if key not in graphs[1]:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._launch(self.parity, self.GRAPH_STEPS)
            graphs[1][key] = g
        graphs[1][key].replay()

    def _plan(self):
        self._plan_id = getattr(self, "_plan_id", 0) + 1
        mem = list(self.members.values())
        segs = (AdrSegment * len(mem))()
        for i, m in enumerate(mem):
            T = m.T
            segs[i] = AdrSegment(ctypes.pointer(T.params), *T.dims, m.phase, ptr(T.table),
                                 T.n_pore, ptr(T.c[0]), ptr(T.c[1]))
        dsz = lib().prt_adr_batch_desc_size()
        desc = np.zeros(dsz * len(mem), np.uint8)
        nb, nt = ctypes.c_int64(), ctypes.c_int64()
        nsm = ctypes.c_int()
        call("prt_adr_batch_plan_shaped", segs, len(mem), self.shape, desc.ctypes.data, None, 0,
             ctypes.byref(nb), ctypes.byref(nt), ctypes.byref(nsm))
        blk = np.zeros(nb.value, np.int32)
        call("prt_adr_batch_plan_shaped", segs, len(mem), self.shape, desc.ctypes.data,
             blk.ctypes.data, nb.value, ctypes.byref(nb), ctypes.byref(nt), ctypes.byref(nsm))
        self.desc = torch.from_numpy(desc).to(self.dev)
        self.blk_seg = torch.from_numpy(blk).to(self.dev)
        self.n_blocks = nb.value
        self.max_species = nsm.value
        self._dirty = False
        self.planned_shape = self.shape
        if self._wants_tuning():
            from . import autotune
            if autotune.tune_batch(self):
                self._plan()            # the tuner picked another shape: plan for it

    def advance(self):
        """Run every member to the earliest pending event, handle the events.
        Returns the members whose runs finished."""
        if not self.members:
            return []
        if self._dirty:
            self._plan()
        todo = min(m.run.next_target() - m.run.n for m in self.members.values())
        left = todo
        while self.use_graph and self.n_blocks > 0 and left >= self.GRAPH_STEPS:
            self._graph_chunk()
            self.parity += self.GRAPH_STEPS
            left -= self.GRAPH_STEPS
        self._launch(self.parity, left)
        self.parity += left
        self._sync_latest()
        done = []
        for m in list(self.members.values()):
            n = m.run.n + todo
            m.T.steps_done = n
            if n == m.run.next_target():
                # a member without pore voxels has no device sums to defer
                m.run.arrive(n, defer=m.T.n_pore > 0)
            else:
                m.run.n = n
            if not m.run.finished and (self.eager or len(m.run.pending) > self.check_lag):
                m.run.resolve(block=len(m.run.pending) > self.check_lag)
            if m.run.finished:
                done.append(m)
        return done

    def finish(self, m, info=None):
        da, da_abio, vel = m.args
        out = m.run.result(info, da, da_abio, vel)
        self.remove(m.key)
        return out
