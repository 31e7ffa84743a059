"""Launch-shape autotuners.

tune(s): the single-geometry lattice-Boltzmann step.  Sweeps the block shape
(threads along z and y) and the nodes per thread of prt_lbm_run on the
actual grid, storage and mode.

tune_batch(b): the batched steps (LbmBatch, AdrBatch).  Sweeps threads per
block, list entries per thread and the block order (PRT_BATCH_SHAPE) on the
batch's actual members, restoring their state afterwards.

Both time each candidate with HIP events and keep the fastest.  Results are
cached for the process and, if PRTLB_TUNE_CACHE names a file, across
processes as JSON: per (lattice, storage, mode, grid) for tune, per (kind,
lattice or width, storage, total list entries to half an octave) for
tune_batch.
"""
import json
import math
import os

import torch

from ._lib import BATCH_SHAPES, batch_shape_name, lbm_tune

_CACHE = {}

# (log2 threads along z, log2 threads along y, nodes per thread, nontemporal
# loads, XCD-contiguous block order)
CANDIDATES = [(6, 0, 1, 0, 0), (6, 1, 1, 0, 0), (6, 2, 1, 0, 0), (6, 3, 1, 0, 0), (7, 0, 1, 0, 0),
              (7, 1, 1, 0, 0), (7, 2, 1, 0, 0), (8, 0, 1, 0, 0), (8, 1, 1, 0, 0), (9, 0, 1, 0, 0),
              (6, 2, 2, 0, 0), (7, 1, 2, 0, 0), (8, 0, 2, 0, 0), (6, 3, 2, 0, 0), (6, 2, 1, 1, 0),
              (7, 1, 1, 1, 0), (7, 0, 1, 1, 0), (8, 0, 1, 1, 0), (8, 1, 1, 1, 0),
              (6, 2, 1, 1, 1), (7, 1, 1, 1, 1), (7, 0, 1, 1, 1), (8, 0, 1, 1, 1), (8, 1, 1, 1, 1),
              (6, 2, 1, 0, 1), (7, 1, 1, 0, 1), (8, 0, 1, 0, 1)]


def _key(s):
    return "q%d_%s_%s_%s" % (s.q, s.storage, "fluid" if s.flags else "dense",
                             "x".join(str(d) for d in s.shape))


def _load():
    path = os.environ.get("PRTLB_TUNE_CACHE")
    if path and os.path.exists(path):
        try:
            _CACHE.update(json.load(open(path)))
        except (OSError, ValueError):
            pass


def _save():
    path = os.environ.get("PRTLB_TUNE_CACHE")
    if path:
        with open(path, "w") as fh:
            json.dump(_CACHE, fh, indent=1, sort_keys=True)


def time_shape(s, tune, steps=10):
    """Mean ms per iteration of solver `s` with launch shape `tune`."""
    old = s.tune
    s.tune = tune
    try:
        s.run(2)
        st = torch.cuda.current_stream(s.dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        s.run(steps)
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / steps
    finally:
        s.tune = old


def tune(s, steps=10, candidates=None, force=False, passes=2):
    """Pick the fastest launch shape for StokesSolver `s`; sets s.tune and
    returns {"tune": flags, "ms": ..., "table": {...}}.  Every shape is timed
    in `passes` passes, the second in reverse order, and keeps its faster
    time (a single short timing can be off by a few percent while the top
    shapes lie within one).  Tuning advances the solver's state by the
    iterations it times."""
    if not _CACHE:
        _load()
    key = _key(s)
    if key in _CACHE and not force:
        s.tune = _CACHE[key]["tune"]
        return _CACHE[key]
    order = [("default", 0)]                 # candidates in the order they ran first
    for bz, by, npt, ntl, xcd in (candidates or CANDIDATES):
        if (1 << bz) * (1 << by) > 512 or ((ntl or xcd) and (s.flags or npt != 1)):
            continue
        name = "%dx%dx%d%s%s" % (1 << bz, 1 << by, npt, "nt" if ntl else "", "x" if xcd else "")
        order.append((name, lbm_tune(bz, by, npt, bool(ntl), bool(xcd))))
    table = {}
    for p in range(max(1, int(passes))):
        for name, t in (order if p % 2 == 0 else order[::-1]):
            ms = time_shape(s, t, steps)
            table[name] = min(ms, table.get(name, ms))
    best_name = min(table, key=lambda n: (table[n], n != "default"))
    best = dict(order)[best_name]
    s.tune = best
    _CACHE[key] = {"tune": best, "ms": table[best_name], "table": table, "order": order}
    _save()
    return _CACHE[key]


def batch_entries(b):
    """List entries one launch of batch b updates (fluid nodes or box nodes
    of Stokes members, pore voxels of transport members)."""
    if hasattr(b, "q"):
        return sum(m.nodes for m in b.members.values())
    return sum(m.T.n_pore for m in b.members.values())


def _batch_key(b):
    if hasattr(b, "q"):
        head = "batch_lbm_q%d_%s_%s" % (b.q, b.storage, "list" if b.sparse else "dense")
    else:
        head = "batch_adr_ns%d" % b.max_species
    return "%s_n%d" % (head, int(round(math.log2(max(batch_entries(b), 1)))))


def _batch_buffers(b):
    """The device buffers a batched launch writes, member by member."""
    if hasattr(b, "q"):
        return [buf for m in b.members.values() for buf in m.f]
    return [buf for m in b.members.values() for buf in m.T.c]


def batch_candidates(b):
    """The launch shapes batch b can take: dense Stokes members only 256
    threads with one node each."""
    if hasattr(b, "q") and not b.sparse:
        return [c for c in BATCH_SHAPES if (c & 0xFF) == 8]
    return list(BATCH_SHAPES)


def _time_batch(b, steps):
    """Mean ms per launch of batch b at its planned shape (2 warm-up
    launches, then `steps`)."""
    def launches(n):
        if hasattr(b, "q"):
            b._launch(b.parity, n, False)
        else:
            b._launch(b.parity, n)
    launches(2)
    st = torch.cuda.current_stream(b.dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    launches(steps)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / steps


# tuning copies the members' buffers: at most this fraction of free memory
TUNE_MEMORY_FRACTION = 0.5

# a shape replaces the default only when it is at least this much faster:
# differences below it are within the run-to-run spread of a short timing
BATCH_MARGIN = 0.02


def tune_batch(b, steps=8, passes=2, candidates=None, force=False):
    """Pick the launch shape of batch b for its current members.  Every
    candidate is planned and timed over `steps` launches on the members'
    own buffers, in `passes` interleaved passes (the fastest pass counts),
    which are copied before and restored after, so tuning leaves every
    member's state as it found it (and a shape never changes results).  The
    batch's default shape is kept unless another is BATCH_MARGIN faster.
    Sets b.shape; returns True when it differs from the shape of b's
    current plan (the caller then plans again)."""
    if not _CACHE:
        _load()
    key = _batch_key(b)
    if key in _CACHE and not force:
        b.shape = _CACHE[key]["shape"]
        return b.shape != b.planned_shape
    default = b.default_shape
    cands = list(candidates or batch_candidates(b))
    bufs = _batch_buffers(b)
    # the copy of the members' state must fit beside them: a batch that
    # fills the device keeps its shape (and is tuned when it is smaller)
    need = sum(t.numel() * t.element_size() for t in bufs)
    free, _ = torch.cuda.mem_get_info(b.dev)
    if need > TUNE_MEMORY_FRACTION * free:
        b.shape = b.planned_shape
        return False
    saved = [t.clone() for t in bufs]
    auto, b.autotune = b.autotune, False
    table = {}
    try:
        for _ in range(passes):
            for shape in cands:
                b.shape = shape
                b._plan()
                ms = _time_batch(b, steps)
                name = batch_shape_name(shape)
                table[name] = min(ms, table.get(name, ms))
    finally:
        b.autotune = auto
        for t, v in zip(bufs, saved):
            t.copy_(v)
        torch.cuda.current_stream(b.dev).synchronize()
        del saved
    best = min(cands, key=lambda c: table[batch_shape_name(c)])
    base = table.get(batch_shape_name(default))
    if base is not None and table[batch_shape_name(best)] > (1.0 - BATCH_MARGIN) * base:
        best = default
    b.shape = best
    _CACHE[key] = {"shape": best, "name": batch_shape_name(best),
                   "ms": table[batch_shape_name(best)], "table": table,
                   "entries": batch_entries(b)}
    _save()
    return b.shape != b.planned_shape
