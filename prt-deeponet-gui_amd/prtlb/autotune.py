"""Launch-shape autotuner for the lattice-Boltzmann step.

Sweeps the block shape (threads along z and y) and the nodes per thread of
prt_lbm_run on the actual grid, storage and mode, times each candidate with
HIP events, and keeps the fastest.  Results are cached per (lattice,
storage, mode, grid) for the process and, if PRTLB_TUNE_CACHE names a file,
across processes as JSON.
"""
import json
import os

import torch

from ._lib import lbm_tune

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


def tune(s, steps=10, candidates=None, force=False):
    """Pick the fastest launch shape for StokesSolver `s`; sets s.tune and
    returns {"tune": flags, "ms": ..., "table": {...}}.  Tuning advances the
    solver's state by the iterations it times."""
    if not _CACHE:
        _load()
    key = _key(s)
    if key in _CACHE and not force:
        s.tune = _CACHE[key]["tune"]
        return _CACHE[key]
    table = {"default": time_shape(s, 0, steps)}
    order = [("default", 0)]                 # candidates in the order they ran
    best, best_ms = 0, table["default"]
    for bz, by, npt, ntl, xcd in (candidates or CANDIDATES):
        if (1 << bz) * (1 << by) > 512 or ((ntl or xcd) and (s.flags or npt != 1)):
            continue
        t = lbm_tune(bz, by, npt, bool(ntl), bool(xcd))
        ms = time_shape(s, t, steps)
        name = "%dx%dx%d%s%s" % (1 << bz, 1 << by, npt, "nt" if ntl else "", "x" if xcd else "")
        table[name] = ms
        order.append((name, t))
        if ms < best_ms:
            best, best_ms = t, ms
    s.tune = best
    _CACHE[key] = {"tune": best, "ms": best_ms, "table": table, "order": order}
    _save()
    return _CACHE[key]
