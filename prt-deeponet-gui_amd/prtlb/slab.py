"""A box split along x into slabs, stepped with halo exchanges.

This is what each GPU of a multi-GPU Stokes solve holds -- one slab per GPU
-- shown with every slab on one device.  A slab of m planes is stepped as the
(m + 2, ny, nz) grid with PRT_LBM_SLAB: planes 0 and m + 1 are halo planes.
Before every step each slab packs the populations that leave it across its
x faces (e_x = +1 of its last own plane, e_x = -1 of its first), the
messages go to the neighbours (here device copies between the slabs'
buffers; across GPUs an RCCL send/recv or a peer copy over xGMI), and each
slab unpacks what it receives into its halo planes.  The result equals the
undivided box's bit for bit (tests/test_gpu_slabs.py).  The reference solves
one box in one process (3D/tools/prtlb_3d.py:348-456); it has no
decomposition.
"""
import numpy as np
import torch

from ._lib import PRT_LBM_SKIP_SOLID, PRT_LBM_SLAB, PRT_STORE, call, lib
from .device import device as _device
from .device import ptr, stream_ptr
from .geometry import PORE
from .stokes import DTYPES, trt_minus


def split_planes(nx, parts):
    """[x0, x1) of each of `parts` slabs of nx planes, the first nx % parts
    one plane longer."""
    if not 1 <= parts <= nx:
        raise ValueError("cannot split %d planes into %d slabs" % (nx, parts))
    base, extra = divmod(nx, parts)
    bounds, x = [], 0
    for p in range(parts):
        m = base + (1 if p < extra else 0)
        bounds.append((x, x + m))
        x += m
    return bounds


class SlabSet:
    """The slabs of grid g (2D (nx, ny) for D2Q9 or 3D (nx, ny, nz) for
    D3Q19), periodic along x like the box, on one device."""

    def __init__(self, g, parts, force=2e-6, tau_lb=0.8, magic=3.0 / 16.0, storage="fp64",
                 skip_solid=False, device=None, fp16_scale=1024.0):
        g = np.asarray(g)
        if g.ndim not in (2, 3):
            raise ValueError("grid must be 2D or 3D, got %dD" % g.ndim)
        self.dev = _device(device)
        self.shape, self.nd = g.shape, g.ndim
        self.q = 9 if g.ndim == 2 else 19
        self.storage, self.st = storage, PRT_STORE[storage]
        self.force, self.tau = float(force), float(tau_lb)
        self.om_p = 1.0 / self.tau
        self.om_m = trt_minus(self.om_p, magic)
        self.flags = (PRT_LBM_SKIP_SOLID if skip_solid else 0) | PRT_LBM_SLAB
        self.fp16_scale = float(fp16_scale)
        self.bounds = split_planes(g.shape[0], parts)
        s = stream_ptr(self.dev)
        self.slabs = []
        for x0, x1 in self.bounds:
            sl = {}
            # own planes x0 .. x1-1 between the planes on either side (periodic)
            local = np.take(g, np.arange(x0 - 1, x1 + 1), axis=0, mode="wrap")
            sl["m"] = x1 - x0
            sl["dims"] = (local.shape[0], local.shape[1], 1 if self.nd == 2 else local.shape[2])
            sl["solid"] = torch.from_numpy(np.ascontiguousarray(local != PORE).astype(np.uint8)
                                           .reshape(-1)).to(self.dev)
            elems = int(lib().prt_lbm_pop_elems(self.q, self.st, *sl["dims"]))
            sl["f"] = [torch.empty(elems, dtype=DTYPES[storage], device=self.dev) for _ in range(2)]
            for buf in sl["f"]:
                call("prt_lbm_fill_rest", self.q, self.st, *sl["dims"], ptr(buf), s)
            n_msg = int(lib().prt_lbm_halo_elems(self.q, self.st, *sl["dims"]))
            # what leaves across the high face (e_x = +1) and the low face (-1)
            sl["send"] = {d: torch.empty(n_msg, dtype=DTYPES[storage], device=self.dev)
                          for d in (1, -1)}
            sl["n"] = local.size
            self.slabs.append(sl)
        self.msg_bytes = self.slabs[0]["send"][1].numel() * self.slabs[0]["send"][1].element_size()
        self.latest = 0
        self.iters = 0

    def exchange(self, which):
        """Fill the halo planes of buffer `which` of every slab from its
        neighbours' own planes of the same buffer."""
        s = stream_ptr(self.dev)
        P = len(self.slabs)
        for sl in self.slabs:
            f = ptr(sl["f"][which])
            call("prt_lbm_halo_pack", self.q, self.st, *sl["dims"], f, sl["m"], 1,
                 ptr(sl["send"][1]), s)
            call("prt_lbm_halo_pack", self.q, self.st, *sl["dims"], f, 1, -1,
                 ptr(sl["send"][-1]), s)
        for p, sl in enumerate(self.slabs):
            left, right = self.slabs[(p - 1) % P], self.slabs[(p + 1) % P]
            f = ptr(sl["f"][which])
            call("prt_lbm_halo_unpack", self.q, self.st, *sl["dims"], f, 0, 1,
                 ptr(left["send"][1]), s)
            call("prt_lbm_halo_unpack", self.q, self.st, *sl["dims"], f, sl["m"] + 1, -1,
                 ptr(right["send"][-1]), s)

    def step_slab(self, p, src, dst):
        """One iteration of slab p alone (its halo planes already filled)."""
        sl = self.slabs[p]
        call("prt_lbm_step", self.q, self.st, *sl["dims"], ptr(sl["solid"]),
             ptr(sl["f"][src]), ptr(sl["f"][dst]), self.om_p, self.om_m, self.force,
             self.fp16_scale, self.flags, None, stream_ptr(self.dev))

    def run(self, steps):
        """`steps` iterations of every slab, each after a halo exchange."""
        for _ in range(int(steps)):
            src, dst = self.latest, 1 - self.latest
            self.exchange(src)
            for p in range(len(self.slabs)):
                self.step_slab(p, src, dst)
            self.latest = dst
            self.iters += 1

    def populations(self):
        """Own planes of every slab's newest buffer, joined: (nx, rows, q,
        zs) in the box's z-row layout."""
        parts = []
        for sl in self.slabs:
            X, Y, Z = sl["dims"] if self.nd == 3 else (sl["dims"][0], 1, sl["dims"][1])
            f = sl["f"][self.latest]
            parts.append(f.reshape(X, Y, self.q, -1)[1:1 + sl["m"]])
        return torch.cat(parts, 0)

    def velocity(self):
        """(3, *grid) float64 velocity of the joined slabs (zero in solid)."""
        self.exchange(self.latest)
        s = stream_ptr(self.dev)
        out = []
        for sl in self.slabs:
            v64 = torch.zeros((3, sl["n"]), dtype=torch.float64, device=self.dev)
            call("prt_lbm_velocity", self.q, self.st, *sl["dims"], ptr(sl["solid"]),
                 ptr(sl["f"][self.latest]), ptr(sl["f"][1 - self.latest]), self.flags, self.force,
                 self.fp16_scale, ptr(v64), None, None, s)
            loc = (3, sl["m"] + 2) + self.shape[1:]
            out.append(v64.reshape(loc)[:, 1:1 + sl["m"]])
        return torch.cat(out, 1)
