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

from ._lib import PRT_LBM_SKIP_SOLID, PRT_LBM_SLAB, PRT_STORE, PrtError, call, lib
from .device import device as _device
from .device import ptr, stream_ptr
from .geometry import PORE
from .stokes import DTYPES, FlowConvergence, trt_minus


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


def slab_dims(shape, x_planes):
    """Kernel grid (nx, ny, nz) of a slab of x_planes own planes of a box of
    this shape, halo planes included."""
    return (x_planes + 2, shape[1], 1 if len(shape) == 2 else shape[2])


def parts_needed(shape, storage):
    """The fewest slabs whose population buffers all fit the kernels'
    32-bit element index (1 when the whole box does)."""
    q = 9 if len(shape) == 2 else 19
    st = PRT_STORE[storage]
    dims = (shape[0], shape[1], 1 if len(shape) == 2 else shape[2])
    if lib().prt_lbm_pop_elems(q, st, *dims) > 0 or min(dims) < 1:
        return 1          # fits (or is refused as a bad grid by the box solver)
    for parts in range(2, shape[0] + 1):
        m = -(-shape[0] // parts)
        if lib().prt_lbm_pop_elems(q, st, *slab_dims(shape, m)) > 0:
            return parts
    raise PrtError("even single x-planes of a %s grid exceed the 32-bit element index"
                   % (shape,))


class SlabSet:
    """The slabs of grid g (2D (nx, ny) for D2Q9 or 3D (nx, ny, nz) for
    D3Q19), periodic along x like the box, on one device.  parts=None takes
    the fewest slabs the kernels' 32-bit element index allows, so one GPU
    can hold a box whose populations exceed 2^31 elements (up to the
    device's memory: 288 GB holds about 0.9G FP64 or 1.8G FP32 nodes)."""

    def __init__(self, g, parts=None, force=2e-6, tau_lb=0.8, magic=3.0 / 16.0, storage="fp64",
                 skip_solid=False, device=None, fp16_scale=1024.0):
        g = np.asarray(g)
        if g.ndim not in (2, 3):
            raise ValueError("grid must be 2D or 3D, got %dD" % g.ndim)
        if parts is None:
            parts = parts_needed(g.shape, storage)
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
            sl["n_fluid"] = int((g[x0:x1] == PORE).sum())
            nblk = int(lib().prt_lbm_launch_blocks(self.q, *sl["dims"], self.flags))
            sl["partials"] = torch.zeros(max(nblk, int(lib().prt_lbm_partials_count(self.q, *sl["dims"]))),
                                         dtype=torch.float64, device=self.dev)
            sl["nblk"] = nblk
            sl["scalar"] = torch.zeros(1, dtype=torch.float64, device=self.dev)
            self.slabs.append(sl)
        self.n = int(g.size)
        self.n_fluid = sum(sl["n_fluid"] for sl in self.slabs)
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

    def step_slab(self, p, src, dst, reduce=False):
        """One iteration of slab p alone (its halo planes already filled);
        with reduce, its own planes' |u_x| block sums go to its partials."""
        sl = self.slabs[p]
        call("prt_lbm_step", self.q, self.st, *sl["dims"], ptr(sl["solid"]),
             ptr(sl["f"][src]), ptr(sl["f"][dst]), self.om_p, self.om_m, self.force,
             self.fp16_scale, self.flags, ptr(sl["partials"]) if reduce else None,
             stream_ptr(self.dev))

    def run(self, steps, reduce=False):
        """`steps` iterations of every slab, each after a halo exchange;
        with reduce, the last one leaves the |u_x| sums (mean_abs_ux)."""
        steps = int(steps)
        for k in range(steps):
            src, dst = self.latest, 1 - self.latest
            self.exchange(src)
            for p in range(len(self.slabs)):
                self.step_slab(p, src, dst, reduce and k == steps - 1)
            self.latest = dst
            self.iters += 1

    def _sum(self, key_count):
        """FP64 sum over the slabs of their partials (count per slab)."""
        s = stream_ptr(self.dev)
        tot = 0.0
        for sl in self.slabs:
            call("prt_sum_f64", ptr(sl["partials"]), key_count(sl), ptr(sl["scalar"]), s)
            tot += float(sl["scalar"].item())
        return tot

    def mean_abs_ux(self):
        """Mean |u_x| over fluid nodes from the last reducing iteration."""
        return self._sum(lambda sl: sl["nblk"]) / self.n_fluid if self.n_fluid else 0.0

    def solve(self, nit=30000, tol=1e-6, info=None, every=50, name=None):
        """Iterate until FlowConvergence says stop or nit iterations ran (as
        StokesSolver.solve)."""
        conv = FlowConvergence(tol=tol, every=every, max_it=nit,
                               name=name or "%dD flow" % self.nd)
        it, nit = 0, int(nit)
        if conv.tol <= 0:
            self.run(nit)
            conv.iters = nit
        else:
            while it < nit:
                chunk = min(conv.every - (it % conv.every), nit - it)
                at_check = (it + chunk) % conv.every == 0
                self.run(chunk, reduce=at_check)
                it += chunk
                if at_check:
                    if conv.done(it - 1, self.mean_abs_ux()):
                        break
                else:
                    conv.iters = it
        conv.report(info)
        return conv

    def velocity32(self):
        """(float32 (3, *grid) device field, zero in solid; FP64 sum of u_x
        over every node) of the joined slabs."""
        self.exchange(self.latest)
        s = stream_ptr(self.dev)
        v = torch.empty((3,) + self.shape, dtype=torch.float32, device=self.dev)
        sux = 0.0
        for sl, (x0, x1) in zip(self.slabs, self.bounds):
            v32 = torch.empty((3, sl["n"]), dtype=torch.float32, device=self.dev)
            call("prt_lbm_velocity", self.q, self.st, *sl["dims"], ptr(sl["solid"]),
                 ptr(sl["f"][self.latest]), ptr(sl["f"][1 - self.latest]), self.flags, self.force,
                 self.fp16_scale, None, ptr(v32), ptr(sl["partials"]), s)
            v[:, x0:x1] = v32.reshape((3, sl["m"] + 2) + self.shape[1:])[:, 1:1 + sl["m"]]
            nblk = int(lib().prt_lbm_launch_blocks(self.q, *sl["dims"], PRT_LBM_SLAB))
            call("prt_sum_f64", ptr(sl["partials"]), nblk, ptr(sl["scalar"]), s)
            sux += float(sl["scalar"].item())
            del v32
        return v, sux

    def permeability(self, sum_ux):
        """Darcy permeability (lattice units): nu <u_x>_box / force."""
        return (self.tau - 0.5) / 3.0 * (sum_ux / self.n) / self.force

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
