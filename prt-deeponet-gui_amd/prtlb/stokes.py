"""D2Q9 / D3Q19 TRT Stokes on the GPU behind the reference's entry points.

stokes_d2q9 / stokes_d3q19 keep the signatures, defaults, return type
(float32 (3, *grid), zero in solid), info keys, exceptions and stdout
warnings of 3D/tools/prtlb_2d.py:348 and 3D/tools/prtlb_3d.py:348.  The
iteration runs in libprtlb_hip.so (fused pull-stream + TRT collide + Guo
force + solid reversal, one launch per iteration); the host only steers the
convergence check every 50 iterations.
"""
import numpy as np
import torch

from ._lib import PRT_LBM_SKIP_SOLID, PRT_STORE, PrtError, call, lib
from .device import device as _device
from .device import ptr, stream_ptr
from .geometry import PORE

DTYPES = {"fp64": torch.float64, "fp32": torch.float32, "fp16": torch.float16}


def trt_minus(om_p, magic=3.0 / 16.0):
    """omega- such that (1/omega+ - 1/2)(1/omega- - 1/2) = magic.
    Raises ValueError when tau <= 1/2 (3D/tools/prtlb_3d.py:130-139)."""
    lam_p = 1.0 / float(om_p) - 0.5
    if lam_p <= 0:
        raise ValueError("tau must exceed 0.5; the viscosity (tau-1/2)/3 is "
                         "negative below it and the solver cannot run.")
    return 1.0 / (float(magic) / lam_p + 0.5)


class FlowConvergence:
    """Stop when the mean pore |u_x| changes by at most `tol` (relative) over
    `every` iterations; a tolerance <= 0 never stops early.  Same rule, info
    keys and cap warning as 3D/tools/prtlb_3d.py:168-252."""

    def __init__(self, tol=1e-6, every=50, max_it=100000, name="flow"):
        self.tol = float(tol)
        self.every = max(int(every), 1)
        self.max_it = int(max_it)
        self.name = name
        self.prev = None
        self.iters = 0
        self.converged = False
        self.change = float("inf")

    def done(self, it, umean):
        self.iters = it + 1
        if self.tol <= 0 or (it + 1) % self.every:
            return False
        u = float(umean)
        if self.prev is not None:
            self.change = abs(u - self.prev) / max(abs(u), 1e-30)
            if self.change <= self.tol:
                self.converged = True
                self.prev = u
                return True
        self.prev = u
        return False

    def report(self, info=None):
        if info is not None:
            info["flow_iters"] = self.iters
            info["flow_converged"] = bool(self.converged)
            info["flow_change"] = self.change
        if not self.converged and self.tol > 0:
            print("     WARNING: the %s solve stopped at its %d iteration cap "
                  "with the mean\n              speed still changing by %.3g "
                  "per %d steps, against a\n              tolerance of %.3g. "
                  "The field is not settled and the\n              "
                  "permeability read from it will be too low. Raise\n"
                  "              ns_max_iT." % (self.name, self.iters, self.change,
                                                self.every, self.tol), flush=True)
        return self.converged


class StokesSolver:
    """Device-resident body-force TRT Stokes state for one geometry.

    g         material grid (nx, ny) -> D2Q9 or (nx, ny, nz) -> D3Q19; any
              code other than PORE is solid (3D/tools/prtlb_3d.py:368)
    storage   "fp64" (the reference's arithmetic), "fp32" (f - w stored in
              float32, float32 collide) or "fp16" ((f - w) s / w stored in
              IEEE half, float32 collide)
    skip_solid  update fluid nodes only; a fluid node takes its bounce-back
              populations from its own slot two iterations back, which is
              exactly what the reference's stored solid nodes hand back
    """

    def __init__(self, g, force=2e-6, tau_lb=0.8, magic=3.0 / 16.0,
                 storage="fp64", skip_solid=False, device=None, fp16_scale=1024.0):
        self.dev = _device(device)
        if isinstance(g, torch.Tensor):
            solid_t = (g != PORE).to(self.dev, torch.uint8)
            shape = tuple(g.shape)
        else:
            g = np.asarray(g)
            shape = g.shape
            solid_t = torch.from_numpy(np.ascontiguousarray(g != PORE).astype(np.uint8)).to(self.dev)
        if len(shape) not in (2, 3):
            raise ValueError("grid must be 2D or 3D, got %dD" % len(shape))
        if storage not in PRT_STORE:
            raise ValueError("storage must be one of %s" % sorted(PRT_STORE))
        self.shape = shape
        self.nd = len(shape)
        self.q = 9 if self.nd == 2 else 19
        self.dims = (shape[0], shape[1], 1 if self.nd == 2 else shape[2])
        self.n = int(np.prod(shape))
        self.storage = storage
        self.st = PRT_STORE[storage]
        self.solid = solid_t.reshape(-1).contiguous()
        self.n_fluid = int(self.n - int(self.solid.sum().item()))
        self.force = float(force)
        self.tau = float(tau_lb)
        self.om_p = 1.0 / self.tau
        self.om_m = trt_minus(self.om_p, magic)
        self.flags = PRT_LBM_SKIP_SOLID if skip_solid else 0
        self.fp16_scale = float(fp16_scale)
        dt = DTYPES[storage]
        elems = int(lib().prt_lbm_pop_elems(self.q, self.st, *self.dims))
        if elems < 1:
            raise PrtError("prt_lbm_pop_elems: " + lib().prt_last_error().decode(errors="replace"))
        self.f = [torch.empty(elems, dtype=dt, device=self.dev) for _ in range(2)]
        # z-row layout: rows (x, y) of the kernel's (X, Y, Z) grid, each
        # holding q runs of zs elements (a 2D grid is (nx, 1, ny))
        self.rows = (shape[0], 1, shape[1]) if self.nd == 2 else tuple(shape)
        self.zs = elems // (self.rows[0] * self.rows[1] * self.q)
        self.tune = 0
        self.latest = 0          # index of the buffer holding the newest state
        self.iters = 0
        npart = lib().prt_lbm_partials_count(self.q, *self.dims)
        self.partials = torch.zeros(npart, dtype=torch.float64, device=self.dev)
        self.scalar = torch.zeros(4, dtype=torch.float64, device=self.dev)
        s = self._stream()
        for buf in self.f:
            call("prt_lbm_fill_rest", self.q, self.st, *self.dims, ptr(buf), s)

    def _stream(self):
        return stream_ptr(self.dev)

    def run(self, steps, reduce=False):
        """Advance `steps` iterations; with reduce, the last one leaves the
        per-block sums of |u_x| over fluid nodes in self.partials."""
        if steps <= 0:
            return
        a, b = self.f[self.latest], self.f[1 - self.latest]
        call("prt_lbm_run", self.q, self.st, *self.dims, ptr(self.solid), ptr(a), ptr(b),
             int(steps), self.om_p, self.om_m, self.force, self.fp16_scale,
             self.flags | self.tune, ptr(self.partials) if reduce else None, self._stream())
        if steps % 2:
            self.latest = 1 - self.latest
        self.iters += steps

    def _launch(self, steps, reduce):
        a, b = self.f[self.latest], self.f[1 - self.latest]
        call("prt_lbm_run", self.q, self.st, *self.dims, ptr(self.solid), ptr(a), ptr(b),
             int(steps), self.om_p, self.om_m, self.force, self.fp16_scale,
             self.flags | self.tune, ptr(self.partials) if reduce else None, self._stream())

    def tick_graph(self, every=50):
        """`every` iterations (even) ending in the |u_x| reduction and its
        device-side sum, replayed from a captured HIP graph: one host launch
        per tick instead of every + 1.  Returns nothing; read mean_abs_ux_ready()."""
        if every % 2:
            raise ValueError("a captured tick must have an even length")
        key = (every, self.latest, self.tune)
        if getattr(self, "_graph_key", None) != key:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._launch(every, True)
                call("prt_sum_f64", ptr(self.partials), self.partials.numel(), ptr(self.scalar),
                     self._stream())
            self._graph, self._graph_key = g, key
        self._graph.replay()
        self.iters += every

    def mean_abs_ux_ready(self):
        """Mean |u_x| over fluid nodes from the sum a captured tick left."""
        s = float(self.scalar[0].item())
        return s / self.n_fluid if self.n_fluid else 0.0

    def mean_abs_ux(self):
        """Mean |u_x| over fluid nodes from the last reducing iteration."""
        call("prt_sum_f64", ptr(self.partials), self.partials.numel(), ptr(self.scalar),
             self._stream())
        s = float(self.scalar[0].item())
        return s / self.n_fluid if self.n_fluid else 0.0

    def solve(self, nit=30000, tol=1e-6, info=None, every=50, name=None, graph=None):
        """Iterate until FlowConvergence says stop or nit iterations ran.

        graph: replay each convergence interval from a captured HIP graph
        (default: for grids under 2^20 nodes, where launches dominate)."""
        conv = FlowConvergence(tol=tol, every=every, max_it=nit,
                               name=name or "%dD flow" % self.nd)
        if graph is None:
            graph = self.n < (1 << 20) and conv.every % 2 == 0
        it = 0
        nit = int(nit)
        if conv.tol <= 0:
            self.run(nit)
            conv.iters = nit
        else:
            while it < nit:
                chunk = min(conv.every - (it % conv.every), nit - it)
                at_check = (it + chunk) % conv.every == 0
                if graph and at_check and chunk == conv.every:
                    self.tick_graph(chunk)
                    it += chunk
                    um = self.mean_abs_ux_ready() if self.n_fluid else 0.0
                    if conv.done(it - 1, um):
                        break
                    continue
                self.run(chunk, reduce=at_check)
                it += chunk
                if at_check:
                    um = self.mean_abs_ux() if self.n_fluid else 0.0
                    if conv.done(it - 1, um):
                        break
                else:
                    conv.iters = it
        conv.report(info)
        return conv

    def velocity(self, f64=False):
        """(v32, v64-or-None, sum of u_x over every node) of the current state."""
        v32 = torch.empty((3, self.n), dtype=torch.float32, device=self.dev)
        v64 = torch.empty((3, self.n), dtype=torch.float64, device=self.dev) if f64 else None
        call("prt_lbm_velocity", self.q, self.st, *self.dims, ptr(self.solid),
             ptr(self.f[self.latest]), ptr(self.f[1 - self.latest]), self.flags, self.force,
             self.fp16_scale, ptr(v64), ptr(v32), ptr(self.partials), self._stream())
        call("prt_sum_f64", ptr(self.partials), self.partials.numel(), ptr(self.scalar),
             self._stream())
        sux = float(self.scalar[0].item())
        shp = (3,) + self.shape
        return v32.reshape(shp), (v64.reshape(shp) if v64 is not None else None), sux

    def permeability(self, sum_ux):
        """Darcy permeability (lattice units): nu <u_x>_box / force."""
        nu = (self.tau - 0.5) / 3.0
        return nu * (sum_ux / self.n) / self.force

    def populations(self):
        """The newest post-collision populations as stored, gathered from
        the z-row layout into a (q, *grid) tensor."""
        X, Y, Z = self.rows
        f = self.f[self.latest].view(X, Y, self.q, self.zs)[..., :Z]
        return f.permute(2, 0, 1, 3).reshape((self.q,) + self.shape)


def checked_flow(v, solid, tau_lb, force, name="flow"):
    """Refuse a non-finite, empty or dead field; warn above Mach 0.05.

    Same refusals (SystemExit) and warning keyword as
    3D/tools/prtlb_3d.py:267-316, evaluated on the device."""
    m = ~solid
    fin = torch.isfinite(v)
    if not bool(fin.all()):
        raise SystemExit(
            "the %s solver produced %d non-finite values. The driving force "
            "of %g is too large for tau = %g: lower it, or raise tau. Nothing "
            "downstream can use this field." % (name, int((~fin).sum()), force, tau_lb))
    if not bool(m.any()):
        raise SystemExit("the %s solver was given a geometry with no pore "
                         "space at all." % name)
    sp = torch.sqrt((v * v).sum(0))[m]
    umax = float(sp.max())
    umean = float(sp.double().mean())
    if umean <= 0:
        raise SystemExit(
            "the %s solver produced a field that is zero everywhere. Either "
            "the driving force (%g) is too small to survive rounding, or the "
            "pore space does not connect inlet to outlet." % (name, force))
    ma = umax / np.sqrt(1.0 / 3.0)
    if ma > 0.05:
        print("     WARNING: peak lattice speed %.4g gives a Mach number of "
              "%.3f.\n              Lattice Boltzmann is a low-Mach "
              "expansion and is only\n              reliable below about "
              "0.05. Halve the pressure drop and the\n              flow "
              "halves with it, so nothing is lost but time." % (umax, ma), flush=True)
    return v


def _stokes(g, nd, nit, force, tau_lb, magic, tol, info, storage, skip_solid,
            return_f64, as_tensor, device, graph=None):
    if isinstance(g, torch.Tensor):
        if g.dim() != nd:
            raise ValueError("stokes_d%s needs a %dD grid" % ("2q9" if nd == 2 else "3q19", nd))
    elif np.asarray(g).ndim != nd:
        raise ValueError("stokes_d%s needs a %dD grid" % ("2q9" if nd == 2 else "3q19", nd))
    shape = tuple(g.shape)
    if not isinstance(g, torch.Tensor):
        from .slab import parts_needed
        if parts_needed(shape, storage) > 1:
            return _stokes_slabs(np.asarray(g), nd, nit, force, tau_lb, magic, tol, info, storage,
                                 skip_solid, return_f64, as_tensor, device)
    s = StokesSolver(g, force=force, tau_lb=tau_lb, magic=magic, storage=storage,
                     skip_solid=skip_solid, device=device)
    s.solve(nit=nit, tol=tol, info=info, name="%dD flow" % nd, graph=graph)
    v32, v64, sux = s.velocity(f64=return_f64)
    solid = s.solid.reshape(s.shape).bool()
    checked_flow(v32, solid, tau_lb, force)
    if info is not None:
        info["permeability"] = s.permeability(sux)
        info["porosity"] = s.n_fluid / s.n
        info["mean_ux_pore"] = (sux / s.n_fluid) if s.n_fluid else 0.0
        info["storage"] = storage
    if as_tensor:
        return (v32, v64) if return_f64 else v32
    out = v32.cpu().numpy()
    if return_f64:
        return out, v64.cpu().numpy()
    return out


def _stokes_slabs(g, nd, nit, force, tau_lb, magic, tol, info, storage, skip_solid, return_f64,
                  as_tensor, device):
    """A grid whose populations exceed the kernels' 32-bit element index:
    the same solve on x-slabs of one device (prtlb.slab.SlabSet), the same
    bits per iteration as one box; float32 output only."""
    from .slab import SlabSet
    if return_f64:
        raise PrtError("return_f64 is not offered for grids beyond the 32-bit element index")
    S = SlabSet(g, None, force=force, tau_lb=tau_lb, magic=magic, storage=storage,
                skip_solid=skip_solid, device=device)
    S.solve(nit=nit, tol=tol, info=info, name="%dD flow" % nd)
    v32, sux = S.velocity32()
    solid = torch.from_numpy(np.ascontiguousarray(g != PORE)).to(S.dev)
    checked_flow(v32, solid, tau_lb, force)
    del solid
    if info is not None:
        info["permeability"] = S.permeability(sux)
        info["porosity"] = S.n_fluid / S.n
        info["mean_ux_pore"] = (sux / S.n_fluid) if S.n_fluid else 0.0
        info["storage"] = storage
        info["slabs"] = len(S.slabs)
    return v32 if as_tensor else v32.cpu().numpy()


def stokes_d2q9(g, nit=30000, force=2e-6, tau_lb=0.8, magic=3.0 / 16.0, tol=1e-6,
                info=None, *, storage="fp64", skip_solid=False, return_f64=False,
                as_tensor=False, device=None, graph=None):
    """D2Q9 TRT Stokes flow on the GPU; drop-in for 3D/tools/prtlb_2d.py:348.

    Returns v float32 (3, nx, ny) with v[2] = 0 and zero in solid.  Extra
    keyword-only options: storage ("fp64" default, "fp32", "fp16"),
    skip_solid, return_f64 (also return the FP64 field), as_tensor (device
    tensors instead of NumPy).  info gains permeability, porosity,
    mean_ux_pore and storage."""
    return _stokes(g, 2, nit, force, tau_lb, magic, tol, info, storage, skip_solid,
                   return_f64, as_tensor, device, graph)


def stokes_d3q19(g, nit=30000, force=2e-6, tau_lb=0.8, magic=3.0 / 16.0, tol=1e-6,
                 info=None, *, storage="fp64", skip_solid=False, return_f64=False,
                 as_tensor=False, device=None, graph=None):
    """D3Q19 TRT Stokes flow on the GPU; drop-in for 3D/tools/prtlb_3d.py:348.

    Returns v float32 (3, nx, ny, nz), zero in solid.  Same extra options as
    stokes_d2q9."""
    return _stokes(g, 3, nit, force, tau_lb, magic, tol, info, storage, skip_solid,
                   return_f64, as_tensor, device, graph)


def plane_flux(v64):
    """Sum of u_x over each x-plane of a (3, nx, ...) FP64 device field."""
    if not isinstance(v64, torch.Tensor) or v64.dtype != torch.float64 or not v64.is_cuda:
        raise PrtError("plane_flux takes the FP64 device field")
    nx = v64.shape[1]
    rest = int(np.prod(v64.shape[2:]))
    out = torch.empty(nx, dtype=torch.float64, device=v64.device)
    call("prt_plane_flux", ptr(v64.contiguous()), nx, rest, 1, ptr(out), stream_ptr(v64.device))
    return out
