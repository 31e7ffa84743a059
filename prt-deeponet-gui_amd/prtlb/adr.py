"""Advection-diffusion-reaction on the GPU behind the reference's solve_adr.

solve_adr keeps the signature, defaults, return values (out float32
(n_t, ns, *grid), t_norm float32), info keys, exceptions and stdout warnings
of 3D/tools/prtlb_3d.py:781-1595.  The set-up scalars (velocity scale,
D, k, ka, kd, dt, face diffusivities) are formed on the host with exactly
the float32 NumPy expressions the reference uses -- they are part of the
interface, and forming them identically is what makes the field bit-exact --
and every step runs in libprtlb_hip.so as one fused kernel.  The stop rule,
the snapshot ladder and the snapshot pick are the reference's, driven from
the host; the per-species pore means they read are NumPy-ordered float32
sums evaluated on the device, so an auto-stopped run stops at the same step.
"""
import ctypes

import numpy as np
import torch

from ._lib import AdrParams, PrtError, call, lib, pairwise_plan
from .chemistry import default_chemistry
from .device import device as _device
from .device import ptr, stream_ptr
from .geometry import PORE

ATTACHED = 3


def _name(ch, j):
    try:
        return str(ch.names[j])
    except Exception:
        return ("the donor", "the acceptor", "the product",
                "the biomass")[j] if j < 4 else "chemical %d" % j


def nearest_unused(i, used, n):
    """Closest index to i not yet taken (3D/tools/prtlb_3d.py:716-722)."""
    for d in range(n):
        for j in (i - d, i + d):
            if 0 <= j < n and j not in used:
                return j
    return i


def wanted_steps(total, n_t):
    """Log-spaced snapshot steps 0..total (3D/tools/prtlb_3d.py:725-737)."""
    if n_t <= 1:
        return np.array([total], int)
    if n_t == 2:
        return np.array([0, total], int)
    tail = np.geomspace(max(total // 400, 1), max(total, 1), n_t - 1)
    return np.round(np.concatenate([[0.0], tail])).astype(int)


def pick_snapshots(keep_steps, total, n_t):
    """Indices into the ladder for the n_t stored snapshots."""
    ks = np.asarray(keep_steps)
    used, picked = set(), []
    for w in wanted_steps(total, n_t):
        i = int(np.abs(ks - w).argmin())
        if i in used:
            i = nearest_unused(i, used, len(ks))
        used.add(i)
        picked.append(i)
    return picked


def ladder(limit, ratio):
    """Steps a run of exactly `limit` steps records (3D/tools/prtlb_3d.py:1488-1546)."""
    out, nxt = [0], 1
    n = nxt
    while n <= limit:
        out.append(n)
        nxt = max(n + 1, int(np.ceil(n * ratio)))
        n = nxt
    if out[-1] != limit:
        out.append(limit)
    return out


class TransportSolver:
    """Device-resident state of one transport solve (one geometry, one flow).

    precision  "fp32" (the reference's arithmetic, bit-exact) or "fp64"
    sorption   None, or dict(da_sorb=..., da_desorb=..., c=0, s=1): reversible
               sorption of mobile chemical c onto immobile s (build-defined;
               the reference has no reversible sorption)
    """

    def __init__(self, mat, vel, pe, da, n_species=2, chem=None, da_abio=0.0,
                 precision="fp32", sorption=None, device=None):
        self._setup_host(mat, vel, pe, da, n_species, chem, da_abio, precision, sorption)
        self._setup_device(device)

    def _setup_host(self, mat, vel, pe, da, n_species, chem, da_abio, precision, sorption):
        """Everything solve_adr forms once, on the host, before the loop."""
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be fp32 or fp64")
        R = np.float32 if precision == "fp32" else np.float64
        self.R, self.precision = R, precision

        # ---- 9.1 the pore space (3D/tools/prtlb_3d.py:831-865)
        mat = np.asarray(mat)
        bf_code = getattr(chem, "biofilm_code", None) if chem is not None else None
        m = mat == PORE
        biofilm = np.zeros(m.shape, bool)
        if bf_code is not None:
            biofilm = mat == bf_code
            m = m | biofilm
        nd = m.ndim
        if nd not in (2, 3):
            raise ValueError("grid must be 2D or 3D, got %dD" % nd)
        nx = m.shape[0]
        vel = np.asarray(vel, R)
        sp = np.sqrt((vel[:nd] ** 2).sum(0))
        mean_axial = float(vel[0][m].mean()) if m.any() else 0.0
        mean_speed = float(sp[m].mean()) if m.any() else 0.0
        u_scale = mean_axial if mean_axial > 0.1 * mean_speed else mean_speed
        u = vel[:nd] / max(u_scale, 1e-30)
        u = np.where(m[None], u, 0.0)

        ns = int(n_species)
        ch = default_chemistry(ns) if chem is None else chem.truncate(ns)
        if len(ch.names) != ns:
            raise ValueError("the chemistry describes %d chemicals but %d were "
                             "asked for" % (len(ch.names), ns))
        if ns > 8:
            raise ValueError("at most 8 chemicals are supported, got %d" % ns)

        # ---- 9.2 lattice quantities (3D/tools/prtlb_3d.py:896-971)
        l_ref = float(getattr(ch, "l_ref_voxels", None) or nx)
        D = R(l_ref / max(pe, 1e-6))
        pe_grid = float(pe) / l_ref
        if pe_grid > 2.0:
            print("     WARNING: Peclet %.4g on %d voxels is a grid Peclet number "
                  "of %.1f.\n              The front is thinner than the mesh; the "
                  "sample will behave\n              as though its Peclet number "
                  "were lower than %.4g. Use a\n              longer domain, or a "
                  "lower Peclet number." % (pe, int(l_ref), pe_grid, pe), flush=True)
        k = R(da / l_ref)
        ka = R(max(da_abio, 0.0) / l_ref)
        j_ab = getattr(ch, "abiotic_index", 2)
        if ka > 0 and j_ab >= ns:
            raise ValueError(
                "an abiotic Damkohler number of %g was given, but it consumes "
                "chemical %d (%s) and this run has only %d chemicals, so the "
                "reaction cannot happen. Ask for at least %d chemicals, or set "
                "the abiotic reactant to one that is in the run, or set the "
                "abiotic Damkohler number to zero."
                % (da_abio, j_ab, _name(ch, j_ab), ns, j_ab + 1))
        kd = R(ch.decay_for(l_ref))
        d_rel = np.asarray(ch.d_rel, R).copy()
        d_rel[ch.immobile] = 0.0
        has_bf = bool(biofilm.any())
        if has_bf:
            d_bf = np.asarray(ch.d_biofilm_rel, R).copy()
            d_bf[ch.immobile] = 0.0
            d_worst = float(max(d_rel.max(), d_bf.max()))
        else:
            d_bf = d_rel
            d_worst = float(d_rel.max())

        # ---- 9.3 the time step (3D/tools/prtlb_3d.py:976-993)
        dmax = float(D) * float(max(d_worst, 1.0))
        umax = float(sum(np.abs(u[ax]).max() for ax in range(nd)))
        k_lin = float(k)
        if ch.ks_donor > 0:
            k_lin = max(k_lin, float(k) / float(ch.ks_donor))
        if ns > 1 and ch.ks_acceptor > 0:
            k_lin = max(k_lin, float(k) / float(ch.ks_acceptor))
        k_sorb = k_des = R(0.0)
        if sorption:
            k_sorb = R(float(sorption.get("da_sorb", 0.0)) / l_ref)
            k_des = R(float(sorption.get("da_desorb", 0.0)) / l_ref)
            dt = R(0.4 / (umax + 2.0 * nd * dmax + k_lin + ka + kd + k_sorb + k_des + 1e-30))
        else:
            # NEP 50: the Python-float partial sum takes the array dtype at `+ ka`
            dt = R(0.4 / (umax + 2.0 * nd * dmax + k_lin + ka + kd + 1e-30))

        fed = np.asarray(ch.inlet, R)
        moving = [j for j in range(ns) if not ch.immobile[j]]
        surface_only = bool(getattr(ch, "abiotic_surface_only", False))
        if ka > 0 and surface_only:
            self.surface_voxels = self._surface_sum(m, R)
        else:
            self.surface_voxels = 0.0

        # ---- kernel parameters
        p = AdrParams()
        p.nd, p.ns = nd, ns
        p.biotic_on = int(bool(getattr(ch, "biotic_on", True)))
        p.coupled = int(bool(ch.biomass_coupled and ns > 3 and ch.c_init[3] > 0))
        p.abio_index = j_ab if (ka > 0 and ns > j_ab) else -1
        p.abio_target = j_ab if j_ab < ns else -1
        mk = getattr(ch, "abiotic_product_index", None)
        p.abio_product = mk if (j_ab < ns and mk is not None and mk < ns and mk != j_ab) else -1
        p.surface_only = int(surface_only)
        p.sorption = int(bool(sorption))
        p.sorb_c = int(sorption.get("c", 0)) if sorption else 0
        p.sorb_s = int(sorption.get("s", 1)) if sorption else 0
        if sorption and not (0 <= p.sorb_c < ns and 0 <= p.sorb_s < ns):
            raise ValueError("sorption chemicals must be in the run")
        p.write_solid = 0
        p.dt, p.k, p.ka, p.kd = float(dt), float(k), float(ka), float(kd)
        p.ks_donor, p.ks_acceptor = float(ch.ks_donor), float(ch.ks_acceptor)
        p.alpha_dt = float(float(ch.alpha) * dt)
        p.neg_alpha = -float(ch.alpha)
        p.alpha_pos = 1.0 if ch.alpha > 0 else 0.0
        p.c2 = 1.0 - (ch.yield_ if ns > 3 else 0.0)
        p.yield_ = float(ch.yield_)
        p.c3_init = float(ch.c_init[3]) if ns > 3 else 1.0
        p.k_sorb, p.k_desorb = float(k_sorb), float(k_des)
        for j in range(ns):
            p.mobile[j] = int(j in moving)
            p.pinned[j] = int(j in moving and fed[j] == fed[j])
            p.fed[j] = float(fed[j]) if fed[j] == fed[j] else 0.0
            a = (d_rel[j] * D, d_bf[j] * D)      # the two cell diffusivities
            for b1 in (0, 1):
                for b2 in (0, 1):
                    a1, a2 = R(a[b1]), R(a[b2])
                    h = (2 * a1 * a2 / (a1 + a2)) if (a1 > 0 and a2 > 0) else R(0.0)
                    p.dface[j][2 * b1 + b2] = float(h)
        self.params = p

        # ---- initial state (3D/tools/prtlb_3d.py:1425-1440)
        C = np.zeros((ns,) + m.shape, R)
        for j in range(ns):
            if not ch.c_init[j]:
                continue
            if j == ATTACHED and ch.seed_in_biofilm and has_bf and ch.immobile[j]:
                C[j] = np.where(biofilm, ch.c_init[j], 0.0)
            else:
                C[j] = np.where(m, ch.c_init[j], 0.0)

        self.m, self.biofilm, self.ch = m, biofilm, ch
        self.shape, self.nd, self.ns, self.nx = m.shape, nd, ns, nx
        self.dims = (m.shape[0], m.shape[1], 1 if nd == 2 else m.shape[2])
        self.n = m.size
        self.dt, self.D, self.k, self.ka, self.kd = dt, D, k, ka, kd
        self.l_ref, self.pe_grid, self.umax, self.dmax = l_ref, pe_grid, umax, dmax
        self.u_scale = u_scale
        self.u_host = np.ascontiguousarray(u, R).reshape(nd, -1)
        self.c_host = C.reshape(ns, -1)
        self.flags_host = (m.astype(np.uint8) | (biofilm.astype(np.uint8) << 1)).reshape(-1)
        self.pore_idx_host = np.flatnonzero(m.reshape(-1)).astype(np.int32)
        self.n_pore = int(self.pore_idx_host.size)

    def _setup_device(self, device):
        """Upload the set-up state; both buffers start from the initial field
        (zero outside the pore, which the kernel then never writes)."""
        self.dev = _device(device)
        self.flags = torch.from_numpy(np.ascontiguousarray(self.flags_host)).to(self.dev)
        self.u = torch.from_numpy(self.u_host).to(self.dev)
        c0 = torch.from_numpy(np.ascontiguousarray(self.c_host)).to(self.dev)
        self.c = [c0, c0.clone()]
        self.latest = 0
        self.steps_done = 0
        self.pore_idx = torch.from_numpy(self.pore_idx_host).to(self.dev)
        # pore table: per pore voxel its index, face bits / reach codes and
        # face velocities, built once from flags and u (an opaque buffer of
        # aligned per-entry records)
        fp64 = self.u.dtype == torch.float64
        nbytes = int(lib().prt_adr_table_bytes(self.nd, int(fp64), self.n_pore))
        self.table = torch.empty(max(nbytes, 16), dtype=torch.uint8, device=self.dev)
        call("prt_adr_pore_table_f64" if fp64 else "prt_adr_pore_table",
             self.nd, *self.dims, ptr(self.flags), ptr(self.u), ptr(self.pore_idx), self.n_pore,
             ptr(self.table), self._stream())
        self._plan = None
        self.means_scratch = None
        self.use_graph = self.n < (1 << 20)
        # immobile chemicals are read and written at their own voxel only:
        # the pore-list launch keeps their rows in list order (contiguous)
        self._immobile = [j for j in range(self.ns) if not self.params.mobile[j]]
        self._pore_list = False
        self.pore_list = True

    @property
    def pore_list(self):
        """True: one thread per pore voxel (the default); False: one thread per
        box voxel.  Switching converts the immobile rows' layout."""
        return self._pore_list

    @pore_list.setter
    def pore_list(self, on):
        on = bool(on)
        if on == self._pore_list:
            return
        box = [self._to_box(c) for c in self.c] if self._pore_list else list(self.c)
        self._pore_list = on
        self.params.immobile_list = int(on and bool(self._immobile))
        self.c = [self._from_box(c) for c in box] if on else box
        self._graph_key = None

    def _from_box(self, c):
        """(ns, n) box field -> the layout the current launch reads."""
        if not (self._pore_list and self._immobile):
            return c
        out = c.clone()
        for j in self._immobile:
            row = c[j][self.pore_idx.long()]
            out[j].zero_()
            out[j][:self.n_pore] = row
        return out

    def _to_box(self, c):
        """The launch's layout -> an (ns, n) box field."""
        if not (self._pore_list and self._immobile):
            return c
        out = c.clone()
        for j in self._immobile:
            row = c[j][:self.n_pore].clone()
            out[j].zero_()
            out[j][self.pore_idx.long()] = row
        return out

    def load_state(self, c):
        """Set both buffers to the (ns, n) box field c."""
        c = torch.as_tensor(c, dtype=self.c[0].dtype).to(self.dev).reshape(self.ns, self.n)
        v = self._from_box(c)
        self.c[0].copy_(v)
        self.c[1].copy_(v)

    @staticmethod
    def _surface_sum(m, R):
        """Sum of the surface weights (info["surface_voxels"], :1000-1018)."""
        nd = m.ndim
        nface = np.zeros(m.shape, R)
        mp = np.pad(m, 1, mode="edge")
        for ax in range(nd):
            lo = [slice(1, -1)] * nd
            hi = [slice(1, -1)] * nd
            lo[ax], hi[ax] = slice(0, -2), slice(2, None)
            nface += (~mp[tuple(lo)]).astype(R)
            nface += (~mp[tuple(hi)]).astype(R)
        return float(np.where(m, nface / float(2 * nd), 0.0).astype(R).sum())

    def _stream(self):
        return stream_ptr(self.dev)

    def current(self):
        """The newest concentration field, (ns, n) on the device (a box
        field; a copy when immobile rows are held in list order)."""
        return self._to_box(self.c[self.latest])

    GRAPH_STEPS = 32

    def _launch_run(self, a, b, steps):
        """`steps` float32 steps a -> b -> ... (pore-list launch by default)."""
        pp = ctypes.byref(self.params)
        if self.pore_list:
            call("prt_adr_run_list", pp, *self.dims, ptr(self.table), self.n_pore, ptr(a), ptr(b),
                 int(steps), stream_ptr(self.dev))
        else:
            call("prt_adr_run", pp, *self.dims, ptr(self.flags), ptr(self.u), ptr(a), ptr(b),
                 int(steps), stream_ptr(self.dev))

    def _graph_steps(self):
        """Replay GRAPH_STEPS (even) steps from a captured HIP graph."""
        key = (self.latest, self.pore_list)
        if getattr(self, "_graph_key", None) != key:
            a, b = self.c[self.latest], self.c[1 - self.latest]
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._launch_run(a, b, self.GRAPH_STEPS)
            self._graph, self._graph_key = g, key
        self._graph.replay()
        self.steps_done += self.GRAPH_STEPS

    def step_n(self, steps):
        """Advance `steps` explicit steps on the device.  On grids under 2^20
        voxels, whole multiples of GRAPH_STEPS replay a captured HIP graph."""
        if steps <= 0:
            return
        if self.use_graph and self.precision == "fp32":
            while steps >= self.GRAPH_STEPS:
                self._graph_steps()
                steps -= self.GRAPH_STEPS
            if steps <= 0:
                return
        a, b = self.c[self.latest], self.c[1 - self.latest]
        pp = ctypes.byref(self.params)
        if self.precision == "fp32":
            self._launch_run(a, b, int(steps))
        else:
            for s in range(steps):
                src, dst = (a, b) if s % 2 == 0 else (b, a)
                call("prt_adr_step_list_f64", pp, *self.dims, ptr(self.table), self.n_pore,
                     ptr(src), ptr(dst), self._stream())
        if steps % 2:
            self.latest = 1 - self.latest
        self.steps_done += steps

    def _pore_sums(self, out=None):
        """Enqueue the float32 pore sums of the newest field, in NumPy's
        pairwise order, into the device tensor out (ns floats; default the
        solver's own); returns out.  A row kept in list order is summed in
        place, through an identity index, instead of through a box copy."""
        if self._plan is None:
            plan, L, I, H = pairwise_plan(self.n_pore)
            self._plan = (torch.from_numpy(plan).to(self.dev), L, I, H)
            self.means_scratch = torch.empty(self.ns * (L + I), dtype=torch.float32,
                                             device=self.dev)
            self.means_out = torch.empty(self.ns, dtype=torch.float32, device=self.dev)
        if out is None:
            out = self.means_out
        plan, L, I, H = self._plan
        if self._pore_list and self._immobile:
            if getattr(self, "_ident", None) is None:
                self._ident = torch.arange(self.n_pore, dtype=torch.int32, device=self.dev)
            c = self.c[self.latest]
            for j in range(self.ns):
                idx = self._ident if j in self._immobile else self.pore_idx
                call("prt_pairwise_sum_f32", ptr(c[j]), self.n, 1, ptr(idx), self.n_pore,
                     ptr(plan), L, I, H, ptr(self.means_scratch), ptr(out[j:j + 1]),
                     self._stream())
            return out
        call("prt_pairwise_sum_f32", ptr(self.current()), self.n, self.ns, ptr(self.pore_idx),
             self.n_pore, ptr(plan),
             L, I, H, ptr(self.means_scratch), ptr(out), self._stream())
        return out

    def _means_of(self, sums):
        """np.mean's float32 result from the float32 pore sums."""
        return np.array([float(np.float32(np.float64(x) / np.float64(self.n_pore))) for x in sums])

    def pore_means(self):
        """float(C[j][m].mean()) for every chemical, as NumPy forms it."""
        if self.n_pore == 0:
            return np.full(self.ns, np.nan)
        if self.precision == "fp64":
            c = self.current()
            return (c[:, self.pore_idx].sum(1) / self.n_pore).cpu().numpy()
        return self._means_of(self._pore_sums().cpu().numpy())

    def check_async(self, keep_state=True):
        """The pore sums of the newest field copied to pinned host memory
        and (keep_state) a copy of the field, all enqueued without waiting:
        returns (host sums, event marking their arrival, field copy or
        None).  float32 only."""
        if self.n_pore == 0 or self.precision == "fp64":
            raise PrtError("check_async: float32 pore-list solves only")
        dev = torch.empty(self.ns, dtype=torch.float32, device=self.dev)
        self._pore_sums(dev)
        host = torch.empty(self.ns, dtype=torch.float32, pin_memory=True)
        host.copy_(dev, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.dev))
        return host, ev, (self.c[self.latest].clone() if keep_state else None)

    def restore(self, state):
        """Make a field copy from check_async the newest field again."""
        self.c[self.latest].copy_(state)

    def snapshot(self):
        """A copy of the newest field in box layout (one copy: the layout
        change already makes one when immobile rows are in list order)."""
        c = self.c[self.latest]
        box = self._to_box(c)
        return box if box is not c else c.clone()

    def outlet_concentration(self, vel=None):
        """Build-defined outlet concentration (FP64) of the current field:
        the mean over the pore voxels of the x = nx-1 slice and, given the
        velocity, the flux-weighted mean sum(u_x C)/sum(u_x) there."""
        c = self.current().reshape((self.ns,) + self.shape)[:, -1].double()
        msk = torch.from_numpy(np.ascontiguousarray(self.m[-1])).to(self.dev)
        cnt = int(msk.sum())
        mean = (c * msk).reshape(self.ns, -1).sum(1) / max(cnt, 1)
        out = {"mean": mean.cpu().numpy()}
        if vel is not None:
            ux = torch.as_tensor(np.asarray(vel)[0, -1], device=self.dev).double() * msk
            den = float(ux.sum())
            out["flux"] = ((c * ux).reshape(self.ns, -1).sum(1) / den).cpu().numpy() \
                if den != 0 else np.zeros(self.ns)
        return out


def adr_host_setup(mat, vel, pe, da, n_species=2, chem=None, da_abio=0.0,
                   precision="fp32", sorption=None):
    """The host half of TransportSolver (no device needed): the kernel
    parameters, normalised velocity and initial field solve_adr would use."""
    T = TransportSolver.__new__(TransportSolver)
    T._setup_host(mat, vel, pe, da, n_species, chem, da_abio, precision, sorption)
    return T


class AdrRun:
    """The reference's time loop around a TransportSolver, cut at events.

    3D/tools/prtlb_3d.py:1460-1595: record at the geometric ladder (ratio
    1.12), check every max(int(1/dt), 100) steps with the per-chemical
    geometric-tail stop rule, then pick the n_t snapshots.  The device is
    advanced by whoever owns it (one solver, or a batch of them); this object
    only says how far it may go before the next event and what to do there.
    """

    def __init__(self, T, n_t, steps=None, tol=2e-3, max_steps=100000, ratio=1.12,
                 on_progress=None):
        self.T, self.n_t, self.steps, self.tol = T, int(n_t), steps, float(tol)
        self.max_steps, self.ratio, self.on_progress = int(max_steps), ratio, on_progress
        self.limit = int(steps) if steps is not None else self.max_steps
        self.keep_only = None
        if steps is not None:
            lad = ladder(self.limit, ratio)
            self.keep_only = {lad[i] for i in pick_snapshots(lad, max(self.limit, 1), self.n_t)}
        self.keep_steps, self.keep = [], []
        self.check = max(int(1.0 / float(T.dt)), 100)
        ns = T.ns
        self.prev = np.zeros(ns)
        self.scale = np.full(ns, 1e-6)
        self.prev_change = np.zeros(ns)
        self.n = 0
        self.nxt = 1
        self.pending = []                    # deferred checks (arrive(defer=True))
        self.finished = self.limit <= 0      # the reference's loop never runs
        self._record(0)

    def _record(self, s):
        self.keep_steps.append(s)
        self.keep.append(self.T.snapshot() if (self.keep_only is None or s in self.keep_only)
                         else None)

    def next_target(self):
        """The step the device may run to before this run needs attention."""
        nxt_check = (self.n // self.check + 1) * self.check
        return min(self.limit, self.nxt, nxt_check)

    def arrive(self, n, defer=False):
        """The device field is now at step n (== next_target()); handle the
        record / check there.  Sets self.finished when the run is over.

        defer: enqueue a check's pore means (and a copy of the field) and
        decide it later, in resolve(), without waiting for the device here.
        A run found settled at a deferred check is put back to that step, so
        the result is the same as deciding at once."""
        self.n = n
        if n >= self.nxt:
            self._record(n)
            self.nxt = max(n + 1, int(np.ceil(n * self.ratio)))
        if n % self.check == 0:
            if defer:
                # only a check that may stop the run needs the field kept
                may_stop = self.steps is None and n >= 10 * self.check
                self.pending.append((n,) + self.T.check_async(keep_state=may_stop))
            elif self._check(n, self.T.pore_means()):
                return
        if n >= self.limit:
            if self.pending:
                self.resolve(block=True)
            self.finished = True

    def resolve(self, block=False):
        """Decide the deferred checks, oldest first; without block only those
        whose means have reached the host.  A settled run is put back to the
        step it settled at."""
        while self.pending:
            n, host, ev, state = self.pending[0]
            if not block and not ev.query():
                return
            ev.synchronize()
            self.pending.pop(0)
            if self._check(n, self.T._means_of(host.numpy())):
                self.pending.clear()
                self.T.restore(state)
                self.T.steps_done = n
                self.n = n
                while self.keep_steps and self.keep_steps[-1] > n:
                    self.keep_steps.pop()
                    self.keep.pop()
                return

    def _check(self, n, cur):
        """The reference's convergence check at step n with pore means cur;
        True (and finished) when the run stops here."""
        self.scale = np.maximum(self.scale, np.abs(cur))
        change = float((np.abs(cur - self.prev) / self.scale).max())
        step_ch = np.abs(cur - self.prev) / self.scale
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(self.prev_change > 0, step_ch / self.prev_change, 0.0)
        rest = np.where(r < 1.0, step_ch * r / np.maximum(1.0 - r, 1e-12), np.inf)
        rest = float(np.nanmax(np.where(np.isfinite(rest), rest, np.inf)))
        if self.on_progress is not None:
            self.on_progress(n, self.limit, float(cur[0]),
                             rest if np.isfinite(rest) else change)
        if self.steps is None and n >= 10 * self.check and rest <= self.tol:
            self.finished = True
            return True
        self.prev, self.prev_change = cur, step_ch
        return False

    def result(self, info=None, da=0.0, da_abio=0.0, vel=None):
        """(out, t_norm) exactly as the reference assembles them, plus info."""
        T, n, ns, n_t = self.T, self.n, self.T.ns, self.n_t
        hit_cap = self.steps is None and n >= self.max_steps
        if self.keep_steps[-1] != n:
            self._record(n)
        total = int(max(n, 1))
        picked = pick_snapshots(self.keep_steps, total, n_t)
        repeated = n_t - len(set(picked))
        out = np.zeros((n_t, ns) + T.shape, T.R)
        for j, i in enumerate(picked):
            snap = self.keep[i]
            if snap is None:
                raise PrtError("internal: snapshot %d was not kept" % self.keep_steps[i])
            out[j] = snap.reshape((ns,) + T.shape).cpu().numpy()
        ks = np.asarray(self.keep_steps)
        t_norm = (ks[picked].astype(np.float32) / total).astype(np.float32)
        if info is not None:
            nd = T.nd
            info["converged"] = not (hit_cap or self.steps is not None)
            info["steps"] = total
            info["dt"] = float(T.dt)
            info["ladder"] = len(ks)
            info["chemistry"] = T.ch.source
            info["da_bio"] = float(da)
            info["da_abio"] = float(da_abio)
            info["dt_limited_by"] = ("advection" if T.umax > 2.0 * nd * T.dmax
                                     else "diffusion")
            info["biofilm_voxels"] = int(T.biofilm.sum())
            info["grid_peclet"] = float(T.pe_grid)
            info["l_ref_voxels"] = float(T.l_ref)
            info["surface_voxels"] = float(T.surface_voxels)
            info["repeated_snapshots"] = int(max(repeated, 0))
            oc = T.outlet_concentration(vel)
            info["outlet_concentration"] = oc["mean"]
            info["outlet_flux_concentration"] = oc.get("flux")
        self.keep = []
        return out, t_norm


def solve_adr(mat, vel, pe, da, n_t, n_species=2, steps=None, tol=2e-3,
              max_steps=100000, info=None, ratio=1.12, chem=None, da_abio=0.0,
              on_progress=None, *, precision="fp32", sorption=None, device=None,
              return_state=False):
    """Transport solve on the GPU; drop-in for 3D/tools/prtlb_3d.py:781.

    Returns (out, t_norm): out (n_t, n_species, *grid) float32 (float64 with
    precision="fp64"), zero outside the pore; t_norm float32.  Extra keyword-
    only options: precision, sorption (reversible, see TransportSolver),
    return_state (also return the solver, whose field stays on the device).
    info gains outlet_concentration (pore mean of the x = nx-1 slice, FP64,
    per chemical) and outlet_flux_concentration."""
    T = TransportSolver(mat, vel, pe, da, n_species, chem, da_abio,
                        precision=precision, sorption=sorption, device=device)
    run = AdrRun(T, n_t, steps=steps, tol=tol, max_steps=max_steps, ratio=ratio,
                 on_progress=on_progress)
    # float32 checks are decided without waiting for the device: at most two
    # stay undecided, and a run found settled late is rolled back (AdrRun)
    defer = T.precision == "fp32" and T.n_pore > 0
    while not run.finished:
        target = run.next_target()
        T.step_n(target - run.n)
        run.arrive(target, defer=defer)
        if defer and not run.finished:
            run.resolve(block=len(run.pending) > 2)
    out, t_norm = run.result(info, da, da_abio, vel)
    if return_state:
        return out, t_norm, T
    return out, t_norm
