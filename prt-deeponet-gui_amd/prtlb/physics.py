"""Post-solve validation and the derived outputs (permeability, outlet
concentration).

check_physics keeps the reference's rules, argument list, return value and
the keywords of its complaints (3D/tools/prtlb_3d.py:1632-1843).  It reads
the stored snapshots once; it is host-side NumPy because the snapshots are
already on the host when a caller validates them.
"""
import numpy as np

from .geometry import PORE


def _name(chem, j):
    try:
        return str(chem.names[j])
    except Exception:
        return ("the donor", "the acceptor", "the product",
                "the biomass")[j] if j < 4 else "chemical %d" % j


def _start(chem, j):
    if chem is None:
        return 1.0
    vals = [float(v) for v in (chem.inlet[j], chem.c_init[j]) if v == v]
    return max(vals) if vals else 0.0


class _HostStats:
    """The numbers check_physics' rules read, from host snapshots (NumPy,
    the reference's own expressions)."""

    def __init__(self, out, m):
        self.out, self.m = out, m
        self.n_t, self.ns = out.shape[0], out.shape[1]
        self.nx = m.shape[0]
        self.n_nan = int(np.isnan(out).sum())
        self.n_inf = int(np.isinf(out).sum())
        self.any_pore = bool(m.any())

    def pore_max_min(self, j):
        fj = self.out[:, j][:, self.m]
        return float(fj.max()), float(fj.min())

    def last_absmax(self, j):
        return float(np.abs(self.out[-1, j][self.m]).max())

    def all_absmax(self, j):
        return float(np.abs(self.out[:, j][:, self.m]).max())

    def checker(self):
        m = self.m
        f = self.out[-1, 0]
        d = np.where(m[1:] & m[:-1], f[1:] - f[:-1], 0.0)
        pair = m[2:] & m[1:-1] & m[:-2]
        npair = int(pair.sum())
        if npair <= 100:
            return npair, 0
        s1, s2 = np.sign(d[1:]), np.sign(d[:-1])
        return npair, int(((s1 * s2 < 0) & pair).sum())

    def planes(self):
        donor = self.out[:, 0]
        a = donor[:, self.nx // 2].reshape(donor.shape[0], -1).max(1)
        b = donor[:, self.nx - 1].reshape(donor.shape[0], -1).max(1)
        return a, b

    def copy_stats(self, j):
        """(std of chemical 0, std of j, correlation, residual(sc)) over
        every (snapshot, pore voxel), float32 NumPy as the reference."""
        a0 = self.out[:, 0][:, self.m].ravel()
        aj = self.out[:, j][:, self.m].ravel()
        s0, sj = a0.std(), aj.std()
        if s0 < 1e-9 or sj < 1e-9:
            return s0, sj, None, None
        r = float(np.corrcoef(a0, aj)[0, 1])
        return s0, sj, r, lambda sc: float(np.abs(aj - sc * a0).max())


def _rules(st, verbose, chem, da, da_abio):
    """The reference's seven rules over precomputed statistics."""
    bad = []
    ns = st.ns
    if st.n_nan or st.n_inf:
        bad.append("the field is not finite: %d NaN and %d infinite values -- "
                   "the integration diverged" % (st.n_nan, st.n_inf))
        if verbose:
            print("     PHYSICS: " + bad[0], flush=True)
        return bad
    caps = [None] * ns
    if chem is not None:
        cap0 = _start(chem, 0)
        caps[0] = cap0
        if ns > 1:
            caps[1] = _start(chem, 1)
        if ns > 2:
            y = 0.0
            if ns > 3 and getattr(chem, "biotic_on", True):
                y = float(getattr(chem, "yield_", 0.0))
            caps[2] = _start(chem, 2) + max(cap0 * (1.0 - y), 0.0)
    else:
        for j in range(min(ns, 3)):
            caps[j] = 1.0
    if st.any_pore:
        for j in range(ns):
            hi, lo = st.pore_max_min(j)
            if caps[j] is not None and hi > 1.02 * max(caps[j], 1e-6):
                bad.append("%s reaches %.3f, above the %.3f it could be made "
                           "from -- the time step is unstable"
                           % (_name(chem, j), hi, caps[j]))
            elif caps[j] is None and hi > 1e3 * max(_start(chem, j), 1e-3):
                bad.append("%s reaches %.4g, a thousand times what it started "
                           "from -- this is runaway growth, not a solution"
                           % (_name(chem, j), hi))
            if lo < -1e-6:
                bad.append("%s reaches %.3g, which is negative" % (_name(chem, j), lo))
        for j in range(ns):
            expect = 1.0
            if chem is not None:
                fed = float(chem.inlet[j]) if chem.inlet[j] == chem.inlet[j] else 0.0
                expect = max(fed, float(chem.c_init[j]))
                if j > 1 and expect <= 0:
                    continue
            if expect > 0 and st.last_absmax(j) <= 0.0:
                bad.append("%s is exactly zero in every voxel at the end of the "
                           "run, though %.3g of it was fed in or present at the "
                           "start -- it has not been solved, it has been lost"
                           % (_name(chem, j), expect))
        if chem is not None:
            for rate, jr, what in ((da, 0, "biotic"),
                                   (da_abio, getattr(chem, "abiotic_index", 2), "abiotic")):
                if rate is None or float(rate) <= 0 or not (0 <= jr < ns):
                    continue
                if st.all_absmax(jr) <= 0.0:
                    bad.append("the %s reaction was given a Damkohler number of "
                               "%.3g but the chemical it consumes, %s, is zero in "
                               "every voxel at every time -- the reaction never "
                               "happened" % (what, float(rate), _name(chem, jr)))
    npair, nflip = st.checker()
    if npair > 100:
        flips = float(nflip) / float(npair)
        if flips > 0.45:
            bad.append("%.0f%% of neighbouring pore voxels along the flow "
                       "disagree in sign -- this is a checkerboard, not a "
                       "concentration field" % (100 * flips))
    a, b = st.planes()
    early = np.where((b > 1e-3) & (a < 1e-6))[0]
    if chem is not None and chem.c_init[0] > 0:
        early = np.array([], int)
    if early.size:
        bad.append("solute is at the outlet while the mid-plane is still empty "
                   "at snapshot %d -- something is wrapping around" % early[0])
    if ns > 1 and st.any_pore:
        for j in range(1, ns):
            s0, sj, r, resid = st.copy_stats(j)
            if r is None:
                continue
            sc = sj / max(s0, 1e-12)
            if r > 0.9999 and resid(sc) < 1e-4:
                bad.append("species %d is a rescaled copy of species 0 "
                           "(correlation %.5f) -- it is not a second chemical" % (j, r))
    if verbose:
        for s in bad:
            print("     PHYSICS: " + s, flush=True)
    return bad


def check_physics(out, mat, verbose=False, chem=None, da=None, da_abio=None):
    """Complaints about a transport result; an empty list when it is sound.

    Rules: finite; per-chemical maximum (donor/acceptor by feed, product by
    feed + donor (1 - yield), biomass only for runaway) and floor; nothing fed
    is lost entirely; a reaction with a Damkohler number has a reactant; no
    checkerboard along x; causality (no solute at the outlet before the
    mid-plane); no chemical is a rescaled copy of the donor.  Host NumPy over
    host snapshots; check_physics_device runs the same rules with the data
    passes on the GPU."""
    mat = np.asarray(mat)
    return _rules(_HostStats(np.asarray(out), mat == PORE), verbose, chem, da, da_abio)


class _DeviceStats:
    """The same statistics from device snapshots: one pass per statistic,
    block partials folded on the host (prt_check_*)."""

    def __init__(self, out, m):
        import torch
        from ._lib import call, lib
        from .device import ptr, stream_ptr
        self._call, self._ptr = call, ptr
        self.out = out
        self.n_t, self.ns = int(out.shape[0]), int(out.shape[1])
        self.nx = int(out.shape[2])
        self.nvox = int(np.prod(out.shape[2:]))
        self.plane = self.nvox // self.nx
        self.dev = out.device
        self.s = stream_ptr(self.dev)
        self.pore = m.reshape(-1).to(torch.uint8).contiguous()
        self.any_pore = bool(self.pore.any().item())
        B = self.blocks = int(lib().prt_check_blocks(self.n_t * self.ns * self.nvox))
        cnt = torch.zeros(2 * B, dtype=torch.int64, device=self.dev)
        call("prt_check_nonfinite", ptr(out), self.n_t * self.ns * self.nvox, B, ptr(cnt), self.s)
        c = cnt.view(B, 2).sum(0).cpu().numpy()
        self.n_nan, self.n_inf = int(c[0]), int(c[1])
        self._range = None

    def _ranges(self):
        import torch
        if self._range is None:
            B = self.blocks
            mx = torch.empty(self.n_t * self.ns * B, dtype=torch.float32, device=self.dev)
            mn = torch.empty_like(mx)
            self._call("prt_check_pore_range", self._ptr(self.out), self.n_t, self.ns, self.nvox,
                       self._ptr(self.pore), B, self._ptr(mx), self._ptr(mn), self.s)
            self._range = (mx.view(self.n_t, self.ns, B).amax(2).cpu().numpy(),
                           mn.view(self.n_t, self.ns, B).amin(2).cpu().numpy())
        return self._range

    def pore_max_min(self, j):
        mx, mn = self._ranges()
        return float(mx[:, j].max()), float(mn[:, j].min())

    def last_absmax(self, j):
        mx, mn = self._ranges()
        return float(max(abs(mx[-1, j]), abs(mn[-1, j])))

    def all_absmax(self, j):
        mx, mn = self._ranges()
        return float(np.maximum(np.abs(mx[:, j]), np.abs(mn[:, j])).max())

    def checker(self):
        import torch
        B = self.blocks
        cnt = torch.zeros(2 * B, dtype=torch.int64, device=self.dev)
        last = self.out[-1, 0]
        self._call("prt_check_checker", self._ptr(last), self._ptr(self.pore), self.nx, self.plane, B,
                   self._ptr(cnt), self.s)
        c = cnt.view(B, 2).sum(0).cpu().numpy()
        return int(c[0]), int(c[1])

    def planes(self):
        import torch
        B = self.blocks
        part = torch.empty(2 * self.n_t * B, dtype=torch.float32, device=self.dev)
        self._call("prt_check_plane_max", self._ptr(self.out), self.n_t, self.ns, self.nx, self.plane, B,
                   self._ptr(part), self.s)
        p = part.view(2, self.n_t, B).amax(2).cpu().numpy()
        return p[0], p[1]

    def copy_stats(self, j):
        """FP64 two-pass moments on the device; the residual pass in float32
        with the float32 scale, as the reference evaluates it."""
        import torch
        if not hasattr(self, "_mom"):
            B = self.blocks
            M = float(self.n_t * int(self.pore.sum().item()))
            mom = torch.empty((self.ns - 1) * B * 5, dtype=torch.float64, device=self.dev)
            self._call("prt_check_moments", self._ptr(self.out), self.n_t, self.ns, self._ptr(self.pore),
                       self.nvox, B, None, self._ptr(mom), self.s)
            sums = mom.view(self.ns - 1, B, 5).cpu().numpy().sum(1)
            mean = np.zeros(self.ns)
            mean[0] = sums[0, 0] / M
            mean[1:] = sums[:, 1] / M
            mean_d = torch.from_numpy(mean).to(self.dev)
            self._call("prt_check_moments", self._ptr(self.out), self.n_t, self.ns, self._ptr(self.pore),
                       self.nvox, B, self._ptr(mean_d), self._ptr(mom), self.s)
            self._mom = (mom.view(self.ns - 1, B, 5).cpu().numpy().sum(1), M)
        c, M = self._mom
        s0 = np.float32(np.sqrt(c[j - 1, 2] / M))
        sj = np.float32(np.sqrt(c[j - 1, 3] / M))
        if s0 < 1e-9 or sj < 1e-9:
            return s0, sj, None, None
        r = float(c[j - 1, 4] / np.sqrt(c[j - 1, 2] * c[j - 1, 3]))

        def resid(sc):
            B = self.blocks
            part = torch.empty(B, dtype=torch.float32, device=self.dev)
            self._call("prt_check_copy_residual", self._ptr(self.out), self.n_t, self.ns, j,
                       float(np.float32(sc)), self._ptr(self.pore), self.nvox, B, self._ptr(part), self.s)
            return float(part.max().item())
        return s0, sj, r, resid


def check_physics_device(out, mat, verbose=False, chem=None, da=None, da_abio=None, device=None):
    """check_physics with its data passes on the GPU: out is the snapshot
    array (n_t, ns, *grid) float32 as a device tensor (or a NumPy array,
    uploaded once), mat the material grid.  Same rules and messages as
    check_physics (3D/tools/prtlb_3d.py:1632-1843); the rescaled-copy rule
    uses FP64 moments instead of NumPy's float32 ones, which can move its
    printed correlation in the last digit, not its decision, except within
    rounding of the 0.9999 threshold."""
    import torch
    from .device import device as _dev
    if not torch.is_tensor(out):
        out = torch.from_numpy(np.ascontiguousarray(out, np.float32)).to(_dev(device))
    out = out.contiguous()
    if out.dtype != torch.float32:
        raise ValueError("check_physics_device: snapshots must be float32")
    if torch.is_tensor(mat):
        m = (mat == PORE).to(out.device)
    else:
        m = torch.from_numpy(np.ascontiguousarray(np.asarray(mat) == PORE)).to(out.device)
    return _rules(_DeviceStats(out, m), verbose, chem, da, da_abio)


def permeability(v, tau_lb, force):
    """Darcy permeability in lattice units from a velocity field (3, *grid):
    k = nu <u_x>_box / force, nu = (tau - 1/2)/3, <u_x>_box the superficial
    mean over every node (solid nodes carry zero).  FP64 reduction."""
    v = np.asarray(v, np.float64)
    nu = (float(tau_lb) - 0.5) / 3.0
    return nu * float(v[0].sum() / v[0].size) / float(force)


def outlet_concentration(C, mat, vel=None):
    """Pore mean of the x = nx-1 slice of C (ns, *grid) per chemical, and the
    flux-weighted mean sum(u_x C)/sum(u_x) there when vel is given (FP64)."""
    C = np.asarray(C, np.float64)
    sl = np.asarray(mat)[-1] == PORE
    mean = np.array([C[j, -1][sl].mean() if sl.any() else 0.0 for j in range(C.shape[0])])
    if vel is None:
        return mean, None
    ux = np.asarray(vel, np.float64)[0, -1] * sl
    den = ux.sum()
    flux = np.array([(ux * C[j, -1]).sum() / den if den != 0 else 0.0
                     for j in range(C.shape[0])])
    return mean, flux
