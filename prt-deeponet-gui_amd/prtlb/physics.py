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


def check_physics(out, mat, verbose=False, chem=None, da=None, da_abio=None):
    """Complaints about a transport result; an empty list when it is sound.

    Rules: finite; per-chemical maximum (donor/acceptor by feed, product by
    feed + donor (1 - yield), biomass only for runaway) and floor; nothing fed
    is lost entirely; a reaction with a Damkohler number has a reactant; no
    checkerboard along x; causality (no solute at the outlet before the
    mid-plane); no chemical is a rescaled copy of the donor."""
    mat = np.asarray(mat)
    m = mat == PORE
    out = np.asarray(out)
    bad = []
    ns = out.shape[1]
    if not np.isfinite(out).all():
        bad.append("the field is not finite: %d NaN and %d infinite values -- "
                   "the integration diverged"
                   % (int(np.isnan(out).sum()), int(np.isinf(out).sum())))
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
    if m.any():
        for j in range(ns):
            fj = out[:, j][:, m]
            hi = float(fj.max())
            if caps[j] is not None and hi > 1.02 * max(caps[j], 1e-6):
                bad.append("%s reaches %.3f, above the %.3f it could be made "
                           "from -- the time step is unstable"
                           % (_name(chem, j), hi, caps[j]))
            elif caps[j] is None and hi > 1e3 * max(_start(chem, j), 1e-3):
                bad.append("%s reaches %.4g, a thousand times what it started "
                           "from -- this is runaway growth, not a solution"
                           % (_name(chem, j), hi))
            lo = float(fj.min())
            if lo < -1e-6:
                bad.append("%s reaches %.3g, which is negative" % (_name(chem, j), lo))
        for j in range(ns):
            expect = 1.0
            if chem is not None:
                fed = float(chem.inlet[j]) if chem.inlet[j] == chem.inlet[j] else 0.0
                expect = max(fed, float(chem.c_init[j]))
                if j > 1 and expect <= 0:
                    continue
            if expect > 0 and float(np.abs(out[-1, j][m]).max()) <= 0.0:
                bad.append("%s is exactly zero in every voxel at the end of the "
                           "run, though %.3g of it was fed in or present at the "
                           "start -- it has not been solved, it has been lost"
                           % (_name(chem, j), expect))
        if chem is not None:
            for rate, jr, what in ((da, 0, "biotic"),
                                   (da_abio, getattr(chem, "abiotic_index", 2), "abiotic")):
                if rate is None or float(rate) <= 0 or not (0 <= jr < ns):
                    continue
                if float(np.abs(out[:, jr][:, m]).max()) <= 0.0:
                    bad.append("the %s reaction was given a Damkohler number of "
                               "%.3g but the chemical it consumes, %s, is zero in "
                               "every voxel at every time -- the reaction never "
                               "happened" % (what, float(rate), _name(chem, jr)))
    f = out[-1, 0]
    d = np.where(m[1:] & m[:-1], f[1:] - f[:-1], 0.0)
    pair = m[2:] & m[1:-1] & m[:-2]
    if pair.sum() > 100:
        s1, s2 = np.sign(d[1:]), np.sign(d[:-1])
        flips = float(((s1 * s2 < 0) & pair).sum()) / float(pair.sum())
        if flips > 0.45:
            bad.append("%.0f%% of neighbouring pore voxels along the flow "
                       "disagree in sign -- this is a checkerboard, not a "
                       "concentration field" % (100 * flips))
    nx = m.shape[0]
    donor = out[:, 0]
    a = donor[:, nx // 2].reshape(donor.shape[0], -1).max(1)
    b = donor[:, nx - 1].reshape(donor.shape[0], -1).max(1)
    early = np.where((b > 1e-3) & (a < 1e-6))[0]
    if chem is not None and chem.c_init[0] > 0:
        early = np.array([], int)
    if early.size:
        bad.append("solute is at the outlet while the mid-plane is still empty "
                   "at snapshot %d -- something is wrapping around" % early[0])
    if ns > 1:
        a0 = out[:, 0][:, m].ravel()
        for j in range(1, ns):
            aj = out[:, j][:, m].ravel()
            if a0.std() < 1e-9 or aj.std() < 1e-9:
                continue
            r = float(np.corrcoef(a0, aj)[0, 1])
            sc = aj.std() / max(a0.std(), 1e-12)
            if r > 0.9999 and float(np.abs(aj - sc * a0).max()) < 1e-4:
                bad.append("species %d is a rescaled copy of species 0 "
                           "(correlation %.5f) -- it is not a second chemical" % (j, r))
    if verbose:
        for s in bad:
            print("     PHYSICS: " + s, flush=True)
    return bad


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
