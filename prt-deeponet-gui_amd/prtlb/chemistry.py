"""What the transport solve reads about the chemicals.

solve_adr reads a chemistry object by attribute only (3D/tools/prtlb_3d.py:
831-1018): per-chemical columns (feed at the inlet, NaN for none; initial
value; diffusivity relative to the solute's and inside biofilm; whether it
moves) and run-wide constants (half-saturations, stoichiometry, yield,
decay, the reaction switches and indices, the reference length).  So a
caller's object -- this package's Chemistry or the reference's own
settings_and_units.Chemistry (3D/tools/settings_and_units.py:1834) -- works
unchanged.  This Chemistry takes the same constructor arguments with the
same defaults and exposes the same attributes, plus the two methods the
solve calls, decay_for and truncate.
"""
import copy

import numpy as np

# per-chemical columns and the dtype each is held in
_COLUMNS = (("inlet", np.float32), ("c_init", np.float32), ("d_rel", np.float32),
            ("immobile", bool))


def _column(name, values, dtype, n):
    col = np.asarray(values, dtype)
    if col.shape != (n,):
        raise ValueError("Chemistry: %s must give one value per chemical (%d chemicals, got "
                         "shape %s)" % (name, n, col.shape))
    return col


def _opt(kind, v):
    return None if v is None else kind(v)


class Chemistry:
    """Dimensionless chemistry: one column entry per chemical, then the
    kinetic constants and switches of the run."""

    def __init__(self, names, inlet, c_init, d_rel, immobile, ks_donor=0.0,
                 ks_acceptor=0.15, alpha=0.4, yield_=0.04, da_decay=None,
                 decay_lattice=None, source="built in", biomass_coupled=False,
                 abiotic_surface_only=False, d_biofilm_rel=None,
                 biofilm_code=None, seed_in_biofilm=False, abiotic_index=2,
                 biotic_on=True, l_ref_voxels=None, abiotic_product_index=None):
        self.names = list(names)
        n = len(self.names)
        given = {"inlet": inlet, "c_init": c_init, "d_rel": d_rel, "immobile": immobile}
        for key, dtype in _COLUMNS:
            setattr(self, key, _column(key, given[key], dtype, n))
        # inside biofilm the pore diffusivities apply unless told otherwise
        self.d_biofilm_rel = (self.d_rel.copy() if d_biofilm_rel is None
                              else _column("d_biofilm_rel", d_biofilm_rel, np.float32, n))
        # kinetic constants (floats) and switches
        self.ks_donor, self.ks_acceptor = float(ks_donor), float(ks_acceptor)
        self.alpha, self.yield_ = float(alpha), float(yield_)
        self.da_decay, self.decay_lattice = _opt(float, da_decay), _opt(float, decay_lattice)
        self.biotic_on, self.biomass_coupled = bool(biotic_on), bool(biomass_coupled)
        self.abiotic_index = int(abiotic_index)
        self.abiotic_product_index = _opt(int, abiotic_product_index)
        self.abiotic_surface_only = bool(abiotic_surface_only)
        self.biofilm_code, self.seed_in_biofilm = _opt(int, biofilm_code), bool(seed_in_biofilm)
        self.l_ref_voxels = _opt(float, l_ref_voxels)
        self.source = source

    def decay_for(self, l_ref):
        """Biomass decay per lattice step for reference length l_ref voxels:
        the lattice value when one is set, else the decay Damkohler number
        over the length (3D/tools/settings_and_units.py:1903)."""
        if self.decay_lattice is not None:
            return float(self.decay_lattice)
        da = 0.0 if self.da_decay is None else self.da_decay
        return float(da / max(l_ref, 1))

    def truncate(self, ns):
        """A copy holding only the first ns chemicals (at least one, at
        most all); the run-wide constants are shared."""
        keep = max(1, min(int(ns), len(self.names)))
        out = copy.copy(self)
        out.names = self.names[:keep]
        for key in ("inlet", "c_init", "d_rel", "immobile", "d_biofilm_rel"):
            setattr(out, key, getattr(self, key)[:keep].copy())
        return out


# The built-in chemicals, in order (3D/tools/settings_and_units.py:1934-1951):
# name, feed at the inlet (NaN: none), initial value, relative diffusivity
# (pore and biofilm), immobile.
_BUILT_IN = (
    ("Ac", 1.0, 0.0, 1.0, False),       # electron donor
    ("A", 1.0, 0.0, 1.0, False),        # electron acceptor
    ("P", 0.0, 0.0, 1.0, False),        # product, held at zero at the inlet
    ("Bio", np.nan, 0.1, 0.0, True),    # attached biomass, not fed
)


def default_chemistry(n_species=2):
    """The first n_species (1..4) built-in chemicals with the built-in
    constants: Ks_A 0.15, alpha 0.4, yield 0.04, lattice decay 0.006."""
    rows = _BUILT_IN[:max(1, min(int(n_species), len(_BUILT_IN)))]
    names, feed, start, drel, fixed = (list(c) for c in zip(*rows))
    return Chemistry(names, feed, start, drel, fixed, d_biofilm_rel=drel, ks_donor=0.0,
                     ks_acceptor=0.15, alpha=0.4, yield_=0.04, decay_lattice=0.006,
                     source="built in")


def first_order(da_surface_only=True):
    """One mobile chemical fed at 1 and lost by a first-order (optionally
    surface-weighted) abiotic reaction: the 2D disk-pack workload."""
    return Chemistry(names=["C"], inlet=[1.0], c_init=[0.0], d_rel=[1.0],
                     immobile=[False], biotic_on=False, abiotic_index=0,
                     abiotic_surface_only=da_surface_only)


def monod(ks=0.5):
    """One mobile chemical consumed at k C/(Ks + C): the 3D Monod workload."""
    return Chemistry(names=["C"], inlet=[1.0], c_init=[0.0], d_rel=[1.0],
                     immobile=[False], ks_donor=ks, biotic_on=True)


def sorption_irreversible(surface_only=True):
    """Mobile C sorbing irreversibly onto immobile S at rate ka C (x surface
    weight): the reference's abiotic route with product S."""
    return Chemistry(names=["C", "S"], inlet=[1.0, np.nan], c_init=[0.0, 0.0],
                     d_rel=[1.0, 0.0], immobile=[False, True], biotic_on=False,
                     abiotic_index=0, abiotic_product_index=1,
                     abiotic_surface_only=surface_only)


def sorption_reversible(surface_only=True):
    """Mobile C and immobile S for reversible sorption; pass the rates to
    solve_adr(..., sorption=dict(da_sorb=..., da_desorb=...))."""
    return Chemistry(names=["C", "S"], inlet=[1.0, np.nan], c_init=[0.0, 0.0],
                     d_rel=[1.0, 0.0], immobile=[False, True], biotic_on=False,
                     abiotic_index=0, abiotic_surface_only=surface_only)
