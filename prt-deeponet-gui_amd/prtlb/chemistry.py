"""What the transport solve reads about the chemicals.

`Chemistry` keeps the constructor, attributes, `decay_for` and `truncate` of
the reference's settings_and_units.Chemistry (3D/tools/settings_and_units.py:
1834-1931) so that a caller's chemistry object works unchanged; solve_adr only
reads attributes, so the reference's own Chemistry objects work too.
"""
import numpy as np


class Chemistry:
    """Dimensionless chemistry: feeds, initial values, relative diffusivities,
    mobility, kinetic constants and the switches the transport step reads."""

    def __init__(self, names, inlet, c_init, d_rel, immobile, ks_donor=0.0,
                 ks_acceptor=0.15, alpha=0.4, yield_=0.04, da_decay=None,
                 decay_lattice=None, source="built in", biomass_coupled=False,
                 abiotic_surface_only=False, d_biofilm_rel=None,
                 biofilm_code=None, seed_in_biofilm=False, abiotic_index=2,
                 biotic_on=True, l_ref_voxels=None, abiotic_product_index=None):
        self.names = list(names)
        n = len(self.names)
        self.inlet = np.asarray(inlet, np.float32)
        self.c_init = np.asarray(c_init, np.float32)
        self.d_rel = np.asarray(d_rel, np.float32)
        self.immobile = np.asarray(immobile, bool)
        for arr, nm in ((self.inlet, "inlet"), (self.c_init, "c_init"),
                        (self.d_rel, "d_rel"), (self.immobile, "immobile")):
            if arr.shape != (n,):
                raise ValueError("chemistry: %s has %d entries but there are %d "
                                 "chemicals" % (nm, arr.size, n))
        self.ks_donor = float(ks_donor)
        self.ks_acceptor = float(ks_acceptor)
        self.alpha = float(alpha)
        self.yield_ = float(yield_)
        self.da_decay = None if da_decay is None else float(da_decay)
        self.decay_lattice = None if decay_lattice is None else float(decay_lattice)
        self.source = source
        self.biomass_coupled = bool(biomass_coupled)
        self.abiotic_surface_only = bool(abiotic_surface_only)
        self.d_biofilm_rel = (self.d_rel.copy() if d_biofilm_rel is None
                              else np.asarray(d_biofilm_rel, np.float32))
        self.biofilm_code = None if biofilm_code is None else int(biofilm_code)
        self.seed_in_biofilm = bool(seed_in_biofilm)
        self.abiotic_index = int(abiotic_index)
        self.biotic_on = bool(biotic_on)
        self.l_ref_voxels = None if l_ref_voxels is None else float(l_ref_voxels)
        self.abiotic_product_index = (None if abiotic_product_index is None
                                      else int(abiotic_product_index))

    def decay_for(self, nx):
        """Biomass decay in lattice units for a reference length nx."""
        if self.decay_lattice is not None:
            return float(self.decay_lattice)
        return float((self.da_decay or 0.0) / max(nx, 1))

    def truncate(self, ns):
        """The same chemistry restricted to its first ns chemicals."""
        ns = max(1, min(int(ns), len(self.names)))
        return Chemistry(self.names[:ns], self.inlet[:ns], self.c_init[:ns],
                         self.d_rel[:ns], self.immobile[:ns], self.ks_donor,
                         self.ks_acceptor, self.alpha, self.yield_,
                         self.da_decay, self.decay_lattice, self.source,
                         self.biomass_coupled, self.abiotic_surface_only,
                         self.d_biofilm_rel[:ns], self.biofilm_code,
                         self.seed_in_biofilm, self.abiotic_index,
                         self.biotic_on, self.l_ref_voxels,
                         self.abiotic_product_index)

    def describe(self):
        rows = ["  %-6s %9s %9s %8s %9s" % ("name", "feed", "at t=0", "D/D_ref", "moves")]
        for i, nm in enumerate(self.names):
            fed = self.inlet[i]
            rows.append("  %-6s %9s %9.4g %8.3g %9s"
                        % (nm, "no feed" if fed != fed else "%.4g" % fed,
                           self.c_init[i], self.d_rel[i],
                           "no" if self.immobile[i] else "yes"))
        return "\n".join(rows)


def default_chemistry(n_species=2):
    """Donor, acceptor, product, attached biomass with the built-in constants
    (3D/tools/settings_and_units.py:1934-1951)."""
    ns = max(1, min(int(n_species), 4))
    return Chemistry(
        names=["Ac", "A", "P", "Bio"][:ns],
        inlet=np.array([1.0, 1.0, 0.0, np.nan], np.float32)[:ns],
        c_init=np.array([0.0, 0.0, 0.0, 0.1], np.float32)[:ns],
        d_rel=np.array([1.0, 1.0, 1.0, 0.0], np.float32)[:ns],
        d_biofilm_rel=np.array([1.0, 1.0, 1.0, 0.0], np.float32)[:ns],
        immobile=np.array([False, False, False, True])[:ns],
        ks_donor=0.0, ks_acceptor=0.15, alpha=0.4, yield_=0.04,
        decay_lattice=0.006, source="built in")


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
