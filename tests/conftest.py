import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "prt-deeponet-gui_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: long CPU test")


def golden(name):
    import json

    import numpy as np
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files if k != "meta"}
    d["meta"] = json.loads(str(z["meta"]))
    return d


def oracle_chem(ch):
    """The oracle's own Chemistry with the attributes of a package one."""
    from oracle import prtlb_oracle as O
    return O.Chemistry(ch.names, ch.inlet, ch.c_init, ch.d_rel, ch.immobile, ks_donor=ch.ks_donor,
                       ks_acceptor=ch.ks_acceptor, alpha=ch.alpha, yield_=ch.yield_,
                       decay_lattice=ch.decay_lattice, biotic_on=ch.biotic_on,
                       abiotic_index=ch.abiotic_index,
                       abiotic_surface_only=ch.abiotic_surface_only,
                       abiotic_product_index=ch.abiotic_product_index)
