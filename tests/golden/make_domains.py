"""Copy released 2D domains of the reference into one fixture.

The reference ships the domains of its three 2D cases as data files
(2D/geometries/Domain_Monod.npz, Domain_Irreversible_Sorption.npz and
Domain_Reversible_Sorption.npz; named by the cases in
2D_scripts/prt2d_model.py:72-109): each a flat uint8 array of 64 x 148
values under the key 'm', codes 0 solid, 1 pore, 2 pore-solid interface.
It also takes two of the 3000 training domains in 2D/Domains/
Input_domains.zip (domain_1.dat and domain_3000.dat, porosity 0.49 and
0.74: one value per line, 148 x 64 in C order, 0 solid, 1 pore).  This
script reads them with NumPy's loaders (no pickles) and stores the arrays,
values as released, in reference_domains.npz beside it (the .dat values as
uint8, keys dat_1 and dat_3000).  Only this script reads the reference;
the fixture is data.

    python tests/golden/make_domains.py
"""
import os
import zipfile

import numpy as np

SRC = "/root/reference/2D/geometries"
ZIP = "/root/reference/2D/Domains/Input_domains.zip"
DAT = ("domain_1.dat", "domain_3000.dat")
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = {"monod": "Domain_Monod.npz",
         "irreversible_sorption": "Domain_Irreversible_Sorption.npz",
         "reversible_sorption": "Domain_Reversible_Sorption.npz"}


def main():
    out = {}
    for case, name in CASES.items():
        with np.load(os.path.join(SRC, name), allow_pickle=False) as z:
            m = np.asarray(z["m"])
        assert m.dtype == np.uint8 and m.shape == (64 * 148,), (name, m.dtype, m.shape)
        assert set(np.unique(m).tolist()) <= {0, 1, 2}, name
        out[case] = m
    with zipfile.ZipFile(ZIP) as z:
        for name in DAT:
            with z.open(name) as f:
                v = np.loadtxt(f).ravel()
            assert v.size == 148 * 64 and set(np.unique(v).tolist()) <= {0.0, 1.0}, name
            out["dat_" + name[len("domain_"):-len(".dat")]] = v.astype(np.uint8)
    np.savez_compressed(os.path.join(HERE, "reference_domains.npz"), **out)
    print({k: int((v > 0).sum()) for k, v in out.items()})


if __name__ == "__main__":
    main()
