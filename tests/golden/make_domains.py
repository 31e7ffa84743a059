"""Copy the reference's three released 2D domains into one fixture.

The reference ships the domains of its three 2D cases as data files
(2D/geometries/Domain_Monod.npz, Domain_Irreversible_Sorption.npz and
Domain_Reversible_Sorption.npz; named by the cases in
2D_scripts/prt2d_model.py:72-109): each a flat uint8 array of 64 x 148
values under the key 'm', codes 0 solid, 1 pore, 2 pore-solid interface.
This script reads them with NumPy's loader (no pickles) and stores the
three arrays, exactly as released, in reference_domains.npz beside it.
Only this script reads the reference; the fixture is data.

    python tests/golden/make_domains.py
"""
import os

import numpy as np

SRC = "/root/reference/2D/geometries"
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
    np.savez_compressed(os.path.join(HERE, "reference_domains.npz"), **out)
    print({k: int((v > 0).sum()) for k, v in out.items()})


if __name__ == "__main__":
    main()
