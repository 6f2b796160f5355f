"""Writes the small fixtures under tests/golden/ (run by hand, not by the tests).

1. etaPm.npy, orb_prod.npy: copied as data from the reference's own test
   fixture tests/data/eri_thc.h5 (datasets 'etaPm', 140 x 10, and 'orb_prod',
   140 x 100, whose row p holds the products etaPm[p, i] * etaPm[p, j]).
   Needs h5py, which is only in the conda interpreter of the build image:
       /opt/conda/bin/python3.9 tests/golden/make_golden.py
2. mt19937_seed5489_10000.txt: the value the C++ standard fixes for the
   10000th output of a default-seeded std::mt19937 ([rand.predef]).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/tests/data/eri_thc.h5"


def main():
    import h5py
    with h5py.File(SRC, "r") as f:
        np.save(os.path.join(HERE, "etaPm.npy"), np.asarray(f["etaPm"], dtype=np.float64))
        np.save(os.path.join(HERE, "orb_prod.npy"), np.asarray(f["orb_prod"], dtype=np.float64))
    with open(os.path.join(HERE, "mt19937_seed5489_10000.txt"), "w") as fh:
        fh.write("4123659995\n")


if __name__ == "__main__":
    sys.exit(main())
