"""Writes the R-MAT fixture tests/golden/rmat_s10_ef8_seed1.npz.

The fixture freezes the seeded R-MAT definition shared by the oracle
(oracle/gre_oracle.c) and the device generator (graphine_amd/csrc/graph.hip):
any change to either generator shows up as a mismatch in tests.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def main():
    g = oracle.rmat_graph(10, 8, seed=1)
    np.savez_compressed(os.path.join(HERE, "rmat_s10_ef8_seed1.npz"), off=g.off, idx=g.idx)
    print("nV", g.nV, "nE", g.nE)


if __name__ == "__main__":
    main()
