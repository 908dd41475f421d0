"""Writes adjoint_8x8_d6_chi36.npz: the oracle's environment gradient of the
log norm for the 8x8 Z2 PEPS with D = 6 as (3, 3), seed 7 (random_state), at
chi = 36, at six sites, plus the log norm. The oracle needs several minutes
for this size, so the GPU test reads the stored vectors.

    python tests/golden/make_adjoint_8x8.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.model import Lattice, random_state  # noqa: E402
from oracle import adjoint as oa  # noqa: E402

SITES = [(0, 0), (7, 0), (3, 4), (4, 3), (5, 6), (7, 7)]


def main():
    state = random_state(Lattice(8, 8), (3, 3), "peps", seed=7)
    grads, ln = oa.lognorm_gradient(state, 36)
    out = {f"g_{x}_{y}": grads[(x, y)].a for (x, y) in SITES}
    out["sites"] = np.array(SITES)
    out["log_norm"] = np.array(ln["log_norm"])
    np.savez_compressed(os.path.join(HERE, "adjoint_8x8_d6_chi36.npz"), **out)


if __name__ == "__main__":
    main()
