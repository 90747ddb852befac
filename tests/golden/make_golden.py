"""Writes the committed fixtures of tests/golden/ (run from the repository root, CPU only).

The reference holds no golden vectors, KATs or sample inputs of its own, and it cannot be built
in this image (its library needs nvcc and cuda_runtime.h).  The fixtures here are therefore
regression pins of this project's own components, not reference outputs:
  synth_hashes.json   FNV-1a hashes of the generator's four output planes for a few
                      (seed, rows, cols); pins the seeded inputs every test and benchmark uses;
  oracle_60x80.npz    the oracle's final flow for one seeded 60x80 pair (rows=60, cam_mode=8)
                      with the reference schedule, plus the inputs; lets a GPU run be compared
                      with a fixed expected output without recomputing it.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "pd-flow_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from pdflow import synth_hash, synth_pair, synth_pair_float  # noqa: E402
from oracle import Oracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    hashes = {}
    for seed, rows, cols in ((1, 120, 160), (2, 60, 80), (3, 480, 640), (4, 240, 320)):
        hashes[f"{seed}_{rows}_{cols}"] = [synth_hash(a) for a in synth_pair(seed, rows, cols)]
    with open(os.path.join(HERE, "synth_hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1)

    I1, Z1, I2, Z2 = synth_pair_float(2024, 60, 80)
    o = Oracle(60, 8)
    dx, dy, dz = o.solve_pair(I1, Z1, I2, Z2)
    np.savez_compressed(os.path.join(HERE, "oracle_60x80.npz"), seed=2024, I1=I1, Z1=Z1, I2=I2,
                        Z2=Z2, dx=dx, dy=dy, dz=dz)


if __name__ == "__main__":
    main()
