"""Writes the golden fixtures under tests/golden/ from the scalar oracle.

    python tests/golden/make_golden.py

workload_hashes.json: SHA-256 of the oracle's fp32 disparity map (raw little-
endian bits) for each named workload, so a GPU run at full size is checked bit
for bit without re-running the oracle.  small_cases.npz: inputs and expected
volumes/maps of tiny pairs for 2, 4 and 8 paths.  The reference itself cannot
run here (Metal + Rust), so these vectors come from the oracle, which the CPU
tests pin independently (hand-derived vectors, closed forms, invariants).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "vx-rs_amd"),
                os.path.join(ROOT, "tests")]

import pyoracle as O  # noqa: E402
import vxsgm  # noqa: E402
from sgm_cases import random_pair  # noqa: E402


def main():
    hashes = {}
    for name in vxsgm.WORKLOADS:
        left, right, _, cfg = vxsgm.workload_pair(name)
        disp = O.sgm(left, right, cfg.num_disparities, num_paths=cfg.num_paths,
                     nthreads=os.cpu_count())
        hashes[name] = hashlib.sha256(disp.astype(np.float32).view(np.uint32)
                                      .tobytes()).hexdigest()
        print(name, hashes[name], flush=True)
    with open(os.path.join(HERE, "workload_hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1, sort_keys=True)

    arrays = {}
    for i, (w, h, d) in enumerate([(24, 16, 8), (40, 30, 16)]):
        left, right = random_pair(w, h, seed=1000 + i)
        arrays[f"left_{i}"] = left
        arrays[f"right_{i}"] = right
        for paths in (2, 4, 8):
            S = O.sgm_volume(left, right, d, num_paths=paths)
            arrays[f"vol_{i}_{paths}"] = S
            arrays[f"disp_{i}_{paths}"] = O.wta(S)
    np.savez_compressed(os.path.join(HERE, "small_cases.npz"), **arrays)


if __name__ == "__main__":
    main()
