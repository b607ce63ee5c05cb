"""Writes the golden fixtures under tests/golden/ from the scalar oracle.

    python tests/golden/make_golden.py [workloads] [small] [mixtures]

(no argument: all three).

workload_hashes.json: SHA-256 of the oracle's fp32 disparity map (raw little-
endian bits) for each named workload, and for the two 8-path ones also with
the LR check (name + "_lr"), so a GPU run at full size is checked bit for bit
without re-running the oracle.  small_cases.npz: inputs and expected
volumes/maps of tiny pairs for 2, 4 and 8 paths.  mixture_hashes.json: the
same SHA-256 for every pair of the batched workload (the 64-pair mixture of
sizes {320x240, 640x480, 1280x720, 1920x1080} x D {64, 128, 256}, 8 paths)
and of the serving workload's 32-pair request pool, each entry with its
(w, h, D, pair seed) so a test can check that it regenerates the same pair.
The reference itself cannot run here (Metal + Rust), so these vectors come
from the oracle, which the CPU tests pin independently (hand-derived vectors,
closed forms, invariants).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "vx-rs_amd"),
                os.path.join(ROOT, "tests")]

import pyoracle as O  # noqa: E402
import vxsgm  # noqa: E402
from sgm_cases import random_pair  # noqa: E402

# The batched workload and the serving workload's request pool: (pairs, seed)
# of vxsgm.mixture.
MIXTURES = {"batch_64": (64, 0x5347_0003), "serve_pool_32": (32, 0x5347_0004)}


def map_sha256(disp):
    return hashlib.sha256(np.ascontiguousarray(disp, np.float32).view(np.uint32)
                          .tobytes()).hexdigest()


def workloads():
    hashes = {}
    for name in vxsgm.WORKLOADS:
        left, right, _, cfg = vxsgm.workload_pair(name)
        disp = O.sgm(left, right, cfg.num_disparities, num_paths=cfg.num_paths,
                     nthreads=os.cpu_count())
        hashes[name] = map_sha256(disp)
        print(name, hashes[name], flush=True)
    # The two 8-path workloads with the LR check (tolerance 1): name + "_lr".
    for name in ("720p_d128_8path", "1080p_d256_8path"):
        left, right, _, cfg = vxsgm.workload_pair(name)
        disp = O.sgm(left, right, cfg.num_disparities, num_paths=cfg.num_paths, lr_check=True,
                     lr_tol=1, nthreads=os.cpu_count())
        hashes[name + "_lr"] = map_sha256(disp)
        print(name + "_lr", hashes[name + "_lr"], flush=True)
    with open(os.path.join(HERE, "workload_hashes.json"), "w") as f:
        json.dump(hashes, f, indent=1, sort_keys=True)


def small():
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


def mixtures():
    out = {}
    for name, (n, seed) in MIXTURES.items():
        rows = []
        t0 = time.perf_counter()
        for w, h, d, pair_seed in vxsgm.mixture(n, seed=seed):
            left, right, cfg = vxsgm.mixture_pair(w, h, d, pair_seed)
            disp = O.sgm(left, right, d, p1=cfg.p1, p2=cfg.p2, num_paths=cfg.num_paths,
                         nthreads=os.cpu_count())
            rows.append({"w": w, "h": h, "d": d, "seed": pair_seed, "paths": cfg.num_paths,
                         "sha256": map_sha256(disp)})
        out[name] = {"seed": seed, "pairs": rows}
        print(f"{name}: {n} pairs in {time.perf_counter() - t0:.1f} s", flush=True)
    with open(os.path.join(HERE, "mixture_hashes.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main(parts):
    parts = parts or ["workloads", "small", "mixtures"]
    for p in parts:
        {"workloads": workloads, "small": small, "mixtures": mixtures}[p]()


if __name__ == "__main__":
    main(sys.argv[1:])
