"""Seeded sparse-stereo inputs shared by the oracle and GPU tests: keypoints on
a few rows of a rectified pair, right keypoints shifted by a disparity and
jittered in y, descriptors of the right view copied from the left with a few
bits flipped, plus distractors (other rows, other descriptors, NaN
positions)."""
import numpy as np

CORNER = np.dtype([("x", "<f4"), ("y", "<f4"), ("response", "<f4"), ("pyramid_level", "<u4")])


def stereo_case(n_left, n_right, seed, n_rows=12, flips=6):
    rng = np.random.default_rng(seed)
    lk = np.zeros(n_left, CORNER)
    rows = rng.uniform(0, 480, n_rows).astype(np.float32)
    lk["y"] = rows[rng.integers(0, n_rows, n_left)] + rng.uniform(-0.5, 0.5, n_left)
    lk["x"] = rng.uniform(0, 640, n_left)
    lk["response"] = rng.uniform(0, 1, n_left)
    lk["pyramid_level"] = rng.integers(0, 4, n_left)
    ld = rng.integers(0, 2 ** 32, (n_left, 8), dtype=np.uint64).astype(np.uint32)
    rk = np.zeros(n_right, CORNER)
    rd = rng.integers(0, 2 ** 32, (n_right, 8), dtype=np.uint64).astype(np.uint32)
    m = min(n_left, n_right)
    src = rng.permutation(n_left)[:m] if n_left else np.zeros(0, np.int64)
    slots = rng.permutation(n_right)[: len(src)]
    for s, r in zip(src, slots):
        rk[r]["x"] = lk[s]["x"] - rng.choice([0.5, 1.0, 7.25, 60.0, 119.5, 120.0, 130.0])
        rk[r]["y"] = lk[s]["y"] + rng.uniform(-2.5, 2.5)
        d = ld[s].copy()
        for _ in range(rng.integers(0, flips + 1)):
            w, b = rng.integers(0, 8), rng.integers(0, 32)
            d[w] ^= np.uint32(1 << int(b))
        rd[r] = d
    rest = np.setdiff1d(np.arange(n_right), slots)
    rk["x"][rest] = rng.uniform(0, 640, len(rest))
    rk["y"][rest] = rows[rng.integers(0, n_rows, len(rest))] + rng.uniform(-3, 3, len(rest))
    if n_right > 3:
        rk["y"][rest[:1]] = np.nan
    if n_right > 8:   # duplicates of a right keypoint: ties broken by the lower index
        rk[-2] = rk[slots[0]] if len(slots) else rk[-2]
        rd[-2] = rd[slots[0]] if len(slots) else rd[-2]
    return lk, rk, ld, rd
