"""Pins the sparse stereo matcher oracle (oracle/stereo_match_oracle.cpp),
which restates vx-vision/shaders/StereoMatch.metal:32-115.  The reference has
no test of its StereoMatcher, so the pins are independent restatements and
closed forms: Python's int.bit_count on every pair with the band and range
masks evaluated in numpy float32; identical descriptors at distance 0 and
complementary ones at 256; numpy's argmin (first occurrence) for the
first-of-ties rule (a scan keeping the last tie fails it); the 16-bit
truncation of max_hamming; triangulation against float64."""
import numpy as np
import pytest

import pyoracle as O
from stereo_cases import CORNER, stereo_case

CFG = dict(max_epipolar=2.0, min_disparity=1.0, max_disparity=120.0, max_hamming=50,
           ratio_thresh=0.8, fx=600.0, fy=600.0, cx=320.0, cy=240.0, baseline=0.12)


def brute_matrix(lk, rk, ld, rd, cfg):
    out = np.full((len(lk), len(rk)), 0xFFFF, np.uint16)
    for i in range(len(lk)):
        for j in range(len(rk)):
            dy = np.abs(np.float32(lk[i]["y"]) - np.float32(rk[j]["y"]))
            d = np.float32(lk[i]["x"]) - np.float32(rk[j]["x"])
            if dy <= np.float32(cfg["max_epipolar"]) and np.float32(cfg["min_disparity"]) <= d \
                    <= np.float32(cfg["max_disparity"]):
                out[i, j] = sum(int(a ^ b).bit_count() for a, b in zip(ld[i], rd[j]))
    return out


@pytest.mark.parametrize("nl,nr,seed", [(1, 1, 0), (7, 5, 1), (40, 50, 2), (33, 9, 3)])
def test_matrix_against_bit_count(nl, nr, seed):
    lk, rk, ld, rd = stereo_case(nl, nr, seed)
    assert (O.stereo_matrix(lk, rk, ld, rd, CFG) == brute_matrix(lk, rk, ld, rd, CFG)).all()


def test_matrix_closed_forms():
    lk = np.zeros(3, CORNER)
    lk["x"], lk["y"] = [100, 100, 100], [50, 50, 50]
    rk = np.zeros(3, CORNER)
    rk["x"], rk["y"] = [90, 90, 90], [51.5, 52.5, 50]
    rk["x"][2] = 100.5                      # disparity -0.5: outside [1, 120]
    d = np.random.default_rng(0).integers(0, 2 ** 32, (3, 8), dtype=np.uint64).astype(np.uint32)
    rd = np.stack([d[0], ~d[1], d[2]])
    m = O.stereo_matrix(lk, rk, d, rd, CFG)
    assert m[0, 0] == 0 and m[1, 0] == int(sum(int(a ^ b).bit_count() for a, b in zip(d[1], d[0])))
    assert m[1, 1] == 0xFFFF                # |dy| = 2.5 > 2
    assert (m[:, 2] == 0xFFFF).all()
    lk2 = lk[:1].copy()
    rk2 = rk[:1].copy()
    assert O.stereo_matrix(lk2, rk2, d[1:2], ~d[1:2], CFG)[0, 0] == 256


def brute_match(lk, rk, ld, rd, cfg):
    mat = brute_matrix(lk, rk, ld, rd, cfg)
    out = []
    f32 = np.float32
    for i in range(len(lk)):
        j = int(np.argmin(mat[i]))            # first occurrence of the minimum
        best = int(mat[i, j])
        if best > (cfg["max_hamming"] & 0xFFFF):
            continue
        disp = f32(lk[i]["x"]) - f32(rk[j]["x"])
        Z = f32(cfg["fx"]) * f32(cfg["baseline"]) / disp
        X = Z * (f32(lk[i]["x"]) - f32(cfg["cx"])) / f32(cfg["fx"])
        Y = Z * (f32(lk[i]["y"]) - f32(cfg["cy"])) / f32(cfg["fy"])
        out.append((i, j, disp, X, Y, Z))
    return out


def as_tuples(m):
    return [(int(r["left_idx"]), int(r["right_idx"]), r["disparity"], *r["point_3d"]) for r in m]


@pytest.mark.parametrize("nl,nr,seed,maxh", [(40, 50, 4, 50), (64, 64, 5, 3), (30, 80, 6, 256),
                                             (25, 25, 7, 65536 + 10), (12, 20, 8, 65535)])
def test_match_against_brute_force(nl, nr, seed, maxh):
    lk, rk, ld, rd = stereo_case(nl, nr, seed)
    cfg = dict(CFG, max_hamming=maxh)
    got = O.stereo_match(lk, rk, ld, rd, cfg)
    want = brute_match(lk, rk, ld, rd, cfg)
    assert len(got) == len(want)
    for g, w in zip(as_tuples(got), want):
        assert g[:2] == w[:2]
        assert np.array(g[2:], np.float32).view(np.uint32).tolist() == \
            np.array(w[2:], np.float32).view(np.uint32).tolist()
    if maxh == 65535:   # every row is accepted, all-excluded rows at right index 0
        assert len(got) == nl


def test_ties_take_the_first_right_index():
    lk = np.zeros(1, CORNER)
    lk["x"], lk["y"] = 200, 10
    rk = np.zeros(4, CORNER)
    rk["x"], rk["y"] = [150, 160, 170, 180], [10, 10, 10, 10]
    d = np.arange(8, dtype=np.uint32)[None]
    rd = np.repeat(d, 4, axis=0)
    rd[0, 0] ^= 1                            # index 0 at distance 1, 1..3 at 0
    m = O.stereo_match(lk, rk, d, rd, CFG)
    assert len(m) == 1 and m[0]["right_idx"] == 1 and m[0]["disparity"] == 40


def test_triangulation_against_float64():
    lk, rk, ld, rd = stereo_case(60, 70, 9)
    cfg = dict(CFG, fx=458.654, fy=457.296, cx=367.215, cy=248.375, baseline=0.110)
    got = O.stereo_match(lk, rk, ld, rd, cfg)
    assert len(got) > 10
    for r in got:
        lx, ly = float(lk[r["left_idx"]]["x"]), float(lk[r["left_idx"]]["y"])
        disp = lx - float(rk[r["right_idx"]]["x"])
        Z = cfg["fx"] * cfg["baseline"] / disp
        want = [Z * (lx - cfg["cx"]) / cfg["fx"], Z * (ly - cfg["cy"]) / cfg["fy"], Z]
        np.testing.assert_allclose(r["point_3d"], want, rtol=2e-6)


def test_empty_inputs():
    lk, rk, ld, rd = stereo_case(5, 5, 10)
    assert len(O.stereo_match(lk[:0], rk, ld[:0], rd, CFG)) == 0
    assert len(O.stereo_match(lk, rk[:0], ld, rd[:0], CFG)) == 0
