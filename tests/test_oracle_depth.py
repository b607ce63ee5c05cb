"""Pins the depth-filter and disparity-to-depth oracle (oracle/depth_oracle.cpp)
without trusting it: the reference's own two DepthFilter tests
(vx-vision/tests/test_kernels.rs:839-902) at their exact inputs, closed forms,
and independent naive restatements (pure Python sorting for the median, a
float64 numpy bilateral) on small maps.  A plausible mistake (median of all
values instead of the first 121 > 0, the lower instead of the upper middle,
invalid neighbours counted, exp of the wrong sign, weights not normalised)
fails one of them."""
import math

import numpy as np
import pytest

import pyoracle as O


def ref_edge_map(w=32, h=32):
    # test_kernels.rs:845-849: left half 1.0, right half 5.0.
    return np.where(np.arange(w)[None, :] < w // 2, 1.0, 5.0).astype(np.float32) * np.ones(
        (h, 1), np.float32)


def test_reference_bilateral_preserves_edges():
    """test_kernels.rs:839-876: radius 2, sigma_spatial 3, sigma_depth 0.1 keeps
    both halves (the cross-edge weights underflow to 0, so each side averages
    equal values and stays exact)."""
    w = h = 32
    out = O.depth_bilateral(ref_edge_map(), 2, 3.0, 0.1)
    assert abs(out[h // 2, w // 4] - 1.0) < 0.5
    assert abs(out[h // 2, 3 * w // 4] - 5.0) < 0.5
    assert out[h // 2, w // 4] == np.float32(1.0)
    assert out[h // 2, 3 * w // 4] == np.float32(5.0)


def test_reference_median_removes_noise():
    """test_kernels.rs:878-902: a 2.0 map with two outliers, default radius 1."""
    d = np.full((32, 32), 2.0, np.float32)
    d.flat[100] = 99.0
    d.flat[200] = 0.5
    out = O.depth_median(d, 1)
    assert abs(out.flat[100] - 2.0) < 1.0
    assert out.flat[100] == np.float32(2.0) and out.flat[200] == np.float32(2.0)


def naive_median(d, r):
    """DepthFilter.metal:66-120 restated with Python lists and sorted()."""
    h, w = d.shape
    out = np.zeros_like(d)
    for y in range(h):
        for x in range(w):
            vals = []
            for dy in range(-r, r + 1):
                for dx in range(-r, r + 1):
                    v = float(d[min(max(y + dy, 0), h - 1), min(max(x + dx, 0), w - 1)])
                    if v > 0.0 and len(vals) < 121:
                        vals.append(v)
            out[y, x] = sorted(vals)[len(vals) // 2] if vals else 0.0
    return out


def small_map(w, h, seed, invalid=0.3):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.2, 9.0, (h, w)).astype(np.float32)
    d[rng.random((h, w)) < invalid] = 0.0
    d[rng.random((h, w)) < 0.05] = -1.0
    d[rng.random((h, w)) < 0.1] = 3.5   # ties
    return d


@pytest.mark.parametrize("r", [-1, 0, 1, 2, 3, 6])
def test_median_against_naive(r):
    """Includes radius 6 (169 window values: only the first 121 valid ones in
    scan order count) and the empty window of a negative radius."""
    d = small_map(13, 9, 10 + r)
    if r == 6:
        d[:] = np.abs(d) + 0.5   # all valid: the 121 cap binds everywhere
        d[0, :] = 50.0
    np.testing.assert_array_equal(O.depth_median(d, r), naive_median(d, r))


def test_median_cap_changes_result():
    """Row r of a 13 x 13 map holds r + 1, all valid.  The centre's radius-6
    window is the whole map; the first 121 values in scan order are rows 0-8
    and 4 of row 9, whose median (index 60) is 5, while the median of all 169
    (index 84) would be 7."""
    d = np.repeat(np.arange(1, 14, dtype=np.float32)[:, None], 13, axis=1)
    out = O.depth_median(d, 6)[6, 6]
    assert out == np.float32(5.0)
    assert sorted(d.ravel().tolist())[84] == 7.0


def naive_bilateral(d, r, ss, sd):
    h, w = d.shape
    out = np.zeros((h, w), np.float64)
    i2s, i2d = 1.0 / (2 * ss * ss), 1.0 / (2 * sd * sd)
    for y in range(h):
        for x in range(w):
            c = float(d[y, x])
            if c <= 0:
                continue
            num = den = 0.0
            for dy in range(-r, r + 1):
                for dx in range(-r, r + 1):
                    n = float(d[min(max(y + dy, 0), h - 1), min(max(x + dx, 0), w - 1)])
                    if n <= 0:
                        continue
                    wt = math.exp(-(dx * dx + dy * dy) * i2s) * math.exp(-(n - c) ** 2 * i2d)
                    num += n * wt
                    den += wt
            out[y, x] = num / den if den > 1e-8 else c
    return out


@pytest.mark.parametrize("r,ss,sd", [(0, 5.0, 0.05), (1, 5.0, 0.05), (3, 3.0, 0.1),
                                     (2, 1.0, 2.0), (-1, 5.0, 0.05)])
def test_bilateral_against_float64(r, ss, sd):
    d = small_map(11, 8, 20 + r)
    got = O.depth_bilateral(d, r, ss, sd)
    want = naive_bilateral(d, r, ss, sd)
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=0)
    assert (got[d <= 0] == 0).all()


def test_bilateral_closed_forms():
    """A constant valid map is a fixed point: exactly for a power of two (every
    n * w and partial sum scales the weight sums exactly), to rounding
    otherwise; a negative radius keeps every valid centre (no weights) and
    zeroes the rest."""
    c = np.full((6, 7), 2.0, np.float32)
    np.testing.assert_array_equal(O.depth_bilateral(c, 3, 5.0, 0.05), c)
    c = np.full((6, 7), 2.25, np.float32)
    np.testing.assert_allclose(O.depth_bilateral(c, 3, 5.0, 0.05), c, rtol=1e-6)
    d = small_map(9, 6, 31)
    np.testing.assert_array_equal(O.depth_bilateral(d, -1, 5.0, 0.05), np.where(d > 0, d, 0))


def test_disparity_to_depth_closed_form():
    """(fx * baseline) / d for d > 0 in fp32, else 0 (0, negatives, the LR
    sentinel -1, NaN)."""
    disp = np.array([2.0, 4.0, 0.0, -1.0, np.nan, 0.5, 64.0], np.float32)
    got = O.disparity_to_depth(disp, 500.0, 0.1)
    fb = np.float32(500.0) * np.float32(0.1)
    want = np.array([fb / np.float32(2.0), fb / np.float32(4.0), 0, 0, 0, fb / np.float32(0.5),
                     fb / np.float32(64.0)], np.float32)
    np.testing.assert_array_equal(got, want)
    assert got[0] == np.float32(25.0)
