"""Pins the remap and rectify-map oracle (oracle/rectify_oracle.cpp) without
trusting it.  The reference ships no test and no fixture for its Undistorter
(vx-vision/src/kernels/undistort.rs has none), so the pins come from what the
sampling and the camera model fix:

- identity maps reproduce the input bit for bit, integer shifts move it with
  repeated edges, half-texel maps give exact midpoints, NaN and far-out
  coordinates take the edge texels (a wrong texel-centre offset, a missing
  clamp or swapped axes fails these);
- an independent float64 numpy bilinear agrees to fp32 rounding on random
  maps (a wrong weight or lerp order fails it);
- unit-K, zero-distortion maps are exactly x + 0.5;
- the maps agree with forward projection through the camera: the output pixel
  back-projected through Knew, rotated by R^T (the inverse of a rotation,
  computed another way than inverse(Knew R)), distorted with the textbook
  Brown-Conrady model and projected through K (a transposed R, swapped p1/p2,
  a wrong sign or Knew/K mixed up fails it)."""
import math

import numpy as np
import pytest

import pyoracle as O


def grid(h, w):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    return xx, yy


def test_identity_maps_reproduce_input():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (13, 17), dtype=np.uint8)
    xx, yy = grid(13, 17)
    assert (O.remap(img, xx + 0.5, yy + 0.5) == img).all()
    f = rng.standard_normal((13, 17)).astype(np.float32)
    assert (O.remap(f, xx + 0.5, yy + 0.5) == f).all()
    # gray8 in, float out: exactly k / 255 in fp32
    got = O.remap(img, xx + 0.5, yy + 0.5, out_gray8=False)
    assert (got == img.astype(np.float32) / np.float32(255.0)).all()


@pytest.mark.parametrize("sx,sy", [(3, 0), (-2, 1), (0, -4), (40, 0)])
def test_integer_shift_repeats_edges(sx, sy):
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (9, 11), dtype=np.uint8)
    xx, yy = grid(9, 11)
    got = O.remap(img, xx + 0.5 + sx, yy + 0.5 + sy)
    cols = np.clip(np.arange(11) + sx, 0, 10)
    rows = np.clip(np.arange(9) + sy, 0, 8)
    assert (got == img[rows][:, cols]).all()


def test_half_texel_midpoints():
    img = np.arange(20, dtype=np.float32).reshape(4, 5) * 4.0
    xx, yy = grid(4, 5)
    got = O.remap(img, xx + 1.0, yy + 0.5)          # halfway to the right neighbour
    right = img[:, np.clip(np.arange(5) + 1, 0, 4)]
    assert (got == (img + right) / 2).all()
    got = O.remap(img, xx + 1.0, yy + 1.0)          # centre of four texels
    nxt = np.clip(np.arange(4) + 1, 0, 3)
    quad = (img + right + img[nxt] + right[nxt]) / 4
    assert (got == quad).all()


def test_nan_and_far_coordinates_take_edges():
    img = np.arange(30, dtype=np.float32).reshape(5, 6)
    mx = np.array([[np.nan, 1e30, -1e30, 3.5]], np.float32)
    my = np.array([[np.nan, -1e30, 1e30, np.inf]], np.float32)
    got = O.remap(img, mx, my)
    assert list(got[0]) == [img[0, 0], img[0, 5], img[4, 0], img[4, 3]]


def bilinear64(img, mx, my):
    """Bilinear sampling with texel centres at i + 0.5 and clamped indices,
    evaluated in float64 with weights (1 - a), a."""
    h, w = img.shape
    u = mx.astype(np.float64) - 0.5
    v = my.astype(np.float64) - 0.5
    i0 = np.floor(u).astype(np.int64)
    j0 = np.floor(v).astype(np.int64)
    a, b = u - i0, v - j0

    def t(i, j):
        return img[np.clip(j, 0, h - 1), np.clip(i, 0, w - 1)].astype(np.float64)

    return ((1 - b) * ((1 - a) * t(i0, j0) + a * t(i0 + 1, j0)) +
            b * ((1 - a) * t(i0, j0 + 1) + a * t(i0 + 1, j0 + 1)))


def test_random_maps_against_float64():
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, (21, 29)).astype(np.float32)
    mx = rng.uniform(-3, 33, (15, 19)).astype(np.float32)
    my = rng.uniform(-3, 25, (15, 19)).astype(np.float32)
    got = O.remap(img, mx, my)
    np.testing.assert_allclose(got, bilinear64(img, mx, my), rtol=0, atol=4e-7)
    # gray8 output: the float64 value quantised, except within rounding of a .5 tie
    g8 = rng.integers(0, 256, (21, 29), dtype=np.uint8)
    got8 = O.remap(g8, mx, my).astype(np.int64)
    want = bilinear64(g8.astype(np.float64) / 255.0, mx, my) * 255.0
    near_tie = np.abs(want - np.floor(want) - 0.5) < 1e-4
    assert (np.abs(got8 - np.rint(want)) <= near_tie).all()
    assert (got8[~near_tie] == np.rint(want[~near_tie])).all()


def base_params(**kw):
    p = dict(fx=1.0, fy=1.0, cx=0.0, cy=0.0, k1=0.0, k2=0.0, p1=0.0, p2=0.0, k3=0.0,
             R=[1, 0, 0, 0, 1, 0, 0, 0, 1], new_fx=1.0, new_fy=1.0, new_cx=0.0, new_cy=0.0)
    p.update(kw)
    return p


def test_unit_camera_maps_are_pixel_centres():
    mx, my, m = O.rectify_maps(base_params(), 23, 7)
    xx, yy = grid(7, 23)
    assert (mx == xx + 0.5).all() and (my == yy + 0.5).all()
    assert (m == np.eye(3, dtype=np.float32).ravel()).all()


def rot(ax, ay, az):
    cx, sx, cy, sy, cz, sz = (math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay),
                              math.cos(az), math.sin(az))
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def forward_model(p, xs, ys):
    """Output pixel -> ray in the rectified camera (inverse of Knew) -> R^T ->
    Brown-Conrady distortion -> K, all in float64."""
    R = np.array(p["R"], np.float64).reshape(3, 3)
    rays = np.stack([(xs - p["new_cx"]) / p["new_fx"], (ys - p["new_cy"]) / p["new_fy"],
                     np.ones_like(xs)])
    P = R.T @ rays.reshape(3, -1)
    x, y = P[0] / P[2], P[1] / P[2]
    r2 = x * x + y * y
    radial = 1 + p["k1"] * r2 + p["k2"] * r2 ** 2 + p["k3"] * r2 ** 3
    xd = x * radial + 2 * p["p1"] * x * y + p["p2"] * (r2 + 2 * x * x)
    yd = y * radial + p["p1"] * (r2 + 2 * y * y) + 2 * p["p2"] * x * y
    return p["fx"] * xd + p["cx"], p["fy"] * yd + p["cy"]


@pytest.mark.parametrize("case", range(4))
def test_maps_match_forward_projection(case):
    """KITTI-like intrinsics (fx ~ 720 px on a 1242 x 375 image)."""
    rng = np.random.default_rng(10 + case)
    R = rot(*rng.uniform(-0.03, 0.03, 3)) if case else np.eye(3)
    p = dict(fx=721.5, fy=718.9, cx=609.6, cy=172.9,
             k1=[0, -0.37, 0.2, -0.28][case], k2=[0, 0.2, -0.1, 0.07][case],
             p1=[0, 1e-3, -2e-3, 4e-4][case], p2=[0, -5e-4, 1e-3, 2e-4][case],
             k3=[0, 0.0, 0.01, -0.02][case], R=list(R.ravel()),
             new_fx=707.0, new_fy=707.0, new_cx=604.1, new_cy=180.5)
    w, h = 1242, 375
    mx, my, _ = O.rectify_maps(p, w, h)
    ys, xs = np.mgrid[0:h:7, 0:w:11].astype(np.float64)
    wx, wy = forward_model(p, xs, ys)
    gx = mx[0:h:7, 0:w:11].astype(np.float64) - 0.5
    gy = my[0:h:7, 0:w:11].astype(np.float64) - 0.5
    np.testing.assert_allclose(gx.ravel(), wx.ravel(), atol=2e-3, rtol=0)
    np.testing.assert_allclose(gy.ravel(), wy.ravel(), atol=2e-3, rtol=0)


def test_singular_rectification_rejected():
    with pytest.raises(ValueError):
        O.rectify_maps(base_params(new_fx=0.0), 4, 4)
