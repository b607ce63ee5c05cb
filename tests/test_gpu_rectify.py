"""GPU parity of the rectification path (include/vxh/rectify.h) against the
oracle, bit for bit: the remap of vx-vision's Undistorter
(vx-vision/shaders/undistort.metal:4-20) in all four format combinations on
random, out-of-range and NaN maps of ragged sizes, input and output of
different sizes; the rectify maps over distortion and rotation cases; the
texture calls and their checks; and a calibrated pair rectified and matched
in one pipeline (maps -> remap left/right -> SGM), equal to the oracle's
chain."""
import numpy as np
import pytest

import pyoracle as O
import vxsgm
from sgm_cases import random_pair
from test_oracle_rectify import rot
from vxsgm import hip

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def random_maps(w, h, in_w, in_h, seed):
    rng = np.random.default_rng(seed)
    mx = rng.uniform(-4, in_w + 4, (h, w)).astype(np.float32)
    my = rng.uniform(-4, in_h + 4, (h, w)).astype(np.float32)
    flat = rng.random((h, w))
    mx[flat < 0.03] = np.nan
    my[flat > 0.97] = np.inf
    mx[(flat > 0.5) & (flat < 0.52)] = -1e30
    # exact texel centres and halves: the ties of the gray8 rounding
    k = (flat > 0.6) & (flat < 0.7)
    mx[k] = np.floor(mx[k]) + 0.5
    my[k] = np.floor(my[k]) + 1.0
    return mx, my


def run_remap(img, mx, my, out_fmt):
    h_in, w_in = img.shape
    h, w = mx.shape
    in_fmt = 1 if img.dtype == np.uint8 else 2
    din = hip.DeviceBuffer.from_array(np.ascontiguousarray(img))
    dmx, dmy = hip.DeviceBuffer.from_array(mx), hip.DeviceBuffer.from_array(my)
    dout = hip.DeviceBuffer(w * h * (1 if out_fmt == 1 else 4))
    dout.fill(0xA5)
    vxsgm.remap_enqueue(din.ptr, in_fmt, w_in, h_in, dmx.ptr, dmy.ptr, dout.ptr, out_fmt, w, h)
    hip.synchronize()
    return dout.download(np.uint8 if out_fmt == 1 else np.float32).reshape(h, w)


@pytest.mark.parametrize("in_fmt,out_fmt", [(1, 1), (1, 2), (2, 1), (2, 2)])
@pytest.mark.parametrize("shape", [(1, 1, 1, 1), (7, 5, 9, 4), (97, 61, 64, 80), (640, 33, 600, 40)])
def test_remap_matches_oracle(in_fmt, out_fmt, shape):
    w, h, in_w, in_h = shape
    rng = np.random.default_rng(w * 7 + in_fmt)
    img = (rng.integers(0, 256, (in_h, in_w), dtype=np.uint8) if in_fmt == 1
           else rng.uniform(-0.2, 1.2, (in_h, in_w)).astype(np.float32))
    mx, my = random_maps(w, h, in_w, in_h, seed=w + 13 * out_fmt)
    got = run_remap(img, mx, my, out_fmt)
    want = O.remap(img, mx, my, out_gray8=(out_fmt == 1))
    assert (bits(got) == bits(want)).all()


PARAMS = [
    vxsgm.RectifyParams(fx=1.0, fy=1.0),
    vxsgm.RectifyParams(fx=721.5, fy=718.9, cx=609.6, cy=172.9, k1=-0.37, k2=0.2, p1=1e-3,
                        p2=-5e-4, k3=0.01, R=tuple(rot(0.01, -0.02, 0.005).ravel()),
                        new_fx=707.0, new_fy=707.0, new_cx=604.1, new_cy=180.5),
    vxsgm.RectifyParams(fx=300.0, fy=310.0, cx=80.0, cy=60.0, k1=0.12, k2=-0.05, k3=0.002),
]


@pytest.mark.parametrize("pi", range(len(PARAMS)))
@pytest.mark.parametrize("size", [(1, 1), (160, 120), (1242, 375)])
def test_rectify_maps_match_oracle(pi, size):
    w, h = size
    p = PARAMS[pi]
    dmx, dmy = hip.DeviceBuffer(w * h * 4), hip.DeviceBuffer(w * h * 4)
    vxsgm.rectify_maps_enqueue(p, w, h, dmx.ptr, dmy.ptr)
    hip.synchronize()
    wx, wy, _ = O.rectify_maps(p.as_dict(), w, h)
    assert (bits(dmx.download(np.float32)) == bits(wx.ravel())).all()
    assert (bits(dmy.download(np.float32)) == bits(wy.ravel())).all()


def test_undistorter_texture_calls(ctx):
    u = vxsgm.Undistorter(ctx)
    try:
        w, h = 120, 80
        img = np.random.default_rng(3).integers(0, 256, (h, w), dtype=np.uint8)
        p = vxsgm.RectifyParams(fx=150.0, fy=150.0, cx=60.0, cy=40.0, k1=-0.2, k2=0.05)
        mx_t, my_t = ctx.texture_output_r32float(w, h), ctx.texture_output_r32float(w, h)
        vxsgm.rectify_maps(ctx, p, mx_t, my_t)
        mx, my, _ = O.rectify_maps(p.as_dict(), w, h)
        assert (bits(mx_t.read_r32float()) == bits(mx.ravel())).all()
        tin = ctx.texture_gray8(img, w, h)
        out8, outf = ctx.texture_output_gray8(w, h), ctx.texture_output_r32float(w, h)
        u.apply(ctx, tin, mx_t, my_t, out8)
        u.apply(ctx, tin, mx_t, my_t, outf)
        assert (out8.read_gray8() == O.remap(img, mx, my).ravel()).all()
        assert (bits(outf.read_r32float()) ==
                bits(O.remap(img, mx, my, out_gray8=False).ravel())).all()
        small = ctx.texture_output_r32float(w - 1, h)
        with pytest.raises(vxsgm.TextureSizeMismatch):
            u.apply(ctx, tin, small, my_t, out8)
        with pytest.raises(vxsgm.InvalidConfig):
            u.apply(ctx, tin, out8, my_t, outf)           # a gray8 map
        with pytest.raises(vxsgm.InvalidConfig):
            u.apply(ctx, ctx.texture_output_rgba8(w, h), mx_t, my_t, out8)
    finally:
        u.close()


def test_rectify_then_sgm_in_one_pipeline(ctx, sgm):
    """A distorted, rotated stereo pair: both images rectified on the device,
    then matched, in one pipeline; equal to oracle maps -> oracle remap ->
    oracle SGM."""
    w, h, D = 256, 144, 48
    left, right = random_pair(w, h, seed=91)
    pl = vxsgm.RectifyParams(fx=200.0, fy=200.0, cx=128.0, cy=72.0, k1=-0.08, k2=0.01,
                             R=tuple(rot(0.004, 0.01, -0.003).ravel()),
                             new_fx=190.0, new_fy=190.0, new_cx=126.0, new_cy=70.0)
    pr = vxsgm.RectifyParams(fx=201.0, fy=199.0, cx=129.0, cy=73.0, k1=-0.07, p1=2e-4,
                             R=tuple(rot(0.003, 0.008, 0.002).ravel()),
                             new_fx=190.0, new_fy=190.0, new_cx=126.0, new_cy=70.0)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8)
    u = vxsgm.Undistorter(ctx)
    maps = [ctx.texture_output_r32float(w, h) for _ in range(4)]
    vxsgm.rectify_maps(ctx, pl, maps[0], maps[1])
    vxsgm.rectify_maps(ctx, pr, maps[2], maps[3])
    tl, tr = ctx.texture_gray8(left, w, h), ctx.texture_gray8(right, w, h)
    rl, rr = ctx.texture_output_gray8(w, h), ctx.texture_output_gray8(w, h)
    disp = ctx.texture_output_r32float(w, h)
    pipe = vxsgm.Pipeline.begin(ctx)
    try:
        u.encode(pipe, tl, maps[0], maps[1], rl)
        u.encode(pipe, tr, maps[2], maps[3], rr)
        sgm.encode(pipe, rl, rr, disp, cfg)
        pipe.commit_and_wait()
        mlx, mly, _ = O.rectify_maps(pl.as_dict(), w, h)
        mrx, mry, _ = O.rectify_maps(pr.as_dict(), w, h)
        want_l, want_r = O.remap(left, mlx, mly), O.remap(right, mrx, mry)
        assert (rl.read_gray8() == want_l.ravel()).all()
        assert (rr.read_gray8() == want_r.ravel()).all()
        want = O.sgm(want_l, want_r, D, num_paths=8).ravel()
        assert (bits(disp.read_r32float()) == bits(want)).all()
    finally:
        pipe.close()
        u.close()


def test_kitti_sequence_to_depth(ctx, sgm, tmp_path):
    """A KITTI-layout gray sequence written here (synthetic pairs with a known
    disparity field): frames decoded by the loader, matched on the device,
    converted to depth with the calibration's fx * baseline; equal to the
    oracle on the same decoded pair."""
    from test_datasets import encode_png
    w, h, D = 160, 96, 32
    left, right, _ = vxsgm.synth_pair(w, h, D, 6, 0, 0x4B495454)
    for cam, img in (("image_0", left), ("image_1", right)):
        (tmp_path / cam).mkdir()
        encode_png(tmp_path / cam / "000000.png", img)
    P0 = "700 0 80 0 0 700 48 0 0 0 1 0"
    P1 = "700 0 80 -378 0 700 48 0 0 0 1 0"
    (tmp_path / "calib.txt").write_text(f"P0: {P0}\nP1: {P1}\n")
    ds = vxsgm.KittiStereo(tmp_path)
    cal = ds.calib
    assert cal.baseline == 0.54
    l8, r8 = ds.read_pair(0)
    assert (l8 == left).all() and (r8 == right).all()
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8)
    disp = sgm.compute_host(l8, r8, w, h, cfg)
    want = O.sgm(left, right, D, num_paths=8).ravel()
    assert (bits(disp) == bits(want)).all()
    ddisp = hip.DeviceBuffer.from_array(disp)
    ddepth = hip.DeviceBuffer(w * h * 4)
    vxsgm.disparity_to_depth(ddisp.ptr, ddepth.ptr, w, h, cal.fx, cal.baseline)
    hip.synchronize()
    assert (bits(ddepth.download(np.float32)) ==
            bits(O.disparity_to_depth(want, cal.fx, cal.baseline))).all()
    ds.close()
