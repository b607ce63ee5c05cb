"""GPU parity of the depth post-processing (include/vxh/depth.h) against the
oracle: the median filter and the disparity-to-depth conversion bit for bit,
the bilateral filter within the last-ulp error of expf (relative 4e-6), on
maps with invalid values (0, negatives, NaN), ties and tiny or ragged sizes,
through the stream-ordered calls and the texture-level DepthFilter calls."""
import numpy as np
import pytest

import pyoracle as O
import vxsgm
from sgm_cases import random_pair
from vxsgm import hip

pytestmark = pytest.mark.gpu

BILATERAL_RTOL = 4e-6   # expf on the device vs the host's, a few ulp per weight


def noisy_map(w, h, seed):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.2, 40.0, (h, w)).astype(np.float32)
    d[rng.random((h, w)) < 0.25] = 0.0
    d[rng.random((h, w)) < 0.05] = -1.0
    d[rng.random((h, w)) < 0.02] = np.nan
    d[rng.random((h, w)) < 0.1] = 7.25
    return d


def run(fn, d, *args):
    h, w = d.shape
    din = hip.DeviceBuffer.from_array(np.ascontiguousarray(d, np.float32))
    dout = hip.DeviceBuffer(w * h * 4)
    dout.fill(0xFF)
    fn(din.ptr, dout.ptr, w, h, *args)
    hip.synchronize()
    return dout.download(np.float32).reshape(h, w)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("shape", [(1, 1), (3, 2), (7, 5), (97, 61), (640, 33)])
@pytest.mark.parametrize("r", [-1, 0, 1, 2, 3, 5, 6])
def test_median_matches_oracle(shape, r):
    w, h = shape
    if r >= 5 and w * h > 4000:
        pytest.skip("large windows on the small maps only")
    d = noisy_map(w, h, 1000 + w + r)
    got = run(vxsgm.DepthFilter.enqueue_median, d, vxsgm.DepthMedianConfig(r))
    assert (bits(got) == bits(O.depth_median(d, r))).all()


@pytest.mark.parametrize("shape", [(1, 1), (7, 5), (97, 61), (640, 33)])
@pytest.mark.parametrize("cfg", [(3, 5.0, 0.05), (0, 5.0, 0.05), (1, 3.0, 0.1), (2, 1.0, 2.0),
                                 (7, 4.0, 1.5), (-1, 5.0, 0.05)])
def test_bilateral_matches_oracle(shape, cfg):
    w, h = shape
    d = noisy_map(w, h, 2000 + w + cfg[0])
    got = run(vxsgm.DepthFilter.enqueue_bilateral, d, vxsgm.DepthFilterConfig(*cfg))
    want = O.depth_bilateral(d, *cfg)
    invalid = d <= 0   # a NaN centre is not skipped by the shader's test: it stays NaN
    assert (got[invalid] == 0).all() and (want[invalid] == 0).all()
    assert (np.isnan(got) == np.isnan(want)).all()
    np.testing.assert_allclose(got, want, rtol=BILATERAL_RTOL, atol=0)


def test_texture_filters_reference_cases(ctx):
    """The reference's own two tests (test_kernels.rs:839-902) through the
    texture-level DepthFilter calls."""
    f = vxsgm.DepthFilter(ctx)
    try:
        w = h = 32
        edge = np.where(np.arange(w)[None, :] < w // 2, 1.0, 5.0).astype(np.float32) * np.ones(
            (h, 1), np.float32)
        tin, tout = ctx.texture_r32float(edge, w, h), ctx.texture_output_r32float(w, h)
        f.apply_bilateral(ctx, tin, tout, vxsgm.DepthFilterConfig(2, 3.0, 0.1))
        r = tout.read_r32float().reshape(h, w)
        assert r[h // 2, w // 4] == np.float32(1.0) and r[h // 2, 3 * w // 4] == np.float32(5.0)
        noisy = np.full((h, w), 2.0, np.float32)
        noisy.flat[100], noisy.flat[200] = 99.0, 0.5
        tin2 = ctx.texture_r32float(noisy, w, h)
        f.apply_median(ctx, tin2, tout, vxsgm.DepthMedianConfig())
        r = tout.read_r32float()
        assert r[100] == np.float32(2.0) and r[200] == np.float32(2.0)
        assert (bits(r) == bits(O.depth_median(noisy, 1).ravel())).all()
    finally:
        f.close()


def test_texture_filters_reject_bad_textures(ctx):
    f = vxsgm.DepthFilter(ctx)
    try:
        a = ctx.texture_r32float(np.ones(12, np.float32), 4, 3)
        b = ctx.texture_output_r32float(3, 4)
        with pytest.raises(vxsgm.TextureSizeMismatch):
            f.apply_median(ctx, a, b, vxsgm.DepthMedianConfig())
        g = ctx.texture_gray8(np.zeros(12, np.uint8), 4, 3)
        c = ctx.texture_output_r32float(4, 3)
        with pytest.raises(vxsgm.InvalidConfig):
            f.apply_bilateral(ctx, g, c, vxsgm.DepthFilterConfig())
        with pytest.raises(vxsgm.InvalidConfig):
            f.apply_median(ctx, a, c, vxsgm.DepthMedianConfig(256))
    finally:
        f.close()


def test_disparity_to_depth_after_sgm(ctx, sgm):
    """End to end: the SGM map (LR check on, so the -1 sentinel appears) to
    depth, bit for bit against the oracle's map converted by the oracle."""
    w, h, D = 160, 90, 48
    left, right = random_pair(w, h, seed=77)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8, lr_check=True)
    disp = sgm.compute_host(left, right, w, h, cfg)
    want_disp = O.sgm(left, right, D, num_paths=8, lr_check=True).ravel()
    assert (bits(disp) == bits(want_disp)).all()
    assert (disp == -1.0).any()
    got = run(lambda a, b, ww, hh: vxsgm.disparity_to_depth(a, b, ww, hh, 718.856, 0.5372),
              disp.reshape(h, w))
    want = O.disparity_to_depth(want_disp, 718.856, 0.5372).reshape(h, w)
    assert (bits(got) == bits(want)).all()
    assert (got[disp.reshape(h, w) <= 0] == 0).all()
