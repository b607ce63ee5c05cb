"""GPU parity: every stage of the HIP path against the scalar CPU oracle, bit
for bit (census words, the u16 summed volume, and the fp32 disparity map
compared as raw bits), through the C ABI."""
import ctypes
import hashlib
import json
import os
import threading

import numpy as np
import pytest

import pyoracle as O
import vxsgm
from sgm_cases import bench_pair, checkerboard_pair, random_pair
from vxsgm import hip

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def gpu_census(img, rx=4, ry=3):
    h, w = img.shape
    d_img = hip.DeviceBuffer.from_array(img)
    out = hip.DeviceBuffer(h * w * 8)
    vxsgm.check(vxsgm.lib.vxh_sgm_census(d_img.ptr, w, h, rx, ry, out.ptr, None))
    hip.synchronize()
    return out.download(np.uint32).reshape(h, w, 2)


def gpu_volume(lc, rc, cfg, layout=(0, 0)):
    h, w = lc.shape[:2]
    D = cfg.num_disparities
    Dp = vxsgm.lib.vxh_sgm_volume_stride(D)
    d_lc = hip.DeviceBuffer.from_array(lc)
    d_rc = hip.DeviceBuffer.from_array(rc)
    vol = hip.DeviceBuffer(h * w * Dp * 2)
    c = cfg.to_c()
    vxsgm.check(vxsgm.lib.vxh_sgm_aggregate_layout(d_lc.ptr, d_rc.ptr, w, h, ctypes.byref(c),
                                                   layout[0], layout[1], vol.ptr, None))
    hip.synchronize()
    return vol.download(np.uint16).reshape(h, w, Dp)[:, :, :D]


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_same_disp(got, want):
    got = np.asarray(got, dtype=np.float32).reshape(want.shape)
    diff = bits(got) != bits(want)
    if diff.any():
        idx = np.argwhere(diff)[:5]
        raise AssertionError(f"{diff.sum()} of {diff.size} pixels differ, first at "
                             f"{idx.tolist()}: got {got[diff][:5]} want {want[diff][:5]}")


@pytest.mark.parametrize("w,h,rx,ry", [(64, 48, 4, 3), (33, 7, 4, 3), (5, 3, 4, 3),
                                       (100, 40, 1, 1), (90, 20, 31, 0), (40, 90, 0, 31),
                                       (77, 33, 3, 4), (1, 1, 4, 3), (128, 2, 2, 2)])
def test_census_matches_oracle(w, h, rx, ry):
    img = np.random.default_rng(w * 1000 + h).integers(0, 256, (h, w), dtype=np.uint8)
    np.testing.assert_array_equal(gpu_census(img, rx, ry), O.census(img, rx, ry))


VOLUME_CASES = [
    # (w, h, D, paths)
    (24, 16, 8, 2), (24, 16, 8, 4), (24, 16, 8, 8),
    (40, 30, 16, 8), (64, 9, 1, 8), (31, 17, 2, 8), (70, 5, 7, 4),
    (65, 12, 64, 8), (80, 11, 65, 8), (150, 9, 100, 8), (130, 13, 128, 8),
    (140, 6, 129, 4), (200, 7, 192, 8), (256, 5, 255, 8), (300, 6, 256, 8),
    (257, 1, 64, 8), (3, 40, 3, 8), (1, 1, 1, 8),
    # Rows long enough for the cost kernel's unpredicated runs (a run of 64
    # pixels inside the row with x - d >= 0 for every d), then a partial run.
    (330, 4, 64, 4), (333, 5, 128, 8), (600, 3, 256, 8),
]


@pytest.mark.parametrize("w,h,D,paths", VOLUME_CASES)
def test_volume_matches_oracle(w, h, D, paths):
    left, right = random_pair(w, h, seed=w * 7 + h * 13 + D)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=paths)
    lc, rc = O.census(left), O.census(right)
    want = O.sgm_volume(left, right, D, num_paths=paths)
    got = gpu_volume(lc, rc, cfg)
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        raise AssertionError(f"{len(bad)} volume entries differ; first {bad[:5].tolist()}")


# Lane layouts by class: D <= 64, <= 128, <= 256 (lanes own a scanline pair).
CLASS_LAYOUTS = [[(32, 2), (16, 4), (8, 8)], [(64, 2), (32, 4), (16, 8)],
                 [(64, 4), (32, 8)]]


def layout_class(D):
    return 0 if D <= 64 else (1 if D <= 128 else 2)


@pytest.mark.parametrize("li", [0, 1, 2])
@pytest.mark.parametrize("D", [1, 5, 16, 32, 63, 64, 100, 128, 200, 256])
def test_volume_every_layout(li, D):
    """Every lane layout the autotuner may pick gives the oracle's volume."""
    lays = CLASS_LAYOUTS[layout_class(D)]
    if li >= len(lays):
        pytest.skip("class has two layouts")
    L, K = lays[li]
    w, h = max(D, 48) + 7, 11
    left, right = random_pair(w, h, seed=L * 31 + K * 7 + D)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8)
    want = O.sgm_volume(left, right, D, num_paths=8)
    got = gpu_volume(O.census(left), O.census(right), cfg, (L, K))
    np.testing.assert_array_equal(got, want)


def test_unsupported_layout_rejected():
    left, right = random_pair(64, 4, seed=1)
    cfg = vxsgm.SGMStereoConfig(num_disparities=64, num_paths=2)
    for bad in ((16, 2), (64, 2), (16, 16), (32, 1), (4, 16)):
        with pytest.raises(vxsgm.InvalidConfig):
            gpu_volume(O.census(left), O.census(right), cfg, bad)


@pytest.mark.parametrize("p1,p2", [(10, 120), (0, 0), (1, 2), (200, 1000),
                                   (65535, 65535), (10.9, 120.99), (3000, 60000)])
def test_volume_penalties(p1, p2):
    left, right = random_pair(96, 20, seed=int(p1 + p2))
    cfg = vxsgm.SGMStereoConfig(num_disparities=48, p1=p1, p2=p2, num_paths=8)
    want = O.sgm_volume(left, right, 48, p1=p1, p2=p2, num_paths=8)
    got = gpu_volume(O.census(left), O.census(right), cfg)
    np.testing.assert_array_equal(got, want)


def test_end_to_end_qvga_d64_4path(sgm):
    """Workload qvga_d64_4path: 320x240, D=64, 4 paths, bit-exact."""
    left, right, _, cfg = vxsgm.workload_pair("qvga_d64_4path")
    got = sgm.compute_host(left, right, 320, 240, cfg)
    assert_same_disp(got, O.sgm(left, right, 64, num_paths=4, nthreads=8))


@pytest.mark.parametrize("paths", [2, 4, 8])
def test_reference_test_image(ctx, sgm, paths):
    """The reference's own SGM test input (test_kernels.rs:780-834)."""
    left, right = checkerboard_pair()
    cfg = vxsgm.SGMStereoConfig.new(32)
    cfg.num_paths = paths
    lt = ctx.texture_gray8(left, 64, 64)
    rt = ctx.texture_gray8(right, 64, 64)
    out = ctx.texture_output_r32float(64, 64)
    sgm.compute_disparity(ctx, lt, rt, out, cfg)
    result = out.read_r32float()
    assert result.size == 64 * 64
    assert np.isfinite(result).any()
    assert_same_disp(result, O.sgm(left, right, 32, num_paths=paths))


@pytest.mark.parametrize("w,h", [(320, 240), (640, 480)])
def test_reference_bench_images(sgm, w, h):
    """The reference's criterion inputs (kernel_benchmarks.rs:189-228), D=64."""
    left, right = bench_pair(w, h)
    cfg = vxsgm.SGMStereoConfig.new(64)
    got = sgm.compute_host(left, right, w, h, cfg)
    assert_same_disp(got, O.sgm(left, right, 64, num_paths=2, nthreads=8))


def test_constant_image_gives_zero(sgm):
    img = np.full((37, 90), 117, dtype=np.uint8)
    for paths in (2, 4, 8):
        cfg = vxsgm.SGMStereoConfig(num_disparities=40, num_paths=paths)
        got = sgm.compute_host(img, img, 90, 37, cfg)
        assert (bits(got) == bits(np.zeros_like(got))).all()


@pytest.mark.parametrize("tol", [0, 1, 3])
def test_lr_check_matches_oracle(sgm, tol):
    left, right, _ = vxsgm.synth_pair(160, 96, 48, 5, 1, 99 + tol)
    cfg = vxsgm.SGMStereoConfig(num_disparities=48, num_paths=8, lr_check=True,
                                lr_tolerance=tol)
    got = sgm.compute_host(left, right, 160, 96, cfg)
    want = O.sgm(left, right, 48, num_paths=8, lr_check=True, lr_tol=tol)
    assert_same_disp(got, want)
    assert (want == -1.0).any()


def test_size_mismatch_and_invalid_config(ctx, sgm):
    a = ctx.texture_gray8(np.zeros((20, 30), np.uint8), 30, 20)
    b = ctx.texture_gray8(np.zeros((21, 30), np.uint8), 30, 21)
    out = ctx.texture_output_r32float(30, 20)
    with pytest.raises(vxsgm.TextureSizeMismatch):
        sgm.compute_disparity(ctx, a, b, out, vxsgm.SGMStereoConfig.new(16))
    with pytest.raises(vxsgm.InvalidConfig):
        sgm.compute_disparity(ctx, a, a, out, vxsgm.SGMStereoConfig.new(31))
    with pytest.raises(vxsgm.InvalidConfig):
        ctx.texture_gray8(np.zeros(10, np.uint8), 3, 3)


def test_compute_returns_depth_map(ctx, sgm):
    left, right = checkerboard_pair()
    lt, rt = ctx.texture_gray8(left, 64, 64), ctx.texture_gray8(right, 64, 64)
    k = vxsgm.CameraIntrinsics(500.0, 500.0, 32.0, 32.0, 64, 64)
    dm = sgm.compute(ctx, lt, rt, vxsgm.SGMStereoConfig.new(32), k, 0.1)
    assert dm.min_depth == 0.0 and dm.max_depth == np.finfo(np.float32).max
    assert_same_disp(dm.depth.read_r32float(), O.sgm(left, right, 32))


def test_concurrent_calls_match_serial(sgm):
    """One handle shared by threads (the reference is Send + Sync)."""
    pairs = [vxsgm.synth_pair(200, 120, 64, 4, 1, s)[:2] for s in range(6)]
    cfg = vxsgm.SGMStereoConfig(num_disparities=64, num_paths=8)
    serial = [sgm.compute_host(l, r, 200, 120, cfg) for l, r in pairs]
    results = [None] * len(pairs)

    def run(i):
        l, r = pairs[i]
        results[i] = sgm.compute_host(l, r, 200, 120, cfg)

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(pairs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for a, b in zip(results, serial):
        assert_same_disp(a, b)


@pytest.mark.parametrize("fmt", ["u8", "u16"])
@pytest.mark.parametrize("paths", [2, 4, 8])
@pytest.mark.parametrize("lr", [False, True])
def test_volume_formats_match_oracle(ctx, fmt, paths, lr):
    """Both per-direction volume formats, forced, give the oracle's map, with
    and without the LR check (which reads the summed volume)."""
    D = 40
    left, right, _ = vxsgm.synth_pair(150, 70, D, 5, 1, 700 + paths + (0 if fmt == "u8" else 50))
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=paths, lr_check=lr)
    s = vxsgm.SGMStereo(ctx)
    try:
        s.set_volume_format(D, fmt)
        got = s.compute_host(left, right, 150, 70, cfg)
    finally:
        s.close()
    want = O.sgm(left, right, D, num_paths=paths, lr_check=lr)
    assert_same_disp(got, want)


@pytest.mark.parametrize("p1,p2", [(10, 300), (200, 1000), (3000, 60000), (65535, 65535)])
def test_u8_volumes_with_large_penalties(ctx, p1, p2):
    """u8 volumes forced with P2 far above 191: most slices exceed a byte and
    take the high-byte path, which must still reproduce every u16 sum."""
    D = 48
    left, right, _ = vxsgm.synth_pair(120, 50, D, 4, 1, 4000 + p2)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, p1=p1, p2=p2, num_paths=8, lr_check=True)
    s = vxsgm.SGMStereo(ctx)
    try:
        s.set_volume_format(D, "u8")
        got = s.compute_host(left, right, 120, 50, cfg)
    finally:
        s.close()
    assert_same_disp(got, O.sgm(left, right, D, p1=p1, p2=p2, num_paths=8, lr_check=True))


@pytest.mark.parametrize("hi", [0, 1, 2])
@pytest.mark.parametrize("oi", [0, 1, 2])
def test_schedules_match_oracle(ctx, hi, oi):
    """Every (horizontal, other) layout pair of a class gives the oracle's map."""
    for D, w, h in ((64, 90, 40), (128, 150, 30), (256, 270, 20)):
        lay = CLASS_LAYOUTS[layout_class(D)]
        if hi >= len(lay) or oi >= len(lay):
            continue
        left, right, _ = vxsgm.synth_pair(w, h, D, 3, 1, 1000 + 10 * hi + oi + D)
        cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8)
        s = vxsgm.SGMStereo(ctx)
        try:
            s.set_schedule(D, (lay[hi], lay[oi]))
            assert s.schedule(D) == (lay[hi], lay[oi])
            got = s.compute_host(left, right, w, h, cfg)
        finally:
            s.close()
        assert_same_disp(got, O.sgm(left, right, D, num_paths=8))


@pytest.mark.parametrize("D,w,h", [(1, 9, 5), (7, 40, 9), (130, 150, 12), (200, 230, 20),
                                   (256, 260, 9)])
def test_formats_disparity_ranges(ctx, D, w, h):
    """The summing winner-take-all at slice widths below one chunk, straddling
    two chunks and at the maximum D, in both volume formats."""
    left, right, _ = vxsgm.synth_pair(w, h, D, 3, 1, 900 + D)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8)
    want = O.sgm(left, right, D, num_paths=8)
    s = vxsgm.SGMStereo(ctx)
    try:
        for fmt in ("auto", "u8", "u16"):
            s.set_volume_format(D, fmt)
            assert_same_disp(s.compute_host(left, right, w, h, cfg), want)
    finally:
        s.close()


def test_enqueue_workspace(ctx):
    """enqueue runs in a workspace of vxh_sgm_workspace_bytes (either format)
    and rejects a smaller one."""
    w, h, D = 128, 64, 48
    left, right, _ = vxsgm.synth_pair(w, h, D, 4, 1, 811)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8)
    want = O.sgm(left, right, D, num_paths=8)
    ws = vxsgm.workspace_bytes(w, h, cfg)
    s = vxsgm.SGMStereo(ctx)
    try:
        d_l = hip.DeviceBuffer.from_array(left)
        d_r = hip.DeviceBuffer.from_array(right)
        d_ws = hip.DeviceBuffer(ws)
        st = hip.Stream()
        for fmt in ("u8", "u16"):
            s.set_volume_format(D, fmt)
            out = hip.DeviceBuffer(w * h * 4)
            out.fill(0xFF)
            s.enqueue(d_l.ptr, d_r.ptr, w, h, cfg, out.ptr, d_ws.ptr, ws, st.handle)
            st.synchronize()
            assert_same_disp(out.download(np.float32), want)
        with pytest.raises(vxsgm.InvalidConfig):
            s.enqueue(d_l.ptr, d_r.ptr, w, h, cfg, out.ptr, d_ws.ptr, ws // 4, st.handle)
        st.close()
    finally:
        s.close()


def test_schedule_setting_validated(ctx):
    s = vxsgm.SGMStereo(ctx)
    try:
        default = s.schedule(64)
        with pytest.raises(vxsgm.InvalidConfig):
            s.set_schedule(64, (64, 2))
        with pytest.raises(vxsgm.InvalidConfig):
            s.set_schedule(128, ((64, 2), (8, 8)))
        with pytest.raises(vxsgm.InvalidConfig):
            s.set_schedule(0, (32, 2))
        with pytest.raises(vxsgm.InvalidConfig):
            s.set_volume_format(64, "u32")
        s.set_schedule(64, (16, 4))
        assert s.schedule(64) == ((16, 4), (16, 4))
        s.set_schedule(64, None)
        assert s.schedule(64) == default
    finally:
        s.close()


def test_default_schedule_by_path_count(ctx):
    """A lone pair's default schedule depends on its path count and size;
    what the handle is set to overrides the layouts and the split level
    separately, and None restores each default.  Maps stay the oracle's
    under every default."""
    s = vxsgm.SGMStereo(ctx)
    cfg = lambda D, paths: vxsgm.SGMStereoConfig(num_disparities=D, num_paths=paths)
    try:
        assert s.schedule_for(1280, 720, cfg(128, 8)) == ((32, 4), (16, 8), 0)
        assert s.schedule_for(1280, 720, cfg(128, 2)) == ((32, 4), (32, 4), 1)
        assert s.schedule_for(1920, 1080, cfg(128, 4)) == ((16, 8), (16, 8), 1)
        assert s.schedule_for(1920, 1080, cfg(256, 4)) == ((64, 4), (32, 8), 0)
        assert s.schedule_for(1920, 1080, cfg(256, 2)) == ((32, 8), (32, 8), 1)
        # D <= 64, 8 paths: split only where level 1 splits the longest passes alone.
        assert s.schedule_for(1280, 720, cfg(64, 8)) == ((16, 4), (16, 4), 1)
        assert s.schedule_for(320, 240, cfg(64, 8)) == ((16, 4), (16, 4), 1)
        assert s.schedule_for(640, 480, cfg(64, 8)) == ((16, 4), (16, 4), 0)
        assert s.schedule_for(480, 640, cfg(64, 8)) == ((16, 4), (16, 4), 0)
        assert s.schedule_for(240, 400, cfg(64, 8)) == ((16, 4), (16, 4), 1)
        # The 1280x720 8-path view of the query by D alone.
        assert s.schedule(128) == ((32, 4), (16, 8)) and s.split(128) == 0
        s.set_split(128, 0)
        assert s.schedule_for(1280, 720, cfg(128, 2)) == ((32, 4), (32, 4), 0)
        s.set_schedule(128, (64, 2))
        assert s.schedule_for(1280, 720, cfg(128, 2)) == ((64, 2), (64, 2), 0)
        s.set_split(128, None)
        assert s.schedule_for(1280, 720, cfg(128, 2)) == ((64, 2), (64, 2), 1)
        s.set_schedule(128, None)
        assert s.schedule_for(1280, 720, cfg(128, 2)) == ((32, 4), (32, 4), 1)
        with pytest.raises(vxsgm.InvalidConfig):
            s.set_split(128, 3)
        with pytest.raises(vxsgm.InvalidConfig):
            s.schedule_for(1280, 720, cfg(300, 8))
        for (w, h, D, paths) in [(600, 300, 64, 8), (330, 250, 64, 8), (520, 200, 128, 2),
                                 (300, 520, 128, 4), (400, 260, 200, 2)]:
            left, right, _ = vxsgm.synth_pair(w, h, D, 4, 1, 7100 + w + D + paths)
            c = cfg(D, paths)
            want = O.sgm(left, right, D, num_paths=paths)
            assert_same_disp(s.compute_host(left, right, w, h, c), want)
    finally:
        s.close()


@pytest.mark.parametrize("w,h,D", [(600, 80, 64), (110, 420, 100), (512, 96, 128), (300, 40, 256),
                                   (257, 30, 32), (96, 96, 16), (300, 290, 48), (270, 700, 96)])
@pytest.mark.parametrize("fmt", ["u8", "u16"])
def test_split_scanlines_match_oracle(ctx, w, h, D, fmt):
    """Scanlines of 256 pixels or more (level 1: in the directions whose
    scanlines are longest; level 2: in every direction, diagonals included)
    run as a head and a warmed-up tail at once; the head's check and any
    overwrite of the tail's first steps must leave the oracle's map, with and
    without the LR check, and identical to the unsplit run (96 x 96: no
    scanline is split)."""
    left, right, _ = vxsgm.synth_pair(w, h, D, 4, 1, 5000 + w + h + D)
    s = vxsgm.SGMStereo(ctx)
    try:
        s.set_volume_format(D, fmt)
        for lr in (False, True):
            cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8, lr_check=lr)
            want = O.sgm(left, right, D, num_paths=8, lr_check=lr)
            for level in (1, 2):
                s.set_split(D, level)
                assert s.split(D) == level
                assert_same_disp(s.compute_host(left, right, w, h, cfg), want)
            s.set_split(D, False)
            assert_same_disp(s.compute_host(left, right, w, h, cfg), want)
    finally:
        s.close()


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("p2", [300, 2000])
def test_split_scanlines_with_wraps(ctx, p2, level):
    """u8 volumes forced with a large P2: the tail's and head's exception flags
    and high bytes in the re-run steps must resolve to the oracle's sums (a
    flagged step never counts as met, so heads run on past the early range and
    wait for their tails to end)."""
    D, w, h = 64, 560, 110
    left, right, _ = vxsgm.synth_pair(w, h, D, 5, 1, 6000 + p2)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, p2=p2, num_paths=8)
    s = vxsgm.SGMStereo(ctx)
    try:
        s.set_volume_format(D, "u8")
        s.set_split(D, level)
        got = s.compute_host(left, right, w, h, cfg)
    finally:
        s.close()
    assert_same_disp(got, O.sgm(left, right, D, p2=p2, num_paths=8))


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("fmt", ["u8", "u16", "auto"])
def test_split_slow_convergence(ctx, fmt, level):
    """Penalties far above the costs (P1 = 1000, P2 = 2000) keep a fresh
    start's state apart from the true one for longer than the tail's warm-up:
    heads then continue through the tail's early range, wait for the tail to
    end and overwrite its values until the states meet."""
    D, w, h = 64, 700, 64
    left, right, _ = vxsgm.synth_pair(w, h, D, 6, 1, 6100)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, p1=1000, p2=2000, num_paths=8)
    s = vxsgm.SGMStereo(ctx)
    try:
        s.set_volume_format(D, fmt)
        s.set_split(D, level)
        got = s.compute_host(left, right, w, h, cfg)
    finally:
        s.close()
    assert_same_disp(got, O.sgm(left, right, D, p1=1000, p2=2000, num_paths=8))


def test_720p_d128_8path_bit_exact(ctx, sgm):
    """Workload 720p_d128_8path at full size against the threaded oracle, in
    both volume formats and with the narrowest and widest layouts."""
    left, right, _, cfg = vxsgm.workload_pair("720p_d128_8path")
    want = O.sgm(left, right, 128, num_paths=8, nthreads=16)
    got = sgm.compute_host(left, right, 1280, 720, cfg)
    assert_same_disp(got, want)
    s = vxsgm.SGMStereo(ctx)
    try:
        s.set_volume_format(128, "u16")
        assert_same_disp(s.compute_host(left, right, 1280, 720, cfg), want)
        s.set_volume_format(128, "auto")
        s.set_schedule(128, ((32, 4), (64, 2)))
        assert_same_disp(s.compute_host(left, right, 1280, 720, cfg), want)
        # Split scanlines at full size (the tuner tries them on this pair).
        s.set_split(128, True)
        assert s.split(128)
        assert_same_disp(s.compute_host(left, right, 1280, 720, cfg), want)
        s.set_schedule(128, ((16, 8), (16, 8)))
        assert s.split(128)
        assert_same_disp(s.compute_host(left, right, 1280, 720, cfg), want)
        s.set_split(128, 2)
        s.set_schedule(128, ((32, 4), (16, 8)))
        assert_same_disp(s.compute_host(left, right, 1280, 720, cfg), want)
    finally:
        s.close()


@pytest.mark.parametrize("workload,D", [("720p_d128_8path", 128), ("1080p_d256_8path", 256),
                                        ("720p_d128_8path", 64)])
def test_default_config_two_paths_full_size(sgm, workload, D):
    """The reference's default configuration (2 paths, P1 10, P2 120, 9x7
    census) at full size, whose default schedule splits the 1280- and
    1920-pixel scanlines, against the threaded oracle, with and without the
    LR check."""
    left, right, _, _ = vxsgm.workload_pair(workload)
    h, w = left.shape
    for lr in (False, True):
        cfg = vxsgm.SGMStereoConfig(num_disparities=D, lr_check=lr)
        assert cfg.num_paths == 2
        assert sgm.schedule_for(w, h, cfg)[2] == 1
        want = O.sgm(left, right, D, num_paths=2, lr_check=lr, nthreads=16)
        assert_same_disp(sgm.compute_host(left, right, w, h, cfg), want)


def test_1080p_d256_8path_matches_golden_hash(ctx, sgm):
    """Workload 1080p_d256_8path: the GPU map's SHA-256 equals the oracle's,
    recorded in tests/golden/workload_hashes.json, in both volume formats."""
    with open(os.path.join(GOLDEN, "workload_hashes.json")) as f:
        golden = json.load(f)
    left, right, _, cfg = vxsgm.workload_pair("1080p_d256_8path")
    got = sgm.compute_host(left, right, 1920, 1080, cfg)
    assert hashlib.sha256(bits(got).tobytes()).hexdigest() == golden["1080p_d256_8path"]
    s = vxsgm.SGMStereo(ctx)
    try:
        s.set_volume_format(256, "u16")
        got = s.compute_host(left, right, 1920, 1080, cfg)
        assert hashlib.sha256(bits(got).tobytes()).hexdigest() == golden["1080p_d256_8path"]
        s.set_volume_format(256, "auto")
        for level in (1, 2):
            s.set_split(256, level)
            got = s.compute_host(left, right, 1920, 1080, cfg)
            assert hashlib.sha256(bits(got).tobytes()).hexdigest() == golden["1080p_d256_8path"]
    finally:
        s.close()


@pytest.mark.parametrize("workload", ["720p_d128_8path", "1080p_d256_8path"])
def test_lr_check_full_size_matches_golden_hash(ctx, workload):
    """The two 8-path workloads with the LR check (tolerance 1): the map's
    SHA-256 equals the oracle's (tests/golden/workload_hashes.json, name +
    "_lr"), in both volume formats, through the plain call and through the
    batch's persistent form (multi-pair winner kernel)."""
    with open(os.path.join(GOLDEN, "workload_hashes.json")) as f:
        want = json.load(f)[workload + "_lr"]
    left, right, _, cfg = vxsgm.workload_pair(workload)
    cfg.lr_check = True
    h, w = left.shape
    D = cfg.num_disparities
    s = vxsgm.SGMStereo(ctx)
    try:
        for fmt in ("auto", "u16"):
            s.set_volume_format(D, fmt)
            got = s.compute_host(left, right, w, h, cfg)
            assert hashlib.sha256(bits(got).tobytes()).hexdigest() == want, fmt
    finally:
        s.close()
    dl, dr = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
    do = hip.DeviceBuffer(w * h * 4)
    b = vxsgm.SGMBatch([(dl.ptr, dr.ptr, do.ptr, w, h, cfg, None, "auto", False)],
                       options={"form": "persistent"})
    try:
        b.run()
        got = do.download(np.float32)
        assert hashlib.sha256(bits(got).tobytes()).hexdigest() == want
    finally:
        b.close()


def test_enqueue_reports_expired_split_wait(ctx):
    """A device-resident caller learns of an expired split-scanline wait: with
    the bounded wait forced to zero polls from the host (every head gives up at
    once, no kernel waits at all), vxh_sgm_workspace_status reports Gpu after
    the enqueue; with the default wait the same call is VXH_OK and exact, and
    the next call clears the word."""
    w, h, D = 600, 48, 64
    left, right, _ = vxsgm.synth_pair(w, h, D, 4, 1, 9100)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8)
    want = O.sgm(left, right, D, num_paths=8)
    ws = vxsgm.workspace_bytes(w, h, cfg)
    assert vxsgm.lib.vxh_sgm_workspace_error_offset() == 4
    s = vxsgm.SGMStereo(ctx)
    try:
        s.set_split(D, True)
        d_l = hip.DeviceBuffer.from_array(left)
        d_r = hip.DeviceBuffer.from_array(right)
        d_ws = hip.DeviceBuffer(ws)
        out = hip.DeviceBuffer(w * h * 4)
        st = hip.Stream()
        s.set_split_wait(0)
        s.enqueue(d_l.ptr, d_r.ptr, w, h, cfg, out.ptr, d_ws.ptr, ws, st.handle)
        st.synchronize()
        with pytest.raises(vxsgm.Gpu):
            s.workspace_status(d_ws.ptr)
        with pytest.raises(vxsgm.Gpu):
            s.compute_host(left, right, w, h, cfg)
        s.set_split_wait(None)
        s.enqueue(d_l.ptr, d_r.ptr, w, h, cfg, out.ptr, d_ws.ptr, ws, st.handle)
        st.synchronize()
        s.workspace_status(d_ws.ptr)
        assert_same_disp(out.download(np.float32), want)
        st.close()
    finally:
        s.close()


# Seeded random configurations through the whole pipeline, bit for bit against
# the oracle: size, D, penalties, census window, path count and LR check drawn
# together (windows of at most 64 bits, D <= width).
WINDOWS = [(4, 3), (3, 3), (3, 4), (1, 1), (2, 2), (0, 5), (5, 2), (31, 0)]


@pytest.mark.parametrize("seed", range(24))
def test_random_configs_match_oracle(sgm, seed):
    rng = np.random.default_rng(0x5347_1000 + seed)
    w, h = int(rng.integers(1, 300)), int(rng.integers(1, 48))
    D = int(rng.integers(1, min(w, 256) + 1))
    paths = int(rng.choice([2, 4, 8]))
    rx, ry = WINDOWS[int(rng.integers(len(WINDOWS)))]
    p1, p2 = float(rng.integers(0, 260)), float(rng.integers(0, 1500))
    lr, tol = bool(rng.integers(0, 2)), int(rng.integers(0, 4))
    left, right = random_pair(w, h, seed=1000 + seed)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, p1=p1, p2=p2, census_radius_x=rx,
                                census_radius_y=ry, num_paths=paths, lr_check=lr,
                                lr_tolerance=tol)
    got = sgm.compute_host(left, right, w, h, cfg)
    want = O.sgm(left, right, D, p1, p2, rx, ry, num_paths=paths, lr_check=lr, lr_tol=tol)
    assert_same_disp(got, want)


def random_case(seed, max_w, max_h):
    """A seeded random configuration and pair, as test_random_configs_match_oracle
    draws them, up to max_w x max_h; returns (left, right, cfg, oracle map)."""
    rng = np.random.default_rng(0x5347_2000 + seed)
    w, h = int(rng.integers(1, max_w)), int(rng.integers(1, max_h))
    D = int(rng.integers(1, min(w, 256) + 1))
    paths = int(rng.choice([2, 4, 8]))
    rx, ry = WINDOWS[int(rng.integers(len(WINDOWS)))]
    p1, p2 = float(rng.integers(0, 260)), float(rng.integers(0, 1500))
    lr, tol = bool(rng.integers(0, 2)), int(rng.integers(0, 4))
    left, right = random_pair(w, h, seed=5000 + seed)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, p1=p1, p2=p2, census_radius_x=rx,
                                census_radius_y=ry, num_paths=paths, lr_check=lr,
                                lr_tolerance=tol)
    want = O.sgm(left, right, D, p1, p2, rx, ry, num_paths=paths, lr_check=lr, lr_tol=tol,
                 nthreads=8)
    return left, right, cfg, want


@pytest.mark.parametrize("seed", range(24))
def test_random_larger_configs_match_oracle(sgm, seed):
    """Random configurations up to 720 x 320: scanlines of 256 pixels or more
    in either axis and the diagonals, so the default schedules' split
    scanlines (2 and 4 paths, and 8 paths at D <= 64) run, with and without
    the LR check."""
    left, right, cfg, want = random_case(seed, 720, 320)
    h, w = left.shape
    assert_same_disp(sgm.compute_host(left, right, w, h, cfg), want)


@pytest.mark.parametrize("form", ["persistent", "streams"])
def test_random_configs_in_batch_forms(form):
    """Sixteen of those random configurations in one batch, in the persistent
    form (multi-pair winner kernels of every kind: one or two chunks, u8 or
    u16 volumes, LR checks) and the streams form."""
    cases = [random_case(seed, 720, 320) for seed in range(24, 40)]
    keep, descs = [], []
    for left, right, cfg, _ in cases:
        h, w = left.shape
        dl, dr = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
        do = hip.DeviceBuffer(w * h * 4)
        keep.append((dl, dr, do))
        descs.append((dl.ptr, dr.ptr, do.ptr, w, h, cfg, None, "auto", False))
    b = vxsgm.SGMBatch(descs, options={"form": form})
    try:
        b.run()
        for (left, _, _, want), (_, _, do) in zip(cases, keep):
            assert_same_disp(do.download(np.float32), want)
    finally:
        b.close()
