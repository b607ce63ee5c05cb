"""Pipelines (include/vxh/pipeline.h; vx-vision/src/pipeline.rs:12-80 and the
encode() half of every reference kernel, docs/src/adding-a-kernel.md:101-138).

SGM and the depth filters encoded into one pipeline give, after the commit,
bit for bit what the oracle gives (the bilateral filter within the expf
tolerance of the depth tests); nothing encoded runs before the commit or when
the pipeline is dropped uncommitted; a committed pipeline relaunches on the
textures' current contents; retained textures come back; a failed encode
leaves the pipeline usable; scratch returns to the context's pool."""
import numpy as np
import pytest

import pyoracle as O
import vxsgm
from sgm_cases import random_pair

pytestmark = pytest.mark.gpu

BILATERAL_RTOL = 4e-6   # expf on the device vs the host's, a few ulp per weight
SENTINEL = -7.5


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def sentinel_map(ctx, w, h):
    return ctx.texture_r32float(np.full(w * h, SENTINEL, np.float32), w, h)


def pair_textures(ctx, w, h, seed):
    left, right = random_pair(w, h, seed=seed)
    return left, right, ctx.texture_gray8(left, w, h), ctx.texture_gray8(right, w, h)


@pytest.mark.parametrize("paths,lr", [(2, False), (4, True), (8, False), (8, True)])
def test_sgm_encode_matches_oracle(ctx, sgm, paths, lr):
    w, h, D = 160, 90, 48
    left, right, tl, tr = pair_textures(ctx, w, h, 11 + paths)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=paths, lr_check=lr)
    out = ctx.texture_output_r32float(w, h)
    pipe = vxsgm.Pipeline.begin(ctx)
    try:
        assert pipe.stream and not pipe.is_committed()
        sgm.encode(pipe, tl, tr, out, cfg)
        pipe.retain_all([tl, tr])
        kept = pipe.commit_and_wait()
        assert pipe.is_committed() and pipe.stream is None
        assert [(t.width, t.height, t.format) for t in kept] == [(w, h, 1)] * 2
        assert (kept[0].read_gray8() == left.ravel()).all()
        assert pipe.node_count() >= 6   # clear + census x2 + cost + paths + wta (+ lr, + or)
        want = O.sgm(left, right, D, num_paths=paths, lr_check=lr).ravel()
        assert (bits(out.read_r32float()) == bits(want)).all()
    finally:
        pipe.close()


def test_nothing_runs_before_commit(ctx, sgm):
    w, h, D = 96, 64, 32
    left, right, tl, tr = pair_textures(ctx, w, h, 21)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=4)
    out = sentinel_map(ctx, w, h)
    pipe = vxsgm.Pipeline.begin(ctx)
    try:
        sgm.encode(pipe, tl, tr, out, cfg)
        assert (out.read_r32float() == SENTINEL).all()   # recorded, not run
        pipe.commit()
        pipe.wait()
        pipe.wait()                                       # a second wait returns at once
        pipe.commit()                                     # a second commit does nothing
        want = O.sgm(left, right, D, num_paths=4).ravel()
        assert (bits(out.read_r32float()) == bits(want)).all()
    finally:
        pipe.close()


def test_uncommitted_pipeline_is_discarded(ctx, sgm):
    w, h, D = 64, 48, 16
    _, _, tl, tr = pair_textures(ctx, w, h, 22)
    out = sentinel_map(ctx, w, h)
    pipe = vxsgm.Pipeline.begin(ctx)
    sgm.encode(pipe, tl, tr, out, vxsgm.SGMStereoConfig(num_disparities=D))
    pipe.wait()                                           # not committed: returns at once
    pipe.close()
    assert (out.read_r32float() == SENTINEL).all()


def test_sgm_median_bilateral_depth_chain(ctx, sgm):
    """SGM -> median -> bilateral -> depth in one pipeline, the last stage a
    plain stream-ordered call on the pipeline's stream."""
    w, h, D = 200, 120, 64
    left, right, tl, tr = pair_textures(ctx, w, h, 31)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8, lr_check=True)
    disp, med, bil, depth = (ctx.texture_output_r32float(w, h) for _ in range(4))
    filt = vxsgm.DepthFilter(ctx)
    pipe = vxsgm.Pipeline.begin(ctx)
    try:
        sgm.encode(pipe, tl, tr, disp, cfg)
        filt.encode_median(pipe, disp, med, vxsgm.DepthMedianConfig(1))
        filt.encode_bilateral(pipe, med, bil, vxsgm.DepthFilterConfig(2, 3.0, 1.5))
        vxsgm.disparity_to_depth(bil.device_ptr, depth.device_ptr, w, h, 718.856, 0.5372,
                                 stream=pipe.stream)
        pipe.commit_and_wait()
        want_disp = O.sgm(left, right, D, num_paths=8, lr_check=True)
        assert (bits(disp.read_r32float()) == bits(want_disp.ravel())).all()
        want_med = O.depth_median(want_disp, 1)
        assert (bits(med.read_r32float()) == bits(want_med.ravel())).all()
        got_bil = bil.read_r32float()
        np.testing.assert_allclose(got_bil, O.depth_bilateral(want_med, 2, 3.0, 1.5).ravel(),
                                   rtol=BILATERAL_RTOL, atol=0)
        want_depth = O.disparity_to_depth(got_bil, 718.856, 0.5372)
        assert (bits(depth.read_r32float()) == bits(want_depth)).all()
    finally:
        pipe.close()
        filt.close()


def test_relaunch_reads_current_contents(ctx, sgm):
    w, h, D = 128, 72, 32
    left, right, tl, tr = pair_textures(ctx, w, h, 41)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8)
    out = ctx.texture_output_r32float(w, h)
    pipe = vxsgm.Pipeline.begin(ctx)
    try:
        with pytest.raises(vxsgm.InvalidConfig):
            pipe.relaunch()                               # nothing committed yet
        sgm.encode(pipe, tl, tr, out, cfg)
        pipe.commit_and_wait()
        assert (bits(out.read_r32float()) ==
                bits(O.sgm(left, right, D, num_paths=8).ravel())).all()
        for seed in (42, 43):
            left2, right2 = random_pair(w, h, seed=seed)
            tl.write_gray8(left2)
            tr.write_gray8(right2)
            pipe.relaunch()
            pipe.wait()
            assert (bits(out.read_r32float()) ==
                    bits(O.sgm(left2, right2, D, num_paths=8).ravel())).all()
    finally:
        pipe.close()


def test_write_refused_while_relaunch_in_flight(ctx, sgm):
    """A host write into an input texture of a relaunched pipeline that has
    not finished is refused, as UnifiedBuffer refuses mutable access while in
    flight (vx-core/src/buffer.rs:59-72); reads wait for the pipeline; after
    wait() the write goes through and the next relaunch reads it.  Five
    relaunches of a 1920x1080 D=256 pair keep the stream busy for several
    milliseconds, far longer than the host calls that follow them."""
    w, h, D = 1920, 1080, 256
    left, right, _ = vxsgm.synth_pair(w, h, D, 12, 2, 7700)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8)
    tl, tr = ctx.texture_gray8(left, w, h), ctx.texture_gray8(right, w, h)
    out = ctx.texture_output_r32float(w, h)
    pipe = vxsgm.Pipeline.begin(ctx)
    try:
        sgm.encode(pipe, tl, tr, out, cfg)
        assert not tl.is_in_flight()          # encoded, not launched
        pipe.commit_and_wait()
        assert not tl.is_in_flight()
        first = out.read_r32float()
        for _ in range(5):
            pipe.relaunch()
        assert tl.is_in_flight() and out.is_in_flight()
        with pytest.raises(vxsgm.InvalidConfig):
            tl.write_gray8(right)
        again = out.read_r32float()            # waits for the relaunches
        assert not out.is_in_flight()
        assert (bits(again) == bits(first)).all()
        pipe.wait()
        l2, r2, _ = vxsgm.synth_pair(w, h, D, 6, 0, 7701)
        tl.write_gray8(l2)
        tr.write_gray8(r2)
        pipe.relaunch()
        pipe.wait()
        # The plain path on the new pair (pinned to the oracle by the parity
        # tests at this size) is the expected map.
        want = sgm.compute_host(l2, r2, w, h, cfg)
        assert (bits(out.read_r32float()) == bits(want)).all()
    finally:
        pipe.close()
        for t in (tl, tr, out):
            t.close()


def test_several_pairs_and_split_in_one_pipeline(ctx):
    """Three pairs of different sizes, one with split scanlines, each with its
    own scratch; then the same pipeline again reuses the pooled scratch."""
    cases = [(160, 90, 48, 8, 0), (110, 420, 100, 8, 1), (257, 30, 32, 4, 0)]
    s = vxsgm.SGMStereo(ctx)
    ctx.pool_trim()   # idle blocks of earlier tests would make best fit order-dependent
    try:
        s.set_split(100, 1)
        held = []
        for rep in range(2):
            pipe = vxsgm.Pipeline.begin(ctx)
            outs = []
            for i, (w, h, D, paths, _) in enumerate(cases):
                left, right, tl, tr = pair_textures(ctx, w, h, 50 + i)
                out = ctx.texture_output_r32float(w, h)
                s.encode(pipe, tl, tr, out, vxsgm.SGMStereoConfig(num_disparities=D,
                                                                  num_paths=paths))
                pipe.retain_all([tl, tr])
                outs.append((left, right, D, paths, out))
            pipe.commit_and_wait()
            in_use = ctx.pool_stats()[2]
            assert in_use > 0
            for left, right, D, paths, out in outs:
                want = O.sgm(left, right, D, num_paths=paths).ravel()
                assert (bits(out.read_r32float()) == bits(want)).all()
            pipe.close()
            blocks, nbytes, in_use_after = ctx.pool_stats()
            assert in_use_after == 0
            held.append((blocks, nbytes))
        assert held[0] == held[1]    # the second pipeline allocated nothing new
    finally:
        s.close()


def test_failed_encodes_leave_pipeline_usable(ctx, sgm):
    w, h, D = 96, 64, 32
    left, right, tl, tr = pair_textures(ctx, w, h, 61)
    cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=4)
    out = ctx.texture_output_r32float(w, h)
    small = ctx.texture_output_r32float(w - 1, h)
    spare = ctx.texture_output_r32float(w, h)
    gray_out = ctx.texture_output_gray8(w, h)
    filt = vxsgm.DepthFilter(ctx)
    pipe = vxsgm.Pipeline.begin(ctx)
    try:
        with pytest.raises(vxsgm.TextureSizeMismatch):
            sgm.encode(pipe, tl, tr, small, cfg)
        with pytest.raises(vxsgm.InvalidConfig):
            sgm.encode(pipe, tl, tr, gray_out, cfg)
        with pytest.raises(vxsgm.InvalidConfig):
            sgm.encode(pipe, tl, tr, out, vxsgm.SGMStereoConfig(num_disparities=300))
        with pytest.raises(vxsgm.TextureSizeMismatch):
            filt.encode_median(pipe, out, small, vxsgm.DepthMedianConfig(1))
        with pytest.raises(vxsgm.InvalidConfig):
            filt.encode_median(pipe, out, spare, vxsgm.DepthMedianConfig(256))
        sgm.encode(pipe, tl, tr, out, cfg)
        pipe.commit_and_wait()
        assert (bits(out.read_r32float()) ==
                bits(O.sgm(left, right, D, num_paths=4).ravel())).all()
        with pytest.raises(vxsgm.InvalidConfig):
            sgm.encode(pipe, tl, tr, out, cfg)            # already committed
        with pytest.raises(vxsgm.InvalidConfig):
            filt.encode_median(pipe, out, out, vxsgm.DepthMedianConfig(1))
    finally:
        pipe.close()
        filt.close()


def test_textures_made_while_encoding(ctx, sgm):
    """Uploads, allocations and read-backs between begin and commit run at
    once (they are not part of the pipeline)."""
    w, h, D = 80, 40, 16
    pipe = vxsgm.Pipeline.begin(ctx)
    try:
        left, right, tl, tr = pair_textures(ctx, w, h, 71)
        pool = vxsgm.TexturePool()
        out = pool.acquire_r32float(ctx, w, h)
        sgm.encode(pipe, tl, tr, out, vxsgm.SGMStereoConfig(num_disparities=D))
        probe = ctx.texture_rgba8(np.arange(w * h * 4, dtype=np.uint8), w, h)
        assert (probe.read_rgba8() == np.arange(w * h * 4, dtype=np.uint8)).all()
        pipe.commit_and_wait()
        assert (bits(out.read_r32float()) == bits(O.sgm(left, right, D).ravel())).all()
        pool.release(out)
        pool.close()
    finally:
        pipe.close()


@pytest.mark.parametrize("fmt,split,lr", [("u16", 0, True), ("u8", 1, True), ("u16", 1, False)])
def test_encode_volume_formats_split_and_lr(ctx, fmt, split, lr):
    """The encode path with u16 or forced-u8 per-direction volumes, split
    scanlines and the LR check (D = 96 on a tall pair, so the vertical
    directions qualify for splitting), bit for bit against the oracle."""
    w, h, D = 120, 400, 96
    left, right, tl, tr = pair_textures(ctx, w, h, 81 + len(fmt) + 2 * split)
    s = vxsgm.SGMStereo(ctx)
    try:
        s.set_volume_format(D, fmt)
        s.set_split(D, split)
        cfg = vxsgm.SGMStereoConfig(num_disparities=D, num_paths=8, lr_check=lr, p1=7.0,
                                    p2=200.0)
        out = ctx.texture_output_r32float(w, h)
        pipe = vxsgm.Pipeline.begin(ctx)
        try:
            s.encode(pipe, tl, tr, out, cfg)
            pipe.commit_and_wait()
            want = O.sgm(left, right, D, p1=7.0, p2=200.0, num_paths=8, lr_check=lr).ravel()
            assert (bits(out.read_r32float()) == bits(want)).all()
        finally:
            pipe.close()
    finally:
        s.close()
