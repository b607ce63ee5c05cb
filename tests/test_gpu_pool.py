"""TexturePool (include/vxh/pipeline.h; vx-vision/src/pool.rs:17-147): the
reference's six pool tests (vx-vision/tests/test_pool.rs:6-103) at their exact
sizes and expectations, plus what they leave open: a hit returns the very
texture released (same device memory, contents kept), a full bucket destroys
what it cannot keep, clear empties every bucket, and the gray8 / RGBA8 texture
calls the pool builds on round-trip their bytes."""
import numpy as np
import pytest

import vxsgm

pytestmark = pytest.mark.gpu


def test_pool_acquire_and_release(ctx):
    pool = vxsgm.TexturePool()
    tex = pool.acquire_gray8(ctx, 64, 64)
    assert (tex.width, tex.height) == (64, 64)
    assert pool.total_acquires() == 1 and pool.total_hits() == 0
    pool.release(tex)
    assert pool.cached_count() == 1
    tex2 = pool.acquire_gray8(ctx, 64, 64)
    assert pool.total_acquires() == 2 and pool.total_hits() == 1
    assert tex2.width == 64
    pool.close()


def test_pool_different_sizes_no_hit(ctx):
    pool = vxsgm.TexturePool()
    pool.release(pool.acquire_gray8(ctx, 64, 64))
    _tex2 = pool.acquire_gray8(ctx, 32, 32)
    assert pool.total_hits() == 0
    pool.close()


def test_pool_different_formats(ctx):
    pool = vxsgm.TexturePool()
    t1 = pool.acquire_r32float(ctx, 64, 64)
    t2 = pool.acquire_rgba8(ctx, 64, 64)
    t3 = pool.acquire_gray8(ctx, 64, 64)
    assert [t.format for t in (t1, t2, t3)] == [2, 3, 1]
    assert pool.total_acquires() == 3 and pool.total_hits() == 0
    pool.release_all([t1, t2, t3])
    assert pool.cached_count() == 3
    pool.close()


def test_pool_clear(ctx):
    pool = vxsgm.TexturePool()
    pool.release(pool.acquire_gray8(ctx, 64, 64))
    assert pool.cached_count() == 1
    pool.clear()
    assert pool.cached_count() == 0
    pool.close()


def test_pool_hit_rate(ctx):
    pool = vxsgm.TexturePool()
    assert pool.hit_rate() == 0.0
    pool.release(pool.acquire_gray8(ctx, 64, 64))
    pool.release(pool.acquire_gray8(ctx, 64, 64))
    assert abs(pool.hit_rate() - 0.5) < 0.01
    pool.close()


def test_pool_capacity_limit(ctx):
    pool = vxsgm.TexturePool.with_capacity(2)
    ts = [pool.acquire_gray8(ctx, 64, 64) for _ in range(3)]
    pool.release_all(ts)   # the third is dropped: the bucket holds two
    assert pool.cached_count() == 2
    pool.close()


def test_hit_returns_released_texture_with_contents(ctx):
    pool = vxsgm.TexturePool()
    data = (np.arange(40 * 30) % 251).astype(np.uint8)
    t = ctx.texture_gray8(data, 40, 30)
    ptr = t.device_ptr
    pool.release(t)
    with pytest.raises(vxsgm.InvalidConfig):
        pool.release(t)    # moved into the pool already
    other = pool.acquire_r32float(ctx, 40, 30)   # same size, other format: a miss
    assert other.device_ptr != ptr and pool.total_hits() == 0
    back = pool.acquire_gray8(ctx, 40, 30)
    assert back.device_ptr == ptr and pool.total_hits() == 1
    assert (back.read_gray8() == data).all()     # contents are not cleared
    pool.close()


def test_zero_capacity_keeps_nothing(ctx):
    pool = vxsgm.TexturePool.with_capacity(0)
    pool.release(pool.acquire_r32float(ctx, 8, 8))
    assert pool.cached_count() == 0
    assert pool.acquire_r32float(ctx, 8, 8).width == 8 and pool.total_hits() == 0
    pool.close()


def test_rgba8_and_gray8_round_trip(ctx):
    w, h = 33, 17
    px = np.random.default_rng(5).integers(0, 256, w * h * 4, dtype=np.uint8)
    t = ctx.texture_rgba8(px, w, h)
    assert t.format == 3 and (t.read_rgba8() == px).all()
    with pytest.raises(vxsgm.InvalidConfig):
        ctx.texture_rgba8(px[:-1], w, h)
    with pytest.raises(vxsgm.InvalidConfig):
        t.read_gray8()
    g = ctx.texture_output_gray8(w, h)
    g.write_gray8(px[: w * h])
    assert g.format == 1 and (g.read_gray8() == px[: w * h]).all()
    assert ctx.texture_output_rgba8(w, h).read_rgba8().size == w * h * 4
