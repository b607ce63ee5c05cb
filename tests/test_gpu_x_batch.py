"""GPU parity of the batched form: every pair of a persistent, graph-captured
batch equals the plain per-pair path and the oracle bit for bit, for mixed
sizes, disparity ranges, path counts, schedules and repeated graph replays."""
import numpy as np
import pytest

import pyoracle as O
import vxsgm
from sgm_cases import random_pair
from vxsgm import hip

pytestmark = pytest.mark.gpu


def run_batch(pairs, scheds=None, replays=1):
    """pairs: [(left, right, cfg)] host arrays.  Returns list of host maps."""
    bufs, descs = [], []
    for i, (l, r, cfg) in enumerate(pairs):
        h, w = l.shape
        dl, dr = hip.DeviceBuffer.from_array(l), hip.DeviceBuffer.from_array(r)
        do = hip.DeviceBuffer(h * w * 4)
        bufs.append((dl, dr, do))
        sched = scheds[i] if scheds else None
        descs.append((dl.ptr, dr.ptr, do.ptr, w, h, cfg, sched))
    b = vxsgm.SGMBatch(descs)
    outs = []
    for _ in range(replays):
        for _, _, do in bufs:
            do.fill(0xFF)   # NaN pattern: every map must be rewritten
        b.run()
        outs.append([do.download(np.float32) for _, _, do in bufs])
    b.close()
    return outs


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_batch_small_mix_matches_oracle():
    pairs = []
    for i, (w, h, d, paths) in enumerate([(64, 40, 16, 2), (100, 31, 48, 4), (80, 80, 64, 8),
                                          (129, 17, 100, 8), (300, 20, 256, 8), (33, 3, 8, 8)]):
        l, r = random_pair(w, h, seed=100 + i)
        cfg = vxsgm.SGMStereoConfig(num_disparities=d, num_paths=paths,
                                    lr_check=(i % 2 == 1))
        pairs.append((l, r, cfg))
    outs = run_batch(pairs, replays=3)
    for rep in outs:
        for (l, r, cfg), got in zip(pairs, rep):
            want = O.sgm(l, r, cfg.num_disparities, num_paths=cfg.num_paths,
                         lr_check=cfg.lr_check, lr_tol=cfg.lr_tolerance)
            assert (bits(got) == bits(want.ravel())).all()


@pytest.mark.parametrize("sched", [(64, 2), (32, 4), (16, 8)])
def test_batch_schedules_equal_plain(sgm, sched):
    pairs = [vxsgm.mixture_pair(256, 96, 128, s)[:2] for s in range(3)]
    cfg = vxsgm.SGMStereoConfig(num_disparities=128, num_paths=8)
    outs = run_batch([(l, r, cfg) for l, r in pairs], scheds=[sched] * 3)[0]
    for (l, r), got in zip(pairs, outs):
        plain = sgm.compute_host(l, r, 256, 96, cfg)
        assert (bits(got) == bits(plain)).all()


def test_batch_mixture_equals_plain(sgm):
    """A scaled-down draw of the batched workload mixture (sizes / 4)."""
    mix = vxsgm.mixture(12, sizes=[(80, 60), (160, 120), (320, 180), (480, 270)])
    pairs = []
    for w, h, d, seed in mix:
        d = min(d, w)
        l, r, cfg = vxsgm.mixture_pair(w, h, d, seed)
        pairs.append((l, r, cfg))
    outs = run_batch(pairs, replays=2)
    for rep in outs:
        for (l, r, cfg), got in zip(pairs, rep):
            h, w = l.shape
            plain = sgm.compute_host(l, r, w, h, cfg)
            assert (bits(got) == bits(plain)).all()


def test_batch_rejects_bad_pair():
    l = hip.DeviceBuffer(64 * 8)
    o = hip.DeviceBuffer(64 * 8 * 4)
    bad = vxsgm.SGMStereoConfig(num_disparities=65)
    with pytest.raises(vxsgm.InvalidConfig):
        vxsgm.SGMBatch([(l.ptr, l.ptr, o.ptr, 64, 8, bad, None)])
    ok = vxsgm.SGMStereoConfig(num_disparities=64)
    with pytest.raises(vxsgm.InvalidConfig):
        vxsgm.SGMBatch([(l.ptr, l.ptr, o.ptr, 64, 8, ok, (16, 2))])
