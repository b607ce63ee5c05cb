"""GPU parity of the sparse stereo matcher (include/vxh/stereo_match.h,
vx-vision/shaders/StereoMatch.metal:32-115) against the oracle, bit for bit:
the distance matrix, the fused best-match + compaction path (matches in
ascending left index, the count), and StereoMatcher::run on host arrays;
sizes from 1 x 1 to the reference benchmark's 512 x 512 keypoints and a
ragged 1000 x 37, with band/range exclusions, NaN positions, ties, the
16-bit max_hamming truncation and empty inputs."""
import numpy as np
import pytest

import pyoracle as O
import vxsgm
from stereo_cases import stereo_case
from vxsgm import hip

pytestmark = pytest.mark.gpu

SIZES = [(1, 1, 0), (7, 5, 1), (100, 130, 2), (512, 512, 3), (1000, 37, 4), (37, 1000, 5)]
CONFIGS = [vxsgm.StereoConfig(),
           vxsgm.StereoConfig(max_hamming=65536 + 40, fx=458.654, fy=457.296, cx=367.215,
                              cy=248.375, baseline=0.110),
           vxsgm.StereoConfig(max_epipolar=0.75, min_disparity=0.0, max_disparity=64.0,
                              max_hamming=65535)]


def device_inputs(lk, rk, ld, rd):
    return [hip.DeviceBuffer.from_array(np.ascontiguousarray(a)) for a in (lk, rk, ld, rd)]


@pytest.mark.parametrize("nl,nr,seed", SIZES)
def test_matrix_matches_oracle(nl, nr, seed):
    lk, rk, ld, rd = stereo_case(nl, nr, seed)
    cfg = CONFIGS[seed % 3]
    bufs = device_inputs(lk, rk, ld, rd)
    dm = hip.DeviceBuffer(nl * nr * 2)
    dm.fill(0x5A)
    vxsgm.stereo_hamming_matrix_enqueue(*[b.ptr for b in bufs], nl, nr, cfg, dm.ptr)
    hip.synchronize()
    got = dm.download(np.uint16).reshape(nl, nr)
    assert (got == O.stereo_matrix(lk, rk, ld, rd, cfg.as_dict())).all()


def check_matches(got, want):
    assert len(got) == len(want)
    assert (got["left_idx"] == want["left_idx"]).all()
    assert (got["right_idx"] == want["right_idx"]).all()
    assert (got["disparity"].view(np.uint32) == want["disparity"].view(np.uint32)).all()
    assert (got["point_3d"].view(np.uint32) == want["point_3d"].view(np.uint32)).all()


@pytest.mark.parametrize("ci", range(len(CONFIGS)))
@pytest.mark.parametrize("nl,nr,seed", SIZES)
def test_match_enqueue_matches_oracle(nl, nr, seed, ci):
    lk, rk, ld, rd = stereo_case(nl, nr, seed)
    cfg = CONFIGS[ci]
    bufs = device_inputs(lk, rk, ld, rd)
    dmatch = hip.DeviceBuffer(nl * vxsgm.MATCH_DTYPE.itemsize)
    dcount = hip.DeviceBuffer(4)
    dcount.fill(0xFF)
    vxsgm.stereo_match_enqueue(*[b.ptr for b in bufs], nl, nr, cfg, dmatch.ptr, dcount.ptr)
    hip.synchronize()
    n = int(dcount.download(np.uint32)[0])
    got = dmatch.download(np.uint8).view(vxsgm.MATCH_DTYPE)[:n]
    check_matches(got, O.stereo_match(lk, rk, ld, rd, cfg.as_dict()))


def test_matcher_run_host_arrays(ctx):
    m = vxsgm.StereoMatcher(ctx)
    try:
        for nl, nr, seed in [(512, 512, 11), (64, 3, 12)]:
            lk, rk, ld, rd = stereo_case(nl, nr, seed)
            ldesc = np.zeros(nl, vxsgm.ORB_DTYPE)
            rdesc = np.zeros(nr, vxsgm.ORB_DTYPE)
            ldesc["desc"], rdesc["desc"] = ld, rd
            ldesc["angle"] = 1.25
            got = m.run(ctx, lk, rk, ldesc, rdesc, vxsgm.StereoConfig())
            want = O.stereo_match(lk, rk, ld, rd, vxsgm.StereoConfig().as_dict())
            assert len(want) > 0
            check_matches(got, want)
        lk, rk, ld, rd = stereo_case(8, 8, 13)
        empty = np.zeros(0, vxsgm.ORB_DTYPE)
        assert len(m.run(ctx, lk[:0], rk, empty, np.zeros(8, vxsgm.ORB_DTYPE),
                         vxsgm.StereoConfig())) == 0
        with pytest.raises(vxsgm.InvalidConfig):
            m.run(ctx, lk, rk, np.zeros(7, vxsgm.ORB_DTYPE), np.zeros(8, vxsgm.ORB_DTYPE),
                  vxsgm.StereoConfig())
    finally:
        m.close()


def test_empty_enqueue_writes_zero_count():
    dcount = hip.DeviceBuffer(4)
    dcount.fill(0xFF)
    d = hip.DeviceBuffer(256)
    vxsgm.stereo_match_enqueue(d.ptr, d.ptr, d.ptr, d.ptr, 0, 5, vxsgm.StereoConfig(), d.ptr,
                               dcount.ptr)
    hip.synchronize()
    assert int(dcount.download(np.uint32)[0]) == 0
