"""GPU check of vxh_sgm_enqueue_static, the entry point that allocates
nothing (its workspace is the library's static device scratch): through the C
ABI its maps equal vxh_sgm_enqueue's with a caller workspace bit for bit, on
two calls in a row (the second on a smaller pair, so stale scratch from the
first would show), and a pair whose workspace exceeds the scratch is
refused."""
import ctypes

import numpy as np
import pytest

import vxsgm
from vxsgm import hip

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def run_static(s, left, right, cfg, stream):
    h, w = left.shape
    d_l, d_r = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
    out = hip.DeviceBuffer(w * h * 4)
    out.fill(0xFF)
    c = cfg.to_c()
    vxsgm.check(vxsgm.lib.vxh_sgm_enqueue_static(s._h, d_l.ptr, d_r.ptr, w, h, ctypes.byref(c),
                                                 out.ptr, stream.handle))
    stream.synchronize()
    return out.download(np.float32)


def run_workspace(s, left, right, cfg, stream):
    h, w = left.shape
    d_l, d_r = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
    out = hip.DeviceBuffer(w * h * 4)
    out.fill(0xFF)
    ws = vxsgm.workspace_bytes(w, h, cfg)
    d_ws = hip.DeviceBuffer(ws)
    s.enqueue(d_l.ptr, d_r.ptr, w, h, cfg, out.ptr, d_ws.ptr, ws, stream.handle)
    stream.synchronize()
    return out.download(np.float32)


def test_static_scratch_matches_workspace_entry_point(ctx):
    cap = vxsgm.lib.vxh_sgm_static_scratch_bytes()
    assert cap == 96 << 20
    s = vxsgm.SGMStereo(ctx)
    st = hip.Stream()
    try:
        cases = [(320, 240, 64, 8, 120.0, True), (200, 90, 48, 4, 900.0, False)]
        for w, h, d, paths, p2, lr in cases:
            left, right, _ = vxsgm.synth_pair(w, h, d, 6, 1, 0x5347 + w)
            cfg = vxsgm.SGMStereoConfig(num_disparities=d, num_paths=paths, p2=p2, lr_check=lr)
            assert vxsgm.workspace_bytes(w, h, cfg) <= cap
            got = run_static(s, left, right, cfg, st)
            want = run_workspace(s, left, right, cfg, st)
            assert (bits(got) == bits(want)).all(), f"{w}x{h} D={d}: static scratch differs"
        big = vxsgm.SGMStereoConfig(num_disparities=128, num_paths=8)
        left, right, _ = vxsgm.synth_pair(640, 480, 128, 6, 1, 77)
        assert vxsgm.workspace_bytes(640, 480, big) > cap
        with pytest.raises(vxsgm.InvalidConfig):
            run_static(s, left, right, big, st)
    finally:
        st.close()
        s.close()
