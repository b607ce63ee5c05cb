"""GPU checks of the serving loop: every request's map equals the plain path
and the oracle whether it runs alone or in a mix with others in flight, and
admission never holds more cost-volume bytes than the HBM budget."""
import threading

import numpy as np
import pytest

import pyoracle as O
import vxsgm

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def small_mix(n, seed=7):
    mix = vxsgm.mixture(n, seed=seed, sizes=[(96, 64), (160, 120), (320, 180), (400, 240)],
                        disparities=[32, 64, 96])
    return [vxsgm.mixture_pair(w, h, d, s) for w, h, d, s in mix]


def test_alone_equals_mix_equals_oracle():
    pairs = small_mix(10)
    srv = vxsgm.SGMServer(hbm_budget_bytes=1 << 30, max_inflight=8, num_streams=4)
    alone = []
    for l, r, cfg in pairs:
        h, w = l.shape
        out = np.empty(w * h, np.float32)
        srv.wait(srv.submit(l, r, w, h, cfg, out))
        alone.append(out)
    outs, ids = [], []
    for _ in range(3):
        for l, r, cfg in pairs:
            h, w = l.shape
            out = np.empty(w * h, np.float32)
            ids.append(srv.submit(l, r, w, h, cfg, out))
            outs.append(out)
    for i in ids:
        srv.wait(i)
    st = srv.stats()
    srv.close()
    assert st["peak_inflight"] > 1, "requests never overlapped on the device"
    for k, out in enumerate(outs):
        assert (bits(out) == bits(alone[k % len(pairs)])).all()
    for (l, r, cfg), a in zip(pairs, alone):
        want = O.sgm(l, r, cfg.num_disparities, num_paths=8)
        assert (bits(a) == bits(want.ravel())).all()


def test_budget_bounds_admitted_bytes():
    pairs = small_mix(8, seed=11)
    biggest = max(vxsgm.request_bytes(l.shape[1], l.shape[0], c) for l, _, c in pairs)
    budget = 2 * biggest
    srv = vxsgm.SGMServer(hbm_budget_bytes=budget, max_inflight=16, num_streams=4)
    outs = []
    for _ in range(4):
        for l, r, cfg in pairs:
            h, w = l.shape
            out = np.empty(w * h, np.float32)
            srv.submit(l, r, w, h, cfg, out)
            outs.append(out)
    srv.drain()
    st = srv.stats()
    assert st["completed"] == len(outs)
    assert st["peak_admitted_bytes"] <= budget
    big_l, big_r, big_cfg = vxsgm.mixture_pair(640, 480, 256, 3)
    tiny = vxsgm.SGMServer(hbm_budget_bytes=1 << 20, max_inflight=2, num_streams=1)
    with pytest.raises(vxsgm.BufferAlloc):
        tiny.submit(big_l, big_r, 640, 480, big_cfg, np.empty(640 * 480, np.float32))
    assert tiny.stats()["rejected"] == 1
    tiny.close()
    srv.close()


def test_concurrent_clients():
    pairs = small_mix(6, seed=23)
    srv = vxsgm.SGMServer(hbm_budget_bytes=1 << 30, max_inflight=6, num_streams=3)
    results = {}

    def client(k):
        l, r, cfg = pairs[k]
        h, w = l.shape
        out = np.empty(w * h, np.float32)
        srv.wait(srv.submit(l, r, w, h, cfg, out))
        results[k] = out

    th = [threading.Thread(target=client, args=(k,)) for k in range(len(pairs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    srv.close()
    for k, (l, r, cfg) in enumerate(pairs):
        want = O.sgm(l, r, cfg.num_disparities, num_paths=8)
        assert (bits(results[k]) == bits(want.ravel())).all()
