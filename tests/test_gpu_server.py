"""GPU checks of the serving loop: every request's map equals the plain path
and the oracle whether it runs alone or in a mix with others in flight, and
admission never holds more cost-volume bytes than the HBM budget."""
import hashlib
import json
import os
import threading
import time

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


def test_work_budgets_bound_sharing():
    """The work cap bounds the pixel-disparities admitted at once, a request
    above it runs alone, and the busy cap applies only while a backlog waits;
    every map stays the one the pair gets alone."""
    l, r, cfg = vxsgm.mixture_pair(320, 180, 64, 5)
    h, w = l.shape
    P = w * h * cfg.num_disparities
    ref = np.empty(w * h, np.float32)
    solo = vxsgm.SGMServer(hbm_budget_bytes=1 << 30, max_inflight=1, num_streams=1)
    solo.wait(solo.submit(l, r, w, h, cfg, ref))
    solo.close()

    def burst(n, **kw):
        srv = vxsgm.SGMServer(hbm_budget_bytes=4 << 30, max_inflight=16, num_streams=4, **kw)
        outs = [np.empty(w * h, np.float32) for _ in range(n)]
        for o in outs:
            srv.submit(l, r, w, h, cfg, o)
        srv.drain()
        st = srv.stats()
        srv.close()
        for o in outs:
            assert (bits(o) == bits(ref)).all()
        return st

    st = burst(24, work_budget_pxd=int(2.5 * P))
    assert st["peak_inflight"] <= 2 and st["peak_inflight_pxd"] <= 2.5 * P
    st = burst(12, work_budget_pxd=P // 2)
    assert st["peak_inflight"] == 1, "a request above the cap shares the GPU"
    # 24 queued at once: the backlog of at least 4 opens the busy cap.
    st = burst(24, work_budget_pxd=int(1.5 * P), busy_work_budget_pxd=int(6.5 * P),
               backlog_requests=4)
    assert 2 <= st["peak_inflight"] <= 6 and st["peak_inflight_pxd"] <= 6.5 * P
    assert st["completed"] == 24


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


def test_poisson_arrivals_full_mixture_match_oracle_hashes():
    """The serving workload: Poisson arrivals at 1000 req/s of the 32-pair
    request pool (the batched mixture's sizes and disparity ranges, 8 paths),
    96 requests through a server with a 64 GB budget.  Requests overlap on the
    device, and every map's SHA-256 equals the oracle's for its pair
    (tests/golden/mixture_hashes.json), so each result is the same alone and in
    the mix."""
    with open(os.path.join(os.path.dirname(__file__), "golden", "mixture_hashes.json")) as f:
        g = json.load(f)["serve_pool_32"]
    rows = [(r["w"], r["h"], r["d"], r["seed"], r["sha256"]) for r in g["pairs"]]
    assert [r[:4] for r in rows] == [tuple(m) for m in vxsgm.mixture(len(rows), seed=g["seed"])]
    pool = [vxsgm.mixture_pair(w, h, d, s) for w, h, d, s, _ in rows]
    srv = vxsgm.SGMServer(hbm_budget_bytes=64 << 30, max_inflight=32, num_streams=4)
    try:
        rng = np.random.default_rng(1000)
        gaps = rng.exponential(1.0 / 1000.0, 96)
        outs, ids = [], []
        t = time.perf_counter()
        for i in range(96):
            t += gaps[i]
            while time.perf_counter() < t:
                pass
            l, r, cfg = pool[i % len(pool)]
            h, w = l.shape
            out = np.full(w * h, np.nan, np.float32)
            ids.append(srv.submit(l, r, w, h, cfg, out))
            outs.append(out)
        for rid in ids:
            srv.wait(rid)
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == 96 and st["peak_inflight"] > 1
    assert st["peak_admitted_bytes"] <= 64 << 30
    bad = [i for i, o in enumerate(outs) if hashlib.sha256(bits(o).tobytes()).hexdigest()
           != rows[i % len(rows)][4]]
    assert not bad, f"requests {bad} differ from the oracle"


def test_reserved_slots_serve_smaller_requests():
    """Slots reserved for the largest shape serve every smaller request
    without allocating (all hits), and the results stay exact."""
    pairs = small_mix(6, seed=31)
    srv = vxsgm.SGMServer(hbm_budget_bytes=1 << 30, max_inflight=4, num_streams=2)
    try:
        cfg = vxsgm.SGMStereoConfig(num_disparities=max(p[2].num_disparities for p in pairs),
                                    num_paths=8)
        w = max(p[0].shape[1] for p in pairs)
        h = max(p[0].shape[0] for p in pairs)
        srv.reserve(w, h, cfg, 4)
        outs = []
        for l, r, c in pairs:
            out = np.empty(l.size, np.float32)
            srv.submit(l, r, l.shape[1], l.shape[0], c, out)
            outs.append(out)
        srv.drain()
        st = srv.stats()
    finally:
        srv.close()
    assert st["slot_misses"] == 0 and st["slot_hits"] == len(pairs)
    for (l, r, c), out in zip(pairs, outs):
        want = O.sgm(l, r, c.num_disparities, num_paths=8)
        assert (bits(out) == bits(want.ravel())).all()
