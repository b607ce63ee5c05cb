"""GPU parity of the batched form: every pair of a graph-captured batch (the
persistent form, the per-pair pipelines over several streams, and the
two-sweep form) equals the plain per-pair path and the oracle bit for bit,
for mixed sizes, disparity ranges, path counts, schedules, split settings,
stream counts and repeated graph replays."""
import hashlib
import json
import os

import numpy as np
import pytest

import pyoracle as O
import vxsgm
from sgm_cases import random_pair
from vxsgm import hip

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def golden_mixture(name):
    """(w, h, D, seed, sha256) rows of a mixture in tests/golden/mixture_hashes.json,
    checked to be the draw vxsgm.mixture makes with the recorded seed."""
    with open(os.path.join(GOLDEN, "mixture_hashes.json")) as f:
        g = json.load(f)[name]
    rows = [(r["w"], r["h"], r["d"], r["seed"], r["sha256"]) for r in g["pairs"]]
    assert [r[:4] for r in rows] == [tuple(m) for m in vxsgm.mixture(len(rows), seed=g["seed"])]
    return rows


def map_sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).view(np.uint32).tobytes()).hexdigest()


def run_batch(pairs, scheds=None, replays=1, formats=None, splits=None, options=None):
    """pairs: [(left, right, cfg)] host arrays.  Returns list of host maps."""
    bufs, descs = [], []
    for i, (l, r, cfg) in enumerate(pairs):
        h, w = l.shape
        dl, dr = hip.DeviceBuffer.from_array(l), hip.DeviceBuffer.from_array(r)
        do = hip.DeviceBuffer(h * w * 4)
        bufs.append((dl, dr, do))
        sched = scheds[i] if scheds else None
        fmt = formats[i] if formats else "auto"
        split = splits[i] if splits else False
        descs.append((dl.ptr, dr.ptr, do.ptr, w, h, cfg, sched, fmt, split))
    b = vxsgm.SGMBatch(descs, options=options)
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


@pytest.mark.parametrize("options", [None, {"form": "streams"},
                                     {"form": "persistent", "blocks_per_cu": 1,
                                      "budget_bytes": 24 << 20}])
def test_batch_small_mix_matches_oracle(options):
    """Mixed sizes, D, path counts, LR checks, volume formats (u8 with a P2
    that overflows bytes, u16) and penalties, over repeated replays, in the
    persistent form (one chunk; several small chunks with one aggregation
    block per CU) and the streams form."""
    pairs, formats = [], []
    for i, (w, h, d, paths) in enumerate([(64, 40, 16, 2), (100, 31, 48, 4), (80, 80, 64, 8),
                                          (129, 17, 100, 8), (300, 20, 256, 8), (33, 3, 8, 8),
                                          (90, 30, 32, 8), (70, 25, 24, 4)]):
        l, r = random_pair(w, h, seed=100 + i)
        p2 = 900.0 if i >= 6 else 120.0
        cfg = vxsgm.SGMStereoConfig(num_disparities=d, num_paths=paths, p2=p2,
                                    lr_check=(i % 2 == 1))
        pairs.append((l, r, cfg))
        formats.append(["auto", "u16", "auto", "u16", "auto", "u8", "u8", "u16"][i])
    outs = run_batch(pairs, replays=3, formats=formats, options=options)
    bad = []
    for k, rep in enumerate(outs):
        for i, ((l, r, cfg), got) in enumerate(zip(pairs, rep)):
            want = O.sgm(l, r, cfg.num_disparities, p2=cfg.p2, num_paths=cfg.num_paths,
                         lr_check=cfg.lr_check, lr_tol=cfg.lr_tolerance).ravel()
            diff = bits(got) != bits(want)
            if diff.any():
                idx = np.flatnonzero(diff)[:4]
                bad.append(f"replay {k} pair {i} ({l.shape[1]}x{l.shape[0]} D="
                           f"{cfg.num_disparities} P={cfg.num_paths} lr={cfg.lr_check} "
                           f"{formats[i]}): {int(diff.sum())} px, at {idx.tolist()} got "
                           f"{got[idx].tolist()} want {want[idx].tolist()}")
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("opts", [None, (1,), (3,), (16,), {"blocks_per_cu": 2}])
def test_batch_launch_schedules_equal_plain(ctx, opts):
    """Both forms (the persistent one by default or with two aggregation
    blocks per CU, the streams one with every stream count) give the plain
    path's maps, and the plain path gives the same maps under each of its lane
    layouts."""
    pairs = [vxsgm.mixture_pair(256, 96, 128, s)[:2] for s in range(3)]
    cfg = vxsgm.SGMStereoConfig(num_disparities=128, num_paths=8)
    bufs, descs = [], []
    for l, r in pairs:
        dl, dr = hip.DeviceBuffer.from_array(l), hip.DeviceBuffer.from_array(r)
        do = hip.DeviceBuffer(256 * 96 * 4)
        bufs.append((dl, dr, do))
        descs.append((dl.ptr, dr.ptr, do.ptr, 256, 96, cfg, (16, 8)))
    b = vxsgm.SGMBatch(descs, options=opts)
    b.run()
    outs = [do.download(np.float32) for _, _, do in bufs]
    b.close()
    s = vxsgm.SGMStereo(ctx)
    try:
        for sched in (((64, 2), (16, 8)), ((32, 4), (32, 4))):
            s.set_schedule(128, sched)
            for (l, r), got in zip(pairs, outs):
                assert (bits(got) == bits(s.compute_host(l, r, 256, 96, cfg))).all()
    finally:
        s.close()


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
    for sched in ((16, 2), (64, 2), ((8, 8), (4, 16))):
        with pytest.raises(vxsgm.InvalidConfig):
            vxsgm.SGMBatch([(l.ptr, l.ptr, o.ptr, 64, 8, ok, sched)])
    with pytest.raises(vxsgm.InvalidConfig):
        vxsgm.SGMBatch([(l.ptr, l.ptr, o.ptr, 64, 8, ok, None)], options=(17,))


def test_batch_schedules_and_splits_per_pair():
    """Each pair keeps its own lane layouts and split setting; streams share
    workspaces between pairs of different sizes, formats and path counts."""
    pairs, scheds, splits, formats = [], [], [], []
    layouts = {64: [(32, 2), (16, 4), (8, 8)], 128: [(64, 2), (32, 4), (16, 8)],
               256: [(64, 4), (32, 8)]}
    for i, (w, h, d, paths) in enumerate([(600, 60, 64, 8), (300, 280, 128, 8), (90, 40, 64, 4),
                                          (520, 50, 256, 8), (270, 260, 128, 2),
                                          (64, 300, 32, 8), (400, 30, 96, 8)]):
        l, r = random_pair(w, h, seed=300 + i)
        cfg = vxsgm.SGMStereoConfig(num_disparities=d, num_paths=paths, lr_check=(i % 3 == 0))
        pairs.append((l, r, cfg))
        cls = 64 if d <= 64 else (128 if d <= 128 else 256)
        lay = layouts[cls]
        scheds.append((lay[i % len(lay)], lay[(i + 1) % len(lay)]))
        splits.append(i % 2 == 0)
        formats.append(["auto", "u16", "u8"][i % 3])
    outs = run_batch(pairs, scheds=scheds, replays=2, formats=formats, splits=splits,
                     options=(3,))
    outs += run_batch(pairs, scheds=scheds, replays=1, formats=formats, splits=splits)
    for rep in outs:
        for (l, r, cfg), got in zip(pairs, rep):
            want = O.sgm(l, r, cfg.num_disparities, num_paths=cfg.num_paths,
                         lr_check=cfg.lr_check).ravel()
            assert (bits(got) == bits(want)).all()


@pytest.mark.parametrize("options", [{"form": "sweep"},
                                     {"form": "sweep", "budget_bytes": 4 << 20}])
def test_sweep_form_small_mix_matches_oracle(options):
    """The two-sweep form (forward sums stored once, backward sweep with the
    winner): odd and tiny heights, widths below and above a wave's lanes,
    disparity ranges not a multiple of 4, every lane-layout class, P2 that
    wraps u16 sums, LR checks, repeated replays, one chunk and many chunks."""
    pairs = []
    for i, (w, h, d, p2, lr) in enumerate([(16, 8, 8, 120.0, False), (33, 3, 8, 120.0, True),
                                           (100, 31, 48, 120.0, False),
                                           (129, 17, 100, 120.0, True),
                                           (300, 20, 256, 120.0, False),
                                           (90, 30, 32, 900.0, True), (50, 9, 7, 120.0, False),
                                           (140, 23, 99, 300.0, False),
                                           (260, 97, 250, 120.0, True),
                                           (200, 61, 128, 120.0, False), (97, 1, 50, 120.0, False),
                                           (64, 2, 64, 65535.0, False)]):
        l, r = random_pair(w, h, seed=500 + i)
        pairs.append((l, r, vxsgm.SGMStereoConfig(num_disparities=d, num_paths=8, p2=p2,
                                                  lr_check=lr)))
    outs = run_batch(pairs, replays=2, options=options)
    bad = []
    for k, rep in enumerate(outs):
        for i, ((l, r, cfg), got) in enumerate(zip(pairs, rep)):
            want = O.sgm(l, r, cfg.num_disparities, p2=cfg.p2, num_paths=8,
                         lr_check=cfg.lr_check, lr_tol=cfg.lr_tolerance).ravel()
            diff = bits(got) != bits(want)
            if diff.any():
                bad.append(f"replay {k} pair {i} ({l.shape[1]}x{l.shape[0]} "
                           f"D={cfg.num_disparities}): {int(diff.sum())} px")
    assert not bad, "\n".join(bad)


def test_sweep_form_rejects_other_path_counts():
    l = hip.DeviceBuffer(64 * 8)
    o = hip.DeviceBuffer(64 * 8 * 4)
    for paths in (2, 4):
        cfg = vxsgm.SGMStereoConfig(num_disparities=32, num_paths=paths)
        with pytest.raises(vxsgm.InvalidConfig):
            vxsgm.SGMBatch([(l.ptr, l.ptr, o.ptr, 64, 8, cfg, None)], options={"form": "sweep"})


@pytest.mark.parametrize("options", [None, {"form": "persistent", "budget_bytes": 16 << 30},
                                     {"form": "streams"}, {"form": "sweep"}])
def test_batch_full_mixture_matches_oracle_hashes(options):
    """The batched workload at full size: all 64 pairs of the mixture
    ({320x240, 640x480, 1280x720, 1920x1080} x D {64, 128, 256}, 8 paths) in
    one batch, replayed twice; every map's SHA-256 equals the oracle's
    (tests/golden/mixture_hashes.json).  Persistent form with the default
    budget and with a 16 GiB budget (several chunks), the streams form and the
    two-sweep form."""
    rows = golden_mixture("batch_64")
    pairs = [vxsgm.mixture_pair(w, h, d, seed) for w, h, d, seed, _ in rows]
    outs = run_batch(pairs, replays=2, options=options)
    bad = [f"replay {k} pair {i} ({rows[i][0]}x{rows[i][1]} D={rows[i][2]})"
           for k, rep in enumerate(outs) for i, got in enumerate(rep)
           if map_sha256(got) != rows[i][4]]
    assert not bad, "maps differ from the oracle: " + ", ".join(bad)
