"""Autotuner for the aggregation schedule of each pair class.

A lane layout is (lanes, per_lane): `lanes` lanes of a wave own a pair of
scanlines and each lane owns `per_lane` consecutive disparities of both;
lanes * per_lane is 64, 128 or 256 for D <= 64, <= 128, <= 256, giving three
layouts per class (two for D > 128).  Results never depend on the schedule;
speed does, and differently for the two forms:

  tune_plain(sgm, w, h, D, paths)  a lone pair is bound by the length of its
      longest scanline: the schedule gives one layout to the horizontal
      directions (the longest scanlines; wide groups cut the critical path)
      and one to the others, and every pair of layouts is timed through the
      per-pair pipeline.
  tune_batch(pairs)  a batch either runs multi-pair launches around one
      persistent aggregation launch per layout class (throughput-bound: the
      tuner tries the per-pair layouts "default" and "widest", the one with
      the most disparities per lane and so the fewest instructions per entry,
      and the aggregation blocks per CU) or runs the per-pair pipelines over
      several streams (the tuner tries the stream count).  Timed on the batch
      itself, replayed from its hipGraph.

Timings use HIP events on the launch stream; the median of `reps` replays
after one warm-up wins.  Winners are cached in a JSON file keyed by form and
class so that later runs skip the sweep.
"""
import json
import os
import statistics

import vxsgm
from vxsgm import hip

CLASS_LAYOUTS = [[(32, 2), (16, 4), (8, 8)], [(64, 2), (32, 4), (16, 8)],
                 [(64, 4), (32, 8)]]
# Batch candidates tried: (options, per-pair layout policy).
BATCH_CANDIDATES = ([({"form": "streams", "streams": s}, "default") for s in (2, 4, 6)] +
                    [({"form": "persistent", "blocks_per_cu": b}, lay)
                     for b in (0, 2) for lay in ("default", "widest")] +
                    [({"form": "sweep"}, "default")])
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def layout_class(D):
    return 0 if D <= 64 else (1 if D <= 128 else 2)


def candidates(D):
    """The lane layouts of D's class."""
    return list(CLASS_LAYOUTS[layout_class(D)])


def widest(D):
    """The layout of D's class with the most disparities per lane."""
    return max(candidates(D), key=lambda lk: lk[1])


def apply_layouts(descs, policy):
    """Pair descriptors (as for SGMBatch) with the per-pair schedule of
    `policy`: "default" leaves them as given, "widest" sets every direction of
    every pair to widest(D)."""
    if policy == "default":
        return list(descs)
    out = []
    for d in descs:
        d = list(d)
        lay = widest(d[5].num_disparities)
        d[6] = (lay, lay)
        out.append(tuple(d))
    return out


def schedules(D):
    """Every (horizontal, other) layout pair of D's class."""
    c = candidates(D)
    return [(a, b) for a in c for b in c]


def _key(form, w, h, D, paths):
    return f"{form}:{w}x{h}:D{D}:P{paths}"


def _name(sched):
    (a, b), (c, d) = sched
    return f"{a}x{b}/{c}x{d}"


def _split_level(name):
    """Split level of a schedule name: "/split" 1, "/splitall" 2, none 0."""
    return 2 if name.endswith("/splitall") else (1 if name.endswith("/split") else 0)


def _parse(name):
    hs, os_ = name.split("/")[:2]
    return tuple(int(v) for v in hs.split("x")), tuple(int(v) for v in os_.split("x"))


class Tuner:
    def __init__(self, path=DEFAULT_CACHE):
        self.path = path
        self.table = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    def save(self):
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)

    def lookup(self, form, w, h, D, paths):
        e = self.table.get(_key(form, w, h, D, paths))
        return _parse(e["schedule"]) if e else None

    def _time(self, launch, stream, reps):
        launch()
        ms = []
        for _ in range(reps):
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            launch()
            e1.record(stream)
            e1.synchronize()
            ms.append(e0.elapsed_ms(e1))
        return statistics.median(ms)

    def tune_batch(self, descs, key="batch", reps=5, force=False):
        """Best (options, layout policy) for the batch of pair descriptors
        `descs` (as for SGMBatch); apply the policy with apply_layouts."""
        if not force and key in self.table:
            e = self.table[key]
            return e["options"], e["layouts"]
        stream = hip.Stream()
        times, chosen = {}, {}
        for opt, lay in BATCH_CANDIDATES:
            try:
                b = vxsgm.SGMBatch(apply_layouts(descs, lay), options=opt)
            except vxsgm.InvalidConfig:
                continue   # a form the batch's pairs do not allow (sweeps: 8 paths only)
            name = ",".join(f"{k}={v}" for k, v in sorted(opt.items())) + f",layouts={lay}"
            times[name] = self._time(lambda: b.launch(stream.handle), stream, reps)
            chosen[name] = (opt, lay)
            b.close()
        best = min(times, key=times.get)
        opt, lay = chosen[best]
        self.table[key] = {"options": opt, "layouts": lay, "ms": times}
        return opt, lay

    def tune_plain(self, sgm, w, h, D, paths, reps=5, force=False):
        """Sweeps every schedule of the class for one pair through the
        per-pair pipeline and leaves the winner set on `sgm`; returns it."""
        key = _key("plain", w, h, D, paths)
        if not force and key in self.table:
            name = self.table[key]["schedule"]
            sched = _parse(name)
            sgm.set_schedule(D, sched)
            sgm.set_split(D, _split_level(name))
            return sched
        left, right, cfg = vxsgm.mixture_pair(w, h, D, 0x5452)
        cfg.num_paths = paths
        dl, dr = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
        out = hip.DeviceBuffer(w * h * 4)
        wsb = vxsgm.workspace_bytes(w, h, cfg)
        ws = hip.DeviceBuffer(wsb)
        stream = hip.Stream()
        times = {}
        for split in (2, 1, 0):
            sgm.set_split(D, split)
            for sched in schedules(D):
                sgm.set_schedule(D, sched)
                times[_name(sched) + ("/split" if split == 1 else "/splitall" if split else "")] = self._time(
                    lambda: sgm.enqueue(dl.ptr, dr.ptr, w, h, cfg, out.ptr, ws.ptr, wsb,
                                        stream.handle), stream, reps)
        best = min(times, key=times.get)
        sched = _parse(best)
        sgm.set_split(D, _split_level(best))
        sgm.set_schedule(D, sched)
        self.table[key] = {"schedule": best, "ms": times}
        return sched
