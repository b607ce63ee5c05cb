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
  tune_batch(pairs)  a batch runs every pair's pipeline with its own
      schedule over several streams at once; what is tuned is the number of
      streams.  Timed on the batch itself, replayed from its hipGraph.

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
# Batch launch options tried: (streams,).
BATCH_OPTIONS = [(1,), (2,), (3,), (4,), (6,)]
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def layout_class(D):
    return 0 if D <= 64 else (1 if D <= 128 else 2)


def candidates(D):
    """The lane layouts of D's class."""
    return list(CLASS_LAYOUTS[layout_class(D)])


def schedules(D):
    """Every (horizontal, other) layout pair of D's class."""
    c = candidates(D)
    return [(a, b) for a in c for b in c]


def _key(form, w, h, D, paths):
    return f"{form}:{w}x{h}:D{D}:P{paths}"


def _name(sched):
    (a, b), (c, d) = sched
    return f"{a}x{b}/{c}x{d}"


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

    def tune_batch(self, descs, key="batch", reps=3, force=False):
        """Best launch options (streams,) for the batch of pair descriptors
        `descs` (as for SGMBatch)."""
        if not force and key in self.table:
            return tuple(self.table[key]["options"])
        stream = hip.Stream()
        times = {}
        for opt in BATCH_OPTIONS:
            b = vxsgm.SGMBatch(descs, options=opt)
            times["/".join(str(v) for v in opt)] = self._time(lambda: b.launch(stream.handle),
                                                             stream, reps)
            b.close()
        best = min(times, key=times.get)
        opt = tuple(int(v) for v in best.split("/"))
        self.table[key] = {"options": list(opt), "ms": times}
        return opt

    def tune_plain(self, sgm, w, h, D, paths, reps=5, force=False):
        """Sweeps every schedule of the class for one pair through the
        per-pair pipeline and leaves the winner set on `sgm`; returns it."""
        key = _key("plain", w, h, D, paths)
        if not force and key in self.table:
            name = self.table[key]["schedule"]
            sched = _parse(name)
            sgm.set_schedule(D, sched)
            sgm.set_split(D, name.endswith("/split"))
            return sched
        left, right, cfg = vxsgm.mixture_pair(w, h, D, 0x5452)
        cfg.num_paths = paths
        dl, dr = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
        out = hip.DeviceBuffer(w * h * 4)
        wsb = vxsgm.workspace_bytes(w, h, cfg)
        ws = hip.DeviceBuffer(wsb)
        stream = hip.Stream()
        times = {}
        for split in (True, False):
            sgm.set_split(D, split)
            for sched in schedules(D):
                sgm.set_schedule(D, sched)
                times[_name(sched) + ("/split" if split else "")] = self._time(
                    lambda: sgm.enqueue(dl.ptr, dr.ptr, w, h, cfg, out.ptr, ws.ptr, wsb,
                                        stream.handle), stream, reps)
        best = min(times, key=times.get)
        sched = _parse(best)
        sgm.set_split(D, best.endswith("/split"))
        sgm.set_schedule(D, sched)
        self.table[key] = {"schedule": best, "ms": times}
        return sched
