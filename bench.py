"""SGM stereo benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--forms plain,batched,stream,served,next,config2,lr]

Headline (`value`): the plain per-pair path.  A step is one pass of the SGM
pipeline (census of both images, matching costs, path aggregation over every
direction, sum + winner-take-all) over one synthetic rectified pair already
resident in HBM,
on the workload 720p_d128_8path (1280x720, D=128, 8 paths, P1=10, P2=120,
census 9x7).  value = W*H*D*K / t / 1e9 Gpx*disp/s over K timed steps.

`roofline` prices the dominant kernel, the path aggregation (all directions
in one launch), against HBM.  The pipeline's algorithmic bytes per pair,
W*H*(4*D + 6) (the u16 summed volume written once and read once, both gray8
inputs, the f32 output), split between its two kernels: the aggregation's
share is W*H*(2*D + 2) (the summed volume written once, both inputs read), the
winner-take-all's W*H*(2*D + 4).  The aggregation's share is divided by its
average HIP-event time over the timed steps, recorded on the launch stream
between the cost and winner-take-all launches.  `traffic` is the HBM bytes
per aggregation launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE passes of
this command (scaled by a copy kernel's readings), taken from the committed
summary profiles/paths_kernel_720p_d128_8path.md when its schedule is the
run's, else null.  `cpu_baseline` times the
scalar C++ oracle (the checker) on the same pair with the host's threads and
confirms that the timed GPU map is bit-identical to it.

`batched`: the 64-pair mixture of sizes {320x240, 640x480, 1280x720,
1920x1080} x D {64, 128, 256}, 8 paths, in the tuned batched form (persistent
multi-pair launches, per-pair pipelines over several streams, or two sweeps)
replayed from one hipGraph, against a loop of plain per-pair calls on one
stream; all maps compared bit for bit and with the oracle's hashes.  `stream`: the headline workload as a stream of
identical 720p pairs overlapped over 2 or 4 streams (one graph replay of 8 pairs), per-pair time
and Gpx*disp/s, every map compared with a plain call; `value` stays the plain per-pair figure.
`served`: Poisson arrivals at 200-2000 req/s of the same mixture against the
serving loop and against a FIFO queue into the plain path; p50/p99 latency and
Gpx*disp/s for each, and the rate each sustains within the p99 SLO.
`next_rows`: the kernels around the path (depth conversion, median and
bilateral filters, rectify maps and remap, sparse ORB stereo) and the
SGM-to-depth pipeline relaunch, each timed with its bytes and rate.
`config2`: the plain path on 1920x1080, D=256, 8 paths, checked against the
oracle's hash.
`lr`: the headline pair with the LR check, checked against the oracle's hash.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "vx-rs_amd")]

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="720p_d128_8path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads of the all-cores CPU baseline (0 = every CPU this "
                         "process may run on)")
    ap.add_argument("--forms", default="plain,batched,stream,served,next,config2,lr")
    ap.add_argument("--next-reps", type=int, default=50,
                    help="launches timed per kernel of the next-rows leg")
    ap.add_argument("--stream-pairs", type=int, default=8,
                    help="720p pairs per replay of the stream form")
    ap.add_argument("--batch-pairs", type=int, default=64)
    ap.add_argument("--batch-reps", type=int, default=7)
    ap.add_argument("--batch-layouts", default="default", choices=["default", "widest"],
                    help="per-pair lane layouts of the batched leg with --batch-options")
    ap.add_argument("--rates", default="200,500,1000,1300,1600,2000",
                    help="Poisson arrival rates of the served leg (req/s)")
    ap.add_argument("--slo-ms", type=float, default=25.0,
                    help="p99 latency target of the served leg's capacity figure")
    ap.add_argument("--serve-streams", type=int, default=4,
                    help="HIP streams of the serving loop")
    ap.add_argument("--work-budget-g", type=float, default=0.3,
                    help="the serving loop's work budget: admitted W*H*D while any request "
                         "runs, in G pixel-disparities (0: no cap)")
    ap.add_argument("--busy-work-budget-g", type=float, default=1.0,
                    help="the serving loop's work budget while --backlog requests wait")
    ap.add_argument("--backlog", type=int, default=4,
                    help="queued requests that open the busy work budget")
    ap.add_argument("--requests", type=int, default=300,
                    help="least requests per rate of the served leg")
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="arrival span per rate of the served leg: each rate offers "
                         "max(--requests, rate * this) requests")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB five times (the "
                         "known-traffic reference for HBM byte counters) and exit")
    ap.add_argument("--schedule", default="",
                    help="lane layouts of the plain path, horizontal/other (e.g. "
                         "64x2/16x8) or one LxK for all directions, optionally followed by "
                         "/split or /splitall; default: tuned or the built-in default")
    ap.add_argument("--volume-format", default="auto", choices=["auto", "u8", "u16"],
                    help="per-direction volume storage of the plain path")
    ap.add_argument("--batch-options", default="",
                    help='options of the batched form as JSON, e.g. \'{"form": "streams", '
                         '"streams": 4}\' (default: tuned, or the built-in default with '
                         '--no-tune)')
    return ap.parse_args()


def sched_label(sched, split):
    (a, b), (c, d) = sched
    return f"{a}x{b}/{c}x{d}" + ("/split" if split == 1 else "/splitall" if split else "")


def committed_traffic(schedule, workload):
    """HBM bytes per aggregation launch from the committed rocprofv3 summary
    of this workload, if it was measured with the same schedule."""
    path = os.path.join(ROOT, "profiles", f"paths_kernel_{workload}.md")
    if not os.path.exists(path):
        return None
    got, sched = None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            got = int(line.split(":")[1].split()[0])
        if line.startswith("schedule:"):
            sched = line.split(":", 1)[1].strip()
    return got if sched == schedule else None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def batched_leg(vx, hip, np, args, sgm, tuner):
    """The 64-pair mixture replayed from one graph over several streams vs a
    loop of plain per-pair calls on one stream; every map compared bit for
    bit."""
    mix = vx.mixture(args.batch_pairs)
    host, bufs, descs = [], [], []
    for w, h, d, seed in mix:
        left, right, cfg = vx.mixture_pair(w, h, d, seed)
        host.append((w, h, d, cfg))
        dl, dr = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
        ob, op = hip.DeviceBuffer(w * h * 4), hip.DeviceBuffer(w * h * 4)
        bufs.append((dl, dr, ob, op))
        descs.append((dl.ptr, dr.ptr, ob.ptr, w, h, cfg, None))
    from vxsgm.tuner import apply_layouts
    lay = "default"
    if args.batch_options:
        opts, lay = json.loads(args.batch_options), args.batch_layouts
    else:
        opts, lay = (None, "default") if tuner is None else tuner.tune_batch(descs)
    log(f"batched: {len(host)} pairs, options {opts}, layouts {lay}")
    if tuner is not None:
        log(f"batched candidates (ms): {tuner.table.get('batch', {}).get('ms')}")
    stream = hip.Stream()
    batch = vx.SGMBatch(apply_layouts(descs, lay), options=opts)
    chunks = batch.chunks()
    batch.launch(stream.handle)
    stream.synchronize()
    e0, e1 = hip.Event(), hip.Event()

    def median_ms(run, reps):
        # Each repetition timed on its own (HIP events), the median kept: one
        # slow replay does not move the figure, as in the tuner's timings.
        ts = []
        for _ in range(reps):
            e0.record(stream)
            run()
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_ms(e1))
        return float(np.median(ts)), [round(t, 3) for t in ts]

    ms_batch, batch_ms_all = median_ms(lambda: batch.launch(stream.handle), args.batch_reps)
    log(f"batched: {ms_batch:.3f} ms per batch (median of {batch_ms_all})")
    # The plain loop runs each pair with the handle's schedule for its D.
    wsb = max(vx.workspace_bytes(w, h, cfg) for w, h, _, cfg in host)
    ws = hip.DeviceBuffer(wsb)

    def plain_loop():
        for (w, h, d, cfg), (dl, dr, _, op) in zip(host, bufs):
            sgm.enqueue(dl.ptr, dr.ptr, w, h, cfg, op.ptr, ws.ptr, wsb, stream.handle)

    plain_loop()
    stream.synchronize()
    ms_plain, plain_ms_all = median_ms(plain_loop, args.batch_reps)
    log(f"batched: plain loop {ms_plain:.3f} ms (median of {plain_ms_all})")
    exact = all(np.array_equal(ob.download(np.uint32), op.download(np.uint32))
                for _, _, ob, op in bufs)
    want = golden_hashes("batch_64") if args.batch_pairs == 64 else None
    vs_oracle = None if want is None else all(
        map_sha256(ob.download(np.float32)) == want[k] for k, (_, _, ob, _) in enumerate(bufs))
    pxd = sum(w * h * d for w, h, d, _ in host)
    alg = sum(w * h * (4 * d + 6) for w, h, d, _ in host)
    batch.close()
    return {
        "pairs": len(host), "gpxd_per_batch": round(pxd / 1e9, 4),
        "ms_batch": round(ms_batch, 3), "ms_plain_loop": round(ms_plain, 3),
        "ms_batch_each": batch_ms_all, "ms_plain_loop_each": plain_ms_all,
        "value_batch": round(pxd / (ms_batch * 1e-3) / 1e9, 3),
        "value_plain_loop": round(pxd / (ms_plain * 1e-3) / 1e9, 3),
        "unit": "Gpx*disp/s",
        "roofline_frac_batch": round(alg / (ms_batch * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "bit_exact_vs_plain": bool(exact),
        "bit_exact_vs_oracle": vs_oracle,
        "form": ("persistent: multi-pair census, cost and winner launches around one "
                 "persistent aggregation launch per layout class, replayed from one hipGraph"
                 if not opts or opts.get("form", "persistent") == "persistent" else
                 "streams: per-pair pipelines over several streams, replayed from one hipGraph"),
        "options": opts, "layouts": lay, "chunks": chunks,
        "candidates_ms": None if tuner is None else tuner.table.get("batch", {}).get("ms"),
    }


def stream_leg(vx, hip, np, args, sgm):
    """The headline workload as a stream of pairs: n copies of the 720p pair
    (one input pair, n outputs) replayed from one graph over S streams, so a
    pair's aggregation overlaps other pairs' census, costs and sums.  Every
    map is compared bit for bit with one plain call."""
    left, right, _, cfg = vx.workload_pair("720p_d128_8path")
    h, w = left.shape
    d = cfg.num_disparities
    dl, dr = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
    outs = [hip.DeviceBuffer(w * h * 4) for _ in range(args.stream_pairs)]
    descs = [(dl.ptr, dr.ptr, o.ptr, w, h, cfg, None) for o in outs]
    want = sgm.compute_host(left, right, w, h, cfg)
    stream = hip.Stream()
    e0, e1 = hip.Event(), hip.Event()
    per_streams = {}
    exact = True
    for nstreams in (2, 4, "persistent"):
        opts = {"form": "persistent"} if nstreams == "persistent" else (nstreams,)
        batch = vx.SGMBatch(descs, options=opts)
        batch.launch(stream.handle)
        stream.synchronize()
        e0.record(stream)
        for _ in range(args.batch_reps):
            batch.launch(stream.handle)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / args.batch_reps
        batch.close()
        exact = exact and all(np.array_equal(o.download(np.float32).view(np.uint32),
                                             want.view(np.uint32)) for o in outs)
        per_streams[nstreams] = round(ms / len(outs), 4)
        log(f"stream: {nstreams}{' streams' if nstreams != 'persistent' else ''}, "
            f"{ms / len(outs):.4f} ms per pair")
    best = min(per_streams, key=per_streams.get)
    ms_pair = per_streams[best]
    return {
        "workload": "720p_d128_8path", "pairs_per_replay": len(outs),
        "ms_per_pair": ms_pair, "ms_per_pair_by_streams": per_streams,
        "value": round(w * h * d / (ms_pair * 1e-3) / 1e9, 3), "unit": "Gpx*disp/s",
        "bit_exact_vs_plain": bool(exact),
        "form": "identical 720p pairs in one hipGraph replay: per-pair pipelines over 2 or 4 "
                "streams, or the persistent multi-pair form",
    }


def golden_hashes(name):
    """SHA-256 of the oracle's map for every pair of a mixture
    (tests/golden/mixture_hashes.json), or None when the file lacks it."""
    path = os.path.join(ROOT, "tests", "golden", "mixture_hashes.json")
    try:
        with open(path) as f:
            return [r["sha256"] for r in json.load(f)[name]["pairs"]]
    except (OSError, KeyError):
        return None


def map_sha256(a):
    import hashlib
    return hashlib.sha256(a.view("uint32").tobytes()).hexdigest()


def config2_leg(vx, hip, np, args, tuner):
    """The plain per-pair path on 1920x1080, D=256, 8 paths (the workload whose
    cost volume, 531 MB, exceeds the 256 MB Infinity Cache), timed like the
    headline over min(--steps, 20) steps with the stage marks, its map compared
    with the committed SHA-256 of the oracle's map for the same pair."""
    return hashed_pair_leg(vx, hip, np, args, tuner, "1080p_d256_8path", False)


def lr_leg(vx, hip, np, args, tuner):
    """The headline pair (1280x720, D=128, 8 paths) with the LR check
    (tolerance 1), timed and checked like config2_leg; the winner stage then
    includes the right view's winners and the check."""
    return hashed_pair_leg(vx, hip, np, args, tuner, "720p_d128_8path", True)


def hashed_pair_leg(vx, hip, np, args, tuner, wl, lr):
    w, h, D, paths, _, _, _ = vx.WORKLOADS[wl]
    left, right, _, cfg = vx.workload_pair(wl)
    cfg.lr_check = lr
    ctx = vx.Context(0)
    sgm = vx.SGMStereo(ctx)
    if tuner is not None:
        sched = tuner.lookup("plain", w, h, D, paths)
        if sched:
            sgm.set_schedule(D, sched)
    d_l, d_r = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
    d_o = hip.DeviceBuffer(w * h * 4)
    wsb = vx.workspace_bytes(w, h, cfg)
    ws = hip.DeviceBuffer(wsb)
    stream = hip.Stream()
    ev = [hip.Event() for _ in range(3)]
    e0 = hip.Event()
    for _ in range(2):
        sgm.enqueue(d_l.ptr, d_r.ptr, w, h, cfg, d_o.ptr, ws.ptr, wsb, stream.handle)
    stream.synchronize()
    steps = min(args.steps, 20)
    stage = {"census_and_cost": [], "aggregate_all_paths": [], "sum_wta": [], "total": []}
    for _ in range(steps):
        e0.record(stream)
        sgm.enqueue_marked(d_l.ptr, d_r.ptr, w, h, cfg, d_o.ptr, ws.ptr, wsb, stream.handle, ev)
        ev[2].synchronize()
        stage["census_and_cost"].append(e0.elapsed_ms(ev[0]))
        stage["aggregate_all_paths"].append(ev[0].elapsed_ms(ev[1]))
        stage["sum_wta"].append(ev[1].elapsed_ms(ev[2]))
        stage["total"].append(e0.elapsed_ms(ev[2]))
    got = d_o.download(np.float32)
    path = os.path.join(ROOT, "tests", "golden", "workload_hashes.json")
    try:
        with open(path) as f:
            want = json.load(f)[wl + ("_lr" if lr else "")]
    except (OSError, KeyError):
        want = None
    ms = float(np.mean(stage["total"]))
    agg_ms = float(np.mean(stage["aggregate_all_paths"]))
    alg = w * h * (2 * D + 2)
    out = {"workload": wl, "lr_check": lr, "steps": steps, "ms_per_pair": round(ms, 4),
           "value": round(w * h * D / (ms * 1e-3) / 1e9, 3), "unit": "Gpx*disp/s",
           "stage_ms": {k: round(float(np.mean(v)), 4) for k, v in stage.items() if k != "total"},
           "schedule": {"horizontal": list(sgm.schedule(D)[0]), "other": list(sgm.schedule(D)[1])},
           "roofline": {"kernel": "paths_kernel (aggregation, all directions)",
                        "alg_bytes_per_launch": alg,
                        "achieved_gbs": round(alg / (agg_ms * 1e-3) / 1e9, 2),
                        "frac": round(alg / (agg_ms * 1e-3) / 8e12, 4)},
           "bit_exact_vs_oracle_hash": None if want is None else map_sha256(got) == want}
    sgm.close()
    ctx.close()
    return out


def next_rows_leg(vx, hip, np, args):
    """The kernels after and before the SGM path, each timed over `reps`
    stream-ordered launches with HIP events on its own stream: disparity to
    depth, the median and bilateral depth filters (vx-vision's DepthFilter
    bench sizes: 640x480, radius 1 median, radius 3 bilateral), the gray8
    remap through rectify maps and the map builder, at 640x480 and 1280x720;
    SGM and the three depth stages as one relaunched pipeline against the same
    stages called one by one; and the sparse ORB stereo match at 512 x 512
    keypoints.  Per kernel: us per
    launch, the bytes it must move (one read of its inputs, one write of its
    outputs) and that rate against 8 TB/s.  Outputs are not checked here (the
    GPU tests compare every one of these kernels with the oracle)."""
    reps = args.next_reps
    stream = hip.Stream()
    e0, e1 = hip.Event(), hip.Event()

    def timed(launch):
        launch()
        stream.synchronize()
        e0.record(stream)
        for _ in range(reps):
            launch()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps * 1e3

    out = {"reps": reps, "unit": "us per launch", "peak_gbs": 8000.0}
    rng = np.random.default_rng(0x5347_0006)
    for (w, h) in ((640, 480), (1280, 720)):
        n = w * h
        disp = rng.uniform(0.0, 127.0, n).astype(np.float32)
        disp[rng.random(n) < 0.05] = -1.0      # invalid pixels, as the LR check leaves them
        depth = (0.12 * 700.0 / np.maximum(disp, 1.0)).astype(np.float32)
        img = rng.integers(0, 256, n, dtype=np.uint8)
        d_disp, d_depth = hip.DeviceBuffer.from_array(disp), hip.DeviceBuffer.from_array(depth)
        d_out, d_img, d_img2 = hip.DeviceBuffer(n * 4), hip.DeviceBuffer.from_array(img), \
            hip.DeviceBuffer(n)
        d_mx, d_my = hip.DeviceBuffer(n * 4), hip.DeviceBuffer(n * 4)
        rp = vx.RectifyParams(fx=700.0, fy=700.0, cx=w / 2, cy=h / 2, k1=-0.05, k2=0.01,
                              R=(0.9998, -0.0175, 0.0, 0.0175, 0.9998, 0.0, 0.0, 0.0, 1.0))
        s = stream.handle
        cases = [
            ("disparity_to_depth", 8 * n, lambda: vx.disparity_to_depth(
                d_disp.ptr, d_out.ptr, w, h, 700.0, 0.12, s)),
            ("depth_median_r1", 8 * n, lambda: vx.DepthFilter.enqueue_median(
                d_depth.ptr, d_out.ptr, w, h, vx.DepthMedianConfig(1), s)),
            ("depth_bilateral_r3", 8 * n, lambda: vx.DepthFilter.enqueue_bilateral(
                d_depth.ptr, d_out.ptr, w, h, vx.DepthFilterConfig(3, 5.0, 0.05), s)),
            ("rectify_maps", 8 * n, lambda: vx.rectify_maps_enqueue(
                rp, w, h, d_mx.ptr, d_my.ptr, s)),
            ("remap_gray8", 10 * n, lambda: vx.remap_enqueue(
                d_img.ptr, 1, w, h, d_mx.ptr, d_my.ptr, d_img2.ptr, 1, w, h, s)),
        ]
        for name, nbytes, launch in cases:
            us = timed(launch)
            out[f"{name}_{w}x{h}"] = {"us": round(us, 2), "bytes": nbytes,
                                       "gbs": round(nbytes / us / 1e3, 1),
                                       "frac": round(nbytes / us / 1e3 / 8000.0, 4)}
    # Row f2: SGM -> 3x3 median -> 7x7 bilateral -> depth on the headline pair,
    # encoded once into a pipeline (one graph) and relaunched, against the
    # same four stages enqueued call by call on one stream.
    w, h, D, paths, _, _, _ = vx.WORKLOADS[args.workload]
    left, right, _, cfg = vx.workload_pair(args.workload)
    ctx = vx.Context(0)
    sgm = vx.SGMStereo(ctx)
    filt = vx.DepthFilter(ctx)
    tl, tr = ctx.texture_gray8(left, w, h), ctx.texture_gray8(right, w, h)
    disp, med, bil, dep = (ctx.texture_output_r32float(w, h) for _ in range(4))
    medc, bilc = vx.DepthMedianConfig(1), vx.DepthFilterConfig(3, 5.0, 0.05)
    pipe = vx.Pipeline.begin(ctx)
    sgm.encode(pipe, tl, tr, disp, cfg)
    filt.encode_median(pipe, disp, med, medc)
    filt.encode_bilateral(pipe, med, bil, bilc)
    vx.disparity_to_depth(bil.device_ptr, dep.device_ptr, w, h, 700.0, 0.12,
                          stream=pipe.stream)
    pipe.commit()
    pipe.wait()
    graph_ms = []
    for _ in range(reps // 5 + 3):
        t0 = time.perf_counter()
        pipe.relaunch()
        pipe.wait()
        graph_ms.append((time.perf_counter() - t0) * 1e3)
    wsb = vx.workspace_bytes(w, h, cfg)
    ws = hip.DeviceBuffer(wsb)
    d_l, d_r = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
    d_a, d_b = hip.DeviceBuffer(w * h * 4), hip.DeviceBuffer(w * h * 4)
    s = stream.handle

    def calls():
        sgm.enqueue(d_l.ptr, d_r.ptr, w, h, cfg, d_a.ptr, ws.ptr, wsb, s)
        vx.DepthFilter.enqueue_median(d_a.ptr, d_b.ptr, w, h, medc, s)
        vx.DepthFilter.enqueue_bilateral(d_b.ptr, d_a.ptr, w, h, bilc, s)
        vx.disparity_to_depth(d_a.ptr, d_b.ptr, w, h, 700.0, 0.12, s)

    calls()
    stream.synchronize()
    call_ms = []
    for _ in range(reps // 5 + 3):
        t0 = time.perf_counter()
        calls()
        stream.synchronize()
        call_ms.append((time.perf_counter() - t0) * 1e3)
    out["pipeline_sgm_median_bilateral_depth"] = {
        "workload": args.workload, "unit": "ms per frame (host wall clock, launch to wait)",
        "graph_relaunch_ms": round(float(np.median(graph_ms)), 4),
        "stream_calls_ms": round(float(np.median(call_ms)), 4),
        "graph_nodes": pipe.node_count()}
    pipe.close()
    for t in (tl, tr, disp, med, bil, dep):
        t.close()
    filt.close()
    sgm.close()
    ctx.close()

    nk = 512
    kp = np.zeros(nk, vx.CORNER_DTYPE)
    kp["x"] = rng.uniform(0, 640, nk)
    kp["y"] = rng.uniform(0, 480, nk)
    kq = kp.copy()
    kq["x"] -= rng.uniform(0, 60, nk)
    kq["y"] += rng.uniform(-1, 1, nk)
    dl = np.zeros(nk, vx.ORB_DTYPE)
    dl["desc"] = rng.integers(0, 1 << 32, (nk, 8), dtype=np.uint32)
    dr = dl.copy()
    flip = rng.integers(0, 1 << 32, (nk, 8), dtype=np.uint32) & \
        rng.integers(0, 1 << 32, (nk, 8), dtype=np.uint32) & \
        rng.integers(0, 1 << 32, (nk, 8), dtype=np.uint32)
    dr["desc"] ^= flip
    d_lk, d_rk = hip.DeviceBuffer.from_array(kp), hip.DeviceBuffer.from_array(kq)
    d_ld = hip.DeviceBuffer.from_array(np.ascontiguousarray(dl["desc"]))
    d_rd = hip.DeviceBuffer.from_array(np.ascontiguousarray(dr["desc"]))
    d_m, d_n = hip.DeviceBuffer(nk * vx.MATCH_DTYPE.itemsize), hip.DeviceBuffer(4)
    cfg = vx.StereoConfig()
    us = timed(lambda: vx.stereo_match_enqueue(d_lk.ptr, d_rk.ptr, d_ld.ptr, d_rd.ptr, nk, nk,
                                               cfg, d_m.ptr, d_n.ptr, stream.handle))
    nbytes = 2 * nk * (16 + 32) + nk * vx.MATCH_DTYPE.itemsize
    out["stereo_match_512x512"] = {"us": round(us, 2), "bytes": nbytes,
                                   "note": "launch-bound: the descriptors fit in L2"}
    return out


def served_leg(vx, np, args):
    """Poisson arrivals of the mixture's 32-pair request pool against the
    serving loop and against a FIFO queue into the plain per-pair path (one
    request at a time), over a sweep of rates; both keep their device slots
    reserved before the timed runs.  At every rate the two forms take turns on
    the same arrival sequence of max(--requests, rate * --serve-seconds)
    requests, each timed run after an untimed warm-up burst at its rate.  Each
    form's first map of every pair is checked against the
    oracle's hash for the pair, and every timed map bit for bit against those.
    `capacity_at_slo`: for each form, the highest swept rate whose p99 stays
    within the SLO, and the Gpx*disp/s it sustained."""
    pool = []
    for w, h, d, seed in vx.mixture(32, seed=0x5347_0004):
        left, right, cfg = vx.mixture_pair(w, h, d, seed)
        pool.append((left, right, w, h, cfg))
    want = golden_hashes("serve_pool_32")
    out = {"requests_per_rate": f"max({args.requests}, rate * {args.serve_seconds:g} s)",
           "budget_gb": 64, "work_budget_gpxd": args.work_budget_g,
           "busy_work_budget_gpxd": args.busy_work_budget_g, "backlog_requests": args.backlog,
           "slo_p99_ms": args.slo_ms,
           "bit_exact_vs_oracle": want is not None}
    shapes = sorted({(w, h, cfg.num_disparities) for _, _, w, h, cfg in pool},
                    key=lambda s: s[0] * s[1] * s[2])
    forms = (("served", dict(max_inflight=32, num_streams=args.serve_streams,
                             work_budget_pxd=int(args.work_budget_g * 1e9),
                             busy_work_budget_pxd=int(args.busy_work_budget_g * 1e9),
                             backlog_requests=args.backlog), 2),
             ("plain_fifo", dict(max_inflight=1, num_streams=1), 0))
    servers, ref_maps = {}, {}
    for name, kw, per_shape in forms:
        srv = vx.SGMServer(hbm_budget_bytes=64 << 30, **kw)
        # Resident slots: `per_shape` per shape of the pool plus the largest.
        w, h, d = shapes[-1]
        srv.reserve(w, h, vx.SGMStereoConfig(num_disparities=d, num_paths=8), 4)
        for w, h, d in shapes[:-1]:
            if per_shape:
                srv.reserve(w, h, vx.SGMStereoConfig(num_disparities=d, num_paths=8), per_shape)
        for k, (l, r, w, h, cfg) in enumerate(pool):
            o = np.empty(w * h, np.float32)
            srv.wait(srv.submit(l, r, w, h, cfg, o))
            if want is not None and map_sha256(o) != want[k]:
                out["bit_exact_vs_oracle"] = False
            ref_maps.setdefault(k, o.view(np.uint32))
        servers[name] = srv
        out[name] = {}
    # The two forms take turns at every rate, on the same arrival sequence.
    for rate in [float(x) for x in args.rates.split(",")]:
        n = max(args.requests, int(rate * args.serve_seconds))
        gaps = np.random.default_rng(int(rate)).exponential(1.0 / rate, n)
        for name, _, _ in forms:
            srv = servers[name]
            # An untimed warm-up burst at the same rate first, so that each
            # form starts its timed run in the same (busy) state.
            warm = np.random.default_rng(int(rate) + 1).exponential(1.0 / rate, max(50, n // 10))
            t = time.perf_counter()
            for i, g in enumerate(warm):
                t += g
                while time.perf_counter() < t:
                    pass
                l, r, w, h, cfg = pool[i % len(pool)]
                srv.submit(l, r, w, h, cfg, np.empty(w * h, np.float32))
            srv.drain()
            srv.wait_all()
            # Result buffers are allocated and touched before the timed run
            # (a client's are resident), so the copies into them measure the
            # server, not first-touch page faults.
            outs = [np.zeros(pool[i % len(pool)][2] * pool[i % len(pool)][3], np.float32)
                    for i in range(n)]
            srv.reset_stats()
            t = time.perf_counter()
            for i in range(n):
                t += gaps[i]
                while time.perf_counter() < t:
                    pass
                l, r, w, h, cfg = pool[i % len(pool)]
                srv.submit(l, r, w, h, cfg, outs[i])
            srv.drain()
            srv.wait_all()
            st = srv.stats()
            span = st["last_complete_s"] - st["first_submit_s"]
            if want is not None:
                exact = all(np.array_equal(o.view(np.uint32), ref_maps[i % len(pool)])
                            for i, o in enumerate(outs))
                out["bit_exact_vs_oracle"] = out["bit_exact_vs_oracle"] and exact
            rec = {
                "requests": n,
                "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                "gpxd_per_s": round(st["pxd_total"] / span / 1e9, 3),
                "offered_gpxd_per_s": round(rate * st["pxd_total"] / st["completed"] / 1e9, 3),
                "peak_inflight": st["peak_inflight"],
                "peak_admitted_gb": round(st["peak_admitted_bytes"] / 2**30, 3),
                "slot_misses": st["slot_misses"]}
            out[name][str(int(rate))] = rec
            log(f"served[{name}] {rate:.0f} req/s: {rec}")
    for name, _, _ in forms:
        servers[name].close()
        ok = [(float(r), v) for r, v in out[name].items() if v["p99_ms"] <= args.slo_ms]
        best = max(ok, key=lambda kv: kv[1]["gpxd_per_s"]) if ok else None
        out.setdefault("capacity_at_slo", {})[name] = (
            {"rate": int(best[0]), "gpxd_per_s": best[1]["gpxd_per_s"],
             "p99_ms": best[1]["p99_ms"]} if best else None)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the SGM "
              "path here is a single-device kernel and no multi-GPU figure is produced.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import vxsgm as vx
    from vxsgm import hip

    if vx.device_count() < 1:
        print("bench.py needs a visible HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        n = args.copy_calibration << 20
        src, dst = hip.DeviceBuffer(n), hip.DeviceBuffer(n)
        src.fill(1)
        for _ in range(5):
            vx.check(vx.lib.vxh_util_copy(dst.ptr, src.ptr, n, None))
        hip.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "launches": 5,
                          "read_bytes_per_launch": n, "write_bytes_per_launch": n}))
        return
    w, h, D, paths, _, _, _ = vx.WORKLOADS[args.workload]
    left, right, _, cfg = vx.workload_pair(args.workload)
    ctx = vx.Context(0)
    sgm = vx.SGMStereo(ctx)
    d_left, d_right = hip.DeviceBuffer.from_array(left), hip.DeviceBuffer.from_array(right)
    d_out = hip.DeviceBuffer(w * h * 4)
    ws_bytes = vx.workspace_bytes(w, h, cfg)
    ws = hip.DeviceBuffer(ws_bytes)
    stream = hip.Stream()

    tuner = None
    sgm.set_volume_format(D, args.volume_format)
    if args.schedule:
        toks = args.schedule.lower().split("/")
        parts = [tuple(int(v) for v in p.split("x")) for p in toks if "x" in p]
        sgm.set_schedule(D, (parts[0], parts[-1]))
        sgm.set_split(D, 2 if "splitall" in toks else (1 if "split" in toks else 0))
    elif not args.no_tune:
        from vxsgm.tuner import Tuner
        tuner = Tuner(path=None)
        tuner.tune_plain(sgm, w, h, D, paths)
        log(f"plain schedules: {tuner.table}")
    plain_sched = sgm.schedule(D)
    plain_split = sgm.split(D)

    for _ in range(args.warmup):
        sgm.enqueue(d_left.ptr, d_right.ptr, w, h, cfg, d_out.ptr, ws.ptr, ws_bytes,
                    stream.handle)
    stream.synchronize()
    # Per step: start, after census, after aggregation, end (stream order).
    marks = [[hip.Event() for _ in range(4)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for m in marks:
        m[0].record(stream)
        sgm.enqueue_marked(d_left.ptr, d_right.ptr, w, h, cfg, d_out.ptr, ws.ptr, ws_bytes,
                           stream.handle, m[1:])
    marks[-1][3].synchronize()
    wall = time.perf_counter() - t0
    ms_event = marks[0][0].elapsed_ms(marks[-1][3]) / args.steps
    stage = [sum(m[i].elapsed_ms(m[i + 1]) for m in marks) / args.steps for i in range(3)]
    ms_per_step = wall * 1e3 / args.steps
    gpu_disp = d_out.download(np.float32).reshape(h, w)
    stages = {"census_and_cost": stage[0], "aggregate_all_paths": stage[1],
              "sum_wta": stage[2]}
    log(f"plain: {ms_per_step:.3f} ms/step, stages {stages}")

    pxd = w * h * D
    value = pxd * args.steps / wall / 1e9
    fmt_ran = {1: "u8", 2: "u16"}[sgm.volume_format_in_effect(D, cfg)]
    alg_bytes = w * h * (4 * D + 6)
    agg_bytes = w * h * (2 * D + 2)
    achieved = agg_bytes / (stage[1] * 1e-3) / 1e9
    result = {
        "metric": f"Gpx*disp/s @{w}x{h} D{D} {paths}-path",
        "value": round(value, 3),
        "unit": "Gpx*disp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",   # the recurrence's arithmetic (mod 2^16); volumes stored per config
        "data": "synthetic",
        "config": {"workload": args.workload, "width": w, "height": h,
                   "num_disparities": D, "num_paths": paths, "p1": cfg.p1,
                   "p2": cfg.p2, "census": "9x7", "form": "plain per-pair",
                   "schedule": {"horizontal": list(plain_sched[0]),
                                "other": list(plain_sched[1]), "split": plain_split},
                   "volume_format": args.volume_format, "volume_format_ran": fmt_ran},
        "roofline": {"bound": "hbm", "kernel": "paths_kernel (aggregation, all directions)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": committed_traffic(sched_label(plain_sched, plain_split),
                                                  args.workload),
                     "alg_bytes_per_launch": agg_bytes,
                     "launch_ms_event": round(stage[1], 4),
                     "pipeline_alg_bytes": alg_bytes,
                     "pipeline_ms_event": round(ms_event, 4),
                     "pipeline_frac": round(alg_bytes / (ms_event * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                            4)},
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        want = O.sgm(left, right, D, p1=cfg.p1, p2=cfg.p2, num_paths=paths,
                     nthreads=threads)
        cpu_s = time.perf_counter() - t0
        # One core on a bounded sample: the first rows of the same pair (every
        # path still runs over whole scanlines of the cropped pair).
        rows = max(16, min(h, int(h * min(1.0, 4.0 / max(cpu_s * threads, 1e-3)))))
        t0 = time.perf_counter()
        O.sgm(left[:rows], right[:rows], D, p1=cfg.p1, p2=cfg.p2, num_paths=paths,
              nthreads=1)
        cpu1_s = time.perf_counter() - t0
        model = cpu_model()
        result["cpu_baseline"] = {
            "value": round(pxd / cpu_s / 1e9, 4), "unit": "Gpx*disp/s",
            "cores": threads, "kind": "port", "cpu_model": model,
            "sample": f"one full {w}x{h} D={D} {paths}-path pair, scalar C++ oracle, "
                      f"{threads} threads over independent scanlines, {cpu_s:.2f} s",
            "single_core": {"value": round(w * rows * D / cpu1_s / 1e9, 4),
                            "unit": "Gpx*disp/s", "cores": 1,
                            "sample": f"the first {rows} rows of the same pair "
                                      f"({w}x{rows}, D={D}, {paths} paths), {cpu1_s:.2f} s"}}
        result["bit_exact"] = bool((gpu_disp.view(np.uint32) ==
                                    want.astype(np.float32).view(np.uint32)).all())
        log(f"cpu baseline {cpu_s:.2f} s on {threads} threads, {cpu1_s:.2f} s on one core "
            f"for {rows} rows ({model}); bit_exact {result['bit_exact']}")
    if result.get("bit_exact"):
        result["metric"] += " (bit-exact)"
    forms = set(args.forms.split(","))
    if "batched" in forms:
        log(f"batched leg ({args.batch_pairs} pairs)")
        result["batched"] = batched_leg(vx, hip, np, args, sgm, tuner)
    if "stream" in forms:
        log(f"stream leg ({args.stream_pairs} pairs of 720p)")
        result["stream"] = stream_leg(vx, hip, np, args, sgm)
    if "served" in forms:
        log(f"served leg (rates {args.rates})")
        result["served"] = served_leg(vx, np, args)
    if "next" in forms:
        log("next-rows leg (depth, filters, rectify, sparse stereo)")
        result["next_rows"] = next_rows_leg(vx, hip, np, args)
    if "config2" in forms:
        log("1920x1080 D=256 leg")
        result["config2"] = config2_leg(vx, hip, np, args, tuner)
    if "lr" in forms:
        log("1280x720 D=128 leg with the LR check")
        result["lr"] = lr_leg(vx, hip, np, args, tuner)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
