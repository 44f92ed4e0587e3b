#!/usr/bin/env python
"""bench.py - pixel-iterations/s of newman's per-pixel iteration on one MI355X.

Workload (default line): the plain-double configuration - 640x480, centre
-0.75+0i, width 3.0, N = 1024, no multisampling, smoothing on, bailout 1024 (the
reference's constant, mandelbrot.cpp:58; the radius-2 figure is reported beside
it), the reference's plain path (getIterationsHW, mandelbrot.cpp:231-254, as
useHardware() selects at this zoom).  One frame is ~2 us of GPU work, so one
step = one pass of the hot path over 512 copies of the frame in one launch set
(start kernel: cardioid test and the first updates; persistent iterate kernel;
smoothing), the coordinate tables already resident in HBM.

value = credited pixel-iterations per second: the sum of reported escape counts
over samples not resolved by the cardioid/bulb test, divided by wall time per
step.  roofline = algorithmic FP64 FLOPs (8 per plain update: 3 mul + 5 add) of
the dominant kernel's executed updates over its HIP-event duration, against the
FP64 vector peak; both plain kernels are listed.  cpu_baseline = the CPU oracle
(the reference's plain path with its HP cardioid test) on the same frame.

deep_zoom (extra key) = the deep-zoom FP64-delta configuration: 1920x1080 output
with 4x4 multisampling (7680x4320 samples), width 1e-30, N = 20000, centre from
the seed-1 zoom walk, with the FP32X-delta and learned-skip variants on the same
frame, the colouring and zoom-video legs and its own CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c2|c3|c4|c5]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pixel-iterations/s on 1 MI355X: plain FP64 vs FP32-floatexp+learned-skip"
FP64_VECTOR_PEAK_TFLOPS = 78.6
FLOP_PER_PERT_STEP = 15  # z = X+d (2), |z|^2 (3), w = X+z (2), d*w + e (8)
FLOP_PER_PLAIN_STEP = 8  # 3 mul + 5 add, |z|^2 reusing re^2, im^2


def load_views():
    with open(os.path.join(ROOT, "tests", "golden", "views.json")) as f:
        return json.load(f)


C1_BATCH = 512  # the 640x480 frame is ~2 us of work: frames per launch set
C4_FRAMES = 120


def make_workload(name):
    """(mandelbrot, description, flop per executed step, step(block_k, waves) -> stats)"""
    import newman_amd as nm
    views = load_views()
    if name == "c2":
        v = views["c2_seed1"]
        m = nm.Mandelbrot(1080 * 4, 1920 * 4)
        m.set_view(v["N"], v["re"], v["im"], v["width"])
        desc = {"workload": "deep zoom, FP64 deltas: 1920x1080 output x 4x4 multisampling (7680x4320 samples), "
                            "width 1e-30, N 20000, bailout 1024, tol 1e-10, centre = zoom walk seed 1; the timed "
                            "pass renders the 7680x4320 samples, the 4x4 reduction to 1920x1080 (~0.05 ms) is "
                            "timed in the colour leg",
                "samples": 7680 * 4320, "N": 20000, "width": "1e-30", "precision_bits": m.precision_bits()}

        def step(k, w, reps=1):
            return m.render(k, w, reps, copy_back=False)
        return m, desc, FLOP_PER_PERT_STEP, step
    if name == "c1":
        m = nm.Mandelbrot(480, 640)
        m.set_view(1024, "-0.75", "0", "3.0")
        widths = ["3.0"] * C1_BATCH
        desc = {"workload": "plain double path: 640x480, centre -0.75+0i, width 3.0, N 1024, "
                            "bailout 1024; %d copies of the frame per launch set" % C1_BATCH,
                "samples": 640 * 480 * C1_BATCH, "N": 1024, "frames_per_step": C1_BATCH}

        def step(k, w, reps=1):
            return m.render_frames(widths, k, w, reps, copy_back=False)[1]
        return m, desc, FLOP_PER_PLAIN_STEP, step
    if name == "c3":
        v = views["c3_detail1"]
        m = nm.Mandelbrot(1440, 2560)
        m.set_view(v["N"], v["re"], v["im"], v["width"])
        m.set_series_mode(nm.NM_SERIES_FLOATEXP)
        desc = {"workload": "precision variants: 2560x1440, width 1e-200, N 50000, bailout 1024, tol 1e-10, centre = "
                            "detail zoom walk seed 1; series test and skip in floatexp (the reference's double "
                            "test degenerates at this depth: e^2 underflows), FP64 deltas (variant: FP32X deltas)",
                "samples": 2560 * 1440, "N": 50000, "width": "1e-200", "precision_bits": m.precision_bits(),
                "series": "floatexp"}

        def step(k, w, reps=1):
            return m.render(k, w, reps, copy_back=False)
        return m, desc, FLOP_PER_PERT_STEP, step
    if name == "c5":
        v = views["c5_detail1"]
        m = nm.Mandelbrot(2160, 3840)
        m.set_view(v["N"], v["re"], v["im"], v["width"])
        m.set_series_mode(nm.NM_SERIES_FLOATEXP)
        desc = {"workload": "learned-skip configuration: 3840x2160, width 1e-100, N 100000, bailout 1024, tol 1e-10, "
                            "centre = detail zoom walk seed 1; series test and skip in floatexp; the FP64 "
                            "per-sample path (variants: FP32X deltas, 64x64 tile skips from the probe search and "
                            "from the learned predictor trained on the detail walks of seeds 2-11)",
                "samples": 3840 * 2160, "N": 100000, "width": "1e-100", "precision_bits": m.precision_bits(),
                "series": "floatexp"}

        def step(k, w, reps=1):
            return m.render(k, w, reps, copy_back=False)
        return m, desc, FLOP_PER_PERT_STEP, step
    if name == "c4":
        v = views["c4_seed1"]
        widths = ["%.17e" % (1e-40 * 1.1 ** (C4_FRAMES - 1 - k)) for k in range(C4_FRAMES)]
        m = nm.Mandelbrot(720, 1280)
        m.set_view(v["N"], v["re"], v["im"], widths[-1])  # the deepest frame fixes precision and reference
        desc = {"workload": "batched zoom video: %d frames of 1280x720, widths 1e-40 * 1.1^(119-k), one shared "
                            "reference orbit at the deepest frame, N 60000, bailout 1024, tol 1e-10, centre = zoom "
                            "walk seed 1; all frames in one start launch and one persistent iterate launch, "
                            "repeated passes replayed from one HIP graph" % C4_FRAMES,
                "samples": 1280 * 720 * C4_FRAMES, "N": 60000, "frames": C4_FRAMES,
                "precision_bits": m.precision_bits()}

        def step(k, w, reps=1):
            return m.render_frames(widths, k, w, reps, copy_back=False)[1]
        step.widths = widths
        return m, desc, FLOP_PER_PERT_STEP, step
    raise SystemExit("unknown config " + name)


# HBM bytes per dispatch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB), scaled by
# what the same counters read for a copy kernel of known size (1 GiB read + 1 GiB
# written: FETCH_SIZE 526656 KiB, WRITE_SIZE 1084704 KiB)
FETCH_SCALE = 1048576 / 526656
WRITE_SCALE = 1048576 / 1084704
# (the zoom-video summary averages over the autotuner's subsample dispatches too,
# so it does not give the full frame set's traffic)
PMC_FILES = {"c1": "profiles/plain_640x480x512_pmc.txt",
             "c2": "profiles/series_deepzoom_7680x4320_1e-30_pmc.txt",
             "c3": "profiles/series_precision_2560x1440_1e-200_pmc.txt",
             "c4": "profiles/series_zoomvideo_120x1280x720_pmc.txt",
             "c5": "profiles/series_learnedskip_3840x2160_1e-100_pmc.txt"}


def pmc_traffic(config, kernel):
    """HBM bytes per dispatch of `kernel` (its full template name, e.g.
    k_series_iterate<32,false>) from the committed PMC summary, or None."""
    path = os.path.join(ROOT, PMC_FILES.get(config, ""))
    if not os.path.isfile(path):
        return None
    fetch = write = None
    with open(path) as f:
        for ln in f:
            p = ln.split()
            name = p[0] if p else ""
            if len(p) >= 3 and name == kernel and p[1] == "FETCH_SIZE_KiB":
                fetch = float(p[2])
            if len(p) >= 3 and name == kernel and p[1] == "WRITE_SIZE_KiB":
                write = float(p[2])
    if fetch is None or write is None:
        return None
    return (fetch * FETCH_SCALE + write * WRITE_SCALE) * 1024.0


def roofline(config, flop, s, fxs=False):
    """Roofline of the dominant kernel: algorithmic FLOPs of the steps it executed
    over its HIP-event duration (mean over the timed launches), against the FP64
    vector peak; both kernels are listed."""
    exec_iter = s["executed"] - s["executed_start"]
    K = s["block_k"]
    series = flop == FLOP_PER_PERT_STEP
    fma2 = "true" if s.get("fma2", 0) else "false"
    grouped = "true" if s.get("grouped", 0) else "false"
    kernels = {"iterate": {"name": "k_series_iterate" if series else "k_plain_iterate",
                           "instance": ("k_series_iterate<%d,false>" % K) if series
                           else ("k_plain_iterate<%d,%s,%s>" % (K, fma2, grouped)),
                           "ms": s["ms_iterate"], "steps": exec_iter},
               "start": {"name": "k_series_start" if series else "k_plain_start",
                         "instance": ("k_series_start<%s>" % ("true" if fxs else "false")) if series
                         else ("k_plain_start<%s,%s>" % (fma2, grouped)),
                         "ms": s["ms_start"], "steps": s["executed_start"]}}
    for k in kernels.values():
        k["tflops"] = flop * k["steps"] / (k["ms"] * 1e-3) / 1e12 if k["ms"] > 0 else 0.0
    dom = max(kernels.values(), key=lambda k: k["ms"])
    return {"bound": "valu-fp64", "achieved": dom["tflops"], "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": dom["tflops"] / FP64_VECTOR_PEAK_TFLOPS,
            "traffic": pmc_traffic(config, dom.get("instance", dom["name"])),
            "kernel": dom["name"], "flop_per_step": flop, "executed_steps_per_launch": dom["steps"],
            "ms_per_launch": dom["ms"], "kernels": list(kernels.values())}


def train_skipnet(nm, m, T, factors, epochs=120):
    """The learned tile skip's network, trained on the GPU on the probe searches
    (the reference's scan at each tile's probe) of the same centre at other widths,
    all against this view's reference point."""
    re, im = m.reference_point()
    cre, cim = m.center
    width = m.sz_d()[0] * m.cols()
    xs, ts = [], []
    for f in factors:
        t = nm.Mandelbrot(m.rows(), m.cols())
        t.set_view(m.N, cre, cim, "%.17e" % (width * f))
        t.precompute_at(re, im)
        ua, _, _ = t.tile_skips(T, mode=0)
        xs.append(t.tile_features(T))
        ts.append(t.tile_targets(ua))
    x, t = np.concatenate(xs), np.concatenate(ts)
    net = nm.SkipNet()
    t0 = time.perf_counter()
    curve = net.train(x, t, epochs=epochs, batch=256, lr=2e-3, seed=1)
    return net, {"samples": int(len(t)), "views": len(factors), "epochs": epochs, "loss_first": float(curve[0]),
                 "loss_last": float(curve[-1]), "train_s": time.perf_counter() - t0}


def seeded_views(nm, m, names, T):
    """Per named zoom-walk view (this frame size, N and series mode, its own
    findProbe reference): the tile features and the probe searches' targets."""
    views = load_views()
    out = []
    for name in names:
        v = views[name]
        t = nm.Mandelbrot(m.rows(), m.cols())
        t.set_view(int(v["N"]), v["re"], v["im"], v["width"])
        t.set_series_mode(nm.NM_SERIES_FLOATEXP)
        t.precompute()
        ua, _, _ = t.tile_skips(T, mode=0)
        out.append((name, t, ua, t.tile_features(T), t.tile_targets(ua)))
    return out


TRAIN_SEEDS = list(range(2, 12))   # detail walks c5_train2 .. c5_train11
HELD_OUT_SEEDS = [12, 13]


def train_skipnet_seeded(nm, m, T, epochs=120):
    """The learned tile skip's network trained on the GPU on the probe searches of
    the detail zoom walks of seeds 2-11 (3840x2160 at width 1e-100, 2040 tiles
    each), and its acceptance on held-out views: seeds 12 and 13 and this view
    (seed 1), none of them seen in training."""
    t0 = time.perf_counter()
    train = seeded_views(nm, m, ["c5_train%d" % s for s in TRAIN_SEEDS], T)
    t_data = time.perf_counter() - t0
    x = np.concatenate([r[3] for r in train])
    t = np.concatenate([r[4] for r in train])
    net = nm.SkipNet()
    t0 = time.perf_counter()
    curve = net.train(x, t, epochs=epochs, batch=256, lr=2e-3, seed=1)
    t_train = time.perf_counter() - t0
    held = [(name, tv, ua) for name, tv, ua, _, _ in seeded_views(nm, m, ["c5_train%d" % s for s in HELD_OUT_SEEDS],
                                                                    T)]
    held.append(("c5_detail1", m, m.tile_skips(T, mode=0)[0]))
    acc = {}
    for name, tv, ua in held:
        ul, how, st = tv.tile_skips(T, mode=2, net=net)
        scan = tv.tile_skips(T, mode=0)[2]
        acc[name] = {"tiles": int(len(how)), "accepted": float((how == 2).mean()),
                     "lower_bound": float((how == 4).mean()),
                     "equal_to_scan": bool(np.array_equal(ul, ua)), "exact_tests_learned": int(st["evals"]),
                     "exact_tests_scan": int(scan["evals"])}
    return net, {"samples": int(len(t)), "views": len(train), "seeds": TRAIN_SEEDS, "epochs": epochs,
                 "loss_first": float(curve[0]), "loss_last": float(curve[-1]), "train_s": t_train,
                 "data_s": t_data, "held_out": acc}


def variants(m, step, args, value64, wall64, seeded=False):
    """The same frame with FP32X deltas, with per-tile skips (64 x 64 tiles) from
    the plain probe search (the reference's scan at each tile's probe) and from
    the learned predictor, and with FP32X + learned skips: credited px-it/s, its
    ratio to the FP64 per-sample path, and count agreement with it."""
    import newman_amd as nm
    T = 64
    k, w = args.block_k, args.waves
    m.render(k, w, 1, copy_back=True)
    base = m.grid()["iterations"].copy()
    if seeded:
        net, train = train_skipnet_seeded(nm, m, T)
    else:
        net, train = train_skipnet(nm, m, T, [0.3, 0.5, 0.7, 1.5, 2.2, 3.3])
    out = {"training": train}
    for name, delta, skip in (("fp32x", 1, 0), ("fp64_tile_scan", 0, 1), ("fp64_learned_skip", 0, 3),
                              ("fp32x_learned_skip", 1, 3)):
        m.set_delta_mode(delta)
        m.set_skip_mode(skip, T, net if skip else None)
        step(k, w, 1)  # warm-up
        t0 = time.perf_counter()
        st = step(k, w, args.steps)
        wl = (time.perf_counter() - t0) / args.steps
        m.render(k, w, 1, copy_back=True)
        same = float((m.grid()["iterations"] == base).mean())
        v = {"value": st["credited"] / wl, "ms_per_step": wl * 1e3, "vs_fp64": (st["credited"] / wl) / value64,
             "ms_iterate": st["ms_iterate"], "ms_start": st["ms_start"], "executed": st["executed"],
             "counts_equal_fp64": same}
        if skip:
            v["tiles"] = m.tile_stats()
        out[name] = v
        print("# variant %s: %.3e px-it/s (%.2fx FP64), counts = FP64 on %.4f%%" % (
            name, v["value"], v["vs_fp64"], 100 * same), file=sys.stderr, flush=True)
    m.set_delta_mode(0)
    m.set_skip_mode(0, T, None)
    return out


def fp32x_variant(m, step, args, value64):
    """FP32X deltas on the same frame: value, ratio and per-sample error against
    the FP64 deltas (bit-exact with the oracle's FP64 method)."""
    k, w = args.block_k, args.waves
    m.render(k, w, 1, copy_back=True)
    g64 = m.grid().copy()
    m.set_delta_mode(1)
    step(k, w, 1)
    t0 = time.perf_counter()
    st = step(k, w, args.steps)
    wl = (time.perf_counter() - t0) / args.steps
    m.render(k, w, 1, copy_back=True)
    g32 = m.grid().copy()
    m.set_delta_mode(0)
    same = g32["iterations"] == g64["iterations"]
    ds = np.abs(g32["smoothing"].astype(np.float64) - g64["smoothing"])[same]
    return {"fp32x": {"value": st["credited"] / wl, "ms_per_step": wl * 1e3, "vs_fp64": (st["credited"] / wl) / value64,
                      "ms_iterate": st["ms_iterate"], "counts_equal_fp64": float(same.mean()),
                      "max_count_diff": int(np.abs(g32["iterations"].astype(np.int64) - g64["iterations"]).max()),
                      "smoothing_absdiff_p99": float(np.percentile(ds, 99)) if ds.size else 0.0}}


def colour_leg(m, sc):
    """k_colour on the last render's device grid with the reference's default
    palette: algorithmic HBM bytes = 8 per sample read + 3 per output pixel
    written (the palette stays in L2), over the kernel's HIP-event time."""
    import newman_amd as nm
    pal = nm.Palette(filename=os.path.join(ROOT, "tests", "golden", "default.pal")).cache(m.N)
    m.colour(pal, True, sc)  # warm
    runs = [m.colour(pal, True, sc)[1] for _ in range(5)]
    ms = sum(runs) / len(runs)
    nin = m.rows() * m.cols()
    nout = (m.rows() // sc) * (m.cols() // sc)
    nbytes = 8 * nin + 3 * nout
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "k_colour", "output": "%dx%d RGB from %dx%d samples, sc=%d, smoothing on" % (
        m.cols() // sc, m.rows() // sc, m.cols(), m.rows(), sc), "ms": ms, "bytes": nbytes,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0}}


def video_leg(m, rate):
    """k_video_frames: `rate` in-between frames of a (rows/3) x (cols/3) canvas from
    two keyframes of (rows/2) x (cols/2) (the grid coloured 2 x 2), as the viewer's
    zoom loop makes them; algorithmic bytes per output pixel = 3 written + 8 read
    (one 4-byte pixel of each keyframe: both scaled keyframes cover the canvas, at
    about one source pixel per output pixel), over the kernels' HIP-event time."""
    import newman_amd as nm
    pal = nm.Palette(filename=os.path.join(ROOT, "tests", "golden", "default.pal")).cache(m.N)
    key, _ = m.colour(pal, True, 2)
    nr, nc = m.rows() // 3, m.cols() // 3
    nm.video_interpolate(key, key, nr, nc, rate)  # warm
    runs = [nm.video_interpolate(key, key, nr, nc, rate)[1] for _ in range(3)]
    ms = sum(runs) / len(runs)
    nbytes = 11 * nr * nc * rate
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "k_video_frames", "output": "%d frames of %dx%d RGB from %dx%d keyframes" % (
        rate, nc, nr, key.shape[1], key.shape[0]), "ms": ms, "bytes": nbytes,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0},
        "note": "VALU-bound: the exact float bilinear taps cost ~190 VALU instructions per output pixel"}


def bailout_leg(m, step, k, w, args):
    """The same 512-frame set at bailout radius 2, as the configuration names it
    (the reference's constant is 1024): credited px-it/s and the kernels' times."""
    m.set_bailout(2.0)
    step(k, w, 1)
    t0 = time.perf_counter()
    s = step(k, w, args.steps)
    wall = (time.perf_counter() - t0) / args.steps
    m.set_bailout(1024.0)
    return {"bailout": 2.0, "value": s["credited"] / wall, "ms_per_step": wall * 1e3,
            "credited_px_it": s["credited"], "ms_start": s["ms_start"], "ms_iterate": s["ms_iterate"],
            "roofline": roofline("c1_b2", FLOP_PER_PLAIN_STEP, s)}


def tune(step, ks, waves):
    """Sweep block size and persistent waves per SIMD; kernel time by HIP events."""
    best = None
    rows = []
    for k in ks:
        for w in waves:
            st = step(k, w)
            st = step(k, w)
            rows.append((k, w, st["ms_total"]))
            print("# tune K=%d waves=%d: %.3f ms" % (k, w, st["ms_total"]), file=sys.stderr, flush=True)
            if best is None or st["ms_total"] < best[2]:
                best = (k, w, st["ms_total"])
    return best, rows


def single_frame_leg(reps=50):
    """One 640x480 frame per call, as the viewer renders it: the device pass, and
    the whole call with the grid copied back to the host (what computeRow's
    caller waits for)."""
    import newman_amd as nm
    m = nm.Mandelbrot(480, 640)
    m.set_view(1024, "-0.75", "0", "3.0")
    m.render()
    s = m.render(0, 0, reps, copy_back=False)
    t0 = time.perf_counter()
    for _ in range(reps):
        m.render()
    wall = (time.perf_counter() - t0) / reps
    return {"frame": "640x480, centre -0.75+0i, width 3.0, N 1024", "ms_device_pass": s["ms_total"],
            "ms_call_with_copy_back": wall * 1e3, "calls": reps}


def _one_thread(run, what):
    """The reference renders on one thread: the same oracle on one core, over a
    smaller share of the same sample (`run` returns (credited px-it, samples))."""
    t0 = time.perf_counter()
    credited, n = run()
    t = time.perf_counter() - t0
    return {"value": credited / t, "unit": "pixel-iterations/s", "cores": 1,
            "sample": "%s (%d samples, %.3g credited px-it, %.1f s)" % (what, n, credited, t)}


def host_cpu():
    """The host's CPU model and the cores this process may use."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "cores_available": len(os.sched_getaffinity(0))}


def cpu_baseline(desc, name, nthreads, rows_every):
    """The oracle (the reference's method on the CPU) on a bounded sample of the
    same workload, timed on this host's cores, and on one core (`single_thread`)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_lib as O
    if name == "c1":
        o = O.OView(480, 640)
        o.set_view(1024, "-0.75", "0", "3.0")
        rows = np.arange(480, dtype=np.int32)
        card = np.array([[o.in_cardioid(r, c) for c in range(640)] for r in range(480)])
        reps = 0
        credited = 0
        t0 = time.perf_counter()
        while True:  # repeat the full frame for about 10 s
            out = o.render_rows(rows, nthreads)
            credited += int(out["iterations"].astype(np.int64)[~card].sum())
            reps += 1
            if time.perf_counter() - t0 > 10.0:
                break
        t_run = time.perf_counter() - t0

        def one():
            c = 0
            for _ in range(4):
                c += int(o.render_rows(rows, 1)["iterations"].astype(np.int64)[~card].sum())
            return c, 4 * rows.size * 640
        return {"value": credited / t_run, "unit": "pixel-iterations/s", "cores": nthreads, "kind": "port",
                "sample": "oracle plain-double path (mandelbrot.cpp:231-254 with the HP cardioid test) on the full "
                          "640x480 frame, repeated %d times (%.1f s); credited = iterations of non-cardioid samples"
                          % (reps, t_run),
                "single_thread": _one_thread(one, "the full frame 4 times")}
    views = load_views()
    if name == "c3":
        v = views["c3_detail1"]
        nr, nc = 1440, 2560
        o = O.OView(nr, nc)
        o.set_view(v["N"], v["re"], v["im"], v["width"])
        o.set_series_mode(1)
        t0 = time.perf_counter()
        o.precompute(nthreads)
        t_pre = time.perf_counter() - t0
        rows = np.arange(rows_every // 2, nr, rows_every * 5, dtype=np.int32)
        rc = np.array([(r, c) for r in rows for c in range(nc)], dtype=np.int32)
        t0 = time.perf_counter()
        out = o.render_samples(rc, nthreads, pert=True)
        t_run = time.perf_counter() - t0
        credited = int(out["iterations"].astype(np.int64).sum())
        rc1 = np.array([(r, c) for r in rows[::4] for c in range(nc)], dtype=np.int32)

        def one():
            o1 = o.render_samples(rc1, 1, pert=True)
            return int(o1["iterations"].astype(np.int64).sum()), o1.size
        return {"value": credited / t_run, "unit": "pixel-iterations/s", "cores": nthreads, "kind": "port",
                "sample": "oracle floatexp series start + FP64 perturbation (the method the GPU runs here) on rows "
                          "%d::%d of the 2560x1440 frame (%d samples, %.3g credited px-it, %.1f s); precompute "
                          "%.1f s, not included" % (rows_every // 2, rows_every * 5, out.size, credited, t_run, t_pre),
                "single_thread": _one_thread(one, "rows %d::%d" % (rows_every // 2, rows_every * 20))}
    if name == "c5":
        v = views["c5_detail1"]
        nr, nc = 2160, 3840
        o = O.OView(nr, nc)
        o.set_view(v["N"], v["re"], v["im"], v["width"])
        o.set_series_mode(1)
        t0 = time.perf_counter()
        o.precompute(nthreads)
        t_pre = time.perf_counter() - t0
        rows = np.arange(rows_every // 2, nr, rows_every * 15, dtype=np.int32)
        t0 = time.perf_counter()
        out = o.render_rows(rows, nthreads)
        t_run = time.perf_counter() - t0
        credited = int(out["iterations"].astype(np.int64).sum())

        def one():
            o1 = o.render_rows(rows[::2], 1)
            return int(o1["iterations"].astype(np.int64).sum()), o1.size
        return {"value": credited / t_run, "unit": "pixel-iterations/s", "cores": nthreads, "kind": "port",
                "sample": "oracle reference method (series scan in floatexp + HP continuation, %d-bit) on rows "
                          "%d::%d of the 3840x2160 frame (%d samples, %.3g credited px-it, %.1f s); precompute "
                          "%.1f s, not included" % (o.bits(), rows_every // 2, rows_every * 15, out.size, credited,
                                                    t_run, t_pre),
                "single_thread": _one_thread(one, "rows %d::%d" % (rows_every // 2, rows_every * 30))}
    if name == "c4":
        v = views["c4_seed1"]
        nr, nc = 720, 1280
        total_s = total_c = 0.0
        pre = 0.0
        nsamp = 0
        frames = (0, 60, 119)
        single = None
        for k in frames:  # the reference method precomputes per keyframe (viewer.cpp:263)
            o = O.OView(nr, nc)
            o.set_view(v["N"], v["re"], v["im"], "%.17e" % (1e-40 * 1.1 ** (C4_FRAMES - 1 - k)))
            t0 = time.perf_counter()
            o.precompute(nthreads)
            pre += time.perf_counter() - t0
            rows = np.arange(rows_every // 4, nr, rows_every // 2, dtype=np.int32)
            t0 = time.perf_counter()
            out = o.render_rows(rows, nthreads)
            total_s += time.perf_counter() - t0
            total_c += float(out["iterations"].astype(np.int64).sum())
            nsamp += out.size
            if k == frames[-1]:
                def one():
                    o1 = o.render_rows(rows[::4], 1)
                    return int(o1["iterations"].astype(np.int64).sum()), o1.size
                single = _one_thread(one, "rows %d::%d of frame %d" % (rows_every // 4, rows_every * 2, k))
        return {"value": total_c / total_s, "unit": "pixel-iterations/s", "cores": nthreads, "kind": "port",
                "sample": "oracle reference method (series scan + HP continuation) on rows %d::%d of frames %s "
                          "(%d samples, %.3g credited px-it, %.1f s); per-frame precompute %.1f s, not included"
                          % (rows_every // 4, rows_every // 2, frames, nsamp, total_c, total_s, pre),
                "single_thread": single}
    v = views["c2_seed1"]
    nr, nc = 1080 * 4, 1920 * 4
    o = O.OView(nr, nc)
    o.set_view(v["N"], v["re"], v["im"], v["width"])
    t0 = time.perf_counter()
    o.precompute(nthreads)
    t_pre = time.perf_counter() - t0
    rows = np.arange(rows_every // 2, nr, rows_every, dtype=np.int32)
    t0 = time.perf_counter()
    out = o.render_rows(rows, nthreads)
    t_run = time.perf_counter() - t0
    it = out["iterations"].astype(np.int64)
    credited = int(it.sum())  # deep frame: no cardioid samples

    def one():
        o1 = o.render_rows(rows[::8], 1)
        return int(o1["iterations"].astype(np.int64).sum()), o1.size
    return {"value": credited / t_run, "unit": "pixel-iterations/s", "cores": nthreads, "kind": "port",
            "sample": "oracle reference method (series scan + HP continuation, %d-bit) on rows %d::%d of the same "
                      "7680x4320 frame (%d samples, %.3g credited px-it, %.1f s); precompute (findProbe over "
                      "13680 probes + series) %.1f s, not included" % (
                          o.bits(), rows_every // 2, rows_every, out.size, credited, t_run, t_pre),
            "single_thread": _one_thread(one, "rows %d::%d" % (rows_every // 2, rows_every * 8))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--deep", type=int, default=1,
                    help="c1: also run the deep-zoom configuration (c2) as the deep_zoom key")
    ap.add_argument("--block-k", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--pre-iter", type=int, default=0, help="plain path: updates in the start kernel")
    ap.add_argument("--tile", type=int, default=0, help="sample order: 0 row-major, T = T x T tiles")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--legs", type=int, default=1,
                    help="also time the colouring and video legs (c2) and the single-frame latency (c1); 0 for the "
                         "main line only, as the profiles are taken")
    ap.add_argument("--variants", type=int, default=1,
                    help="also time FP32X deltas and the learned tile skip on the same frame (c2, c5)")
    ap.add_argument("--tune", action="store_true",
                    help="sweep block size and waves per SIMD first (series: then the chunk interleaving)")
    ap.add_argument("--chunk-run", type=int, default=0,
                    help="series path: active-list chunks of 64 consecutive samples (64) or interleaved runs of "
                         "1, 2, 4, 8, 16, 32; 0 = the default (runs of 4 for a floatexp series, else 64)")
    ap.add_argument("--cpu-rows-every", type=int, default=36)
    ap.add_argument("--cpu-threads", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the frame would shard by "
              "rows/tiles with the reference orbit broadcast once)", file=sys.stderr)
        sys.exit(2)

    line = measure(args.config, args)
    line = dict({"metric": METRIC}, **line)
    if args.config == "c1" and args.deep:
        # the deep-zoom configuration under its own key, with its variants, the
        # colouring and video legs and its CPU baseline
        print("# deep_zoom (7680x4320 at 1e-30) ...", file=sys.stderr, flush=True)
        line["deep_zoom"] = measure("c2", args, block_k=0, waves=0)
    print(json.dumps(line), flush=True)


def measure(config, args, block_k=None, waves=None):
    """One workload: warm-up, `steps` timed passes in one engine call, and its
    side legs; returns the bench line's fields (without the metric name)."""
    import newman_amd as nm  # noqa: F401
    m, desc, flop, step = make_workload(config)
    t0 = time.perf_counter()
    if not m.useHardware():
        m.precompute()
    t_pre = time.perf_counter() - t0
    n_hp = m.probe_candidates() if not m.useHardware() else 0
    if not m.useHardware():
        print("# %s: precompute %.2f s (findProbe screened on the GPU, %d probe orbits in HP; series in HP), "
              "len(X) = %s" % (config, t_pre, n_hp, m.reference_length()), file=sys.stderr, flush=True)
    k = args.block_k if block_k is None else block_k
    w = args.waves if waves is None else waves
    if config != "c4":
        m.set_launch(tile=args.tile if config == args.config else 0,
                     pre_iter=args.pre_iter if config == args.config else 0,
                     chunk_run=args.chunk_run if config == args.config else 0)
    tuned = None
    if config == "c4":
        # the project's launch autotuner (sample order x waves x block size), then
        # graph replay of the repeated passes
        t0 = time.perf_counter()
        best, table = m.autotune_frames(step.widths)
        m.set_launch(best["block_k"], best["waves_per_simd"], best["tile"], graph=True, chunk_run=best["chunk_run"])
        k, w = best["block_k"], best["waves_per_simd"]
        tuned = {"best": best, "sweep": table, "seconds": time.perf_counter() - t0}
        print("# autotune: %s (%.1f s)" % (best, tuned["seconds"]), file=sys.stderr, flush=True)
    elif args.tune and config == args.config:
        best, rows = tune(step, [8, 16, 32, 64] if config == "c1" else [12, 16, 32, 64], [2, 4, 6, 8])
        k, w = best[0], best[1]
        tuned = {"best": {"block_k": k, "waves_per_simd": w, "chunk_run": args.chunk_run, "ms": best[2]},
                 "sweep": [{"block_k": a, "waves_per_simd": b, "ms": c} for a, b, c in rows]}
        if config != "c1":  # the series kernel's chunk interleaving at the best setting
            for run in (1, 4, 16, 64):
                m.set_launch(tile=args.tile, chunk_run=run)
                step(k, w)
                ms = step(k, w)["ms_total"]
                print("# tune chunk_run=%d: %.3f ms" % (run, ms), file=sys.stderr, flush=True)
                tuned["sweep"].append({"block_k": k, "waves_per_simd": w, "chunk_run": run, "ms": ms})
                if ms < tuned["best"]["ms"]:
                    tuned["best"] = {"block_k": k, "waves_per_simd": w, "chunk_run": run, "ms": ms}
            m.set_launch(tile=args.tile, chunk_run=tuned["best"]["chunk_run"])

    if args.warmup:
        step(k, w, args.warmup)
    # the timed region: one engine call that runs the whole pipeline `steps`
    # times back to back on the resident inputs (each call is synchronous)
    t0 = time.perf_counter()
    s = step(k, w, args.steps)
    wall = (time.perf_counter() - t0) / args.steps
    credited = s["credited"]
    value = credited / wall
    line = {
        "value": value,
        "unit": "pixel-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (fixed view)" if config == "c1" else "synthetic (seeded zoom-walk view)",
        "config": dict(desc, parallelism="single GPU", block_k=s["block_k"], waves_per_simd=s["waves_per_simd"],
                       tile=s["tile"], graph=s["graph"], pre_iter=m.get_launch()["pre_iter"],
                       chunk_run=s["chunk_run"]),
        "roofline": roofline(config, flop, s, fxs=desc.get("series") == "floatexp"),
        "detail": {"credited_px_it": credited, "executed_px_it": s["executed"],
                   "executed_in_start_kernel": s["executed_start"], "skipped": s["skipped"],
                   "n_iterated": s["n_iterated"], "n_sa_resolved": s["n_sa_resolved"],
                   "n_cardioid": s["n_cardioid"], "n_rebased_blocks": s["n_rebased"],
                   "ms_start": s["ms_start"], "ms_compact": s["ms_compact"], "ms_iterate": s["ms_iterate"],
                   "ms_finalize": s["ms_finalize"], "ms_kernels": s["ms_total"], "precompute_s": t_pre,
                   "probe_orbits_hp": n_hp, "folded_update": bool(s.get("fma2", 0)), "grouped_test": bool(s.get("grouped", 0)),
                   "kernel_px_it_per_s": credited / (s["ms_total"] * 1e-3)},
        "cpu_baseline": None,
    }
    if tuned:
        line["autotune"] = tuned
    if args.legs and config == "c1":
        line["single_frame"] = single_frame_leg()
        line["bailout_2"] = bailout_leg(m, step, k, w, args)
    if args.legs and config == "c2":
        # the viewer's colouring of the rendered grid, on the device copy (no host
        # round trip): 4 x 4 average into 1920x1080 RGB with smoothing; then the
        # zoom video's in-between frames from that grid's 2 x 2 colouring
        line["colour"] = colour_leg(m, 4)
        line["video"] = video_leg(m, 15)
    if args.variants and config in ("c2", "c5"):
        line["variants"] = variants(m, step, args, value, wall, seeded=config == "c5")
    if args.variants and config == "c3":
        line["variants"] = fp32x_variant(m, step, args, value)
    if not args.no_cpu_baseline:
        print("# cpu baseline (%s) ..." % config, file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(desc, config, args.cpu_threads, args.cpu_rows_every)
        line["cpu_baseline"]["host"] = host_cpu()
    return line


if __name__ == "__main__":
    main()
