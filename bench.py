"""Benchmark of the locate path on one MI355X.

Workload: a device-resident movie of 1000 frames of 2048x2048 uint16 (5000
Gaussian spots per frame, sigma 2 px, lognormal peaks with median 1200,
Brownian motion D = 0.5 px^2/frame, background 150 + Poisson noise), located
with diameter 11, separation 11, smoothing_size 11, percentile 64, minmass 0.

A step is one pass of the locate path over the whole movie.  `value` is the
batched path's frames/s with frames and result rows resident in HBM; the plain
per-frame path (launch + completion wait per frame, as the reference's
execute_gpu loop does) is timed on the same movie beside it.  The roofline
entry is for the dominant kernel, measured with HIP events on the stream the
library launches it on, in one extra step that runs the batches on one stream
(so that a kernel's duration is not stretched by the next batch's kernels);
the CPU baseline is the oracle (the CPU restatement of the reference
pipeline, compiled with -march=native on the host it runs on) on the first 64
frames, on one thread and on the 16 threads of one GPU's CPU share.  The
same line carries the single-frame latency of the 512x512 workload (plain GPU
path and CPU oracle), the uneven 4096-frame batch (batched and plain), the
adjoint workload (256 x 512^2, localisation errors) and the locate + link
workload (2000 x 1024^2, 3000 blinking particles, with the C oracle linker
timed on the same rows).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(h, w, n_cand, n_rows, ncols, rint, in_bytes=2):
    """Bytes of the three-kernel locate pipeline per frame, each counted once:
    frame read (in_bytes: 2 for uint16), float32 bandpass written (4) and
    read once (4) per pixel; per candidate its index (4) and one (2r+1)^2 box
    read; 4 bytes per output value."""
    return ((in_bytes + 8.0) * h * w + n_cand * (4.0 + 4.0 * (2 * rint + 1) ** 2)
            + 4.0 * ncols * n_rows)


_NATIVE_ORACLE = None


def native_oracle():
    """The oracle compiled for this host's CPU (gcc -O3 -march=native
    -ffp-contract=off -fopenmp) into a temporary directory: the CPU baseline
    is timed with the build its host would use."""
    global _NATIVE_ORACLE
    from oracle import oracle
    if _NATIVE_ORACLE is None:
        import tempfile
        _NATIVE_ORACLE = oracle.use_native_build(tempfile.mkdtemp(prefix="gt_oracle_native_"))
    return oracle


def host_cpu() -> dict:
    """CPUs this process may run on, the machine's total and its model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpus_available": len(os.sched_getaffinity(0)), "cpus_machine": os.cpu_count(),
            "model": model}


def cpu_baseline(frames_host: np.ndarray, kw: dict, threads: int):
    """Oracle frames/s over the sample, frames spread over `threads` threads,
    one frame per thread at a time (ctypes releases the GIL, so the threads
    run in parallel)."""
    from concurrent.futures import ThreadPoolExecutor
    oracle = native_oracle()
    oracle.set_threads(1)
    p = oracle.Params(**kw)
    n = len(frames_host)
    f32 = [frames_host[i].astype(np.float32) for i in range(n)]

    def one(i):
        return len(oracle.locate_frame(f32[i], p, frame_index=i)[0])

    t0 = time.perf_counter()
    if threads == 1:
        rows = sum(one(i) for i in range(n))
    else:
        with ThreadPoolExecutor(max_workers=threads) as ex:
            rows = sum(ex.map(one, range(n)))
    dt = time.perf_counter() - t0
    return n / dt, rows, dt


def c1_latency(ctx, threads: int, repeats: int = 30) -> dict:
    """Single-frame locate of the 512x512 frame (200 spots, sigma 1.5, peaks
    U(500, 2000), background 100 + Poisson; diameter 9, separation 9,
    smoothing 9, percentile 64, minmass 1000): median wall time of a warm
    call on the plain per-frame GPU path (frame in HBM, rows left in HBM;
    and numpy frame in, rows back on the host) and of the CPU oracle on one
    thread and on `threads` threads (the frame's rows split over them)."""
    import ctypes
    from tracking_optimizations_amd import _lib, synth
    from tracking_optimizations_amd.gpu_tracking import locate, make_params
    L = _lib.load()
    spec = synth.single_512(seed=1)
    frame = synth.generate(ctx, spec)
    host = frame.to_host(np.uint16, (1, 512, 512))
    kw = dict(diameter=9, separation=9, minmass=1000.0, percentile=64.0)
    p = make_params(**kw)

    def dev_call():
        out = ctypes.c_void_p()
        _lib.check(L.gt_locate_u16(ctx.handle, ctypes.c_void_p(frame.ptr), 1, 512, 512,
                                   ctypes.byref(p), _lib.GT_FRAMES_ON_DEVICE |
                                   _lib.GT_RESULT_ON_DEVICE | _lib.GT_PLAIN, ctypes.byref(out)))
        _lib.Result(out)

    def host_call():
        locate(host, p, plain=True, ctx=ctx)

    def median_ms(fn, n):
        for _ in range(3):
            fn()
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return 1000.0 * float(np.median(ts))

    out = {"frame": "512x512 uint16, 200 spots", "params": kw, "repeats": repeats,
           "gpu_plain_device_ms": median_ms(dev_call, repeats),
           "gpu_plain_host_ms": median_ms(host_call, repeats)}
    oracle = native_oracle()
    op = oracle.Params(**kw)
    f32 = host[0].astype(np.float32)
    rows = len(oracle.locate_frame(f32, op)[0])
    old = oracle.set_threads(1)
    out["cpu_1_thread_ms"] = median_ms(lambda: oracle.locate_frame(f32, op), max(20, repeats // 2))
    oracle.set_threads(threads)
    out["cpu_threads"] = threads
    out["cpu_all_threads_ms"] = median_ms(lambda: oracle.locate_frame(f32, op),
                                          max(20, repeats // 2))
    oracle.set_threads(old)
    out["rows"] = rows
    frame.free()
    return out


def uneven_batch(ctx, steps: int, n_frames: int = 4096) -> dict:
    """The uneven batch (4096 frames: 40 % 256x256, 40 % 512x1024, 20 %
    2048x2048, shuffled; one spot per 800 px^2, sigma U(1.2, 2.5), background
    U(50, 300)) device-resident, located in one gt_locate_batch call per step
    on the batched path, then once on the plain per-frame path."""
    import ctypes
    from tracking_optimizations_amd import _lib, synth
    from tracking_optimizations_amd.gpu_tracking import make_params
    kw = dict(diameter=11, separation=11, smoothing_size=11, percentile=64, minmass=0.0)
    params = make_params(**kw)
    L = _lib.load()
    movie = synth.UnevenMovie(ctx, n_frames, seed=3)
    descs = movie.frame_descs()
    ctx.synchronize()

    def run(flags):
        out = ctypes.c_void_p()
        _lib.check(L.gt_locate_batch(ctx.handle, descs, n_frames, 0, ctypes.byref(params), flags,
                                     ctypes.byref(out)))
        r = _lib.Result(out)
        t = ctx.timing()
        del r
        return t

    bflags = _lib.GT_FRAMES_ON_DEVICE | _lib.GT_RESULT_ON_DEVICE
    run(bflags)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = run(bflags)
    ctx.synchronize()
    el = time.perf_counter() - t0
    run(bflags | _lib.GT_PLAIN)
    ctx.synchronize()
    t0 = time.perf_counter()
    run(bflags | _lib.GT_PLAIN)
    ctx.synchronize()
    pel = time.perf_counter() - t0
    fps = n_frames * steps / el
    res = {"frames": n_frames, "steps": steps, "frames_per_s": fps,
           "pixels_per_s": movie.pixels * steps / el, "plain_frames_per_s": n_frames / pel,
           "batched_over_plain": fps * pel / n_frames, "rows": st["rows"],
           "candidates": st["candidates"]}
    movie.free()
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=1000)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--batch", type=int, default=0, help="frames per launch (0 = automatic)")
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-batch figures of the default line")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the adjoint and locate + link figures of the default line")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-percentile", action="store_true",
                    help="locate without the percentile test (the reference's default)")
    ap.add_argument("--seg-rows", type=int, default=0,
                    help="rows per bandpass work item (default: the library's 128)")
    ap.add_argument("--separate-refine", action="store_true",
                    help="walk in a separate refine kernel instead of inside detect")
    ap.add_argument("--no-graph", action="store_true",
                    help="do not also time the steps replaying a captured HIP graph")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the launch-shape autotuner (rows per bandpass item, frames per "
                         "launch, walk placement) that runs before timing")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="also run one device-to-device copy of MB megabytes (a known-size "
                         "kernel for scaling hardware byte counters)")
    ap.add_argument("--input", choices=["u16", "f32"], default="u16",
                    help="frame type of the movie workload: uint16, or float32 frames holding "
                         "the same values (the reference's own frame type)")
    ap.add_argument("--workload", choices=["movie", "uneven", "adjoint", "link", "file", "raw"],
                    default="movie",
                    help="movie: 1000 x 2048^2 (headline); uneven: 4096 mixed-size frames; "
                         "adjoint: 256 x 512^2 localisation errors and Jacobians; "
                         "link: locate + link, 2000 x 1024^2, 3000 blinking particles; "
                         "file: batch_file on a 256 x 2048^2 uint16 TIFF; "
                         "raw: raw signal columns on 200 frames of the headline movie")
    ap.add_argument("--link-cpu-frames", type=int, default=150,
                    help="frames of the link workload also linked by the Python oracle")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": f"--gpus {args.gpus}: this benchmark runs on exactly one GPU; "
                                   "frames are independent, so N GPUs would each run this same "
                                   "path on 1/N of the movie"}))
        return 2

    from tracking_optimizations_amd import _lib, synth
    from tracking_optimizations_amd.gpu_tracking import make_params
    if args.workload == "uneven":
        return bench_uneven(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "link":
        return bench_link(args)
    if args.workload == "file":
        return bench_file(args)
    if args.workload == "raw":
        return bench_raw(args)

    kw = dict(diameter=11, separation=11, smoothing_size=11, percentile=64, minmass=0.0)
    if args.no_percentile:
        del kw["percentile"]
    params = make_params(**kw)
    ctx = _lib.Context(0)
    if args.batch:
        ctx.set_batch(args.batch)
    if args.separate_refine:
        ctx.set_fused_walk(False)
    if args.seg_rows:
        ctx.set_launch_shape(args.seg_rows)
    L = _lib.load()
    import ctypes

    H = W = args.size
    nf = args.frames
    log(f"generating {nf} frames of {H}x{W} uint16 on the device")
    t0 = time.perf_counter()
    spec = synth.movie_2048(seed=2, n_frames=nf, size=H)
    movie = synth.generate(ctx, spec)
    ctx.synchronize()
    log(f"movie ready in {time.perf_counter() - t0:.1f} s ({movie.nbytes / 1e9:.2f} GB)")
    locate_fn, frames_ptr = L.gt_locate_u16, movie.ptr
    if args.input == "f32":
        # float32 frames holding the uint16 values, converted on the host in
        # chunks of 32 frames
        fmovie = synth.DeviceBuffer(ctx, nf * H * W * 4)
        for f0 in range(0, nf, 32):
            k = min(32, nf - f0)
            chunk = movie.to_host(np.uint16, (k, H, W), f0 * H * W * 2).astype(np.float32)
            _lib.check(L.gt_memcpy_h2d(ctx.handle, ctypes.c_void_p(fmovie.ptr + f0 * H * W * 4),
                                       chunk.ctypes.data_as(ctypes.c_void_p), chunk.nbytes))
        ctx.synchronize()
        locate_fn, frames_ptr = L.gt_locate_f32, fmovie.ptr
        log(f"float32 copy of the movie ready ({fmovie.nbytes / 1e9:.2f} GB)")

    def run(flags):
        out = ctypes.c_void_p()
        _lib.check(locate_fn(ctx.handle, ctypes.c_void_p(frames_ptr), nf, H, W,
                             ctypes.byref(params), flags, ctypes.byref(out)))
        r = _lib.Result(out)
        t = ctx.timing()
        del r
        return t

    batched_flags = _lib.GT_FRAMES_ON_DEVICE | _lib.GT_RESULT_ON_DEVICE
    plain_flags = batched_flags | _lib.GT_PLAIN

    def sync():
        ctx.synchronize()

    if args.copy_calibration:
        nb = min(args.copy_calibration << 20, movie.nbytes // 16 * 16)
        dst = synth.DeviceBuffer(ctx, nb)
        _lib.check(L.gt_memcpy_d2d(ctx.handle, ctypes.c_void_p(dst.ptr),
                                   ctypes.c_void_p(movie.ptr), nb))
        log(f"copy calibration: k_copy read {nb} bytes and wrote {nb} bytes")
        dst.free()

    tuning = None
    if not args.no_autotune and not args.seg_rows and not args.batch and not args.separate_refine:
        # the sample spans several full launches, as the timed movie does
        tuning = ctx.autotune(params, H, W, n_frames=min(nf, 256))
        log(f"autotune: {tuning['seg_rows']} rows per bandpass item, "
            f"{tuning['frames_per_launch'] or 'automatic'} frames per launch, walk "
            f"{'in detect' if tuning['fused_walk'] else 'in its own kernel'}, "
            f"{1000 * tuning['device_ms_per_frame']:.1f} us/frame on the tuning sample")

    for i in range(args.warmup):
        t = run(batched_flags)
        log(f"warmup {i}: device {t['device_ms']:.1f} ms, {t['candidates']} candidates, "
            f"{t['rows']} rows")
    sync()
    # timed region: K steps, frames and rows resident in HBM, two streams
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        stats.append(run(batched_flags))
    sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    fps = nf * args.steps / elapsed
    dev_ms = float(np.mean([s["device_ms"] for s in stats]))
    n_cand = stats[-1]["candidates"]
    n_rows = stats[-1]["rows"]
    fps_direct = fps
    log(f"batched: {fps:.0f} frames/s ({ms_per_step:.1f} ms/step, device {dev_ms:.1f} ms), "
        f"{n_cand / nf:.0f} candidates/frame, {n_rows / nf:.0f} rows/frame")
    # one more step with per-stage HIP events recorded by the library on the
    # stream each launch goes to; profiling runs the batches on one stream,
    # so every kernel's duration is its own (no overlap with the next batch)
    ctx.set_profiling(True)
    t1 = time.perf_counter()
    run(batched_flags)
    sync()
    fps_prof = nf / (time.perf_counter() - t1)
    stages = ctx.stage_times()
    ctx.set_profiling(False)
    log(f"profiled step (one stream, stage events): {fps_prof:.0f} frames/s")
    fps_graph = None
    path = "batched, direct launches"
    if not args.no_graph:
        # the same steps replaying the captured launch sequence (one capture,
        # then K replays); the faster of the two is `value`
        ctx.set_graphs(True)
        run(batched_flags)  # capture
        sync()
        t1 = time.perf_counter()
        for i in range(args.steps):
            run(batched_flags)
        sync()
        el_graph = time.perf_counter() - t1
        fps_graph = nf * args.steps / el_graph
        ctx.set_graphs(False)
        log(f"batched, graph replay: {fps_graph:.0f} frames/s")
        if fps_graph > fps:
            fps, elapsed = fps_graph, el_graph
            ms_per_step = 1000.0 * elapsed / args.steps
            path = "batched, graph replay"
    for k, (ms, n) in stages.items():
        if n:
            log(f"  stage {k:9s}: {ms:8.3f} ms/step over {n} launches ({ms / n:.4f} ms/launch)")
    dominant = max(stages, key=lambda k: stages[k][0])
    dom_ms, dom_n = stages[dominant]
    rint = (kw["diameter"] - 1) // 2
    px = float(H) * W * nf
    in_b = 2 if args.input == "u16" else 4
    # algorithmic bytes of each kernel over the whole step
    kernel_bytes = {
        "bandpass": px * (in_b + 4),                    # frame in, float32 out
        # float32 in, candidate slots out, plus (centroid walk fused in
        # detect) one disk box read and one record written per candidate
        "detect": px * 4 + n_cand * (2 + 4 * (2 * rint + 1) ** 2 + 16),
        "refine": n_cand * (2 + 4 * (2 * rint + 1) ** 2 + 12),
        "dedup": n_cand * 16.0,
        "emit": n_rows * 4.0 * 4 + n_rows * 8.0,
    }
    dom_bytes = kernel_bytes.get(dominant, 0.0)
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9 if dom_ms > 0 else 0.0
    per_launch_bytes = dom_bytes / max(dom_n, 1)
    # measured HBM bytes of the dominant kernel (rocprofv3 FETCH_SIZE and
    # WRITE_SIZE per dispatch, scaled by a copy kernel of known size; see
    # profiles/traffic_movie_2048.json) for this launch's frame count
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_movie_2048.json")
    if H == 2048 and W == 2048 and args.input == "u16" and os.path.exists(tpath):
        tk = json.load(open(tpath)).get("kernels", {})
        for name, rec in tk.items():
            if name.startswith(f"gt::k_{dominant}") and rec.get("read_bytes_per_frame") is not None:
                traffic = (rec["read_bytes_per_frame"] + rec["write_bytes_per_frame"]) * nf / max(dom_n, 1)
    pipe_bytes = algorithmic_bytes(H, W, n_cand / nf, n_rows / nf, 4, rint, in_b) * nf
    pipe_gbs = pipe_bytes / (elapsed / args.steps) / 1e9
    log(f"dominant kernel {dominant}: {achieved:.0f} GB/s over {per_launch_bytes / 1e6:.1f} MB "
        f"per launch; pipeline {pipe_gbs:.0f} GB/s of algorithmic bytes "
        f"({100 * pipe_gbs / HBM_PEAK_GBS:.1f} % of 8 TB/s)")

    plain = None
    if not args.no_plain:
        run(plain_flags)
        sync()
        t0 = time.perf_counter()
        tp = run(plain_flags)
        sync()
        pe = time.perf_counter() - t0
        plain = {"frames_per_s": nf / pe, "ms_per_frame": 1000.0 * pe / nf,
                 "launches": tp["launches"]}
        log(f"plain per-frame path: {nf / pe:.0f} frames/s ({1000 * pe / nf:.3f} ms/frame); "
            f"batched / plain = {fps / (nf / pe):.2f}x")

    uneven = None
    if not args.no_uneven:
        uneven = uneven_batch(ctx, args.steps)
        log(f"uneven batch (4096 frames): {uneven['frames_per_s']:.0f} frames/s batched, "
            f"{uneven['plain_frames_per_s']:.0f} plain, {uneven['batched_over_plain']:.2f}x")

    # the adjoint (256 x 512^2) and locate + link (2000 x 1024^2) workloads,
    # each with its own context and movie, summarised on the same line
    adjoint = link = None
    if not args.no_extra:
        a = bench_adjoint(args, summary=True)
        adjoint = {k: a[k] for k in ("value", "unit", "ms_per_step", "with_jacobian_frames_per_s",
                                     "forward_only_frames_per_s")}
        adjoint["workload"] = a["config"]["workload"]
        lk = bench_link(args, summary=True)
        link = {k: lk[k] for k in ("value", "unit", "ms_per_step", "locate_only_frames_per_s",
                                   "link_rust_only_frames_per_s", "cpu_baseline")}
        link["workload"] = lk["config"]["workload"]

    cpu = None
    c1 = None
    if not args.no_cpu and args.cpu_sample > 0:
        hc = host_cpu()
        # the CPU share of one GPU on the box: 16 threads, or fewer CPUs
        threads = max(1, min(hc["cpus_available"], 16))
        c1 = c1_latency(ctx, threads)
        log(f"C1 single frame: GPU plain {c1['gpu_plain_device_ms']:.3f} ms (frame in HBM), "
            f"{c1['gpu_plain_host_ms']:.3f} ms (numpy in, rows out); CPU oracle "
            f"{c1['cpu_1_thread_ms']:.1f} ms on 1 thread, {c1['cpu_all_threads_ms']:.1f} ms on "
            f"{threads}")
        ns = min(args.cpu_sample, nf)
        sample = movie.to_host(np.uint16, (ns, H, W))
        okw = dict(diameter=11, separation=11, smoothing_size=11, minmass=0.0,
                   percentile=None if args.no_percentile else 64.0)
        c1fps, _, c1dt = cpu_baseline(sample, okw, 1)
        cfps, crows, cdt = cpu_baseline(sample, okw, threads)
        log(f"cpu oracle: {cfps:.2f} frames/s over {ns} frames on {threads} threads ({cdt:.1f} s); "
            f"{c1fps:.2f} frames/s on 1 thread ({c1dt:.1f} s); host {hc}")
        cpu = {"value": cfps, "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"first {ns} frames of the same movie, oracle binary32 mode built "
                         f"gcc -O3 -march=native on this host, {threads} threads (one frame "
                         f"per thread)",
               "single_thread_frames_per_s": c1fps, "host": hc}

    line = {
        "metric": "frames/s at 2048^2 uint16 (~5k spots) batched vs plain; HBM GB/s % of peak",
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded device generator), device-resident",
        "config": {"workload": f"{nf} frames {H}x{W} "
                               f"{'uint16' if args.input == 'u16' else 'float32 (uint16 values)'}, "
                               "5000 spots/frame, sigma 2, "
                               "Brownian D=0.5, bg 150 + Poisson",
                   "params": kw, "path": path, "frames": nf},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "bytes_per_launch": per_launch_bytes,
                     "ms_per_launch": dom_ms / max(dom_n, 1)},
        "pipeline": {"algorithmic_GBps": pipe_gbs, "frac": pipe_gbs / HBM_PEAK_GBS,
                     "candidates_per_frame": n_cand / nf, "rows_per_frame": n_rows / nf,
                     "device_ms_per_step": dev_ms,
                     "stage_ms": {k: v[0] for k, v in stages.items() if v[1]}},
        "frames_per_s_direct": fps_direct,
        "frames_per_s_profiled_one_stream": fps_prof,
        "frames_per_s_graph_replay": fps_graph,
        "tuning": tuning,
        "plain": plain,
        "batched_over_plain": (fps / plain["frames_per_s"]) if plain else None,
        "uneven": uneven,
        "adjoint": adjoint,
        "locate_link": link,
        "c1_latency": c1,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_uneven(args) -> int:
    """Uneven batch: 4096 frames, 40 % 256x256, 40 % 512x1024, 20 % 2048x2048
    (shuffled), one spot per 800 px^2, sigma U(1.2, 2.5), background U(50, 300);
    batched vs plain per-frame path on the same device-resident frames."""
    import ctypes
    from tracking_optimizations_amd import _lib, synth
    from tracking_optimizations_amd.gpu_tracking import make_params
    kw = dict(diameter=11, separation=11, smoothing_size=11, percentile=64, minmass=0.0)
    if args.no_percentile:
        del kw["percentile"]
    params = make_params(**kw)
    ctx = _lib.Context(0)
    L = _lib.load()
    nf = 4096 if args.frames == 1000 else args.frames
    log(f"generating {nf} uneven frames on the device")
    movie = synth.UnevenMovie(ctx, nf, seed=3)
    descs = movie.frame_descs()
    ctx.synchronize()

    def run(flags):
        out = ctypes.c_void_p()
        _lib.check(L.gt_locate_batch(ctx.handle, descs, nf, 0, ctypes.byref(params), flags,
                                     ctypes.byref(out)))
        r = _lib.Result(out)
        t = ctx.timing()
        del r
        return t

    bflags = _lib.GT_FRAMES_ON_DEVICE | _lib.GT_RESULT_ON_DEVICE
    for _ in range(args.warmup):
        run(bflags)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = run(bflags)
    ctx.synchronize()
    el = time.perf_counter() - t0
    fps = nf * args.steps / el
    run(bflags | _lib.GT_PLAIN)
    ctx.synchronize()
    t0 = time.perf_counter()
    run(bflags | _lib.GT_PLAIN)
    ctx.synchronize()
    pel = time.perf_counter() - t0
    log(f"uneven batched: {fps:.0f} frames/s ({movie.pixels * args.steps / el / 1e9:.2f} Gpx/s); "
        f"plain {nf / pel:.0f} frames/s; ratio {fps * pel / nf:.2f}x; "
        f"{st['candidates']} candidates, {st['rows']} rows")
    print(json.dumps({"metric": "frames/s, uneven batch (4096 frames: 40% 256^2, 40% 512x1024, "
                                "20% 2048^2), batched vs plain",
                      "value": fps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": 1000 * el / args.steps,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f32", "data": "synthetic (seeded device generator), device-resident",
                      "config": {"workload": "uneven batch", "frames": nf, "params": kw,
                                 "path": "batched"},
                      "pixels_per_s": movie.pixels * args.steps / el,
                      "plain_frames_per_s": nf / pel, "batched_over_plain": fps * pel / nf}),
          flush=True)
    movie.free()
    return 0


def bench_adjoint(args, summary: bool = False):
    """Adjoint workload: 256 frames of 512x512 uint16 (300 spots, sigma 1.5,
    peaks U(200, 3000), background 100 + Poisson), diameter 9, separation 9,
    smoothing 9, percentile 64: forward rows plus per-particle Poisson errors
    (and, second line of the log, with the Jacobian patches kept)."""
    import ctypes
    from tracking_optimizations_amd import _lib, synth
    from tracking_optimizations_amd.gpu_tracking import make_params
    kw = dict(diameter=9, separation=9, smoothing_size=9, percentile=64, minmass=0.0)
    params = make_params(**kw)
    ctx = _lib.Context(0)
    L = _lib.load()
    nf = 256 if args.frames == 1000 else args.frames
    spec = synth.adjoint_512(seed=4, n_frames=nf)
    movie = synth.generate(ctx, spec)
    ctx.synchronize()

    def run(flags):
        out = ctypes.c_void_p()
        _lib.check(L.gt_locate_adjoint(ctx.handle, ctypes.c_void_p(movie.ptr), 0, nf, 512, 512,
                                       ctypes.byref(params), flags, ctypes.byref(out)))
        a = _lib.Adjoint(out, nf, 512, 512)
        return a.nrows

    res = {}
    for name, flags in (("errors", _lib.GT_FRAMES_ON_DEVICE),
                        ("jacobian", _lib.GT_FRAMES_ON_DEVICE | _lib.GT_ADJ_JACOBIAN)):
        for _ in range(args.warmup):
            run(flags)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            nrows = run(flags)
        ctx.synchronize()
        el = time.perf_counter() - t0
        res[name] = (nf * args.steps / el, nrows, 1000 * el / args.steps)
        log(f"adjoint ({name}): {res[name][0]:.0f} frames/s, {nrows} rows, "
            f"{res[name][2]:.1f} ms/step")
    # forward alone on the same frames, for the cost of the adjoint
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ctypes.c_void_p()
        _lib.check(L.gt_locate_u16(ctx.handle, ctypes.c_void_p(movie.ptr), nf, 512, 512,
                                   ctypes.byref(params),
                                   _lib.GT_FRAMES_ON_DEVICE | _lib.GT_RESULT_ON_DEVICE,
                                   ctypes.byref(out)))
        _lib.Result(out)
    ctx.synchronize()
    fwd = nf * args.steps / (time.perf_counter() - t0)
    line = {"metric": "frames/s, locate + adjoint (Poisson localisation errors), "
                                "256 x 512^2 uint16",
                      "value": res["errors"][0], "unit": "frames/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": res["errors"][2], "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                      "data": "synthetic (seeded device generator), device-resident",
                      "config": {"workload": "adjoint 256 x 512x512, 300 spots/frame",
                                 "params": kw, "rows": res["errors"][1]},
                      "with_jacobian_frames_per_s": res["jacobian"][0],
                      "forward_only_frames_per_s": fwd}
    movie.free()
    if summary:
        return line
    print(json.dumps(line), flush=True)
    return 0


def bench_link(args, summary: bool = False):
    """Locate + link: 2000 frames of 1024x1024 uint16, 3000 Brownian particles
    (D = 0.5 px^2/frame, sigma 1.8, lognormal peaks with median 1200,
    background 120 + Poisson), each off in a frame with probability 0.05;
    diameter 9, separation 9, minmass set from frame 0 (a quarter of the
    median mass of its brightest rows, as many rows as spots drawn),
    search_range 5, memory 3.  A step is one locate call with linking (rows
    with the particle column back on the host).  Beside it: locate alone,
    gt_link alone on the same row table, the C oracle linker on every row
    (the CPU baseline) and the Python oracle on the first --link-cpu-frames
    frames."""
    import ctypes
    from oracle.linker import link_rows
    from tracking_optimizations_amd import _lib, synth
    from tracking_optimizations_amd.gpu_tracking import link_rust, make_params
    ctx = _lib.Context(0)
    L = _lib.load()
    nf = 2000 if args.frames == 1000 else args.frames
    H = W = 1024
    log(f"generating {nf} frames of {H}x{W} uint16 (3000 blinking particles)")
    spec = synth.link_1024(seed=5, n_frames=nf)
    movie = synth.generate(ctx, spec)
    ctx.synchronize()

    def locate_rows(params, n):
        out = ctypes.c_void_p()
        _lib.check(L.gt_locate_u16(ctx.handle, ctypes.c_void_p(movie.ptr), n, H, W,
                                   ctypes.byref(params), _lib.GT_FRAMES_ON_DEVICE,
                                   ctypes.byref(out)))
        return _lib.Result(out).rows()

    kw = dict(diameter=9, separation=9)
    r0 = locate_rows(make_params(**kw, minmass=0.0), 1)
    n0 = int(spec.spot_off[1])
    top = np.sort(r0[:, 1])[::-1][:n0]
    kw["minmass"] = float(0.25 * np.median(top))
    log(f"minmass from frame 0: {kw['minmass']:.1f}")
    p_locate = make_params(**kw)
    p_link = make_params(**kw, search_range=5.0, memory=3)

    for _ in range(args.warmup):
        locate_rows(p_link, nf)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rows = locate_rows(p_link, nf)
    el = time.perf_counter() - t0
    st = ctx.link_stats()
    fps = nf * args.steps / el
    log(f"locate + link: {fps:.0f} frames/s, {rows.shape[0]} rows, link search "
        f"{st['search_ms']:.1f} ms, solve {st['solve_ms']:.1f} ms, {st['subnets']} subnets "
        f"(largest {st['max_subnet']})")

    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain = locate_rows(p_locate, nf)
    loc_fps = nf * args.steps / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ids = link_rust(plain, 5.0, 3, ctx=ctx)
    link_el = (time.perf_counter() - t0) / args.steps
    lst = ctx.link_stats()
    log(f"locate alone {loc_fps:.0f} frames/s; link_rust alone {nf / link_el:.0f} frames/s "
        f"({link_el * 1e3:.0f} ms: search {lst['search_ms']:.1f}, solve {lst['solve_ms']:.1f})")

    cpu = None
    if not args.no_cpu:
        # the C restatement of the linker (oracle/gt_linker.c), one thread, on
        # every row of the movie: the CPU baseline of the link step, and a
        # full-size check of the GPU linker's ids
        from oracle.oracle import link_rows_c
        t0 = time.perf_counter()
        ref = link_rows_c(plain, 5.0, 3)
        dt = time.perf_counter() - t0
        same = bool(np.array_equal(ref, ids))
        cpu = {"value": nf / dt, "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"C oracle linker (oracle/gt_linker.c, one thread) on all {nf} frames' "
                         f"rows ({len(plain)} rows), link only; ids identical to the GPU "
                         f"linker: {same}"}
        log(f"CPU C oracle linker: {nf / dt:.0f} frames/s ({dt:.2f} s); ids identical: {same}")
        k = min(args.link_cpu_frames, nf)
        if not summary and k > 0:
            sub = plain[plain[:, 0] < k]
            t0 = time.perf_counter()
            pref = link_rows(sub, 5.0, 3)
            pdt = time.perf_counter() - t0
            log(f"Python oracle linker: {k / pdt:.1f} frames/s on {k} frames; ids identical: "
                f"{bool(np.array_equal(pref, ids[: len(pref)]))}")
    line = {"metric": "frames/s, locate + link (search_range 5, memory 3), "
                                "2000 x 1024^2 uint16, 3000 particles",
                      "value": fps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": 1000 * el / args.steps,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f32",
                      "data": "synthetic (seeded device generator), device-resident frames, "
                              "rows returned to the host",
                      "config": {"workload": f"{nf} x {H}x{W} uint16, 3000 Brownian particles, "
                                             "blink-off p 0.05", "params": kw,
                                 "search_range": 5.0, "memory": 3, "rows": int(rows.shape[0])},
                      "link": {k2: st[k2] for k2 in st},
                      "locate_only_frames_per_s": loc_fps,
                      "link_rust_only_frames_per_s": nf / link_el,
                      "cpu_baseline": cpu}
    movie.free()
    if summary:
        return line
    print(json.dumps(line), flush=True)
    return 0


def bench_raw(args) -> int:
    """Raw signal columns (raw_mass, raw_bg_median, raw_mass_corrected) on
    200 frames of the headline movie (2048x2048 uint16, d = 11, separation 11,
    smoothing 11, percentile 64): signal disk radius 3, background annulus
    radius 8 outside a gap of 1, frames and rows resident in HBM; beside it
    the same call without the background and without the raw columns."""
    import ctypes
    from tracking_optimizations_amd import _lib, synth
    from tracking_optimizations_amd.gpu_tracking import make_params
    ctx = _lib.Context(0)
    L = _lib.load()
    nf = 200 if args.frames == 1000 else args.frames
    movie = synth.generate(ctx, synth.movie_2048(seed=2, n_frames=nf))
    ctx.synchronize()
    base = dict(separation=11, smoothing_size=11, percentile=64, minmass=0.0)
    variants = [("locate", {}), ("signal", dict(sig_radius=3.0)),
                ("signal + background", dict(sig_radius=3.0, bg_radius=8.0, gap_radius=1.0))]
    flags = _lib.GT_FRAMES_ON_DEVICE | _lib.GT_RESULT_ON_DEVICE
    res = {}
    for name, kw in variants:
        p = make_params(11, **base, **kw)

        def run():
            out = ctypes.c_void_p()
            _lib.check(L.gt_locate_u16(ctx.handle, ctypes.c_void_p(movie.ptr), nf, 2048, 2048,
                                       ctypes.byref(p), flags, ctypes.byref(out)))
            r = _lib.Result(out)
            n = r.nrows if hasattr(r, "nrows") else None
            del r
            return n
        for _ in range(max(1, args.warmup)):
            run()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run()
        ctx.synchronize()
        el = (time.perf_counter() - t0) / args.steps
        res[name] = nf / el
        log(f"raw workload, {name}: {nf / el:.0f} frames/s ({el * 1e3:.1f} ms per {nf} frames)")
    print(json.dumps({"metric": "frames/s, locate + raw signal columns (sig 3, bg 8, gap 1), "
                                "2048^2 uint16",
                      "value": res["signal + background"], "unit": "frames/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": 1e3 * nf / res["signal + background"],
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f32", "data": "synthetic (seeded device generator), device-resident",
                      "config": {"workload": f"{nf} x 2048x2048 uint16, raw signal", "params": base},
                      "locate_only_frames_per_s": res["locate"],
                      "signal_only_frames_per_s": res["signal"]}), flush=True)
    movie.free()
    return 0


def _write_tiff_u16(path: str, frames: np.ndarray) -> None:
    """Uncompressed little-endian TIFF, one strip per page."""
    import struct
    n, h, w = frames.shape
    ifd = 2 + 8 * 12 + 4
    data0 = 8 + n * ifd
    page = h * w * 2
    with open(path, "wb") as f:
        f.write(b"II" + struct.pack("<HI", 42, 8))
        for i in range(n):
            nxt = 8 + (i + 1) * ifd if i + 1 < n else 0
            tags = [(256, 4, w), (257, 4, h), (258, 3, 16), (259, 3, 1), (262, 3, 1),
                    (273, 4, data0 + i * page), (277, 3, 1), (279, 4, page)]
            f.write(struct.pack("<H", len(tags)))
            for tag, typ, val in tags:
                f.write(struct.pack("<HHI", tag, typ, 1) +
                        struct.pack("<" + ("H" if typ == 3 else "I"), val).ljust(4, b"\0"))
            f.write(struct.pack("<I", nxt))
        for i in range(n):
            f.write(np.ascontiguousarray(frames[i], "<u2").tobytes())


def bench_file(args) -> int:
    """batch_file end to end on a 256-frame 2048x2048 uint16 TIFF (the
    headline movie's first frames, uncompressed, in the page cache after the
    warm-up): native decode into pinned buffers, upload and locate, rows
    back on the host.  Beside it: batch on the same frames from a pageable
    host array (staged through the same pinned buffers), and the reader alone."""
    import tempfile
    from tracking_optimizations_amd import _lib, synth
    from tracking_optimizations_amd.gpu_tracking import (MovieFile, batch_file_rust, locate,
                                                         make_params)
    ctx = _lib.Context(0)
    nf = 256 if args.frames == 1000 else args.frames
    spec = synth.movie_2048(seed=2, n_frames=nf)
    video = synth.generate_host(ctx, spec)
    kw = dict(separation=11, smoothing_size=11, percentile=64, minmass=0.0)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "movie.tif")
        _write_tiff_u16(path, video)
        log(f"wrote {os.path.getsize(path) / 1e9:.2f} GB TIFF")
        for _ in range(max(1, args.warmup)):
            rows, _ = batch_file_rust(path, 11, ctx=ctx, **kw)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rows, _ = batch_file_rust(path, 11, ctx=ctx, **kw)
        el = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        with MovieFile(path) as m:
            buf = m.read()
        rd = time.perf_counter() - t0
        del buf
    params = make_params(11, **kw)
    ref = locate(video, params, ctx=ctx)[0]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ref = locate(video, params, ctx=ctx)[0]
    hel = (time.perf_counter() - t0) / args.steps
    same = bool(np.array_equal(rows, ref))
    log(f"batch_file: {nf / el:.0f} frames/s ({el * 1e3:.0f} ms); batch from a host array: "
        f"{nf / hel:.0f} frames/s; reader alone {nf / rd:.0f} frames/s "
        f"({video.nbytes / rd / 1e9:.1f} GB/s); rows identical: {same}")
    print(json.dumps({"metric": "frames/s, batch_file (TIFF decode + upload + locate), "
                                "256 x 2048^2 uint16",
                      "value": nf / el, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": el * 1e3, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                      "data": "synthetic movie written as an uncompressed TIFF (page cache)",
                      "config": {"workload": f"{nf} x 2048x2048 uint16 TIFF", "params": kw,
                                 "rows": int(rows.shape[0])},
                      "host_array_frames_per_s": nf / hel, "reader_frames_per_s": nf / rd,
                      "rows_identical_to_batch": same}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
