"""Benchmark of the stereo hot path on one MI355X.

Workload (the headline metric's configuration): one 1280x720 uint8 pair from
the seeded generator (12 rectangles, disparities in [0, 127]), Birchfield-
Tomasi cost, D = 128, 8-path SGM (P1 = 8, P2 = 32), lowest-index WTA. A step
is one pass of the path over that pair with the images already in HBM.
Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--mode fused|staged]
  python bench.py --profile-run          a few plain steps only (for rocprofv3)
  python bench.py --profile-run --shape 640x480x64   the same at another pair shape
  python bench.py --window-only          the 1920x1080 D=256 window-cost / DP / WTA paths
  python bench.py --pairings-only        every cost x optimiser chain on the 720p pair
  python bench.py --copy-calibration N   a plain N-byte device copy (PMC scale)
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def b_alg(W, H, D):
    """Algorithmic bytes per pair of the staged uint16 model: C write + C read
    + S write + S read per cell (8 B) and L, R reads + disparity write per
    pixel (4 B)."""
    return 8 * W * H * D + 4 * W * H


def measured_traffic(kernel):
    """HBM bytes per dispatch of the kernel the library profiles as `kernel`,
    from the committed PMC summary (FETCH_SIZE / WRITE_SIZE scaled by the
    copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_sgm_fused_720p_d128.txt")
    try:
        rows, scale = {}, {}
        for line in open(path):
            if line.startswith("#") or not line.strip():
                continue
            parts = line.split()
            if parts[0].startswith("scale_"):
                scale[parts[0]] = float(parts[1])
            else:
                rows[parts[0]] = (float(parts[1]), float(parts[2]))
        f, w = rows[kernel]
        return int((f * scale["scale_fetch"] + w * scale["scale_write"]) * 1024)
    except (OSError, KeyError, ValueError, IndexError):
        return None


def step_ms_hip_events(step, stream, torch, steps):
    """Mean milliseconds of one step between two hipEvents recorded on the
    stream the kernels are launched on, over `steps` back-to-back steps."""
    step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


MIX_WORKLOAD = ("64 pairs, sizes {320x240,640x480,1280x720,1920x1080}, D in {32,64,128,256} "
                "clamped to W/4, seeds 1000+i; BT+SGM8+WTA")
HD_WORKLOAD = "16 pairs 1280x720, D=128, 12 rectangles, seeds 3000+i; BT+SGM8+WTA"


def hd_specs(n=16):
    return [dict(width=1280, height=720, dmax=127, nrect=12, seed=3000 + i, depth=128)
            for i in range(n)]


def bench_batched(tv, synth, torch, ctx, reps, specs=None, workload=MIX_WORKLOAD):
    """A batch of pairs (default: 64 mixed pairs, 320x240..1920x1080, D in
    {32,64,128,256} clamped to W/4, seeds 1000+i) replayed from a hipGraph,
    against the plain per-pair call over the same pairs."""
    from thundervision_amd import autotune
    specs = specs if specs is not None else synth.batch_specs(64)
    dev = torch.device("cuda:0")
    pairs = []
    for s in specs:
        L, R, _ = synth.pair(s["width"], s["height"], s["dmax"], s["nrect"], s["seed"])
        pairs.append((torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev), s["depth"],
                      torch.empty((s["height"], s["width"]), dtype=torch.int16, device=dev)))
    cells = sum(s["width"] * s["height"] * s["depth"] for s in specs)
    b = tv.Batch(pairs, ctx=ctx)
    tuned = autotune.tune(b, [(s["width"], s["height"], s["depth"]) for s in specs], reps=2)
    ms_graph = autotune.time_batch(b, reps, use_graph=True)
    # plain per-pair calls over the same pairs
    def plain_round():
        for L, R, D, disp in pairs:
            tv.stereo_bt_sgm(L, R, D, out=disp, ctx=ctx)
    plain_round()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_round()
    torch.cuda.synchronize()
    ms_plain = (time.perf_counter() - t0) / reps * 1e3
    # bit-exact check of the last batch run against the plain outputs
    b.run(use_graph=True)
    b.wait()
    outs = [d.clone() for _, _, _, d in pairs]
    plain_round()
    torch.cuda.synchronize()
    exact = all(torch.equal(o, d) for o, (_, _, _, d) in zip(outs, pairs))
    return {"workload": workload,
            "cells": cells, "ms_per_batch": round(ms_graph, 3),
            "value": round(cells / (ms_graph * 1e-3) / 1e6, 1), "unit": "MDE/s",
            "plain_per_pair_ms": round(ms_plain, 3),
            "plain_per_pair_value": round(cells / (ms_plain * 1e-3) / 1e6, 1),
            "speedup_vs_plain": round(ms_plain / ms_graph, 3),
            "tuned": {k: tuned[k] for k in ("fusion_depth", "lines_per_wave", "waves_per_group",
                                              "prefetch", "sum_lag", "cached")},
            "tune_trials_ms": [dict(t, ms=round(t["ms"], 3)) for t in tuned["trials"]],
            "bit_exact_vs_plain": exact}


SERVED_POLICIES = {
    # name: (slots, Scheduler keyword arguments; cell budgets in largest frames)
    "plain_per_pair": (1, {}),
    # frames above a tenth of the largest start only on an idle device
    # (10 s at 600 pairs/s: p99 5.98 ms against 6.12 ms for one slot, where
    # letting every frame join gave 6.38 ms)
    "served": (4, {"cell_budget": 1.3, "backfill_s": 0.003, "join_cells": 0.1}),
}


def served_frames(torch, dev, duration, rate, distinct):
    """The served workload's arrivals (Poisson, seeds 2000+stream) over
    `duration` seconds and the distinct frames they use: frame k of stream s
    is the generator's frame k % distinct of that stream (rectangles drifting
    1-3 px per frame), generated on a thread pool and kept on the device."""
    from concurrent.futures import ThreadPoolExecutor

    from thundervision_amd import serve, synth
    shapes = serve.SERVED_STREAMS
    keys = [(s, k) for s in range(len(shapes)) for k in range(distinct)]

    def gen(key):
        s, k = key
        W, H, D = shapes[s]
        L, R, _ = synth.pair(W, H, D - 1, 12, 2000 + s, frame=k)
        return key, L, R

    pool = {}
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        for key, L, R in ex.map(gen, keys):
            pool[key] = (torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev))
    return serve.poisson_schedule(len(shapes), rate, duration), pool


def bench_served(torch, tv, duration=10.0, rate=600.0, distinct=24, policies=None):
    """16 streams (4 each at 640x480/D64, 1280x720/D128, 1280x720/D192,
    1920x1080/D256), Poisson arrivals at `rate` pairs/s in total, each stream
    cycling through `distinct` generator frames. Every policy serves the same
    arrivals: one slot (the plain per-pair call back to back) and continuous
    batching over concurrent slots under a 64 GB budget. Every served
    disparity map is checked against the same frame computed alone; p50/p99
    come from device completion events."""
    from thundervision_amd import serve
    dev = torch.device("cuda:0")
    shapes = serve.SERVED_STREAMS
    sched, pool = served_frames(torch, dev, duration, rate, distinct)
    alone = {key: tv.stereo_bt_sgm(L, R, shapes[key[0]][2]) for key, (L, R) in pool.items()}
    torch.cuda.synchronize()
    big = max(W * H * D for W, H, D in shapes)
    out = {"workload": "16 streams: 4x 640x480/D64, 4x 1280x720/D128, 4x 1280x720/D192, "
                       "4x 1920x1080/D256; Poisson %.0f pairs/s for %.1f s; %d distinct "
                       "drifting frames per stream; BT+SGM8+WTA" % (rate, duration, distinct),
           "frames": len(sched), "all_frames_exact": True}
    for name, (nslots, kw) in (policies or SERVED_POLICIES).items():
        frames = []
        for t, s, k in sched:
            W, H, D = shapes[s]
            L, R = pool[(s, k % distinct)]
            disp = torch.full((H, W), -1, dtype=torch.int16, device=dev)
            frames.append(serve.Frame(t, s, k, W, H, D, (L, R, disp)))
        ex = serve.GpuExecutor(nslots)
        kw = dict(kw)
        for key in ("cell_budget", "join_cells"):
            if key in kw:
                kw[key] = kw[key] * big
        sch = serve.Scheduler(ex, nslots, 64 << 30, **kw)
        # every slot runs every frame shape once before the clock starts
        ex.warm(frames)
        ex.start()
        done = sch.run(frames)
        torch.cuda.synchronize()
        exact = len(done) == len(frames) and all(
            torch.equal(f.payload[2], alone[(f.stream, f.index % distinct)]) for f in frames)
        out["all_frames_exact"] = out["all_frames_exact"] and exact
        st = serve.latency_stats(done)
        row = {k: round(v, 3) for k, v in st.items()}
        row.update(slots=nslots, exact=exact, max_concurrent=sch.max_concurrent,
                   peak_inflight_GB=round(sch.peak_in_use / 2**30, 2),
                   policy={k: (round(v, 4) if isinstance(v, float) else v) for k, v in kw.items()},
                   by_shape={k: {kk: round(vv, 3) for kk, vv in v.items()}
                             for k, v in serve.latency_by_shape(done).items()})
        out[name] = row
        del frames, done
    out["offered_MDE/s"] = round(sum(W * H * D for t, s, k in sched
                                     for (W, H, D) in [shapes[s]]) / duration / 1e6, 1)
    return out


def bench_window(tv, synth, torch, ctx, reps=5):
    """The window-cost paths at 1920x1080, D = 256 (seed 3): SSD 7x7 (uint32)
    -> DP (P_occ = 20), 9x9 zero-mean NCC (fp32) -> WTA, and the fused SSD +
    WTA without a volume. B_alg: 9 B/cell for SSD -> DP (cost write + read
    at uint32, 1 B path code), 8 B/cell for NCC -> WTA."""
    spec = synth.WINDOW_1080P
    W, H, D = spec["width"], spec["height"], spec["depth"]
    L, R, _ = synth.workload_pair(spec)
    dev = torch.device("cuda:0")
    Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
    cells = W * H * D
    # SSD 7x7 costs (<= 49 * 255^2 < 2^22, invalid exactly where d > x) meet
    # the DP's bounded-cost contract (32-bit sums); the general 64-bit form
    # is timed beside it
    paths = {
        "ssd_dp": (lambda: tv.dp(tv.cost_ssd(Ld, Rd, D, radius=3, ctx=ctx), p_occ=20,
                                 flags=tv.DP_COST_BOUNDED, ctx=ctx), 9),
        "ssd_dp_64bit_sums": (lambda: tv.dp(tv.cost_ssd(Ld, Rd, D, radius=3, ctx=ctx), p_occ=20,
                                            ctx=ctx), 9),
        "ncc_wta": (lambda: tv.wta(tv.cost_ncc(Ld, Rd, D, radius=4, zero_mean=True, ctx=ctx),
                                   ctx=ctx), 8),
        "fused_ssd_wta": (lambda: tv.fast_wta_ssd(Ld, Rd, D, radius=3, ctx=ctx), None),
        # the SSD -> DP path without the 2.12 GB volume (each DP wave computes
        # its row's costs); same disparities as ssd_dp
        "fused_ssd_dp": (lambda: tv.stereo_ssd_dp(Ld, Rd, D, radius=3, p_occ=20, ctx=ctx), None),
        "fused_ncc_wta": (lambda: tv.fast_wta_ncc(Ld, Rd, D, radius=4, zero_mean=True, ctx=ctx),
                          None),
    }
    out = {"workload": "1920x1080 uint8 pair, D=256, seed 3"}
    for name, (fn, bpc) in paths.items():
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        ctx.profile(True)
        ctx.profile_reset()
        fn()
        torch.cuda.synchronize()
        table = ctx.profile_table()
        ctx.profile(False)
        row = {"ms": round(ms, 3), "MDE/s": round(cells / (ms * 1e-3) / 1e6, 1),
               "kernels_ms": {k: round(v[0] / v[1], 4) for k, v in table.items()}}
        if bpc:
            row["frac_of_hbm_roofline"] = round(cells * bpc / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        out[name] = row
    return out


def bench_pairings(tv, torch, ctx, Ld, Rd, D, reps=5):
    """Every cost x optimiser chain of the reference's factory on the 720p
    pair (the fused BT + SGM + WTA is the headline): the cost volume and the
    optimiser (disparities only), with the per-kernel hipEvent times of one
    call."""
    W, H = Ld.shape[1], Ld.shape[0]
    cells = W * H * D
    fp1, fp2 = 30.0 / 255, 150.0 / 255
    chains = {
        "bt_wta": lambda: tv.wta(tv.cost_bt(Ld, Rd, D, ctx=ctx), ctx=ctx),
        "bt_dp": lambda: tv.dp(tv.cost_bt(Ld, Rd, D, ctx=ctx), p_occ=20, ctx=ctx),
        "bt_sgm_volume": lambda: tv.sgm(tv.cost_bt(Ld, Rd, D, ctx=ctx), ctx=ctx),
        "ssd_wta": lambda: tv.wta(tv.cost_ssd(Ld, Rd, D, radius=3, ctx=ctx), ctx=ctx),
        "ssd_dp": lambda: tv.dp(tv.cost_ssd(Ld, Rd, D, radius=3, ctx=ctx), p_occ=20,
                                flags=tv.DP_COST_BOUNDED, ctx=ctx),
        "ssd_sgm": lambda: tv.sgm(tv.cost_ssd(Ld, Rd, D, radius=3, ctx=ctx), ctx=ctx),
        "ncc_wta": lambda: tv.wta(tv.cost_ncc(Ld, Rd, D, radius=4, zero_mean=True, ctx=ctx),
                                  ctx=ctx),
        "ncc_dp": lambda: tv.dp(tv.cost_ncc(Ld, Rd, D, radius=4, zero_mean=True, ctx=ctx),
                                p_occ=0.1, ctx=ctx),
        "ncc_sgm": lambda: tv.sgm(tv.cost_ncc(Ld, Rd, D, radius=4, zero_mean=True, ctx=ctx),
                                  fp1=fp1, fp2=fp2, ctx=ctx),
    }
    out = {"workload": "%dx%d uint8 pair, D=%d (the headline pair); cost volume + optimiser, "
                       "disparities only" % (W, H, D)}
    for name, fn in chains.items():
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        ctx.profile(True)
        ctx.profile_reset()
        fn()
        torch.cuda.synchronize()
        table = ctx.profile_table()
        ctx.profile(False)
        out[name] = {"ms": round(ms, 3), "MDE/s": round(cells / (ms * 1e-3) / 1e6, 1),
                     "kernels_ms": {k: round(v[0] / v[1], 4) for k, v in table.items()}}
    return out


def cpu_threads(L, R, D, want, threads=8, reps=2):
    """The oracle's BT + SGM-8 + WTA on `threads` host threads (each oracle
    call releases the interpreter lock): the cost volume and the plane sum +
    WTA by bands of rows (both are per-row), the SGM as its 8 direction scans
    at once. Checked equal to the single-thread pipeline's disparities."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    import oracle

    H = L.shape[0]
    bands = [(H * i // threads, H * (i + 1) // threads) for i in range(threads)]

    def once(ex):
        C = np.empty((H, L.shape[1], D), np.uint16)

        def bt(b):
            C[b[0]:b[1]] = oracle.cost_bt2(L[b[0]:b[1]], R[b[0]:b[1]], D)
        list(ex.map(bt, bands))
        lrs = list(ex.map(lambda d: oracle.sgm_lr(C, 8, 32, d), range(8)))
        out = np.empty(L.shape, np.int16)

        def sum_wta(b):
            c = C[b[0]:b[1]]
            S = np.zeros(c.shape, np.uint32)
            for lr in lrs:
                S += lr[b[0]:b[1]]
            S = np.where(c == oracle.INV16, np.uint32(oracle.INV16), S).astype(np.uint16)
            out[b[0]:b[1]] = oracle.wta(S)
        list(ex.map(sum_wta, bands))
        return out

    with ThreadPoolExecutor(threads) as ex:
        got = once(ex)
        t0 = time.perf_counter()
        for _ in range(reps):
            once(ex)
        sec = (time.perf_counter() - t0) / reps
    return {"value": round(L.size * D / sec / 1e6, 3), "unit": "MDE/s", "cores": threads,
            "kind": "port", "bit_exact": bool((got == want).all()),
            "sample": "%d full 1280x720 D=128 pairs on %d host threads (BT and plane sum + WTA "
                      "by row bands, the 8 SGM direction scans at once), oracle/tdvref.c, %s"
                      % (reps, threads, cpu_model())}


def bench_plain_384(tv, synth, torch, ctx, with_cpu, reps=50):
    """The 384x288, D = 16 pair (seed 1): the fused GPU call (images
    resident), and with_cpu: the oracle's BT, SGM and WTA timed separately on
    one host thread."""
    spec = synth.PLAIN_384
    W, H, D = spec["width"], spec["height"], spec["depth"]
    L, R, _ = synth.workload_pair(spec)
    dev = torch.device("cuda:0")
    Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
    out = torch.empty((H, W), dtype=torch.int16, device=dev)
    fn = lambda: tv.stereo_bt_sgm(Ld, Rd, D, out=out, ctx=ctx)
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    cells = W * H * D
    row = {"workload": "%dx%d uint8 pair, D=%d, seed 1; BT+SGM8+WTA" % (W, H, D),
           "gpu_ms": round(ms, 4), "gpu_MDE/s": round(cells / (ms * 1e-3) / 1e6, 1)}
    if with_cpu:
        import oracle
        t0 = time.perf_counter()
        C = oracle.cost_bt2(L, R, D)
        t1 = time.perf_counter()
        S = oracle.sgm_std(C, 8, 32)
        t2 = time.perf_counter()
        d = oracle.wta(S)
        t3 = time.perf_counter()
        row["cpu_ms"] = {"bt": round((t1 - t0) * 1e3, 2), "sgm": round((t2 - t1) * 1e3, 2),
                         "wta": round((t3 - t2) * 1e3, 2)}
        row["cpu_MDE/s"] = round(cells / (t3 - t0) / 1e6, 2)
        row["cpu"] = "oracle/tdvref.c gcc -O3 -march=x86-64-v3, 1 thread (%s, nproc %d)" % (cpu_model(),
                                                                           os.cpu_count() or 0)
        row["bit_exact"] = bool((out.cpu().numpy() == d).all())
    return row


def bench_boundary(tv, torch, ctx, L, R, D, reps=20):
    """Steps around the path at the headline shape (device-resident unless
    stated): Gaussian prefilter, erode/dilate postfilter, colour to gray,
    rectification remap, reprojection (per-kernel hipEvent times and HBM
    rates), the graph pipeline vs the same stages called one by one, and the
    PCIe-inclusive rate from pinned host images to pinned host disparities,
    on one stream and alternating over two."""
    import numpy as np
    dev = torch.device("cuda:0")
    H, W = L.shape
    Ld, Rd = torch.from_numpy(L).to(dev), torch.from_numpy(R).to(dev)
    rgb = torch.from_numpy(np.stack([L, R, L], -1)).to(dev)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    mx = torch.from_numpy(0.995 * xx + 0.01 * yy + 2.0).to(dev)
    my = torch.from_numpy(-0.01 * xx + 1.005 * yy - 1.0).to(dev)
    disp = tv.stereo_bt_sgm(Ld, Rd, D, ctx=ctx)
    Q = np.array([[1, 0, 0, -W / 2], [0, 1, 0, -H / 2], [0, 0, 0, 700.0], [0, 0, 8.0, 0]],
                 np.float32)
    px = W * H
    steps = {
        "gauss5_kernel": (lambda: tv.prefilter_gauss5(Ld, ctx=ctx), 2 * px),
        "morph_kernel": (lambda: tv.postfilter_open(disp, 3, ctx=ctx), 8 * px),
        "gray_kernel": (lambda: tv.rgb_to_gray(rgb, ctx=ctx), 4 * px),
        "remap_kernel": (lambda: tv.remap(Ld, mx, my, ctx=ctx), 10 * px),
        "reproject_kernel": (lambda: tv.reproject(disp, Q, 1.0 / D, 24.0, ctx=ctx), 14 * px),
    }
    out = {"workload": "%dx%d uint8 pair, D=%d" % (W, H, D), "kernels": {}}
    for name, (fn, nbytes) in steps.items():
        fn()
        torch.cuda.synchronize()
        ctx.profile(True)
        ctx.profile_reset()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        t = ctx.profile_table()
        ctx.profile(False)
        ms = sum(v[0] for k, v in t.items() if k.startswith(name)) / reps
        out["kernels"][name] = {"ms": round(ms, 4), "bytes": nbytes,
                                "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1)}

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    pipe = tv.Pipeline(W, H, D, flags=tv.PIPE_PREFILTER | tv.PIPE_POSTFILTER)
    res = torch.empty((H, W), dtype=torch.int16, device=dev)
    ms_graph = timed(lambda: pipe.run(Ld, Rd, out=res), reps)
    lf = torch.empty_like(Ld)
    rf = torch.empty_like(Rd)

    def eager():
        tv.prefilter_gauss5(Ld, out=lf, ctx=ctx)
        tv.prefilter_gauss5(Rd, out=rf, ctx=ctx)
        tv.stereo_bt_sgm(lf, rf, D, out=res, ctx=ctx)
        tv.postfilter_open(res, 3, out=res, ctx=ctx)

    ms_eager = timed(eager, reps)
    # PCIe-inclusive: pinned host images in, pinned host disparities out
    hl, hr = tv.PinnedBuffer((H, W), np.uint8), tv.PinnedBuffer((H, W), np.uint8)
    hd = tv.PinnedBuffer((H, W), np.int16)
    hl.array[:], hr.array[:] = L, R
    Lb, Rb = torch.empty_like(Ld), torch.empty_like(Rd)

    def host_frame():
        tv.copy_2d(Lb.data_ptr(), W, hl.array.ctypes.data, W, W, H)
        tv.copy_2d(Rb.data_ptr(), W, hr.array.ctypes.data, W, W, H)
        pipe.run(Lb, Rb, out=res)
        tv.copy_2d(hd.array.ctypes.data, 2 * W, res.data_ptr(), 2 * W, 2 * W, H)

    ms_host = timed(host_frame, reps)
    exact = bool((hd.array == res.cpu().numpy()).all())

    # the same with two streams (a context, pipeline, device buffers and
    # pinned output each): frame i runs on stream i % 2, so one frame's
    # copies overlap the other stream's frame
    lanes = []
    for _ in range(2):
        c2 = tv.Context(0)
        lanes.append(dict(ctx=c2, stream=torch.cuda.Stream(),
                          pipe=tv.Pipeline(W, H, D, flags=tv.PIPE_PREFILTER | tv.PIPE_POSTFILTER,
                                           ctx=c2),
                          L=torch.empty_like(Ld), R=torch.empty_like(Rd),
                          res=torch.empty((H, W), dtype=torch.int16, device=dev),
                          hd=tv.PinnedBuffer((H, W), np.int16)))
    count = [0]

    def two_stream_frame():
        ln = lanes[count[0] % 2]
        count[0] += 1
        c2, st = ln["ctx"], ln["stream"]
        tv.copy_2d(ln["L"].data_ptr(), W, hl.array.ctypes.data, W, W, H, ctx=c2, stream=st)
        tv.copy_2d(ln["R"].data_ptr(), W, hr.array.ctypes.data, W, W, H, ctx=c2, stream=st)
        ln["pipe"].run(ln["L"], ln["R"], out=ln["res"], stream=st)
        tv.copy_2d(ln["hd"].array.ctypes.data, 2 * W, ln["res"].data_ptr(), 2 * W, 2 * W, H,
                   ctx=c2, stream=st)

    two_stream_frame()  # each lane's first run captures its graph before the clock
    ms_two = timed(two_stream_frame, 2 * reps)
    exact = exact and all(bool((ln["hd"].array == hd.array).all()) for ln in lanes)
    for ln in lanes:
        ln["pipe"].close()
        ln["hd"].free()
        ln["ctx"].close()
    for b in (hl, hr, hd):
        b.free()
    pipe.close()
    cells = W * H * D
    out["pipeline"] = {"stages": "gauss5 x2 -> BT+SGM8+WTA -> erode/dilate x3",
                       "graph_ms_per_frame": round(ms_graph, 4),
                       "eager_ms_per_frame": round(ms_eager, 4),
                       "graph_MDE/s": round(cells / (ms_graph * 1e-3) / 1e6, 1),
                       "pcie_inclusive_ms_per_frame": round(ms_host, 4),
                       "pcie_inclusive_MDE/s": round(cells / (ms_host * 1e-3) / 1e6, 1),
                       "pcie_inclusive_two_streams_ms_per_frame": round(ms_two, 4),
                       "pcie_inclusive_two_streams_MDE/s": round(cells / (ms_two * 1e-3) / 1e6, 1),
                       "host_result_exact": exact}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 200 steps (0.1 s of timed work at 720p): over 50 steps after 5 warm-up
    # steps the host clock read ~4 % above the steady per-step time
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mode", choices=["fused", "staged"], default="fused")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--profile-run", action="store_true")
    ap.add_argument("--window-only", action="store_true",
                    help="run only the 1920x1080 D=256 window-cost / DP / WTA paths (for profilers)")
    ap.add_argument("--pairings-only", action="store_true",
                    help="run only the cost x optimiser chains on the 720p pair (for profilers)")
    ap.add_argument("--copy-calibration", type=int, default=0)
    ap.add_argument("--shape", default="",
                    help="WxHxD of the pair for --profile-run (default: the 1280x720x128 workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every pair and stream is independent, so N GPUs "
              "would each run their own share; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import thundervision_amd.tdvh as tv
    from thundervision_amd import synth

    if args.copy_calibration:
        n = args.copy_calibration
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "launches": 3}))
        return

    spec = synth.PLAIN_720P
    if args.shape:
        if not args.profile_run:
            print("--shape applies to --profile-run only", file=sys.stderr)
            sys.exit(2)
        w, h, d = (int(v) for v in args.shape.lower().split("x"))
        spec = dict(width=w, height=h, dmax=d - 1, nrect=12, depth=d, seed=2)
    W, H, D = spec["width"], spec["height"], spec["depth"]
    L, R, _ = synth.workload_pair(spec)
    dev = torch.device("cuda:0")
    Ld = torch.from_numpy(L).to(dev)
    Rd = torch.from_numpy(R).to(dev)
    disp = torch.empty((H, W), dtype=torch.int16, device=dev)
    ctx = tv.Context(0)
    stream = torch.cuda.current_stream()
    fusion = tv.FUSE_COST if args.mode == "fused" else tv.FUSE_STAGED

    def step(fz=fusion):
        tv.stereo_bt_sgm(Ld, Rd, D, fusion=fz, out=disp, ctx=ctx, stream=stream)

    def timed(fz, steps, warmup):
        for _ in range(warmup):
            step(fz)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(fz)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    cells = W * H * D
    if args.window_only:
        print(json.dumps({"window_1080p": bench_window(tv, synth, torch, ctx)}))
        return
    if args.pairings_only:
        print(json.dumps({"pairings_720p": bench_pairings(tv, torch, ctx, Ld, Rd, D)}))
        return
    if args.profile_run:
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        print(json.dumps({"profile_run": args.mode, "steps": args.steps}))
        return
    sec = timed(fusion, args.steps, args.warmup)
    value = cells / sec / 1e6
    other = tv.FUSE_STAGED if fusion == tv.FUSE_COST else tv.FUSE_COST
    sec_other = timed(other, max(5, args.steps // 5), 2)

    # the step on the launch stream's hipEvents, and per-kernel hipEvent
    # timing (the library's profiler, same stream)
    step_ms = step_ms_hip_events(step, stream, torch, args.steps)
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(max(5, args.steps // 5)):
        step()
    torch.cuda.synchronize()
    table = ctx.profile_table()
    ctx.profile(False)
    dom = max(table, key=lambda k: table[k][0])
    dom_ms = table[dom][0] / table[dom][1]
    # roofline of the whole step (three kernels: BT term planes, line scans,
    # plane sum + WTA): the algorithmic bytes of the staged uint16 model over
    # the step's hipEvent time; the measured HBM bytes of the three kernels
    # (PMC) beside them
    achieved = b_alg(W, H, D) / (step_ms * 1e-3) / 1e9
    per_kernel = {}
    traffic = 0
    for k, (tot, cnt) in table.items():
        tb = measured_traffic(k) if args.mode == "fused" else None
        per_kernel[k] = {"avg_ms": round(tot / cnt, 4), "measured_bytes": tb,
                         "measured_GB/s": round(tb / (tot / cnt * 1e-3) / 1e9, 1) if tb else None}
        traffic = traffic + tb if (tb is not None and traffic is not None) else None

    batched = batched_hd = None
    if not args.no_batched:
        breps = max(3, min(10, args.steps // 5))
        batched = bench_batched(tv, synth, torch, ctx, breps)
        batched_hd = bench_batched(tv, synth, torch, ctx, breps,
                                   specs=hd_specs(), workload=HD_WORKLOAD)

    served = None
    if not args.no_served:
        served = bench_served(torch, tv)

    boundary = bench_boundary(tv, torch, ctx, L, R, D)
    window = bench_window(tv, synth, torch, ctx)
    pairings = bench_pairings(tv, torch, ctx, Ld, Rd, D)

    # the 384x288, D = 16 pair (seed 1): GPU call, and the oracle per stage
    plain_384 = bench_plain_384(tv, synth, torch, ctx, None if args.no_cpu_baseline else True)

    cpu = cpu_mt = None
    if not args.no_cpu_baseline:
        import oracle
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            want = oracle.stereo_bt_sgm(L, R, D)
        cs = (time.perf_counter() - t0) / reps
        cpu = {"value": round(cells / cs / 1e6, 3), "unit": "MDE/s", "cores": 1, "kind": "port",
               "sample": "%d full 1280x720 D=128 pairs, BT+SGM8+WTA, oracle/tdvref.c gcc -O3 -march=x86-64-v3 "
                         "single thread (%s)" % (reps, cpu_model())}
        got = disp.cpu().numpy()
        if not (got == want).all():
            print("ERROR: GPU disparities differ from the oracle", file=sys.stderr)
            sys.exit(1)
        cpu_mt = cpu_threads(L, R, D, want)
    if served is not None and not served["all_frames_exact"]:
        print("ERROR: a served frame differs from the same frame computed alone", file=sys.stderr)
        sys.exit(1)
    for bt in (batched, batched_hd):
        if bt is not None and not bt["bit_exact_vs_plain"]:
            print("ERROR: a batched pair differs from the plain per-pair call", file=sys.stderr)
            sys.exit(1)

    out = {
        "metric": "MDE/s (W*H*D cells per second), SGM-BT 8-path 1280x720 D=128",
        "value": round(value, 1),
        "unit": "MDE/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic (seeded generator: blurred noise, 12 rectangles, shift warp, sigma-2 "
                "noise), bit-exact vs oracle",
        "config": {"workload": "plain pair 1280x720 uint8, BT2 cost, D=128, SGM 8 paths "
                               "P1=8 P2=32, lowest-index WTA",
                   "path": args.mode, "width": W, "height": H, "depth": D},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "scope": "whole step: bt_planes_kernel + sgm_lines_kernel + "
                              "sgm_sum_wta_kernel (one pair)",
                     "alg_bytes_per_step": b_alg(W, H, D),
                     "step_ms_hip_events": round(step_ms, 4),
                     "dominant_kernel": dom, "dominant_kernel_avg_ms": round(dom_ms, 4),
                     "kernels": per_kernel},
        "pipeline_frac_of_hbm_roofline": round(b_alg(W, H, D) / (sec * 1e9) / HBM_PEAK_GBS, 4),
        "other_path": {"path": "staged" if other == tv.FUSE_STAGED else "fused",
                       "value": round(cells / sec_other / 1e6, 1),
                       "ms_per_step": round(sec_other * 1e3, 4)},
        "kernels_ms": {k: round(v[0] / v[1], 4) for k, v in table.items()},
        "batched": batched,
        "batched_720p": batched_hd,
        "served": served,
        "boundary": boundary,
        "window_1080p": window,
        "pairings_720p": pairings,
        "plain_384": plain_384,
        "cpu_baseline": cpu,
        "cpu_baseline_8_threads": cpu_mt,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
