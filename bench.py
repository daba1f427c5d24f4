#!/usr/bin/env python3
"""Stitch benchmark: 2 x 5312x2988 -> 8500x3300, Laplacian blend with 6 levels.

One step = one CudaStitchPano::process of one frame pair (inputs already in
HBM), as the reference's `test_cuda_blend --perf` loop measures
(tests/testCudaBlend.cpp:224-241).  Prints ONE JSON line.

  python bench.py [--steps K] [--warmup W] [--levels L] [--no-cpu]

Single GPU only: --gpus N with N > 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("hm-cupano_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import numpy as np  # noqa: E402

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
WORKLOAD = "2 x 5312x2988 uint8 BGRA -> 8500x3300 canvas, batch 1"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(rig, levels, pano_levels, tail_base=2, fuse_up=0):
    """Minimum HBM bytes per frame of each kernel label (what it must read + write).

    Maps are read as one packed 32-bit word per pixel; each source image is
    read once in full (W_in x H_in x 4 B), shared between the two remap
    launches in proportion to their valid map entries; uint8 BGRA pixels are
    4 B, float pyramid pixels 16 B, masks 4 B per float (the level-0 2-image
    mask 1 B).
    """
    cw, ch = rig.canvas_w, rig.canvas_h
    map_px = [m.size for m in rig.map_x]
    valid_px = [int((m != 65535).sum()) for m in rig.map_x]
    src_bytes = rig.src_w * rig.src_h * 4
    out = {}
    if levels == 0:
        # each canvas pixel: seam byte, the selected image's map word + source pixel, canvas write
        sel = 0
        sel_valid = 0
        seam = rig.seam
        for i, lab in ((0, 1), (1, 0)):
            h, w = rig.map_x[i].shape
            sub = seam[rig.pos_y[i]: rig.pos_y[i] + h, rig.pos_x[i]: rig.pos_x[i] + w] == lab
            sel += int(sub.sum())
            sel_valid += int((sub & (rig.map_x[i] != 65535)).sum())
        # source bytes: the part of each image its selected map entries read, in full-image units
        out["hard_seam"] = cw * ch * 1 + sel * 4 + 2 * src_bytes + cw * ch * 4
        return out
    x2 = rig.pos_x[1]
    w1 = rig.map_x[0].shape[1]
    bw = w1 - x2 + 256
    bh = ch
    rx0, rx1 = x2 - 128, x2 - 128 + bw  # canvas columns feeding the blend inputs
    a = [(bw, bh)]
    for _ in range(1, pano_levels):
        a.append(((a[-1][0] + 1) // 2, (a[-1][1] + 1) // 2))
    A = [w * h for w, h in a]
    coarse = sum(A[1:])  # pixels of levels >= 1
    # map words inside the blend columns, and the source bytes that go with them
    roi_map = 0
    roi_src = 0.0
    for i in range(2):
        c0 = max(0, rx0 - rig.pos_x[i])
        c1 = min(rig.map_x[i].shape[1], rx1 - rig.pos_x[i])
        if c1 > c0:
            sub = rig.map_x[i][:, c0:c1]
            roi_map += sub.size
            roi_src += src_bytes * int((sub != 65535).sum()) / max(1, valid_px[i])
    # remap of the blend columns: maps, source, both level-0 images (uint8)
    out["remap_roi"] = int(roi_map * 4 + roi_src + 2 * A[0] * 4)
    # remap of the rest of the canvas: maps, source, canvas writes
    out["remap_canvas"] = int((sum(map_px) - roi_map) * 4 + (2 * src_bytes - roi_src) + (cw - (rx1 - rx0)) * ch * 4)
    # Gaussian levels >= 1 of both images from level 0
    out["down_fused"] = 2 * A[0] * 4 + 2 * coarse * 16
    # reconstruction of levels 1..L-1: per level, the level's images and mask in, the next
    # level's images and R (taps) in, R out.  Levels 2..L-1 run as one tile-fused launch
    # ("up_tail": only R_2 leaves it) when there are more than three levels, level 1 (and
    # the small levels otherwise) one launch each ("up_quad").
    def level_bytes(k, r_out=True):
        b = 2 * A[k] * 16 + A[k] * 4 + (A[k] * 16 if r_out else 0)
        if k + 1 < pano_levels:
            b += 2 * A[k + 1] * 16 + A[k + 1] * 16
        return b
    if tail_base > 0 and pano_levels - 1 > tail_base:
        out["up_tail"] = sum(2 * A[k] * 16 + A[k] * 4 for k in range(tail_base, pano_levels)) + A[tail_base] * 16
        quad_levels = range(1, tail_base)
    else:
        quad_levels = range(1, pano_levels)
    fused = fuse_up and pano_levels >= 2 and 1 in quad_levels
    if fused:  # level 1 runs inside the level-0 launch
        quad_levels = [k for k in quad_levels if k != 1]
    quad = sum(level_bytes(k) for k in quad_levels)
    if quad:
        out["up_quad"] = quad
    if fused:
        # levels 1 and 0 in one launch: both level-0 images, the uint8 mask and the canvas;
        # level 1's images and mask; level 2's images and R (taps); R_1 stays in LDS
        b = 2 * A[0] * 4 + A[0] * 1 + A[0] * 4 + 2 * A[1] * 16 + A[1] * 4
        if pano_levels > 2:
            b += 2 * A[2] * 16 + A[2] * 16
        out["up_l10"] = b
        return out
    # level 0: both level-0 images, the uint8 mask, level-1 taps and R_1 in, canvas out
    out["up_level0"] = 2 * A[0] * 4 + A[0] * 1 + A[0] * 4
    if pano_levels > 1:
        out["up_level0"] += 2 * A[1] * 16 + A[1] * 16
    return out


TRAFFIC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic_8500x3300.json")


def measured_traffic():
    """HBM bytes per frame of each kernel, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes over this workload, scaled by the same counters on a copy kernel of known size."""
    if not os.path.exists(TRAFFIC):
        return None
    with open(TRAFFIC) as f:
        return json.load(f)["kernels"]


def host_threads():
    """Threads for the all-cores CPU baseline: the cores this process may run on, capped by
    OMP_NUM_THREADS when the environment sets it (a shared host gives a job a share)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def time_oracle(hm_oracle, rig, f1, f2, levels, threads, budget_s, min_frames=3):
    """Median seconds per frame of the oracle's stitch on `threads` threads: frames until
    `budget_s` has passed and at least `min_frames` ran."""
    hm_oracle.set_threads(threads)
    hm_oracle.pano2(rig, f1, f2, levels)  # warm-up (page-in of the buffers)
    ts = []
    t0 = time.perf_counter()
    while len(ts) < min_frames or time.perf_counter() - t0 < budget_s:
        t = time.perf_counter()
        hm_oracle.pano2(rig, f1, f2, levels)
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), len(ts)


def cpu_baseline(hm_synth, rig, f1, f2, levels):
    """The CPU oracle on the benchmark workload (8500x3300) and on the 2400x900 workload,
    single-threaded and on every host core (OpenMP over rows), in the same run."""
    import hm_oracle

    hm_oracle.lib()
    cores = host_threads()
    rig0 = hm_synth.config_rig("cfg0")
    g1 = hm_synth.make_frame(rig0.src_h, rig0.src_w, 11)[None]
    g2 = hm_synth.make_frame(rig0.src_h, rig0.src_w, 12)[None]
    runs = {}
    for name, r, a, b, budget in (("8500x3300", rig, f1, f2, 5.0), ("2400x900", rig0, g1, g2, 2.0)):
        for th in (1, cores):
            sec, n = time_oracle(hm_oracle, r, a, b, levels, th, budget)
            runs[f"{name}_threads{th}"] = {"ms_per_frame": round(sec * 1e3, 2), "frames": n,
                                           "Mpix_s": round(r.canvas_w * r.canvas_h / sec / 1e6, 3),
                                           "fps": round(1 / sec, 3)}
            log(f"cpu baseline {name} L={levels}, {th} thread(s): {sec * 1e3:.1f} ms/frame over {n} frames")
    hm_oracle.set_threads(cores)
    main = runs[f"8500x3300_threads{cores}"]
    return {"value": main["Mpix_s"], "unit": "Mpix/s", "cores": cores, "kind": "port",
            "sample": f"the oracle's 8500x3300 L={levels} stitch of the benchmark frames, median of "
                      f"{main['frames']} frames on {cores} threads (OpenMP over rows); the same at 1 thread "
                      f"and for 2 x 1280x720 -> 2400x900 in `runs`",
            "fps": main["fps"], "nproc": os.cpu_count(), "runs": runs}


def kt_of(r):
    return r["kt"]


SERVED_CONFIGS = None  # set from hm_synth in main(): the 32-rig served load
BATCH_MIX = None


def served_benchmark(torch, hm_stitch, hm_synth, seconds, n_streams, frames_closed, seed=7):
    """Plain (process + synchronize per frame, one frame per launch sequence) against served
    (hmServer: rigs resident, concurrent streams, graph replay) on the mixed 32-rig load:
    a closed loop (as fast as possible) for the throughput ratio, then Poisson arrivals at
    30 fps per rig for `seconds` for p50/p99 frame latency."""
    rng = np.random.default_rng(seed)
    rigs = []  # dict(name, kind, pano, inputs=[[tensors]], canvases=[t, t], mpix)
    for name, kind, geo, levels, count, n_sets in SERVED_CONFIGS:
        log(f"served load: generating {name} ({count} rigs)")
        rig = hm_synth.served_rig(name)
        sets = [[torch.from_numpy(f).cuda() for f in hm_synth.served_frames(rig, k)] for k in range(n_sets)]
        tuned = None
        for _ in range(count):
            if kind == "2":
                pano = hm_stitch.StitchPano(1, levels, rig)
            elif kind == "3":
                pano = hm_stitch.StitchPano3(1, levels, rig)
            else:
                pano = hm_stitch.StitchPanoN(1, levels, rig)
            assert pano.status == 0, (name, pano.status)
            assert pano.prepare() == 0
            w, h = pano.canvas_size()
            canvases = [pano.alloc_canvas(), pano.alloc_canvas()]
            if tuned is None:  # autotune the config once, on its first rig
                tuned, _ = pano.autotune(sets[0], canvases[0], frames=5)
            pano.set_tuning(**tuned)
            rigs.append(dict(name=name, kind=kind, pano=pano, inputs=sets, canvases=canvases, mpix=w * h / 1e6,
                             tuning=tuned))
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    def plain_one(r, k):
        if r["kind"] == "2":
            return r["pano"].process(r["inputs"][k % len(r["inputs"])][0], r["inputs"][k % len(r["inputs"])][1],
                                     r["canvases"][k % 2], stream)
        if r["kind"] == "3":
            ins = r["inputs"][k % len(r["inputs"])]
            return r["pano"].process(ins[0], ins[1], ins[2], r["canvases"][k % 2], stream=stream)
        return r["pano"].process(r["inputs"][k % len(r["inputs"])], r["canvases"][k % 2], stream)

    order = [(i % len(rigs), i // len(rigs)) for i in range(frames_closed)]
    mpix_closed = sum(rigs[ri]["mpix"] for ri, _ in order)
    # plain: warm-up, then one frame at a time
    for ri, k in order[: len(rigs) * 2]:
        hm_stitch.check(plain_one(rigs[ri], k), "plain process")
    torch.cuda.synchronize()
    per_cfg = {}
    marks = {}

    def mark():  # both clocks: a kernel trace's timestamps follow one of them
        return [time.clock_gettime_ns(time.CLOCK_MONOTONIC), time.clock_gettime_ns(time.CLOCK_BOOTTIME)]

    marks["plain_start"] = mark()
    t0 = time.perf_counter()
    for ri, k in order:
        t1 = time.perf_counter()
        hm_stitch.check(plain_one(rigs[ri], k), "plain process")
        torch.cuda.synchronize()
        per_cfg.setdefault(rigs[ri]["name"], []).append(time.perf_counter() - t1)
    t_plain = time.perf_counter() - t0
    marks["plain_end"] = mark()
    plain_ref = [r["canvases"][0].clone() for r in rigs]  # each rig's canvas 0 after its last even frame
    torch.cuda.synchronize()

    server = hm_stitch.Server(n_streams=n_streams, device_budget_bytes=200 << 30, use_graphs=True)
    ids = [server.add_rig(r["pano"]) for r in rigs]
    used = server.stats()["device_used_bytes"]

    def submit(ri, k, tag):
        r = rigs[ri]
        return server.submit(ids[ri], r["inputs"][k % len(r["inputs"])], r["canvases"][k % 2], tag=tag)

    # served, for each scheduling mode: warm-up (captures the graphs), then everything at once.
    # "graphs": each request alone on a free stream, its rig's graph replayed; "batchedK":
    # continuous batching, the ready requests of up to K rigs as one batched pass.
    modes = {}
    equal = True
    mode_set = {"graphs": (0, False), "batched8": (8, False), "batched32": (32, False),
                "batched8_same": (8, True), "batched32_same": (32, True)}
    for mode, (max_batch, same) in mode_set.items():
        server.set_batching(max_batch, same)
        for i, (ri, k) in enumerate(order[: len(rigs) * 4]):
            hm_stitch.check(submit(ri, k, i), "submit")
        hm_stitch.check(server.wait_idle(), "served warm-up")
        server.poll(1 << 16)
        st0 = server.stats()
        marks[f"served_{mode}_start"] = mark()
        t0 = time.perf_counter()
        for i, (ri, k) in enumerate(order):
            hm_stitch.check(submit(ri, k, i), "submit")
        hm_stitch.check(server.wait_idle(), "served run")
        dt = time.perf_counter() - t0
        marks[f"served_{mode}_end"] = mark()
        server.poll(1 << 16)
        st1 = server.stats()
        torch.cuda.synchronize()
        equal = equal and all(bool(torch.equal(r["canvases"][0], ref)) for r, ref in zip(rigs, plain_ref))
        passes = st1["batched_passes"] - st0["batched_passes"]
        modes[mode] = {"fps": round(frames_closed / dt, 1), "Mpix_s": round(mpix_closed / dt, 1),
                       "vs_plain": round(t_plain / dt, 3), "batched_passes": passes,
                       "frames_per_pass": round((st1["batched_frames"] - st0["batched_frames"]) / passes, 2)
                       if passes else None,
                       "pass_graph_launches": st1["pass_graph_launches"] - st0["pass_graph_launches"]}
        log(f"served load: closed loop {mode}: {modes[mode]}")
    best = max(modes, key=lambda m: modes[m]["fps"])
    t_served = frames_closed / modes[best]["fps"]
    server.set_batching(*mode_set[best])

    # open loop: Poisson arrivals, 30 fps per rig
    arrivals = []
    for ri in range(len(rigs)):
        t = float(rng.uniform(0, 1 / 30))
        k = 0
        while t < seconds:
            arrivals.append((t, ri, k))
            t += float(rng.exponential(1 / 30))
            k += 1
    arrivals.sort()
    lat = []
    log(f"served load: open loop ({best}), {len(arrivals)} Poisson arrivals over {seconds:.0f} s")
    t_start = time.perf_counter()
    next_log = 10.0
    for tag, (t_arr, ri, k) in enumerate(arrivals):
        if t_arr >= next_log:
            log(f"served load: open loop at {t_arr:.0f} s, {tag} frames submitted")
            next_log += 10.0
        while True:  # sleep (yielding the CPU to the scheduler thread) until the arrival time
            dt = t_arr - (time.perf_counter() - t_start)
            if dt <= 0:
                break
            time.sleep(dt if dt < 2e-4 else dt - 1e-4)
        hm_stitch.check(submit(ri, k, tag), "submit")
    hm_stitch.check(server.wait_idle(), "open loop")
    t_open = time.perf_counter() - t_start
    res = server.poll(len(arrivals) + 16)
    lat = np.array([(r["t_done_ns"] - r["t_submit_ns"]) / 1e6 for r in res])
    mpix_open = sum(rigs[ri]["mpix"] for _, ri, _ in arrivals)
    st = server.stats()
    server.close()
    batched = batched_mixed_benchmark(torch, hm_stitch, rigs, stream)
    plain_fps = frames_closed / t_plain
    served_fps = frames_closed / t_served
    return {
        "rigs": len(rigs), "configs": [c[0] + f" x{c[4]}" for c in SERVED_CONFIGS], "streams": n_streams,
        "tuning": {r["name"]: r["tuning"] for r in rigs},
        "device_bytes_resident": int(used),
        "closed_loop_frames": frames_closed,
        "plain_fps": round(plain_fps, 1), "plain_Mpix_s": round(mpix_closed / t_plain, 1),
        "served_fps": round(served_fps, 1), "served_Mpix_s": round(mpix_closed / t_served, 1),
        "served_vs_plain": round(served_fps / plain_fps, 3),
        "served_mode": best,
        "served_modes": modes,
        "served_equals_plain_output": equal,
        "plain_ms_per_frame": {k: round(1e3 * float(np.median(v)), 3) for k, v in per_cfg.items()},
        "open_loop": {"mode": best, "seconds": seconds, "offered_fps": round(len(arrivals) / seconds, 1),
                      "achieved_fps": round(len(res) / t_open, 1), "Mpix_s": round(mpix_open / t_open, 1),
                      "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
                      "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
                      "latency_ms_max": round(float(lat.max()), 3)},
        "graphs_captured": st["graphs_captured"], "graph_launches": st["graph_launches"],
        "phase_clock_ns": marks,
        "batched_mixed": batched,
    }


def batched_mixed_benchmark(torch, hm_stitch, rigs, stream, passes=20):
    """Plain (one frame per process call, synchronised) against the batched pass
    (hmBatch: the 16 frames' pipelines as one pass of persistent work-list
    launches, grid autotuned, replayed as a hipGraph).  Every frame of the pass is
    compared with the same frame stitched alone."""
    sel = []
    for name, count in BATCH_MIX:
        sel += [r for r in rigs if r["name"] == name][:count]

    def plain_one(r):
        ins, can = r["inputs"][0], r["canvases"][0]
        if r["kind"] == "2":
            return r["pano"].process(ins[0], ins[1], can, stream)
        if r["kind"] == "3":
            return r["pano"].process(ins[0], ins[1], ins[2], can, stream=stream)
        return r["pano"].process(ins, can, stream)

    for r in sel:
        r["canvases"][0].fill_(3)
    for r in sel:
        hm_stitch.check(plain_one(r), "plain process")
    torch.cuda.synchronize()
    alone = [r["canvases"][0].clone() for r in sel]
    t0 = time.perf_counter()
    for _ in range(passes):
        for r in sel:
            hm_stitch.check(plain_one(r), "plain process")
            torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / passes
    batch = hm_stitch.Batch([(r["pano"], r["inputs"][0], r["canvases"][0]) for r in sel])
    grid, _ = batch.autotune(passes=3, stream=stream)
    for r in sel:
        r["canvases"][0].fill_(3)
    torch.cuda.synchronize()
    hm_stitch.check(batch.run(stream, use_graph=True), "batched pass")
    torch.cuda.synchronize()
    identical = all(bool(torch.equal(r["canvases"][0], a)) for r, a in zip(sel, alone))
    t0 = time.perf_counter()
    for _ in range(passes):
        hm_stitch.check(batch.run(stream, use_graph=True), "batched pass")
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / passes
    info = batch.info()
    batch.close()
    return {"frames": len(sel), "mix": [f"{n} x{c}" for n, c in BATCH_MIX],
            "plain_ms_per_16_frames": round(1e3 * t_plain, 3), "batched_ms_per_16_frames": round(1e3 * t_batch, 3),
            "batched_vs_plain": round(t_plain / t_batch, 3), "plain_launches": info["recorded_launches"],
            "batched_launches": info["launches"], "stages": info["stages"], "grid": grid,
            "every_frame_equals_alone": identical}


# Single workloads for profiling: (class, rig geometry, levels).  n3/n8:
# 3 or 8 x 3840x2160 -> 6200x2300 / 14000x2600 at L=6; p3: 3 x 2704x1520 ->
# 6000x2000 at L=4; h4: 4 x 1920x1080 -> 4000x1500 at L=0 (hard seam); mix16:
# the 16-frame mixed batch as one batched pass.
WORKLOADS = {
    "n3": ("N", "cfg2_n3", 6),
    "n8": ("N", "cfg2_n8", 6),
    "p3": ("3", (3, 2704, 1520, 6000, 2000, 0.25, 1500.0, 77), 4),
    "h4": ("N", (4, 1920, 1080, 4000, 1500, 0.25, 1100.0, 6), 0),
    "mix16": ("mix", None, None),
}


def workload_benchmark(torch, hm_stitch, hm_synth, name, steps, warmup):
    """Frames (or batched passes) of one workload back to back on one stream,
    timed over `steps` after `warmup`; prints per-kernel times from HIP events."""
    kind, geo, levels = WORKLOADS[name]
    stream = torch.cuda.current_stream()

    def rig_of(g):
        if isinstance(g, str):
            return hm_synth.config_rig(g)
        return hm_synth.cylinder_rig(g[0], g[1], g[2], g[3], g[4], overlap=g[5], focal=g[6], seed=g[7])

    def frames_of(rig, seed):
        return [torch.from_numpy(hm_synth.make_frame(rig.src_h, rig.src_w, seed + 17 * i)[None]).cuda()
                for i in range(rig.n)]

    if kind == "mix":
        calls = []
        for cname, count in BATCH_MIX:
            _, ckind, cgeo, clevels, _, _ = next(c for c in SERVED_CONFIGS if c[0] == cname)
            rig = hm_synth.config_rig("cfg1") if cgeo is None else rig_of(cgeo)
            for k in range(count):
                cls = {"2": hm_stitch.StitchPano, "3": hm_stitch.StitchPano3, "N": hm_stitch.StitchPanoN}[ckind]
                pano = cls(1, clevels, rig)
                calls.append((pano, frames_of(rig, 100 * k), pano.alloc_canvas()))
        batch = hm_stitch.Batch(calls)
        run = lambda: hm_stitch.check(batch.run(stream, use_graph=True), "batched pass")  # noqa: E731
        mpix = sum(c[0].canvas_size()[0] * c[0].canvas_size()[1] for c in calls) / 1e6
        timer = None
    else:
        rig = rig_of(geo)
        cls = hm_stitch.StitchPanoN if kind == "N" else hm_stitch.StitchPano3
        pano = cls(1, levels, rig)
        fr = frames_of(rig, 0)
        canvas = pano.alloc_canvas()
        if kind == "N":
            run = lambda: hm_stitch.check(pano.process(fr, canvas, stream), "process")  # noqa: E731
        else:
            run = lambda: hm_stitch.check(pano.process(fr[0], fr[1], fr[2], canvas, stream=stream), "process")  # noqa: E731
        w, h = pano.canvas_size()
        mpix = w * h / 1e6
        timer = pano
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    out = {"workload": name, "steps": steps, "ms_per_step": round(ms, 4), "Mpix_s": round(mpix / ms * 1e3, 1)}
    if timer is not None:
        timer.set_timing(True)
        timer.timing_report(True)
        for _ in range(min(steps, 20)):
            run()
        rep = timer.timing_report(True)
        n = min(steps, 20)
        out["kernels_ms"] = {k: round(v[0] / n, 4) for k, v in rep.items()}
        out["levels"] = pano.levels()
    else:
        out["batch"] = batch.info()
    return out


def frame_loop_benchmark(hm_stitch, rig, f1, f2, levels, frames):
    """Host frames in, host frames out (hmFrameLoopRun): BGR over PCIe, the stitch
    on the GPU, 3 pinned slots in flight.  The source fills each slot once and then
    resends the slot's pinned contents (no host-side copy), so the steady rate is
    the transfer + stitch pipeline's; the first fills (BGR copies of the frames into
    fresh pinned memory) are reported apart as the time to the first frame."""
    pano = hm_stitch.StitchPano(1, levels, rig)
    sizes = [(rig.src_w, rig.src_h)] * 2
    first = [f1[0, ..., :3], f2[0, ..., :3]]
    seen = [0]

    def source(k):
        return first if k < 3 else True  # fill each slot once

    def sink(k, bgr):
        seen[0] += 1

    names = {hm_stitch.HM_TRANSFER_DMA: "dma", hm_stitch.HM_TRANSFER_KERNEL: "kernel"}
    runs = {}
    for up_mode in names:
        for down_mode in names:
            kw = dict(slots=3, upload=up_mode, download=down_mode)
            hm_stitch.frame_loop(pano, sizes, source, sink, max_frames=12, **kw)  # warm-up
            st = hm_stitch.frame_loop(pano, sizes, source, sink, max_frames=frames, **kw)
            # the rate once every slot has been filled once (the first fills copy the
            # source frames into fresh pinned memory), and over the whole call
            fps = st["steady_frames"] / st["steady_seconds"] if st["steady_seconds"] > 0 else 0.0
            fps_call = st["frames"] / st["seconds"]
            up, down = st["bytes_up_per_frame"], st["bytes_down_per_frame"]
            key = f"up_{names[up_mode]}_down_{names[down_mode]}"
            log(f"frame loop {key}: {fps:.1f} fps steady, {fps_call:.1f} over the call "
                f"(first frame after {st['first_frame_seconds'] * 1e3:.0f} ms); upload {st['upload_ms']:.3f} ms, stitch {st['stitch_ms']:.3f} ms, "
                f"download {st['download_ms']:.3f} ms per frame")
            runs[key] = {"fps": round(fps, 1), "fps_whole_call": round(fps_call, 1),
                         "first_frame_ms": round(st["first_frame_seconds"] * 1e3, 1), "upload_ms": round(st["upload_ms"], 4),
                         "stitch_ms": round(st["stitch_ms"], 4), "download_ms": round(st["download_ms"], 4),
                         "upload_GBps": round(up / (st["upload_ms"] * 1e-3) / 1e9, 1) if st["upload_ms"] else None,
                         "download_GBps": round(down / (st["download_ms"] * 1e-3) / 1e9, 1) if st["download_ms"]
                         else None}
    best = max(runs, key=lambda k: runs[k]["fps"])
    return {"fps": runs[best]["fps"], "fps_whole_call": runs[best]["fps_whole_call"],
            "first_frame_ms": runs[best]["first_frame_ms"], "best": best, "frames": frames, "slots": 3,
            "bytes_up_per_frame": up, "bytes_down_per_frame": down, "transfers": runs}


def pct(v, q):
    """The q-th percentile of v in ms (4 decimals), or None with no samples."""
    return round(float(np.percentile(v, q)), 4) if len(v) else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--latency-frames", type=int, default=200)
    ap.add_argument("--served-seconds", type=float, default=60.0,
                    help="open-loop duration of the 32-rig served load (0: skip the served measurement)")
    ap.add_argument("--served-streams", type=int, default=4)
    ap.add_argument("--served-frames", type=int, default=640, help="frames of the closed-loop served/plain runs")
    ap.add_argument("--served-only", action="store_true", help="measure only the served load")
    ap.add_argument("--tuning", action="append", default=[], metavar="LEVELS:PX,TAIL,OVERLAP[,XCD]",
                    help="fix the launch-shape knobs for a level count instead of autotuning (e.g. 6:4,2,1,1)")
    ap.add_argument("--loop-frames", type=int, default=120,
                    help="frames of the PCIe-inclusive frame-loop measurement (0: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (scale for the HBM byte counters)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default=None,
                    help="time only one other workload (--steps frames or passes after --warmup) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the stitch path shards by rig/frame with no collectives, "
              "and multi-GPU runs are not part of this build. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    import hm_stitch
    import hm_synth

    global SERVED_CONFIGS, BATCH_MIX
    SERVED_CONFIGS, BATCH_MIX = hm_synth.SERVED_CONFIGS, hm_synth.BATCH_MIX
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.calibrate_copy:
        # A plain elementwise copy kernel of known size: b = a + 0 over 1 GiB of int32
        # reads 1 GiB and writes 1 GiB per launch.
        n = 1 << 28
        a = torch.ones(n, dtype=torch.int32, device="cuda")
        b = torch.empty(n, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        for _ in range(5):
            torch.add(a, 0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "elementwise copy (torch.add, int32)", "bytes_read": 4 * n,
                          "bytes_written": 4 * n, "launches": 5}))
        return
    hm_stitch.lib()
    if args.workload:
        print(json.dumps(workload_benchmark(torch, hm_stitch, hm_synth, args.workload, args.steps, args.warmup)),
              flush=True)
        return
    if args.served_only:
        print(json.dumps(served_benchmark(torch, hm_stitch, hm_synth, args.served_seconds, args.served_streams,
                                          args.served_frames)), flush=True)
        return
    log("generating the 8500x3300 rig and frames")
    rig = hm_synth.config_rig("cfg1")
    f1 = hm_synth.make_frame(rig.src_h, rig.src_w, 101)[None]
    f2 = hm_synth.make_frame(rig.src_h, rig.src_w, 202)[None]
    d1 = torch.from_numpy(f1).cuda()
    d2 = torch.from_numpy(f2).cuda()
    stream = torch.cuda.current_stream()
    results = {}
    for levels in ([args.levels, 0] if args.levels != 0 else [0]):
        pano = hm_stitch.StitchPano(1, levels, rig)
        assert pano.status == 0, pano.status
        canvas = pano.alloc_canvas()
        cw, ch = pano.canvas_size()
        fixed = [t.split(":", 1)[1] for t in args.tuning if int(t.split(":", 1)[0]) == levels]
        if fixed:
            vals = [int(v) for v in fixed[0].split(",")]
            pano.set_tuning(*vals)
            tuning = pano.get_tuning()
            log(f"levels={levels}: fixed tuning {tuning}")
        else:
            # the project's autotuner: launch-shape knobs timed on this frame (output unchanged)
            tuning, tuned_ms = pano.autotune([d1, d2], canvas, frames=10, stream=stream)
            log(f"levels={levels}: autotuned {tuning} ({tuned_ms:.4f} ms/frame while tuning)")
        for _ in range(args.warmup):
            hm_stitch.check(pano.process(d1, d2, canvas, stream), "process")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            hm_stitch.check(pano.process(d1, d2, canvas, stream), "process")
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        ms = el * 1e3 / args.steps
        log(f"levels={levels}: {ms:.4f} ms/frame, {1e3 / ms:.1f} fps")
        # per-frame latency, one frame in flight (process + synchronize)
        lat = []
        for _ in range(args.latency_frames):
            t = time.perf_counter()
            pano.process(d1, d2, canvas, stream)
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - t) * 1e3)
        # per-kernel durations with HIP events on the launch stream
        pano.set_timing(True)
        pano.timing_report(reset=True)
        kt = max(20, args.steps // 3)
        for _ in range(kt):
            pano.process(d1, d2, canvas, stream)
        rep = pano.timing_report(reset=True)
        pano.set_timing(False)
        kernels = {k: {"avg_ms": v[0] / v[1], "per_frame_ms": v[0] / kt, "launches": v[1]} for k, v in rep.items()}
        alg = algorithmic_bytes(rig, levels, len(pano.levels()) if levels else 0, tuning["tail_base"],
                                tuning.get("fuse_up", 0))
        for k, b in alg.items():
            if k in kernels:
                per_frame = kernels[k]["launches"] / kt
                kernels[k]["alg_bytes"] = int(b / per_frame)  # per launch
                kernels[k]["achieved_GBps"] = kernels[k]["alg_bytes"] / (kernels[k]["avg_ms"] * 1e-3) / 1e9
        results[levels] = dict(ms=ms, lat=np.array(lat), kernels=kernels, canvas=(cw, ch), kt=kt, tuning=tuning)
        del pano, canvas
        torch.cuda.synchronize()

    main_l = args.levels
    r = results[main_l]
    cw, ch = r["canvas"]
    mpix = cw * ch / (r["ms"] * 1e-3) / 1e6
    # dominant kernel = largest time per frame among those with an algorithmic byte count
    cands = {k: v for k, v in r["kernels"].items() if "alg_bytes" in v}
    dom = max(cands, key=lambda k: cands[k]["per_frame_ms"])
    dk = cands[dom]
    mt = measured_traffic()
    traffic = None
    traffic_low = None
    if mt and dom in mt:
        traffic = int(mt[dom]["traffic_bytes_per_frame"] / mt[dom]["launches_per_frame"])
        if "traffic_bytes_lower_bound" in mt[dom]:
            traffic_low = int(mt[dom]["traffic_bytes_lower_bound"] / mt[dom]["launches_per_frame"])
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(dk["achieved_GBps"], 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(dk["achieved_GBps"] / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_ms": round(dk["avg_ms"], 5)}
    if traffic:
        # traffic: read requests taken as 128 bytes (the calibrated x2 FETCH_SIZE scale, an upper
        # bound for gathers); traffic_low: max(64-byte requests, algorithmic reads) + writes
        roofline["traffic_GBps"] = round(traffic / (dk["avg_ms"] * 1e-3) / 1e9, 1)
        roofline["traffic_frac"] = round(traffic / (dk["avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
        if traffic_low:
            roofline["traffic_low"] = traffic_low
    # the whole frame: every kernel's bytes over the frame time
    frame_alg = sum(v["alg_bytes"] * v["launches"] / kt_of(r) for v in cands.values())
    frame = {"alg_bytes": int(frame_alg), "alg_GBps": round(frame_alg / (r["ms"] * 1e-3) / 1e9, 1)}
    frame["alg_frac"] = round(frame["alg_GBps"] / HBM_PEAK_GBPS, 4)
    if mt and all(k in mt for k in cands):
        ft = sum(mt[k]["traffic_bytes_per_frame"] for k in cands)
        frame.update({"traffic_bytes": int(ft), "traffic_GBps": round(ft / (r["ms"] * 1e-3) / 1e9, 1),
                      "traffic_frac": round(ft / (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)})
        if all("traffic_bytes_lower_bound" in mt[k] for k in cands):
            fl = sum(mt[k]["traffic_bytes_lower_bound"] for k in cands)
            frame.update({"traffic_bytes_low": int(fl),
                          "traffic_frac_low": round(fl / (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)})
    line = {
        "metric": "pano Mpix/s + fps @8500x3300 L=6; p50/p99 frame latency, served vs plain",
        "value": round(mpix, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded frames, cylindrical uint16 maps, wavy seam)",
        "config": {"workload": WORKLOAD, "levels": main_l, "canvas": f"{cw}x{ch}", "batch": 1,
                   "blend_roi": "minimized (overlap + 2x128)"},
        "fps": round(1e3 / r["ms"], 1),
        "latency_ms_p50": pct(r["lat"], 50),
        "latency_ms_p99": pct(r["lat"], 99),
        "roofline": roofline,
        "hbm_frame": frame,
        "tuning": r["tuning"],
        "reference_published": {"fps": 340, "hardware": "RTX 4090", "levels": 6,
                                "note": "README.md:7; same canvas size, different GPU"},
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in r["kernels"].items()},
    }
    if 0 in results and main_l != 0:
        r0 = results[0]
        line["levels0"] = {"ms_per_step": round(r0["ms"], 5), "fps": round(1e3 / r0["ms"], 1), "tuning": r0["tuning"],
                           "value_Mpix_s": round(cw * ch / (r0["ms"] * 1e-3) / 1e6, 1),
                           "latency_ms_p50": pct(r0["lat"], 50),
                           "latency_ms_p99": pct(r0["lat"], 99),
                           "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                                       for k, v in r0["kernels"].items()}}
    if args.loop_frames > 0:
        line["pcie_frame_loop"] = frame_loop_benchmark(hm_stitch, rig, f1, f2, main_l, args.loop_frames)
    if args.served_seconds > 0:
        line["served"] = served_benchmark(torch, hm_stitch, hm_synth, args.served_seconds, args.served_streams,
                                          args.served_frames)
    if not args.no_cpu:
        log("timing the CPU oracle on the same workload")
        line["cpu_baseline"] = cpu_baseline(hm_synth, rig, f1, f2, main_l)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
