"""Benchmark of the MI355X hot path (one GPU).

Headline: rig frames per second of batched ORB extraction + matching.  One
step = one hipGraph launch over a batch of 64 rig frames (L and R 1280x720, W
640x480; per-image ORB feature counts drawn from U{500..4000}, seed 40 + frame;
8 levels, scale 1.2, FAST 20/7), then brute-force Hamming top-2 + ratio 0.7 +
threshold 50 for L-R, W-L and W-R.  Inputs are resident in HBM before timing
starts.  With --streams S (default 5) consecutive batches rotate over S rigs
on S streams, so each batch's serial octree replay overlaps the other
batches' pixel kernels; the one-stream rate is reported beside it.  Streams
only run side by side on distinct hardware queues: the HIP runtime's
per-process limit (GPU_MAX_HW_QUEUES, 4 unless set) is raised to
--hw-queues (default 16) before HIP starts, since S rig streams, their graphs'
branch streams and torch's streams share that many queues.

Also reported on the same line: the plain path (one rig frame at a time, no
graph), BA linearization throughput (50 keyframes x 3 cameras, 200k points,
~1.5M observations), the roofline of the dominant kernel (HIP events on the
stream it runs on) and the CPU oracle on this host's cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FRAMES = 64
UNIQUE_FRAMES = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def orb_level_pixels(w, h, inv_scale):
    px = []
    for l, s in enumerate(inv_scale):
        lw = w if l == 0 else int(np.rint(np.float32(w) * s))
        lh = h if l == 0 else int(np.rint(np.float32(h) * s))
        px.append(lw * lh)
    return px


def algorithmic_bytes(kernel, orb, frames, counts, octree=None):
    """Compulsory HBM bytes of one launch of `kernel` over the batch (None for
    the latency-bound octree kernels)."""
    levels = [orb_level_pixels(w, h, orb.mvInvScaleFactor) for (w, h) in ((1280, 720), (1280, 720), (640, 480))]
    sizes = [level_sizes(w, h, orb.mvInvScaleFactor) for (w, h) in ((1280, 720), (1280, 720), (640, 480))]
    P = sum(sum(l) for l in levels) * frames          # pyramid pixels
    P0 = sum(l[0] for l in levels) * frames
    Plast = sum(l[-1] for l in levels) * frames
    det = sum(max(w - 38, 0) * max(h - 38, 0) for s in sizes for (w, h) in s) * frames   # FAST cell detection areas
    nkp = int(counts.sum())
    if kernel == "k_blur":
        return 2 * P                                  # read level, write blurred level
    if kernel == "k_fast_score":
        return 2 * P                                  # read level, write score plane
    if kernel == "k_fast_cells":
        return det                                    # read score plane over the detection areas
    if kernel == "k_resize":
        return (P - Plast) + (P - P0)                 # read level l-1, write level l
    if kernel == "k_copy_level0":
        return 2 * P0
    if kernel == "k_keypoints":
        return 56 * nkp                               # keypoint record + descriptor written
    if kernel == "k_octree_heap":
        return 12 * nkp                               # node-table entries of each output node + its slot word
    if kernel == "k_octree_build":
        if octree is None:
            return 4 * nkp
        cand, table, _ = octree
        # candidate words gathered from the cells and copied per level, the
        # sorted index per candidate, the node tables (start + child words)
        return 12 * cand + 8 * table
    if kernel == "k_knn2":
        nq = counts.reshape(-1, 3)
        return int(sum(32 * (nq[:, a] + nq[:, b]) + 16 * nq[:, a] for a, b in ((0, 1), (2, 0), (2, 1))).sum())
    if kernel == "k_filter":
        nq = counts.reshape(-1, 3)
        return int(sum(16 * nq[:, a] for a in (0, 2, 2)).sum())
    return None


def measured_traffic(path=ROOT / "profiles" / "hbm_traffic_rig64.txt"):
    """HBM bytes per rig step (FETCH_SIZE + WRITE_SIZE, scaled by the
    calibration copy) from the committed rocprofv3 --pmc summary of this
    workload, keyed by the names the timed run reports; {} when it is absent.
    The summary lists bytes per dispatch and the dispatch count of the run;
    k_zero_words runs once per step (graph replay or eager pass), so its
    count is the number of steps."""
    rows = {}
    if not path.exists():
        return rows
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 5 and parts[0].startswith("k_") and parts[1].isdigit():
            rows[parts[0]] = (int(parts[1]), int(parts[4]))
    if "k_zero_words" not in rows:
        return {}
    steps = rows["k_zero_words"][0]
    out = {}
    for name, (calls, per_call) in rows.items():
        # launch variants (k_knn2_s / k_knn2_lds, k_keypoints_k) under the
        # name the timed run reports
        key = next((k for k in ("k_knn2", "k_keypoints") if name.startswith(k)), name)
        out[key] = out.get(key, 0) + calls * per_call / steps
    return out


def lib_schur(ba, lam):
    """mcv_ba_schur alone (enqueued on the context stream, no read-back)."""
    import ctypes
    from mcvslam_amd._native import check, lib
    check(ba.ctx.h, lib().mcv_ba_schur(ba.h, ctypes.c_float(lam)))


def level_sizes(w, h, inv_scale):
    out = []
    for l, s in enumerate(inv_scale):
        lw = w if l == 0 else int(np.rint(np.float32(w) * s))
        lh = h if l == 0 else int(np.rint(np.float32(h) * s))
        out.append((lw, lh))
    return out


def cpu_baseline(frames_np, nfeat, budget_s=12.0, match_threads=None):
    """Oracle on the host: per rig frame 3 extraction threads (one per camera,
    as src/Frame.cpp:120-125) then the three brute-force matches, each split
    over the host's cores by query (src/Matcher.cpp:299)."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    match_threads = match_threads or host_threads()
    done = 0
    t0 = time.perf_counter()
    for k in range(1 << 20):
        i = k % len(frames_np)
        f = frames_np[i]
        res = [None] * 3

        def work(k):
            res[k] = oracle.orb_extract(f[k], nfeatures=int(nfeat[3 * i + k]))
        th = [threading.Thread(target=work, args=(k,)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for qa, ta in ((0, 1), (2, 0), (2, 1)):
            # KnnMatch runs cv::parallel_for_ over the queries on every core
            # (src/Matcher.cpp:299): query slices on match_threads threads
            q, t = res[qa][1], res[ta][1]
            cuts = np.linspace(0, len(q), match_threads + 1).astype(int)
            with ThreadPoolExecutor(match_threads) as ex:
                parts = list(ex.map(lambda k: oracle.knn2(q[cuts[k]:cuts[k + 1]], t), range(match_threads)))
            knn = np.concatenate(parts) if parts else np.zeros((0, 4), np.int32)
            oracle.filter_ratio_threshold(knn, 0.7, 50)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return done / dt, done, dt


def _obs_slice(prob, a, b):
    import copy
    sub = copy.copy(prob)
    for f in ("obs_uv", "obs_pt", "obs_frame", "obs_cam", "obs_isig"):
        setattr(sub, f, getattr(prob, f)[a:b])
    return sub


def ba_cpu_baseline(prob, n_obs=1500000, threads=1):
    """FP64 oracle linearization of the first n_obs observations: one thread
    (g2o's per-edge loop with OpenMP off, ThirdParty/g2o/CMakeLists.txt:140)
    or `threads` threads over contiguous slices (the OpenMP build's loop)."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    n = min(n_obs, len(prob.obs_pt))
    cuts = np.linspace(0, n, threads + 1).astype(int)
    subs = [_obs_slice(prob, cuts[k], cuts[k + 1]) for k in range(threads)]
    t0 = time.perf_counter()
    if threads == 1:
        oracle.ba_linearize(subs[0])
    else:
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(oracle.ba_linearize, subs))
    dt = time.perf_counter() - t0
    return n / dt, n, dt


def host_threads() -> int:
    """The host cores this process may use (sched affinity), at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=FRAMES)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ba", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-frame-at-a-time comparison run")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch variants")
    ap.add_argument("--variants", type=str, default=None,
                    help="fixed rig launch variants 'knn,blur,kp,heap' instead of the autotuner's "
                         "(counter runs use the variants the autotuner chose)")
    ap.add_argument("--streams", type=int, default=5, choices=range(1, 9), metavar="S",
                    help="batches in flight: S > 1 rotates S rigs on S streams")
    ap.add_argument("--hw-queues", type=int, default=16, choices=range(1, 33), metavar="Q",
                    help="hardware queues of this process (GPU_MAX_HW_QUEUES, 1..32)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path would shard by rig "
              "frame and by BA observation, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    # read once, when HIP starts (the import below)
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    import torch
    from mcvslam_amd import BALinearizer, Context, ORB, RigBatch, synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()            # default priority
    torch.cuda.set_stream(stream)
    ctx = Context(0)
    ctx.use_torch_stream()                  # library work and torch events share this stream
    orb = ORB(2000, 1.2, 8, 20, 7, ctx=ctx)

    log("generating frames ...")
    t0 = time.perf_counter()
    n = args.frames
    uniq, _, nfeat = synth.rig_batch(n, UNIQUE_FRAMES)
    log(f"  {len(uniq)} unique rig frames in {time.perf_counter() - t0:.1f}s")
    L = torch.empty((n, 720, 1280), dtype=torch.uint8, device=dev)
    R = torch.empty((n, 720, 1280), dtype=torch.uint8, device=dev)
    W = torch.empty((n, 480, 640), dtype=torch.uint8, device=dev)
    for f in range(n):
        l, r, w = uniq[f % len(uniq)]
        L[f].copy_(torch.from_numpy(l))
        R[f].copy_(torch.from_numpy(r))
        W[f].copy_(torch.from_numpy(w))
    ptrs = []
    for f in range(n):
        ptrs += [L[f].data_ptr(), R[f].data_ptr(), W[f].data_ptr()]
    torch.cuda.synchronize()

    rig = RigBatch(orb, n, ratio=0.7, th=50, nfeatures=nfeat)
    tuning = ""
    fixed = tuple(int(v) for v in args.variants.split(",")) if args.variants else None
    if fixed:
        rig.set_variants(*fixed)
        tuning = f"fixed variants {fixed}"
    elif not args.no_autotune:
        log("autotune ...")
        tuning = rig.autotune(ptrs, reps=3)
        log(tuning.strip())
    def one(variants):
        rig.set_variants(*variants)
        for _ in range(args.warmup):
            rig.run(ptrs, graph=True)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(args.steps):
            rig.run(ptrs, graph=True)
        ev1.record()
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / args.steps

    # the blur's place in the graph (before FAST, or on a branch beside the
    # octree heap replay) is timed here as well: how much the branch gains
    # depends on how the hardware queues serve the two branches, which the
    # autotuner's replays do not always show
    log("warmup (graph) ...")
    tuned = rig.variants()
    one_by_blur = {}
    for bv in (tuned[1], tuned[1] ^ 2):
        v = (tuned[0], bv, tuned[2], tuned[3])
        one_by_blur[v] = one(v)
        log(f"batched, one stream, blur {'beside the heap' if bv & 2 else 'before FAST'}: "
            f"{one_by_blur[v]:.3f} ms per {n}-frame step")
    v1 = min(one_by_blur, key=one_by_blur.get)
    again = sorted(one(v1) for _ in range(3))
    ms_step1 = again[1]
    fps1 = n / (ms_step1 / 1e3)
    log(f"batched, one stream: {ms_step1:.3f} ms per {n}-frame step -> {fps1:.0f} rig frames/s")
    _, _, counts, _, mcounts, _ = rig.read()
    ms_step, fps = ms_step1, fps1
    placements = {}
    best_v = None
    if args.streams > 1:
        # consecutive batches rotate over S rigs on S streams, so one batch's
        # serial octree replay overlaps the other batches' pixel kernels; the
        # S rigs of their own (each with its context and stream) take the
        # tuned rig's launch variants
        S = args.streams
        extra = []
        rigs, streams = [], []
        chosen = rig.variants()
        main = torch.cuda.Stream()
        for _ in range(S):
            st = torch.cuda.Stream()
            cx = Context(0)
            cx.set_stream(st.cuda_stream)
            rg = RigBatch(ORB(2000, 1.2, 8, 20, 7, ctx=cx), n, ratio=0.7, th=50, nfeatures=nfeat)
            extra.append((rg, cx))
            rigs.append(rg)
            streams.append(st)

        def multi(variants):
            for rg in rigs:
                rg.set_variants(*variants)
            for k in range(S * args.warmup):
                rigs[k % S].run(ptrs, graph=True)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(main)
            for st in streams:
                st.wait_stream(main)
            for k in range(args.steps):
                rigs[k % S].run(ptrs, graph=True)
            for st in streams:
                main.wait_stream(st)
            e1.record(main)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / args.steps

        # the autotuner times kernels one batch at a time; with S batches in
        # flight the blur's place in the graph (before FAST, or on a branch
        # beside the octree heap) and the top-2 kernel (the matrix-core one
        # competes with the other batches' kernels differently) are measured
        # as well, and value is the best combination (identical outputs)
        knn_alt = 1 if chosen[0] >= 8 else 8
        best_v = chosen
        for kv in (chosen[0], knn_alt):
            for bv in (chosen[1], chosen[1] ^ 2):
                v = (kv, bv, chosen[2], chosen[3])
                ms = multi(v)
                tag = f"knn {kv}, blur {'beside the heap' if bv & 2 else 'before FAST'}"
                placements[tag] = round(ms, 4)
                log(f"batched, {S} streams, {tag}: {ms:.3f} ms per {n}-frame step -> "
                    f"{n / (ms / 1e3):.0f} rig frames/s")
                if ms <= min(placements.values()):
                    best_v = v
        # the chosen combination is timed three more times and value is the
        # median, not the minimum of several samples
        again = sorted(multi(best_v) for _ in range(3))
        log("batched, {} streams, chosen again: {}".format(S, ", ".join(f"{t:.3f}" for t in again)))
        ms_step = again[1]
        fps = n / (ms_step / 1e3)
        log(f"batched, {S} streams: {ms_step:.3f} ms per {n}-frame step -> {fps:.0f} rig frames/s")
        # every rig produced the tuned rig's outputs from the same images
        ka, da, ca, ma, mca, _ = rig.read()
        for rg, cx in extra:
            kb, db, cb, mb, mcb, _ = rg.read()
            assert (ca == cb).all() and (mca == mcb).all(), "the rigs disagree"
            assert np.array_equal(ka, kb) and np.array_equal(da, db) and np.array_equal(ma, mb), "the rigs disagree"
            rg.close()
            cx.close()

    plain_fps = None
    if not args.no_plain:
        # plain path: one rig frame at a time, no graph
        rig.run(ptrs, graph=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plain_reps = 2
        for _ in range(plain_reps):
            rig.run(ptrs, graph=False)
        torch.cuda.synchronize()
        plain_fps = n * plain_reps / (time.perf_counter() - t0)
        log(f"plain: {plain_fps:.0f} rig frames/s (batched one stream / plain = {fps1 / plain_fps:.1f}x)")

    # per-kernel durations (events after every launch, eager sequence)
    per = {}
    reps = 5
    seq = []
    for _ in range(reps):
        run = rig.run_timed(ptrs)
        seq.append([ms for _, ms in run])
        for name, ms in run:
            per.setdefault(name, []).append(ms)
    # every launch of the sequence in order (median over the reps)
    launch_seq = [[name, round(float(np.median([r[i] for r in seq])), 4)] for i, (name, _) in enumerate(run)]
    kern = {k: sum(v) / reps for k, v in per.items()}           # ms per step, summed over launches
    launches = {k: len(v) / reps for k, v in per.items()}
    dom = max(kern, key=kern.get)
    log("per-kernel ms/step: " + ", ".join(f"{k}={v:.3f}" for k, v in sorted(kern.items(), key=lambda x: -x[1])))

    result = {
        "metric": "3-cam rig frames/s (ORB + L-R/W-L/W-R match, bit-exact), batched hipGraph",
        "value": round(fps, 2),
        "unit": "rig frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/int32 (f32 angle, rotation)",
        "data": f"synthetic: seeded blob/rectangle/noise images, {len(uniq)} unique rig frames tiled to {n}",
        "config": {"workload": f"{n} rig frames per step: L,R 1280x720 + W 640x480, ORB features per image "
                               "U{500..4000} (seed 40+frame), 8 levels, scale 1.2, FAST 20/7, match ratio 0.7 "
                               "threshold 50",
                   "frames_per_step": n, "features_per_step": int(nfeat.sum()), "parallelism": "single GPU"},
        "streams": args.streams,
        "hw_queues": args.hw_queues,
        "streams_ms_per_step_by_variant": placements or None,
        "streams_variants": list(best_v) if placements else None,
        "one_stream_frames_per_s": round(fps1, 2),
        "one_stream_ms_per_step": round(ms_step1, 4),
        "one_stream_variants": list(v1),
        "one_stream_ms_by_blur": {("beside the heap" if v[1] & 2 else "before FAST"): round(t, 4)
                                  for v, t in one_by_blur.items()},
        "plain_frames_per_s": round(plain_fps, 2) if plain_fps else None,
        "batched_over_plain": round(fps1 / plain_fps, 2) if plain_fps else None,
        "kernels_ms_per_step": {k: round(v, 4) for k, v in kern.items()},
        "launch_sequence_ms": launch_seq,
        "autotune": [ln for ln in tuning.strip().splitlines()],
        "keypoints_per_step": int(counts.sum()),
        "matches_per_step": int(mcounts.sum()),
    }
    traffic = measured_traffic()
    octree = rig.octree_stats()
    result["octree"] = {"candidates": octree[0], "table_entries": octree[1], "nodes": octree[2]}
    # achievable HBM bandwidth on this box: a device-to-device copy of 1 GiB
    nb = 1 << 30
    ca = torch.empty(nb, dtype=torch.uint8, device=dev)
    cb = torch.empty_like(ca)
    cb.copy_(ca)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        cb.copy_(ca)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * nb * 10 / (e0.elapsed_time(e1) / 1e3) / 1e9
    del ca, cb
    result["hbm_copy_gbs"] = round(copy_gbs, 1)
    log(f"device copy: {copy_gbs:.0f} GB/s")
    table = {}
    for k in kern:
        ab = algorithmic_bytes(k, orb, n, counts, octree)
        if ab is None:
            continue
        per_launch_ms = kern[k] / launches[k]
        per_launch_bytes = ab / launches[k]
        ach = per_launch_bytes / (per_launch_ms / 1e3) / 1e9
        table[k] = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 5), "frac_of_copy": round(ach / copy_gbs, 5),
                    "traffic": int(traffic[k] / launches[k]) if k in traffic else None,
                    "algorithmic_bytes": int(per_launch_bytes), "avg_launch_ms": round(per_launch_ms, 5)}
    result["roofline"] = dict(kernel=dom, **table[dom])
    result["kernels_roofline"] = table

    if not args.no_ba:
        log("BA problem ...")
        t0 = time.perf_counter()
        prob = synth.ba_problem(seed=40, n_kf=50, n_points=200000, target_obs=1500000)
        ba = BALinearizer.from_problem(prob, ctx=ctx)
        log(f"  {len(prob.obs_pt)} observations in {time.perf_counter() - t0:.1f}s")
        ba_tune = ba.autotune(reps=5).strip().splitlines() if not args.no_autotune else []
        for _ in range(3):
            ba.linearize()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        iters = 50
        e0.record()
        for _ in range(iters):
            ba.linearize()
        e1.record()
        torch.cuda.synchronize()
        ba_ms = e0.elapsed_time(e1) / iters
        # plain: one launch per (keyframe, camera) edge group
        ba.linearize_plain()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(iters):
            ba.linearize_plain()
        e1.record()
        torch.cuda.synchronize()
        plain_ms = e0.elapsed_time(e1) / iters
        M = len(prob.obs_pt)
        P = len(prob.points)
        # 20 B in + 88 B out per observation, 12 B of state per point (the
        # kernel reads 32 B per point, the FP64 estimate as a float pair: an
        # implementation choice, not counted as algorithmic traffic)
        ba_bytes = 20 * M + 88 * M + 12 * P
        ach = ba_bytes / (ba_ms / 1e3) / 1e9
        result["ba"] = {"workload": "50 keyframes x 3 cams, 200k points, EdgeSE3XYZ2Camera linearize",
                        "observations": M, "obs_per_s": round(M / (ba_ms / 1e3), 1), "ms_per_launch": round(ba_ms, 5),
                        "plain_obs_per_s": round(M / (plain_ms / 1e3), 1), "plain_ms": round(plain_ms, 5),
                        "plain_launches": int(len(prob.frames) * len(prob.cams)),
                        "batched_over_plain": round(plain_ms / ba_ms, 2), "autotune": ba_tune,
                        "roofline": {"kernel": "k_ba_linearize", "bound": "hbm", "achieved": round(ach, 2),
                                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                                     "frac_of_copy": round(ach / copy_gbs, 4),
                                     "bytes_per_obs": round(ba_bytes / M, 2)}}
        log(f"BA linearize: {ba_ms * 1e3:.1f} us per {M} obs -> {M / (ba_ms / 1e3):.3e} obs/s, {ach:.0f} GB/s; "
            f"plain {plain_ms * 1e3:.1f} us")
        # Schur complement of the 200k points (FP32, H_ll / H_pl / H_pp / S
        # kernels) at a damping in g2o's starting range; the structure is
        # built on the host by the first call
        lam = 1e2
        ba.schur(lam)
        torch.cuda.synchronize()
        sreps = 10
        e0.record()
        for _ in range(sreps):
            lib_schur(ba, lam)
        e1.record()
        torch.cuda.synchronize()
        schur_ms = e0.elapsed_time(e1) / sreps
        # one full Levenberg-Marquardt iteration (OptimizationAlgorithmLevenberg::
        # solve): linearize + rho(chi2) sum + Schur + host FP64 Cholesky of the
        # 294 x 294 system + back-substitution + update + residual pass, trials
        # included
        ba.set_state(prob.points, prob.frames, prob.cams)
        ba.optimize(1)
        lm_iters = 3
        ba.set_state(prob.points, prob.frames, prob.cams)
        t0 = time.perf_counter()
        st, nit = ba.optimize(lm_iters)
        lm_ms = (time.perf_counter() - t0) * 1e3 / max(nit, 1)
        result["ba"]["schur_ms"] = round(schur_ms, 4)
        result["ba"]["schur_size"] = int(6 * int((np.asarray(prob.frame_fixed) == 0).sum()))
        result["ba"]["lm_iteration_ms"] = round(lm_ms, 3)
        result["ba"]["lm_trials_per_iteration"] = float(np.mean(st[:, 3])) if nit else None
        result["ba"]["lm_chi2"] = [round(float(v), 1) for v in (st[:, 1] if nit else [])]
        log(f"BA Schur: {schur_ms:.3f} ms; one LM iteration {lm_ms:.2f} ms ({nit} iterations)")
        ba.close()
        if not args.no_cpu:
            cps, n_cpu, dt = ba_cpu_baseline(prob)
            result["ba"]["cpu_baseline"] = {"value": round(cps, 1), "unit": "obs/s", "cores": 1, "kind": "port",
                                            "sample": f"first {n_cpu} observations of the same problem, FP64 "
                                                      f"oracle edge linearization, {dt:.2f}s"}
            mt = host_threads()
            cpsm, n_cpu, dtm = ba_cpu_baseline(prob, threads=mt)
            result["ba"]["cpu_baseline_all_cores"] = {
                "value": round(cpsm, 1), "unit": "obs/s", "cores": mt, "kind": "port",
                "sample": f"the same {n_cpu} observations over {mt} threads, {dtm:.2f}s"}
            log(f"BA CPU: {cps:.3e} obs/s one thread, {cpsm:.3e} obs/s on {mt} threads")

    if not args.no_cpu:
        log("CPU oracle baseline ...")
        fr = [uniq[i % len(uniq)] for i in range(8)]
        mt = host_threads()
        cfps, done, dt = cpu_baseline(fr, nfeat, match_threads=mt)
        result["cpu_baseline"] = {"value": round(cfps, 3), "unit": "rig frames/s", "cores": max(3, mt),
                                  "kind": "port",
                                  "sample": f"{done} rig frames (3 extraction threads, then 3 brute-force matches "
                                            f"each over {mt} threads by query) in {dt:.1f}s, oracle built -O3, "
                                            f"{os.cpu_count()} host cpus visible"}
        log(f"CPU: {cfps:.2f} rig frames/s")
    rig.close()
    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        a = torch.empty(nb, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        result["calib_copy_bytes"] = nb
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
