"""The batched-throughput workload at its full size against the oracle.

* The 64-frame rig batch the benchmark times (192 images: L, R 1280x720 and
  W 640x480, per-image ORB feature counts U{500..4000} from seed 40 + frame),
  replayed as the captured hipGraph with the autotuner's launch variants:
  every keypoint, descriptor, knn row and filtered match of every image and
  pair, bit for bit (ORBextractor.cc:831-899, src/Matcher.cpp:245-302,
  :100-111, :23-35).
* The large local BA (50 keyframes x 3 cameras, 200k points, ~1.5M
  observations): the FP32 linearization within 1e-5 of the FP64 oracle and
  the FP32 Schur complement against FP64 elimination
  (include/OptimizeTypes.hpp:80-113, core/block_solver.hpp:315-449).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from ba_check import check_against_oracle
from mcvslam_amd import BALinearizer, RigBatch, synth

pytestmark = pytest.mark.gpu

PAIRS = [(0, 1), (2, 0), (2, 1)]
WORKERS = min(16, os.cpu_count() or 1)


def test_rig64_benchmark_batch(orb_default):
    import torch
    uniq, index, nfeat = synth.rig_batch(64, 16)
    n = len(index)
    dev = [[torch.from_numpy(im).cuda() for im in fr] for fr in uniq]
    ptrs = [dev[index[f]][k].data_ptr() for f in range(n) for k in range(3)]
    rig = RigBatch(orb_default, n, ratio=0.7, th=50, nfeatures=nfeat)
    rig.autotune(ptrs, reps=1)
    rig.run(ptrs, graph=True)
    kps, desc, counts, matches, mcounts, knn = rig.read()
    print(f"rig64: {int(counts.sum())} keypoints, {int(mcounts.sum())} matches", flush=True)

    def extract(i):
        return oracle.orb_extract(uniq[index[i // 3]][i % 3], nfeatures=int(nfeat[i]))

    with ThreadPoolExecutor(WORKERS) as ex:
        ref = list(ex.map(extract, range(3 * n)))
    for i, (rk, rd) in enumerate(ref):
        c = counts[i]
        assert c == len(rk), f"image {i}: {c} keypoints, oracle {len(rk)}"
        for f in ("x", "y", "size", "angle", "response", "octave"):
            np.testing.assert_array_equal(kps[i, :c][f], rk[f], err_msg=f"image {i} field {f}")
        np.testing.assert_array_equal(desc[i, :c], rd, err_msg=f"image {i} descriptors")

    def match(p):
        f, k = divmod(p, 3)
        qa, ta = PAIRS[k]
        wk = oracle.knn2(ref[3 * f + qa][1], ref[3 * f + ta][1])
        return wk, oracle.filter_ratio_threshold(wk, 0.7, 50)

    with ThreadPoolExecutor(WORKERS) as ex:
        want = list(ex.map(match, range(3 * n)))
    for p, (wk, wm) in enumerate(want):
        np.testing.assert_array_equal(knn[p, :len(wk)], wk, err_msg=f"pair {p} knn rows")
        assert mcounts[p] == len(wm), f"pair {p}: {mcounts[p]} matches, oracle {len(wm)}"
        np.testing.assert_array_equal(matches[p, :mcounts[p]], wm, err_msg=f"pair {p} matches")
    rig.close()


def test_ba_1p5m_linearize_and_schur(gpu_ctx):
    prob = synth.ba_problem(seed=40, n_kf=50, n_points=200000, target_obs=1500000)
    M = len(prob.obs_pt)
    assert M > 1400000
    b = BALinearizer.from_problem(prob, ctx=gpu_ctx)
    b.autotune(reps=2)             # the launch shape the benchmark times
    b.linearize()
    got = b.results()
    check_against_oracle(prob, got)
    print(f"ba 1.5M: {M} observations linearized within 1e-5", flush=True)
    lam = 1e2
    S, bs = b.schur(lam)
    e, chi2, w, Jp, Jx = oracle.ba_linearize(prob)
    r = oracle.ba_schur(e, w, Jp, Jx, prob.obs_pt, prob.obs_frame, prob.obs_isig, len(prob.points),
                        len(prob.frames), prob.frame_fixed, lam)
    assert r["ok"] and S.shape == r["S"].shape
    es = np.linalg.norm(S - r["S"]) / np.linalg.norm(r["S"])
    eb = np.linalg.norm(bs - r["bs"]) / np.linalg.norm(r["bs"])
    x = b.solve_llt(S, bs)
    xr = np.linalg.solve(r["S"], r["bs"])
    ex = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    print(f"ba 1.5M Schur ({S.shape[0]}x{S.shape[0]}): rel err S {es:.2e}, b {eb:.2e}, step {ex:.2e}", flush=True)
    assert es < 1e-5 and eb < 1e-4 and ex < 1e-3
    b.close()
