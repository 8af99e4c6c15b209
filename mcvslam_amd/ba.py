"""The reference's bundle-adjustment callers over the GPU optimizer.

    bundle_adjustment   PoseEstimation::BoundleAdjustment + FilterCallBack_Chi2
                        (include/PoseEstimation.hpp:35-119,
                        src/PoseEstimation.cpp:132-151)
    pose_optimization   PoseEstimation::PoseOptimization
                        (src/PoseEstimation.cpp:45-130)

Both build nothing themselves: the edges, fixed vertices and state are those
of a BALinearizer (mono edges, stereo edges, fixed frames and points), and
every step runs in the HIP library (mcv_ba_optimize, mcv_ba_outliers).

The reference files every edge under its keyframe's LEFT object
(BAoptimizer::EdgeInfos, src/BAoptimizer.cpp:64, :86, :107), so its filters,
which pick the stereo threshold for edges of the RIGHT object only, test
every edge, stereo ones included, against CHI2_MONO_THRESHOLD; the defaults
below keep that.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .api import BALinearizer

CHI2_MONO_THRESHOLD = 5.991     # include/PoseEstimation.hpp:16
CHI2_STEREO_THRESHOLD = 7.815   # include/PoseEstimation.hpp:15


def local_ba_linearizer(mono, stereo=None, point_fixed=None, ctx=None) -> BALinearizer:
    """A BALinearizer over the edges BAoptimizer::addEdge / addEdgeToBody
    (mono, a synth.BAProblem-like object) and addEdgeStereo (stereo, a
    synth.StereoProblem-like object on the same points and frames) add, with
    mono.frame_fixed and point_fixed as the fixed vertices."""
    lin = BALinearizer.from_problem(mono, ctx=ctx)
    if stereo is not None:
        lin.set_stereo(stereo.obs, stereo.obs_pt, stereo.obs_frame, stereo.obs_isig, stereo.cam)
    if point_fixed is not None:
        lin.set_point_fixed(point_fixed)
    return lin


@dataclass
class BAResult:
    stats: np.ndarray            # per LM iteration: chi2 before, chi2 after, lambda, trials
    iterations: int
    outliers_mono: np.ndarray    # FilterCallBack_Chi2 flags (1 = set to level 1)
    outliers_stereo: np.ndarray | None
    success_edges: int
    points: np.ndarray | None    # the recovered estimates (recovery_all), None when not recovered
    frames: np.ndarray | None


def bundle_adjustment(lin: BALinearizer, n_iter: int, threshold_mono: float = CHI2_MONO_THRESHOLD,
                      threshold_stereo: float = CHI2_MONO_THRESHOLD) -> BAResult:
    """op.optimize(n_iter); op.filter(FilterCallBack_Chi2): every edge with
    chi2 >= threshold at the optimized state is an outlier; recovery_all when
    more than 10 edges remain."""
    stats, n = lin.optimize(n_iter)
    stereo = getattr(lin, "Ms", 0) > 0
    fm, fs = lin.outliers(threshold_mono, threshold_stereo, stereo=stereo)
    n_edges = lin.M + (lin.Ms if stereo else 0)
    n_bad = int(fm.sum()) + (int(fs.sum()) if stereo else 0)
    success = n_edges - n_bad
    pts = frs = None
    if success > 10:
        pts, frs = lin.get_state()
    return BAResult(stats, n, fm, fs, success, pts, frs)


@dataclass
class PoseResult:
    frames: np.ndarray | None    # the optimized keyframe poses, None when discarded
    inliers: int                 # nInitialCorrespondences - nBad of the last round
    active_mono: np.ndarray
    active_stereo: np.ndarray | None


def pose_optimization(lin: BALinearizer, rounds=(15, 15, 15, 15), threshold_mono: float = CHI2_MONO_THRESHOLD,
                      threshold_stereo: float = CHI2_MONO_THRESHOLD) -> PoseResult:
    """Four optimizations of 15 iterations; after each, the edges still at
    level 0 whose chi2 reaches the threshold go to level 1 (out of the next
    optimizations, never back); after the third round the robust kernels are
    dropped; the loop stops when fewer than 10 correspondences survive a
    round."""
    stereo = getattr(lin, "Ms", 0) > 0
    am = np.ones(lin.M, np.uint8)
    ast = np.ones(lin.Ms, np.uint8) if stereo else None
    n_initial = lin.M + (lin.Ms if stereo else 0)
    lin.set_robust(True)
    n_bad = 0
    for it, iters in enumerate(rounds):
        lin.set_active(am, ast)
        lin.optimize(iters)
        fm, fs = lin.outliers(threshold_mono, threshold_stereo, stereo=stereo)
        bad_m = (am == 1) & (fm == 1)
        am[bad_m] = 0
        n_bad = int(bad_m.sum())
        if stereo:
            bad_s = (ast == 1) & (fs == 1)
            ast[bad_s] = 0
            n_bad += int(bad_s.sum())
        if it == 2:
            lin.set_robust(False)
        if n_initial - n_bad < 10:
            break
    frames = None
    if n_initial - n_bad >= 10:
        frames = lin.get_state()[1]
    return PoseResult(frames, n_initial - n_bad, am, ast)
