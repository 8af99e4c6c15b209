"""Local bundle adjustment on the GPU against the FP64 oracle: the
pose-only linearization mode, the Schur complement with stereo edges, fixed
points and inactive edges against the dense damped normal equations, the
robust chi2 reduction, ten Levenberg-Marquardt iterations of
PoseEstimation::BoundleAdjustment with its chi2 outlier pass, and the four
rounds of PoseEstimation::PoseOptimization
(core/optimization_algorithm_levenberg.cpp:56-159, core/block_solver.hpp:
315-449, include/PoseEstimation.hpp:35-119, src/PoseEstimation.cpp:45-151)."""
import numpy as np
import pytest

import oracle
from lm_check import dense_schur, dense_system
from mcvslam_amd import BALinearizer, synth
from mcvslam_amd.ba import bundle_adjustment, local_ba_linearizer, pose_optimization

pytestmark = pytest.mark.gpu


def _lin(lp, ctx):
    return local_ba_linearizer(lp.mono, lp.stereo, lp.point_fixed, ctx=ctx)


def test_pose_only_mode_writes_all_but_j_point(gpu_ctx):
    prob = synth.ba_problem(seed=70, n_kf=4, n_points=2000, target_obs=15000)
    b = BALinearizer.from_problem(prob, ctx=gpu_ctx)
    shifted = prob.points + 0.1
    b.linearize()
    full_a = b.read()
    b.set_state(shifted, prob.frames, prob.cams)
    b.linearize()
    full_b = b.read()
    b.set_state(prob.points, prob.frames, prob.cams)
    b.linearize()
    b.set_state(shifted, prob.frames, prob.cams)
    b.linearize(mode=BALinearizer.POSE_ONLY)
    pose_b = b.read()
    np.testing.assert_array_equal(pose_b[:16], full_b[:16])
    np.testing.assert_array_equal(pose_b[16:], full_a[16:])     # J_point planes untouched
    assert not np.array_equal(full_a[16:], full_b[16:])
    # the stereo edge likewise: J_point rows are planes 23..31
    sp = synth.stereo_problem(seed=70, n_kf=4, n_points=2000, target_obs=3000)
    b2 = BALinearizer.from_problem(synth.ba_problem(seed=70, n_kf=4, n_points=2000, target_obs=5000), ctx=gpu_ctx)
    b2.set_stereo(sp.obs, sp.obs_pt, sp.obs_frame, sp.obs_isig, sp.cam)
    b2.linearize_stereo()
    e_a, _, _, jp_a, jx_a = b2.stereo_results()
    b2.set_state(sp.points + 0.1, sp.frames, b2_cams(sp))
    b2.set_stereo(sp.obs, sp.obs_pt, sp.obs_frame, sp.obs_isig, sp.cam)
    b2.linearize_stereo(mode=BALinearizer.POSE_ONLY)
    e_b, _, _, jp_b, jx_b = b2.stereo_results()
    np.testing.assert_array_equal(jx_b, jx_a)
    assert not np.array_equal(e_a, e_b) and not np.array_equal(jp_a, jp_b)


def b2_cams(sp):
    return synth.ba_problem(seed=70, n_kf=4, n_points=2000, target_obs=5000).cams


@pytest.mark.parametrize("robust", [True, False])
def test_schur_with_stereo_fixed_points_and_levels_vs_dense(gpu_ctx, robust):
    lp = synth.local_ba_problem(seed=71, n_kf=5, n_points=800, target_obs=8000, stereo_frac=0.9)
    rng = np.random.default_rng(3)
    am = (rng.random(len(lp.mono.obs_pt)) > 0.1).astype(np.uint8)
    ast = (rng.random(len(lp.stereo.obs_pt)) > 0.1).astype(np.uint8)
    lin = _lin(lp, gpu_ctx)
    assert lin.Ms > 50 and lp.point_fixed.sum() > 10
    lin.set_active(am, ast)
    lin.set_robust(robust)
    lin.linearize()
    lin.linearize_stereo()
    H, b, lam, free_f, free_p = dense_system(lp, lp.mono.points, lp.mono.frames, active_m=am, active_s=ast,
                                             robust=robust)
    S_ref, bs_ref = dense_schur(H, b, lam, 6 * len(free_f))
    S, bs = lin.schur(lam)
    es = np.linalg.norm(S - S_ref) / np.linalg.norm(S_ref)
    eb = np.linalg.norm(bs - bs_ref) / np.linalg.norm(bs_ref)
    print(f"Schur with stereo/fixed points/levels (robust={robust}): rel err S {es:.2e}, b {eb:.2e}", flush=True)
    assert es < 1e-5 and eb < 1e-4


def test_robust_chi2_vs_oracle(gpu_ctx):
    lp = synth.local_ba_problem(seed=73, n_kf=6, n_points=3000, target_obs=30000, stereo_frac=0.8)
    lin = _lin(lp, gpu_ctx)
    rng = np.random.default_rng(4)
    am = (rng.random(len(lp.mono.obs_pt)) > 0.2).astype(np.uint8)
    for robust in (True, False):
        for act in (None, am):
            lin.set_robust(robust)
            lin.set_active(act, None)
            lin.linearize(residual_only=True)
            lin.linearize_stereo(residual_only=True)
            got = lin.robust_chi2()
            want = oracle.ba_robust_chi2(lp.mono, lp.stereo, robust=robust, active=act)
            assert abs(got - want) < 1e-5 * want, (robust, act is None, got, want)


def _oracle_chi2(lp, pts, frs):
    m = lp.mono
    prob = synth.BAProblem(m.obs_uv, m.obs_pt, m.obs_frame, m.obs_cam, m.obs_level, m.obs_isig, pts, frs, m.cams,
                           m.frame_fixed)
    chi2 = oracle.ba_linearize(prob)[1]
    s = lp.stereo
    schi2 = oracle.ba_stereo_linearize(s.obs, s.obs_pt, s.obs_frame, s.obs_isig, pts, frs, s.cam)[1]
    return chi2, schi2


def test_bundle_adjustment_10_iterations_vs_fp64_oracle(gpu_ctx):
    """The BA linearization configuration (10 keyframes x 3 cameras, 20k
    points, ~150k observations) with LEFT stereo edges, keyframe 0 and a
    tenth of the points fixed, from a perturbed state: ten LM iterations on
    the GPU (FP32 linearization and Schur, FP64 Cholesky) against the FP64
    oracle LM, then FilterCallBack_Chi2."""
    lp = synth.local_ba_problem(seed=72, n_kf=10, n_points=20000, target_obs=150000, stereo_frac=0.5)
    lin = _lin(lp, gpu_ctx)
    res = bundle_adjustment(lin, 10)
    opts, ofrs, ost, on, _ = oracle.ba_lm(lp.mono, lp.stereo, iterations=10, point_fixed=lp.point_fixed)
    print("gpu   :", np.array2string(res.stats, precision=6), flush=True)
    print("oracle:", np.array2string(ost, precision=6), flush=True)
    assert res.iterations == on == 10
    rel = np.abs(res.stats[:, :2] - ost[:, :2]) / ost[:, :2]
    print(f"chi2 trajectory max rel diff {rel.max():.2e}", flush=True)
    assert rel.max() < 1e-4
    assert res.stats[-1, 1] < 0.9 * res.stats[0, 0]
    pts, frs = res.points, res.frames
    dt = np.abs(frs[:, 4:] - ofrs[:, 4:]).max()
    dq = min(np.abs(frs[:, :4] - ofrs[:, :4]).max(), np.abs(frs[:, :4] + ofrs[:, :4]).max())
    dp = np.abs(pts - opts).max(1)
    print(f"state: max |dt| {dt:.2e} m, max |dq| {dq:.2e}, points p99 {np.percentile(dp, 99):.2e} m, "
          f"max {dp.max():.2e} m", flush=True)
    assert dt < 1e-4 and dq < 1e-5 and np.percentile(dp, 99) < 1e-3
    # FilterCallBack_Chi2 at the optimized states
    chi2, schi2 = _oracle_chi2(lp, opts, ofrs)
    thr = 5.991
    for got, c in ((res.outliers_mono, chi2), (res.outliers_stereo, schi2)):
        want = (c.astype(np.float32) >= np.float32(thr)).astype(np.uint8)
        far = np.abs(c - thr) > 1e-3 * thr
        np.testing.assert_array_equal(got[far], want[far])
        assert 0 < got.sum() < len(got)


def _pose_problem(seed):
    """PoseOptimization's graph: one keyframe free, every point fixed, its
    mono and stereo edges."""
    lp = synth.local_ba_problem(seed=seed, n_kf=3, n_points=3000, target_obs=20000, stereo_frac=0.8,
                                perturb_pose=0.02, fixed_point_frac=0.0)
    m, s = lp.mono, lp.stereo
    km, ks = m.obs_frame == 2, s.obs_frame == 2
    mono = synth.BAProblem(m.obs_uv[km], m.obs_pt[km], m.obs_frame[km], m.obs_cam[km], m.obs_level[km],
                           m.obs_isig[km], lp.true_points.copy(), m.frames, m.cams, np.array([1, 1, 0], np.int32))
    stereo = synth.StereoProblem(s.obs[ks], s.obs_pt[ks], s.obs_frame[ks], s.obs_isig[ks], mono.points, mono.frames,
                                 s.cam)
    return mono, stereo, np.ones(len(mono.points), np.int32)


def test_pose_optimization_rounds_vs_oracle(gpu_ctx):
    mono, stereo, pfix = _pose_problem(74)
    lin = local_ba_linearizer(mono, stereo, pfix, ctx=gpu_ctx)
    res = pose_optimization(lin)
    # the oracle runs the same four rounds
    am, ast = np.ones(len(mono.obs_pt), np.uint8), np.ones(len(stereo.obs_pt), np.uint8)
    pts, frs = mono.points, mono.frames
    robust = True
    n_init = len(am) + len(ast)
    lp = synth.LocalBAProblem(mono, stereo, pfix, None, None)
    for it in range(4):
        pts, frs, _, _, _ = oracle.ba_lm(mono, stereo, iterations=15, point_fixed=pfix, active=am, sactive=ast,
                                         robust=robust, points=pts, frames=frs)
        chi2, schi2 = _oracle_chi2(lp, pts, frs)
        bm = (am == 1) & (chi2.astype(np.float32) >= np.float32(5.991))
        bs = (ast == 1) & (schi2.astype(np.float32) >= np.float32(5.991))
        am[bm] = 0
        ast[bs] = 0
        if it == 2:
            robust = False
        if n_init - bm.sum() - bs.sum() < 10:
            break
    assert res.frames is not None
    dt = np.abs(res.frames[2, 4:] - frs[2, 4:]).max()
    dq = min(np.abs(res.frames[2, :4] - frs[2, :4]).max(), np.abs(res.frames[2, :4] + frs[2, :4]).max())
    nd = int((res.active_mono != am).sum() + (res.active_stereo != ast).sum())
    print(f"pose optimization: |dt| {dt:.2e}, |dq| {dq:.2e}, level disagreements {nd} of {n_init}, "
          f"inliers {res.inliers}", flush=True)
    assert dt < 1e-5 and dq < 1e-6
    assert nd <= max(2, n_init // 1000)
    assert (am == 0).sum() > 0
