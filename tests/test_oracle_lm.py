"""Pins of the FP64 Levenberg-Marquardt oracle (g2o's
OptimizationAlgorithmLevenberg with BlockSolver's Schur complement,
ThirdParty/g2o/g2o/core/optimization_algorithm_levenberg.cpp:56-159,
core/block_solver.hpp:315-449): the first step equals the dense solution of
the damped normal equations over every free vertex; noise-free problems
converge to the true state; SE3 exp updates match closed forms."""
import math

import numpy as np

import oracle
from lm_check import dense_system
from mcvslam_amd import synth


def _R(q):
    x, y, z, w = np.asarray(q, float) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _rodrigues(w):
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if th == 0:
        return np.eye(3)
    return np.eye(3) + math.sin(th) / th * K + (1 - math.cos(th)) / th ** 2 * K @ K


def test_se3_oplus_closed_forms():
    yaw = lambda a: np.array([0, 0, math.sin(a / 2), math.cos(a / 2)])
    f = np.concatenate([yaw(0.3), [1.0, -2.0, 0.5]])
    # rotation about z composes the yaw angles and rotates the translation
    g = oracle.se3_oplus(f, [0, 0, 0.2, 0, 0, 0])
    np.testing.assert_allclose(g[:4], yaw(0.5), atol=1e-15)
    np.testing.assert_allclose(g[4:], _rodrigues([0, 0, 0.2]) @ f[4:], atol=1e-14)
    # pure translation increment adds
    g = oracle.se3_oplus(f, [0, 0, 0, 0.1, 0.2, -0.3])
    np.testing.assert_allclose(g, f + [0, 0, 0, 0, 0.1, 0.2, -0.3], atol=1e-15)
    # large rotations (every branch of the matrix -> quaternion conversion,
    # angles near pi): the rotation matrix is exp(w) R
    rng = np.random.default_rng(0)
    for w in [np.array([3.1, 0, 0]), np.array([0, 3.14, 0.01]), np.array([0, 0, -3.1]), rng.normal(size=3),
              np.array([1e-7, -2e-7, 3e-7])]:
        xi = np.concatenate([w, [0.3, -0.2, 0.1]])
        g = oracle.se3_oplus(f, xi)
        np.testing.assert_allclose(_R(g[:4]), _rodrigues(w) @ _R(f[:4]), atol=1e-12)
        assert g[3] >= 0 and abs(np.linalg.norm(g[:4]) - 1) < 1e-15


def test_lm_first_step_is_the_dense_damped_solution():
    lp = synth.local_ba_problem(seed=61, n_kf=4, n_points=500, target_obs=4000, stereo_frac=0.9)
    assert len(lp.stereo.obs_pt) > 20 and lp.point_fixed.sum() > 10
    H, b, lam, free_f, free_p = dense_system(lp, lp.mono.points, lp.mono.frames)
    x = np.linalg.solve(H + lam * np.eye(len(b)), b)
    _, _, st, n, xo = oracle.ba_lm(lp.mono, lp.stereo, iterations=1, point_fixed=lp.point_fixed)
    nf = 6 * len(free_f)
    got = np.concatenate([xo[:nf], xo[nf:].reshape(-1, 3)[free_p].ravel()])
    assert np.linalg.norm(got - x) / np.linalg.norm(x) < 1e-8
    assert np.allclose(xo[nf:].reshape(-1, 3)[lp.point_fixed == 1], 0)
    # lambda after the step follows the accepted-step rule: lam * max(1/3, min(2/3, 1 - (2 rho - 1)^3))
    assert n == 1 and st[0, 3] == 1 and st[0, 1] < st[0, 0]
    assert lam / 3 - 1e-12 <= st[0, 2] <= lam * 2 / 3 + 1e-12


def test_lm_noise_free_converges_to_the_truth():
    lp = synth.local_ba_problem(seed=62, n_kf=5, n_points=600, target_obs=6000, stereo_frac=0.9, noise=0.0,
                                outlier_frac=0.0, perturb_point=0.03, perturb_pose=0.005)
    pts, frs, st, n, _ = oracle.ba_lm(lp.mono, lp.stereo, iterations=30, point_fixed=lp.point_fixed)
    # the observations are float32 (about 3e-5 px at 600 px), which bounds
    # the attainable chi2 and the recovered state
    assert st[0, 0] > 1e3 and st[-1, 1] < 1e-9 * st[0, 0]
    # every free frame and every point seen from two frames or more (a point
    # seen once has no depth) recover the generating state
    frames_of = [set() for _ in range(len(pts))]
    for p, f in zip(np.concatenate([lp.mono.obs_pt, lp.stereo.obs_pt]),
                    np.concatenate([lp.mono.obs_frame, lp.stereo.obs_frame])):
        frames_of[p].add(int(f))
    two = np.array([len(s) >= 2 for s in frames_of])
    assert two.sum() > 400
    np.testing.assert_allclose(pts[two], lp.true_points[two], atol=1e-4)
    for f in range(len(frs)):
        q, qt = frs[f, :4], lp.true_frames[f, :4]
        assert min(np.abs(q - qt).max(), np.abs(q + qt).max()) < 1e-6
        np.testing.assert_allclose(frs[f, 4:], lp.true_frames[f, 4:], atol=1e-6)


def test_lm_rejected_steps_restore_the_state():
    """A tiny tau (huge steps) forces rejected trials: lambda grows by the
    doubling factors and chi2 never increases."""
    lp = synth.local_ba_problem(seed=63, n_kf=4, n_points=400, target_obs=3000, stereo_frac=0.5)
    _, _, st, n, _ = oracle.ba_lm(lp.mono, lp.stereo, iterations=6, tau=1e-12, point_fixed=lp.point_fixed)
    assert (st[:, 1] <= st[:, 0]).all()
    assert (st[:, 3] >= 1).all()
    chi0 = oracle.ba_robust_chi2(lp.mono, lp.stereo)
    assert abs(st[0, 0] - chi0) < 1e-9 * chi0
