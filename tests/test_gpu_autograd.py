"""Reprojection adjoint behind torch.autograd.Function: the GPU backward
kernels against central finite differences of the FP64 oracle residual
(h = 1e-4) and against the FP64 analytic Jacobians (g2o linearizeOplus)."""
import numpy as np
import pytest

import oracle
from mcvslam_amd import BALinearizer, synth
from mcvslam_amd.autograd import apply_increments, reprojection

pytestmark = pytest.mark.gpu


def _problem():
    return synth.ba_problem(seed=34, n_kf=4, n_points=1500, target_obs=12000)


def _oracle_e(prob, dpose, dpoint):
    pts, frs = apply_increments(prob.points, prob.frames, dpose, dpoint)
    p = synth.BAProblem(**{k: np.copy(v) for k, v in prob.__dict__.items()})
    p.points, p.frames = pts, frs
    return oracle.ba_linearize(p)[0]


def test_adjoint_vs_fd_and_jacobians(gpu_ctx):
    import torch
    prob = _problem()
    lin = BALinearizer.from_problem(prob, ctx=gpu_ctx)
    F, P, M = len(prob.frames), len(prob.points), len(prob.obs_pt)
    dpose = torch.zeros((F, 6), dtype=torch.float64, device="cuda", requires_grad=True)
    dpoint = torch.zeros((P, 3), dtype=torch.float64, device="cuda", requires_grad=True)
    e = reprojection(dpose, dpoint, lin, (prob.points, prob.frames, prob.cams))
    rng = np.random.default_rng(0)
    G = rng.normal(size=(M, 2))
    (e * torch.from_numpy(G).float().cuda()).sum().backward()
    gpose = dpose.grad.cpu().numpy()
    gpoint = dpoint.grad.cpu().numpy()
    # FP64 analytic adjoint from the oracle's Jacobians
    _, _, _, Jp, Jx = oracle.ba_linearize(prob)
    want_pose = np.zeros((F, 6))
    want_point = np.zeros((P, 3))
    np.add.at(want_pose, prob.obs_frame, np.einsum("mij,mi->mj", Jp, G))
    np.add.at(want_point, prob.obs_pt, np.einsum("mij,mi->mj", Jx, G))
    for got, want in ((gpose, want_pose), (gpoint, want_point)):
        err = np.linalg.norm(got - want, axis=1) / np.maximum(np.linalg.norm(want, axis=1), 1e-30)
        assert err.max() < 1e-5, err.max()
    # central differences of the FP64 residual along random directions
    h = 1e-4
    for trial in range(3):
        vp = rng.normal(size=(F, 6)) * np.array([1e-2] * 3 + [1e-1] * 3)
        vx = rng.normal(size=(P, 3)) * 1e-1
        fd = ((_oracle_e(prob, h * vp, h * vx) - _oracle_e(prob, -h * vp, -h * vx)) * G).sum() / (2 * h)
        ad = (gpose * vp).sum() + (gpoint * vx).sum()
        assert abs(ad - fd) <= 1e-5 * max(abs(fd), np.abs(gpose * vp).sum() + np.abs(gpoint * vx).sum())
    lin.close()


def test_forward_moves_with_increments(gpu_ctx):
    import torch
    prob = _problem()
    lin = BALinearizer.from_problem(prob, ctx=gpu_ctx)
    F, P = len(prob.frames), len(prob.points)
    rng = np.random.default_rng(1)
    dp = rng.normal(size=(F, 6)) * 1e-3
    dx = rng.normal(size=(P, 3)) * 1e-3
    e = reprojection(torch.from_numpy(dp).cuda(), torch.from_numpy(dx).cuda(), lin,
                     (prob.points, prob.frames, prob.cams)).cpu().numpy()
    re = _oracle_e(prob, dp, dx)
    err = np.linalg.norm(e - re, axis=1) / np.maximum(np.linalg.norm(re, axis=1), 1e-9)
    assert err.max() < 1e-5
    lin.close()


def test_backward_after_relinearization_raises(gpu_ctx):
    """Two forwards on one BALinearizer, then the first one's backward: its
    Jacobians were overwritten by the second forward, so backward refuses
    instead of returning another state's gradient."""
    import torch
    prob = _problem()
    lin = BALinearizer.from_problem(prob, ctx=gpu_ctx)
    F, P = len(prob.frames), len(prob.points)
    d1 = torch.zeros((F, 6), dtype=torch.float64, device="cuda", requires_grad=True)
    x1 = torch.zeros((P, 3), dtype=torch.float64, device="cuda", requires_grad=True)
    e1 = reprojection(d1, x1, lin, (prob.points, prob.frames, prob.cams))
    d2 = torch.full((F, 6), 1e-3, dtype=torch.float64, device="cuda", requires_grad=True)
    e2 = reprojection(d2, x1, lin, (prob.points, prob.frames, prob.cams))
    with pytest.raises(RuntimeError, match="relinearized"):
        (e1.sum() + e2.sum()).backward()
    # the latest forward's backward still works
    x3 = torch.zeros((P, 3), dtype=torch.float64, device="cuda", requires_grad=True)
    e3 = reprojection(d1.detach().requires_grad_(), x3, lin, (prob.points, prob.frames, prob.cams))
    e3.sum().backward()
    assert torch.isfinite(x3.grad).all()
    lin.close()


def test_adjoint_at_150k_observations(gpu_ctx):
    """The adjoint at the BA linearization configuration's size (10 keyframes
    x 3 cameras, 20k points, ~150k observations) against the FP64 analytic
    J^T g."""
    import torch
    prob = synth.ba_problem(seed=30, n_kf=10, n_points=20000, target_obs=150000)
    lin = BALinearizer.from_problem(prob, ctx=gpu_ctx)
    F, P, M = len(prob.frames), len(prob.points), len(prob.obs_pt)
    dpose = torch.zeros((F, 6), dtype=torch.float64, device="cuda", requires_grad=True)
    dpoint = torch.zeros((P, 3), dtype=torch.float64, device="cuda", requires_grad=True)
    e = reprojection(dpose, dpoint, lin, (prob.points, prob.frames, prob.cams))
    G = np.random.default_rng(7).normal(size=(M, 2))
    (e * torch.from_numpy(G).float().cuda()).sum().backward()
    _, _, _, Jp, Jx = oracle.ba_linearize(prob)
    want_pose = np.zeros((F, 6))
    want_point = np.zeros((P, 3))
    np.add.at(want_pose, prob.obs_frame, np.einsum("mij,mi->mj", Jp, G))
    np.add.at(want_point, prob.obs_pt, np.einsum("mij,mi->mj", Jx, G))
    for got, want in ((dpose.grad.cpu().numpy(), want_pose), (dpoint.grad.cpu().numpy(), want_point)):
        err = np.linalg.norm(got - want, axis=1) / np.maximum(np.linalg.norm(want, axis=1), 1e-30)
        print(f"adjoint 150k: max rel err {err.max():.2e}", flush=True)
        assert err.max() < 1e-5
    lin.close()
