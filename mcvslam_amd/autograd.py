"""torch.autograd.Function around the multi-camera reprojection residual.

Reprojection.apply(pose_delta [F, 6], point_delta [P, 3], lin, state) -> e [M, 2]

forward applies the increments the way g2o's vertices do (VertexSE3Expmap:
T <- exp(delta) * T, types_six_dof_expmap.h:98-101; VertexSBAPointXYZ:
P <- P + delta, types_sba.h:146-149) to the FP64 state on the host, loads it,
and evaluates EdgeSE3XYZ2Camera::computeError for every edge on the GPU
(include/OptimizeTypes.hpp:80-86).  backward runs the hand-written adjoint
kernels k_ba_pose_adj / k_ba_point_adj: grad_pose[f] = sum of J_pose^T g over
frame f's edges, grad_point[p] = sum of J_point^T g over point p's edges,
with J = linearizeOplus at the evaluated state (:94-113).  That is the exact
derivative of e with respect to the increments at delta = 0, the point where
g2o linearizes.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .api import BALinearizer


def se3_exp(xi: np.ndarray):
    """SE3Quat::exp (ThirdParty/g2o/g2o/types/slam3d/se3quat.h:218-257)."""
    w, u = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th = float(np.linalg.norm(w))
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    W2 = W @ W
    if th < 1e-5:
        R = np.eye(3) + W + 0.5 * W2
        V = np.eye(3) + 0.5 * W + W2 / 6.0
    else:
        R = np.eye(3) + math.sin(th) / th * W + (1 - math.cos(th)) / th ** 2 * W2
        V = np.eye(3) + (1 - math.cos(th)) / th ** 2 * W + (th - math.sin(th)) / th ** 3 * W2
    return R, V @ u


def quat_to_R(q):
    x, y, z, w = np.asarray(q, float) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def R_to_quat(R):
    tr = R[0, 0] + R[1, 1] + R[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        q = [(R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s, 0.25 * s]
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = math.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        q = [0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s, (R[2, 1] - R[1, 2]) / s]
    elif R[1, 1] > R[2, 2]:
        s = math.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        q = [(R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s, (R[0, 2] - R[2, 0]) / s]
    else:
        s = math.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        q = [(R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s, (R[1, 0] - R[0, 1]) / s]
    q = np.array(q)
    return -q if q[3] < 0 else q


def apply_increments(points, frames, pose_delta, point_delta):
    """FP64 oplus of every vertex: frames (F x 7 qxyzw t), points (P x 3)."""
    fr = np.array(frames, np.float64, copy=True)
    for f in range(len(fr)):
        d = pose_delta[f]
        if not np.any(d):
            continue
        dR, dt = se3_exp(d)
        R = quat_to_R(fr[f, :4])
        fr[f, :4] = R_to_quat(dR @ R)
        fr[f, 4:] = dR @ fr[f, 4:] + dt
    return np.asarray(points, np.float64) + point_delta, fr


class Reprojection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pose_delta: torch.Tensor, point_delta: torch.Tensor, lin: BALinearizer, state):
        points, frames, cams = state
        pts, frs = apply_increments(points, frames, pose_delta.detach().double().cpu().numpy(),
                                    point_delta.detach().double().cpu().numpy())
        lin.ctx.use_torch_stream()
        lin.set_state(pts, frs, cams)
        lin.linearize()
        e = torch.empty((2, lin.M), dtype=torch.float32, device=pose_delta.device)
        lin.copy_planes_dev(0, 2, e.data_ptr())
        ctx.lin = lin
        ctx.gen = lin.generation()
        ctx.dtypes = (pose_delta.dtype, point_delta.dtype)
        return e.t().contiguous()

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        lin = ctx.lin
        if lin.generation() != ctx.gen:
            # the planes were relinearized (another forward on the same
            # linearizer, or a state update) since this forward: its Jacobians
            # are gone, and the adjoint would use another state's
            raise RuntimeError("Reprojection.backward: the BALinearizer was relinearized after this forward; "
                               "use one BALinearizer per forward that is still to be differentiated")
        g = g.contiguous().float()
        gpose = torch.empty((lin.F, 6), dtype=torch.float32, device=g.device)
        gpoint = torch.empty((lin.P, 3), dtype=torch.float32, device=g.device)
        lin.ctx.use_torch_stream()
        lin.adjoint_dev(g.data_ptr(), gpose.data_ptr(), gpoint.data_ptr())
        return gpose.to(ctx.dtypes[0]), gpoint.to(ctx.dtypes[1]), None, None


def reprojection(pose_delta, point_delta, lin: BALinearizer, state):
    return Reprojection.apply(pose_delta, point_delta, lin, state)
