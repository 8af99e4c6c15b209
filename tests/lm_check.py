"""Dense FP64 normal equations of a local BA problem (every free frame and
free point as unknowns), built from the oracle's edge Jacobians with the
Huber weights as constructQuadraticForm sums them
(core/base_binary_edge.hpp:64-136); shared by the LM tests."""
import numpy as np

import oracle
from mcvslam_amd import synth


def dense_system(lp, pts, frs, lam_tau=1e-5, active_m=None, active_s=None, robust=True):
    """H, b over [free frames x 6; free points x 3] from the oracle's edge
    Jacobians, with the Huber weights, as constructQuadraticForm sums them."""
    m, s = lp.mono, lp.stereo
    prob = synth.BAProblem(m.obs_uv, m.obs_pt, m.obs_frame, m.obs_cam, m.obs_level, m.obs_isig, pts, frs, m.cams,
                           m.frame_fixed)
    e, chi2, w, Jp, Jx = oracle.ba_linearize(prob)
    se, schi2, sw, sJp, sJx = oracle.ba_stereo_linearize(s.obs, s.obs_pt, s.obs_frame, s.obs_isig, pts, frs, s.cam)
    F, P = len(frs), len(pts)
    fidx = -np.ones(F, int)
    free_f = np.nonzero(m.frame_fixed == 0)[0]
    fidx[free_f] = np.arange(len(free_f))
    pidx = -np.ones(P, int)
    free_p = np.nonzero(lp.point_fixed == 0)[0]
    pidx[free_p] = np.arange(len(free_p))
    n = 6 * len(free_f) + 3 * len(free_p)
    H = np.zeros((n, n))
    b = np.zeros(n)
    if not robust:
        w, sw = np.ones_like(w), np.ones_like(sw)
    am = np.ones(len(e), bool) if active_m is None else np.asarray(active_m, bool)
    ast = np.ones(len(se), bool) if active_s is None else np.asarray(active_s, bool)
    edges = [(e[i], w[i] * m.obs_isig[i], Jp[i], Jx[i], m.obs_frame[i], m.obs_pt[i]) for i in range(len(e)) if am[i]]
    edges += [(se[i], sw[i] * s.obs_isig[i], sJp[i], sJx[i], s.obs_frame[i], s.obs_pt[i])
              for i in range(len(se)) if ast[i]]
    for ei, W, A, B, f, p in edges:
        J = np.zeros((len(ei), n))
        if fidx[f] >= 0:
            J[:, 6 * fidx[f]:6 * fidx[f] + 6] = A
        if pidx[p] >= 0:
            o = 6 * len(free_f) + 3 * pidx[p]
            J[:, o:o + 3] = B
        H += W * J.T @ J
        b -= W * J.T @ ei
    lam = lam_tau * np.abs(np.diag(H)).max()
    return H, b, lam, free_f, free_p


def dense_schur(H, b, lam, n_pose):
    """Reduced camera system of (H + lam I) x = b after eliminating the
    point unknowns (3x3 diagonal blocks): S, b_schur."""
    Hd = H + lam * np.eye(len(b))
    A, B, D = Hd[:n_pose, :n_pose], Hd[:n_pose, n_pose:], Hd[n_pose:, n_pose:]
    Dinv = np.zeros_like(D)
    for k in range(0, D.shape[0], 3):
        Dinv[k:k + 3, k:k + 3] = np.linalg.inv(D[k:k + 3, k:k + 3])
    return A - B @ Dinv @ B.T, b[:n_pose] - B @ Dinv @ b[n_pose:]
