"""Comparison of the GPU's FP32 BA linearization with the FP64 oracle
(EdgeSE3XYZ2Camera, include/OptimizeTypes.hpp:70-117), shared by the BA
tests."""
import numpy as np

import oracle

TOL = 1e-5


def rel_rows(a, b, floor=0.0):
    """per-row normwise relative error ||a - b|| / max(||b||, floor)"""
    a = a.reshape(len(a), -1).astype(np.float64)
    b = b.reshape(len(b), -1).astype(np.float64)
    num = np.linalg.norm(a - b, axis=1)
    den = np.maximum(np.linalg.norm(b, axis=1), floor)
    return num / den


def check_against_oracle(prob, got):
    e, chi2, w, Jp, Jx = got
    re, rchi2, rw, rJp, rJx = oracle.ba_linearize(prob)
    # residual: normwise relative, floor 1e-9 px only guards e == 0
    assert rel_rows(e, re, 1e-9).max() < TOL
    assert (np.abs(chi2 - rchi2) / np.maximum(rchi2, 1e-12)).max() < TOL
    # Huber boundary (chi2 within 1e-6 of delta^2) may flip the branch
    d2 = prob.cams[prob.obs_cam, 11] ** 2
    ok = np.abs(rchi2 - d2) > 1e-6 * d2
    assert (np.abs(w - rw)[ok] / rw[ok]).max() < TOL
    for r in range(2):
        assert rel_rows(Jp[:, r], rJp[:, r]).max() < TOL, f"J_pose row {r}"
        assert rel_rows(Jx[:, r], rJx[:, r]).max() < TOL, f"J_point row {r}"
