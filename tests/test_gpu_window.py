"""Projection-guided window search on the GPU against the oracle, bit for
bit (src/Object.cpp:182-236, :249-308)."""
import numpy as np
import pytest

import oracle
from mcvslam_amd import synth, window_search

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ratio,th", [(0.6, 46), (0.7, 50)])
def test_projection_scene(gpu_ctx, ratio, th):
    sc = synth.projection_scene(seed=20, n_points=5000, n_kps=2000)
    scales = oracle.orb_tables()["scale"]
    for ci, cam in enumerate(sc.cams):
        K = np.array([cam.fx, cam.fy, cam.cx, cam.cy], np.float32)
        args = (sc.points, sc.pdesc, sc.plevel, sc.R[ci], sc.t[ci], K, scales, 15.0, ratio, th, sc.kxy[ci],
                sc.koct[ci], sc.kdesc[ci], cam.width, cam.height)
        got = window_search(*args, ctx=gpu_ctx)
        want = oracle.project_search(*args)
        np.testing.assert_array_equal(got, want)
        assert (want[:, 5] >= 0).sum() > 100


def test_window_edges(gpu_ctx):
    """Points behind the camera, projections outside the image, keypoints on
    grid-cell rounding boundaries, empty keypoint set."""
    rng = np.random.default_rng(5)
    w, h = 640, 480
    P = 600
    pts = np.stack([rng.uniform(-8, 8, P), rng.uniform(-6, 6, P), rng.uniform(-2, 6, P)], 1).astype(np.float32)
    pdesc = rng.integers(0, 256, (P, 32), dtype=np.uint8)
    plevel = rng.integers(0, 8, P).astype(np.int32)
    # keypoints exactly on multiples of the cell pitch and half pitch
    gx = np.arange(0, w, w / 60.0, dtype=np.float32)
    gy = np.arange(0, h, h / 60.0, dtype=np.float32)
    xx, yy = np.meshgrid(gx[::3], gy[::3])
    kxy = np.stack([xx.ravel(), yy.ravel()], 1).astype(np.float32)
    kdesc = rng.integers(0, 256, (len(kxy), 32), dtype=np.uint8)
    koct = np.zeros(len(kxy), np.int32)
    R = np.eye(3, dtype=np.float32)
    t = np.zeros(3, np.float32)
    K = np.array([300, 300, w / 2, h / 2], np.float32)
    scales = oracle.orb_tables()["scale"]
    for kx, kd in [(kxy, kdesc), (kxy[:0], kdesc[:0])]:
        args = (pts, pdesc, plevel, R, t, K, scales, 15.0, 0.9, 120, kx, koct[:len(kx)], kd, w, h)
        np.testing.assert_array_equal(window_search(*args, ctx=gpu_ctx), oracle.project_search(*args))
