"""Projection-guided window search on the GPU against the oracle, bit for
bit (src/Object.cpp:182-236, :249-308), and the window query by position of
Tracker::Wnd_Track (src/Tracker.cpp:339-353)."""
import numpy as np
import pytest

import oracle
from mcvslam_amd import MCVError, synth, window_search
from mcvslam_amd.api import window_search_dev, window_search_uv_dev

pytestmark = pytest.mark.gpu


def _oracle_args(sc, ci, scales, ratio, th):
    cam = sc.cams[ci]
    K = np.array([cam.fx, cam.fy, cam.cx, cam.cy], np.float32)
    return (sc.points, sc.pdesc, sc.plevel, sc.R[ci], sc.t[ci], K, scales, 15.0, ratio, th, sc.kxy[ci],
            sc.koct[ci], sc.kdesc[ci], cam.width, cam.height)


def _gpu_args(oargs):
    # the GPU entry point takes no keypoint octaves (ProjectBunchMapPoints has no level filter)
    return oargs[:11] + oargs[12:]


@pytest.mark.parametrize("ratio,th", [(0.6, 46), (0.7, 50)])
def test_projection_scene(gpu_ctx, ratio, th):
    sc = synth.projection_scene(seed=20, n_points=5000, n_kps=2000)
    scales = oracle.orb_tables()["scale"]
    for ci in range(len(sc.cams)):
        oargs = _oracle_args(sc, ci, scales, ratio, th)
        got = window_search(*_gpu_args(oargs), ctx=gpu_ctx)
        want = oracle.project_search(*oargs)
        np.testing.assert_array_equal(got, want)
        assert (want[:, 5] >= 0).sum() > 100


def test_projection_scene_three_cameras_one_launch(gpu_ctx):
    """mcv_window_search_dev: the three cameras of the rig in one call."""
    import torch
    sc = synth.projection_scene(seed=20, n_points=5000, n_kps=2000)
    scales = oracle.orb_tables()["scale"]
    Rt = np.stack([np.concatenate([np.asarray(sc.R[k], np.float32).ravel(), np.asarray(sc.t[k], np.float32)])
                   for k in range(3)])
    cams = np.array([[c.fx, c.fy, c.cx, c.cy] for c in sc.cams], np.float32)
    sizes = np.array([[c.width, c.height] for c in sc.cams], np.int32)
    dev = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).cuda()
    out = window_search_dev(dev(sc.points, np.float32), dev(sc.pdesc, np.uint8), dev(sc.plevel, np.int32), Rt, cams,
                            sizes, scales, 15.0, 0.7, 50, [dev(k, np.float32) for k in sc.kxy],
                            [dev(k, np.uint8) for k in sc.kdesc], ctx=gpu_ctx)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for ci in range(3):
        np.testing.assert_array_equal(got[ci], oracle.project_search(*_oracle_args(sc, ci, scales, 0.7, 50)),
                                      err_msg=f"camera {ci}")


def test_window_level_out_of_range(gpu_ctx):
    """A map point level outside the scale table: the host entry point
    rejects it, the device one marks the row (n_candidates = -1)."""
    import torch
    sc = synth.projection_scene(seed=21, n_points=300, n_kps=500)
    scales = oracle.orb_tables()["scale"]
    oargs = list(_oracle_args(sc, 0, scales, 0.7, 50))
    bad = np.array(sc.plevel, np.int32)
    bad[[3, 77]] = [8, -1]
    oargs[2] = bad
    with pytest.raises(MCVError):
        window_search(*_gpu_args(tuple(oargs)), ctx=gpu_ctx)
    c = sc.cams[0]
    dev = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).cuda()
    Rt = np.concatenate([np.asarray(sc.R[0], np.float32).ravel(), np.asarray(sc.t[0], np.float32)])[None]
    out = window_search_dev(dev(sc.points, np.float32), dev(sc.pdesc, np.uint8), dev(bad, np.int32), Rt,
                            np.array([[c.fx, c.fy, c.cx, c.cy]], np.float32), np.array([[c.width, c.height]], np.int32),
                            scales, 15.0, 0.7, 50, [dev(sc.kxy[0], np.float32)], [dev(sc.kdesc[0], np.uint8)],
                            ctx=gpu_ctx)
    got = out.cpu().numpy()[0]
    np.testing.assert_array_equal(got[[3, 77]], [[-1, 999, -1, 999, -1, -1]] * 2)
    good = np.ones(len(bad), bool)
    good[[3, 77]] = False
    want = oracle.project_search(*_oracle_args(sc, 0, scales, 0.7, 50))
    np.testing.assert_array_equal(got[good], want[good])


def _wnd_oracle(uv, qdesc, r, kxy, koct, kdesc, w, h, minL, maxL, ratio, th):
    """GetFeaturesInArea + KnnMatch + FilterRatio + FilterThreshold per query,
    composed from the oracle's restatements."""
    rows = np.zeros((len(uv), 6), np.int32)
    for i, (x, y) in enumerate(uv):
        cand = oracle.window(kxy, koct, w, h, float(x), float(y), r, minL, maxL)
        if len(cand) == 0:
            rows[i] = [-1, 999, -1, 999, 0, -1]
            continue
        knn = oracle.knn2(qdesc[i:i + 1], kdesc[cand])[0]
        i0, d0, i1, d1 = (int(v) for v in knn)
        kp0 = int(cand[i0]) if d0 != 999 else -1
        kp1 = int(cand[i1]) if d1 != 999 else -1
        m = oracle.filter_ratio_threshold(knn[None], ratio, th)
        rows[i] = [kp0, d0, kp1, d1, len(cand), kp0 if len(m) else -1]
    return rows


@pytest.mark.parametrize("r,minL,maxL", [(20.0, -1, -1), (10.0, 1, 3), (10.0, 0, 0), (20.0, 2, -1)])
def test_window_query_by_position(gpu_ctx, r, minL, maxL):
    import torch
    sc = synth.projection_scene(seed=22, n_points=2000, n_kps=2000)
    ci = 2
    c = sc.cams[ci]
    kxy, koct, kdesc = (np.asarray(sc.kxy[ci], np.float32), np.asarray(sc.koct[ci], np.int32),
                        np.asarray(sc.kdesc[ci], np.uint8))
    rng = np.random.default_rng(9)
    n = 700
    sel = rng.integers(0, len(kxy), n)
    uv = (kxy[sel] + rng.normal(0, 3, (n, 2))).astype(np.float32)
    uv[:20] = rng.uniform(-30, 700, (20, 2))      # some outside the image
    qdesc = kdesc[sel].copy()
    flips = rng.integers(0, 256, (n, 32), dtype=np.uint8) & rng.integers(0, 256, (n, 32), dtype=np.uint8)
    qdesc ^= np.where(rng.random((n, 1)) < 0.5, flips, 0).astype(np.uint8)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out = window_search_uv_dev(dev(uv), dev(qdesc), r, dev(kxy), dev(koct), dev(kdesc), c.width, c.height, minL, maxL,
                               0.6, 46, ctx=gpu_ctx)
    got = out.cpu().numpy()
    want = _wnd_oracle(uv, qdesc, r, kxy, koct, kdesc, c.width, c.height, minL, maxL, 0.6, 46)
    np.testing.assert_array_equal(got, want)
    assert (want[:, 5] >= 0).sum() > 30      # the test must see matches, not only empty rows


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
        oargs = (pts, pdesc, plevel, R, t, K, scales, 15.0, 0.9, 120, kx, koct[:len(kx)], kd, w, h)
        np.testing.assert_array_equal(window_search(*_gpu_args(oargs), ctx=gpu_ctx), oracle.project_search(*oargs))
