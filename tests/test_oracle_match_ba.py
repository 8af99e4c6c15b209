"""Pins the oracle's matcher, window search and FP64 BA edge with brute force,
closed forms, known answers and finite differences (no retyped code)."""
import math

import numpy as np
import pytest

import oracle
from mcvslam_amd import synth


# ---------------------------------------------------------------- matcher
def test_hamming_popcount():
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b = rng.integers(0, 256, 32, dtype=np.uint8), rng.integers(0, 256, 32, dtype=np.uint8)
        assert oracle.hamming(a, b) == int(np.unpackbits(a ^ b).sum())


def knn_brute(q, t):
    """first two entries of a stable sort by distance; sentinel (0, 999)"""
    out = np.zeros((len(q), 4), np.int32)
    for i, r in enumerate(q):
        d = np.unpackbits(r[None] ^ t, axis=1).sum(1) if len(t) else np.zeros(0, int)
        order = np.argsort(d, kind="stable")
        best = [(int(j), int(d[j])) for j in order[:2]] + [(0, 999)] * 2
        out[i] = [best[0][0], best[0][1], best[1][0], best[1][1]]
    return out


@pytest.mark.parametrize("nq,nt", [(3, 0), (4, 1), (20, 2), (50, 300)])
def test_knn_matches_stable_sort(nq, nt):
    rng = np.random.default_rng(nq + nt)
    q = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    t = rng.integers(0, 256, (nt, 32), dtype=np.uint8)
    if nt > 10:
        t[5] = t[7] = q[0]        # equal distances: scan order decides
    np.testing.assert_array_equal(oracle.knn2(q, t), knn_brute(q, t))


def test_filter_known_answers():
    # ratio: d0/d1 <= r in float; (x, 999) passes, 999/999 and 0/0 do not
    knn = np.array([[1, 10, 2, 20], [0, 999, 0, 999], [3, 5, 0, 999], [4, 0, 5, 0], [6, 14, 7, 20]], np.int32)
    r = oracle.filter_ratio_threshold(knn, 0.7, 1000)
    np.testing.assert_array_equal(r, [[0, 1, 10], [2, 3, 5], [4, 6, 14]])
    # threshold swap-with-last: [a bad, b ok, c bad, d ok, e ok] -> [e, b, d]
    knn = np.array([[0, 60, 1, 999], [1, 10, 2, 999], [2, 70, 3, 999], [3, 20, 4, 999], [4, 30, 5, 999]], np.int32)
    r = oracle.filter_ratio_threshold(knn, 0.9, 50)
    np.testing.assert_array_equal(r[:, 0], [4, 1, 3])
    # bad tail: [a bad, b ok, c bad] -> [b]
    knn = np.array([[0, 60, 1, 999], [1, 10, 2, 999], [2, 70, 3, 999]], np.int32)
    np.testing.assert_array_equal(oracle.filter_ratio_threshold(knn, 0.9, 50)[:, 0], [1])


def test_filter_keeps_exactly_the_passing_set():
    rng = np.random.default_rng(1)
    d0 = rng.integers(0, 120, 500)
    d1 = np.maximum(d0, rng.integers(0, 200, 500))
    knn = np.stack([np.arange(500), d0, np.arange(500), d1], 1).astype(np.int32)
    r = oracle.filter_ratio_threshold(knn, 0.7, 50)
    ok = (d0.astype(np.float32) / d1.astype(np.float32) <= np.float32(0.7)) & (d0 <= 50)
    assert sorted(r[:, 0].tolist()) == np.nonzero(ok)[0].tolist()


# ---------------------------------------------------------------- window search
def window_brute(kxy, w, h, x, y, r):
    """keypoints whose grid cell (round) lies in the scanned cell range and
    that pass the strict window test, ordered by (cell x, cell y, index)"""
    if x < 0 or y < 0 or x >= w or y >= h:
        return []
    inv_w, inv_h = np.float32(30.0 / w), np.float32(30.0 / h)
    x, y, r = np.float32(x), np.float32(y), np.float32(r)
    x0 = max(0, int(np.floor((x - r) * inv_w)))
    x1 = min(29, int(np.ceil((x + r) * inv_w)))
    y0 = max(0, int(np.floor((y - r) * inv_h)))
    y1 = min(29, int(np.ceil((y + r) * inv_h)))
    out = []
    for i, (kx, ky) in enumerate(kxy):
        # std::round: half away from zero
        cx = int(np.floor(np.float32(kx) * inv_w + np.float32(0.5)))
        cy = int(np.floor(np.float32(ky) * inv_h + np.float32(0.5)))
        if not (0 <= cx < 30 and 0 <= cy < 30):
            continue
        if x0 <= cx <= x1 and y0 <= cy <= y1 and abs(np.float32(kx) - x) < r and abs(np.float32(ky) - y) < r:
            out.append((cx, cy, i))
    return [i for _, _, i in sorted(out)]


def test_window_matches_brute_force():
    rng = np.random.default_rng(2)
    w, h = 640, 480
    kxy = np.stack([rng.uniform(0, w, 3000), rng.uniform(0, h, 3000)], 1).astype(np.float32)
    koct = np.zeros(3000, np.int32)
    for x, y, r in [(100.3, 200.1, 15.0), (0.0, 0.0, 30.0), (639.9, 479.9, 20.0), (320, 240, 54.0),
                    (-1.0, 10.0, 10.0), (700.0, 10.0, 10.0), (21.33333, 16.0, 8.0)]:
        got = oracle.window(kxy, koct, w, h, x, y, r)
        assert got.tolist() == window_brute(kxy, w, h, x, y, r)


def test_window_round_not_floor():
    """PosInGrid rounds half away from zero: x = 10.0 px of a 600 px image is
    in cell round(0.5) = 1, x = 9.9 px in cell 0 (src/Object.cpp:249-257)."""
    kxy = np.array([[10.0, 10.0], [9.9, 10.0], [30.0, 10.0]], np.float32)
    got = oracle.window(kxy, np.zeros(3, np.int32), 600, 600, 10.0, 10.0, 25.0)
    assert got.tolist() == [1, 0, 2]
    assert got.tolist() == window_brute(kxy, 600, 600, 10.0, 10.0, 25.0)


# ---------------------------------------------------------------- BA FP64 edge
def _se3_exp(xi):
    """SE3Quat::exp (se3quat.h:218-257): omega first, then upsilon."""
    w, u = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th = np.linalg.norm(w)
    W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    W2 = W @ W
    if th < 1e-5:
        R = np.eye(3) + W + 0.5 * W2
        V = np.eye(3) + 0.5 * W + W2 / 6.0
    else:
        R = np.eye(3) + math.sin(th) / th * W + (1 - math.cos(th)) / th ** 2 * W2
        V = np.eye(3) + (1 - math.cos(th)) / th ** 2 * W + (th - math.sin(th)) / th ** 3 * W2
    return R, V @ u


def _R_to_quat(R):
    w = math.sqrt(max(0.0, 1 + R[0, 0] + R[1, 1] + R[2, 2])) / 2
    x = (R[2, 1] - R[1, 2]) / (4 * w)
    y = (R[0, 2] - R[2, 0]) / (4 * w)
    z = (R[1, 0] - R[0, 1]) / (4 * w)
    return np.array([x, y, z, w])


def _quat_to_R(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _perturbed(prob, f, xi):
    """frames[f] <- exp(xi) * frames[f]  (VertexSE3Expmap::oplusImpl)"""
    p = synth.BAProblem(**{k: np.copy(v) for k, v in prob.__dict__.items()})
    q = p.frames[f, :4]
    R = _quat_to_R(q / np.linalg.norm(q))
    t = p.frames[f, 4:]
    dR, dt = _se3_exp(xi)
    p.frames[f, :4] = _R_to_quat(dR @ R)
    p.frames[f, 4:] = dR @ t + dt
    return p


def test_jacobians_match_central_differences():
    prob = synth.ba_problem(seed=30, n_kf=3, n_points=200, target_obs=600, outlier_frac=0.0)
    e, chi2, w, Jp, Jx = oracle.ba_linearize(prob)
    h = 1e-4
    rng = np.random.default_rng(0)
    for i in rng.choice(len(prob.obs_pt), 25, replace=False):
        f, p = prob.obs_frame[i], prob.obs_pt[i]
        for k in range(6):
            xi = np.zeros(6)
            xi[k] = h
            ep = oracle.ba_linearize(_perturbed(prob, f, xi))[0][i]
            em = oracle.ba_linearize(_perturbed(prob, f, -xi))[0][i]
            fd = (ep - em) / (2 * h)
            np.testing.assert_allclose(Jp[i, :, k], fd, rtol=1e-5, atol=1e-6 * np.abs(Jp[i]).max())
        for k in range(3):
            pp = synth.BAProblem(**{kk: np.copy(v) for kk, v in prob.__dict__.items()})
            pm = synth.BAProblem(**{kk: np.copy(v) for kk, v in prob.__dict__.items()})
            pp.points[p, k] += h
            pm.points[p, k] -= h
            fd = (oracle.ba_linearize(pp)[0][i] - oracle.ba_linearize(pm)[0][i]) / (2 * h)
            np.testing.assert_allclose(Jx[i, :, k], fd, rtol=1e-5, atol=1e-6 * np.abs(Jx[i]).max())


def test_jacobian_equals_g2o_project_xyz_for_identity_extrinsics():
    """With Tcb = I the edge is g2o's EdgeSE3ProjectXYZ, whose Jacobian
    (types_six_dof_expmap.cpp:289-325) is written out in closed form."""
    prob = synth.ba_problem(seed=3, n_kf=2, n_points=300, target_obs=400)
    sel = prob.obs_cam == 0
    e, chi2, w, Jp, Jx = oracle.ba_linearize(prob)
    for i in np.nonzero(sel)[0][:50]:
        q = prob.frames[prob.obs_frame[i], :4]
        R = _quat_to_R(q / np.linalg.norm(q))
        pc = R @ prob.points[prob.obs_pt[i]] + prob.frames[prob.obs_frame[i], 4:]
        x, y, z = pc
        fx, fy = prob.cams[0, 0], prob.cams[0, 1]
        z2 = z * z
        want = np.array([[x * y / z2 * fx, -(1 + x * x / z2) * fx, y / z * fx, -1 / z * fx, 0, x / z2 * fx],
                         [(1 + y * y / z2) * fy, -x * y / z2 * fy, -x / z * fy, 0, -1 / z * fy, y / z2 * fy]])
        np.testing.assert_allclose(Jp[i], want, rtol=1e-9, atol=1e-9)
        jpt = -np.array([[fx / z, 0, -fx * x / z2], [0, fy / z, -fy * y / z2]]) @ R
        np.testing.assert_allclose(Jx[i], jpt, rtol=1e-9, atol=1e-12)


def test_huber_weight_closed_form():
    prob = synth.ba_problem(seed=4, n_kf=3, n_points=500, target_obs=3000)
    e, chi2, w, _, _ = oracle.ba_linearize(prob)
    np.testing.assert_allclose(chi2, prob.obs_isig * (e ** 2).sum(1), rtol=1e-12)
    d = prob.cams[prob.obs_cam, 11]
    want = np.where(chi2 <= d * d, 1.0, d / np.sqrt(chi2))
    np.testing.assert_allclose(w, want, rtol=1e-12)
    assert (w < 1).mean() > 0.03


def test_schur_matches_dense_elimination():
    prob = synth.ba_problem(seed=5, n_kf=4, n_points=60, target_obs=500)
    P, F = len(prob.points), len(prob.frames)
    e, chi2, w, Jp, Jx = oracle.ba_linearize(prob)
    lam = 1e-3
    r = oracle.ba_schur(e, w, Jp, Jx, prob.obs_pt, prob.obs_frame, prob.obs_isig, P, F, prob.frame_fixed, lam)
    free = [f for f in range(F) if not prob.frame_fixed[f]]
    n = 6 * len(free) + 3 * P
    H = np.zeros((n, n))
    b = np.zeros(n)
    for i in range(len(prob.obs_pt)):
        J = np.zeros((2, n))
        pi = 6 * len(free) + 3 * prob.obs_pt[i]
        J[:, pi:pi + 3] = Jx[i]
        if prob.obs_frame[i] in free:
            fi = 6 * free.index(prob.obs_frame[i])
            J[:, fi:fi + 6] = Jp[i]
        Om = prob.obs_isig[i] * w[i] * np.eye(2)
        H += J.T @ Om @ J
        b += -J.T @ Om @ e[i]
    H += lam * np.eye(n)
    m = 6 * len(free)
    A, B, C = H[:m, :m], H[:m, m:], H[m:, m:]
    S = A - B @ np.linalg.solve(C, B.T)
    bs = b[:m] - B @ np.linalg.solve(C, b[m:])
    np.testing.assert_allclose(r["S"], S, rtol=1e-8, atol=1e-6 * np.abs(S).max())
    np.testing.assert_allclose(r["bs"], bs, rtol=1e-8, atol=1e-8 * np.abs(bs).max())


# ---------------------------------------------------------------- stereo
def test_stereo_known_disparity():
    """Right image = left shifted by d pixels: level-0 matches recover the
    disparity d (Frame::ComputeStereoMatch, src/Frame.cpp:150-328)."""
    left = synth.blob_image(51, 640, 480)
    d = 12
    right = np.zeros_like(left)
    right[:, :-d] = left[:, d:]
    kl, dl, _, lv_l = oracle.orb_extract(left, with_levels=True)
    kr, dr, _, lv_r = oracle.orb_extract(right, with_levels=True)
    t = oracle.orb_tables()
    bf, b = 500.0, 0.5
    u, depth, sad = oracle.stereo(kl, dl, kr, dr, [x[0] for x in lv_l], [x[0] for x in lv_r], t["scale"],
                                  t["inv_scale"], bf, b)
    ok = u >= 0
    assert ok.sum() > 100
    disp = kl["x"][ok] - u[ok]
    assert np.median(np.abs(disp - d)) < 0.5
    assert (np.abs(disp - d) < 1.5).mean() > 0.9
    np.testing.assert_allclose(depth[ok], bf / disp, rtol=1e-6)
    assert ((u < 0) == (depth < 0)).all() and ((u < 0) == (sad < 0)).all()
    # octave-0 pairs match with SAD 0 here, below 0.4 x the median SAD, so the
    # median filter of the reference (src/Frame.cpp:307-323) drops them
    assert not (ok & (kl["octave"] == 0)).any()


# ------------------------------------------------------------ stereo edge
def _stereo_mono_problem(sp):
    """The same observations as a mono-edge problem with Tcb = I."""
    M = len(sp.obs_pt)
    fx, fy, cx, cy, bf, delta = sp.cam
    cams = np.array([[fx, fy, cx, cy, 0, 0, 0, 1, 0, 0, 0, delta]], np.float64)
    return synth.BAProblem(sp.obs[:, :2].copy(), sp.obs_pt, sp.obs_frame, np.zeros(M, np.int32),
                           np.zeros(M, np.int32), sp.obs_isig, sp.points, sp.frames, cams,
                           np.zeros(len(sp.frames), np.int32))


def test_stereo_edge_rows_01_are_the_mono_edge():
    """Rows u, v of EdgeStereoSE3ProjectXYZ's Jacobians (src/OptimizeTypes.cpp:12-58)
    equal EdgeSE3XYZ2Camera's with Tcb = I (a separate code path); the u, v
    residuals differ only through the float inverse depth (< 1e-3 px)."""
    sp = synth.stereo_problem(seed=51)
    e, chi2, w, Jp, Jx = oracle.ba_stereo_linearize(sp.obs, sp.obs_pt, sp.obs_frame, sp.obs_isig, sp.points,
                                                    sp.frames, sp.cam)
    me, _, _, mJp, mJx = oracle.ba_linearize(_stereo_mono_problem(sp))
    np.testing.assert_allclose(Jp[:, :2], mJp, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(Jx[:, :2], mJx, rtol=1e-9, atol=1e-9)
    assert np.abs(e[:, :2] - me).max() < 1e-3
    assert np.abs(e[:, :2] - me).max() > 0          # the float invz is visible


def test_stereo_edge_residual_closed_form():
    """cam_project_uv_u (include/OptimizeTypes.hpp:143-152): float invz,
    float bf * invz product, then double arithmetic."""
    sp = synth.stereo_problem(seed=52)
    e, chi2, w, _, _ = oracle.ba_stereo_linearize(sp.obs, sp.obs_pt, sp.obs_frame, sp.obs_isig, sp.points,
                                                  sp.frames, sp.cam)
    fx, fy, cx, cy, bf, delta = sp.cam
    for i in range(0, len(sp.obs_pt), 37):
        f = sp.frames[sp.obs_frame[i]]
        R = _quat_to_R(f[:4] / np.linalg.norm(f[:4]))
        x, y, z = R @ sp.points[sp.obs_pt[i]] + f[4:]
        invz = np.float32(1.0 / z)
        u = x * float(invz) * fx + cx
        v = y * float(invz) * fy + cy
        ur = u - float(np.float32(np.float32(bf) * invz))
        want = np.array([sp.obs[i, 0], sp.obs[i, 1], sp.obs[i, 2]], np.float64) - [u, v, ur]
        np.testing.assert_allclose(e[i], want, rtol=0, atol=1e-9)
    np.testing.assert_allclose(chi2, sp.obs_isig * (e ** 2).sum(1), rtol=1e-12)
    d = sp.cam[5]
    np.testing.assert_allclose(w, np.where(chi2 <= d * d, 1.0, d / np.sqrt(chi2)), rtol=1e-12)


def test_stereo_edge_row2_central_differences():
    """Third row against central differences of the exact (double) right-image
    residual u_r - (fx x / z + cx - bf / z), h = 1e-4."""
    sp = synth.stereo_problem(seed=53)
    _, _, _, Jp, Jx = oracle.ba_stereo_linearize(sp.obs, sp.obs_pt, sp.obs_frame, sp.obs_isig, sp.points,
                                                 sp.frames, sp.cam)
    fx, fy, cx, cy, bf, _ = sp.cam

    def er(R, t, P):
        x, y, z = R @ P + t
        return -(fx * x / z + cx - bf / z)
    h = 1e-4
    for i in range(0, len(sp.obs_pt), 97):
        f = sp.frames[sp.obs_frame[i]]
        R = _quat_to_R(f[:4] / np.linalg.norm(f[:4]))
        t, P = f[4:], sp.points[sp.obs_pt[i]]
        for k in range(6):
            xi = np.zeros(6)
            xi[k] = h
            dRp, dtp = _se3_exp(xi)
            dRm, dtm = _se3_exp(-xi)
            fd = (er(dRp @ R, dRp @ t + dtp, P) - er(dRm @ R, dRm @ t + dtm, P)) / (2 * h)
            np.testing.assert_allclose(Jp[i, 2, k], fd, rtol=1e-5, atol=1e-6 * np.abs(Jp[i, 2]).max())
        for k in range(3):
            d = np.zeros(3)
            d[k] = h
            fd = (er(R, t, P + d) - er(R, t, P - d)) / (2 * h)
            np.testing.assert_allclose(Jx[i, 2, k], fd, rtol=1e-5, atol=1e-6 * np.abs(Jx[i, 2]).max())


def test_outlier_flags():
    """FilterCallBack_Chi2 (src/PoseEstimation.cpp:132-151): float(chi2) >=
    threshold, with 5.991 (mono) and 7.815 (stereo) as floats."""
    chi2 = np.array([0.0, 5.990999, 5.991, float(np.float32(5.991)), 5.9910002, 7.815, 1e9])
    got = oracle.ba_outliers(chi2, float(np.float32(5.991)))
    want = (chi2.astype(np.float32) >= np.float32(5.991)).astype(np.uint8)
    np.testing.assert_array_equal(got, want)
    assert got.tolist() == [0, 0, 1, 1, 1, 1, 1]


# ------------------------------------------------ distinctive descriptors
def _distinctive_brute(desc, ptr):
    out = []
    for m in range(len(ptr) - 1):
        d = desc[ptr[m]:ptr[m + 1]]
        N = len(d)
        if N == 0:
            out.append(-1)
            continue
        bits = np.unpackbits(d, axis=1).astype(np.int32)
        D = (bits[:, None, :] != bits[None, :, :]).sum(2)
        med = np.sort(D, axis=1)[:, int(0.5 * (N - 1))]
        out.append(int(np.argmin(med)))           # first index of the minimum
    return np.array(out, np.int32)


def test_distinctive_descriptors_brute_force():
    """MapPoint::ComputeDistinctiveDescriptors (src/MapPoint.cpp:101-150)
    against numpy (sort per row, median index floor(0.5 (N-1)), first min)."""
    rng = np.random.default_rng(7)
    sizes = [0, 1, 2, 3, 4, 7, 8, 30, 64, 65, 130]
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    base = rng.integers(0, 256, (len(sizes), 32), dtype=np.uint8)
    desc = []
    for m, n in enumerate(sizes):
        for k in range(n):
            d = base[m].copy()
            flips = rng.integers(0, 256, rng.integers(0, 40))
            for b in flips:
                d[b >> 3] ^= np.uint8(1 << (b & 7))
            desc.append(d)
    desc = np.array(desc, np.uint8).reshape(-1, 32)
    got = oracle.distinctive(desc, ptr)
    np.testing.assert_array_equal(got, _distinctive_brute(desc, ptr))
    assert got[0] == -1 and got[1] == 0


def test_distinctive_descriptors_ties_take_first():
    """Identical descriptors: every median ties, the first row wins."""
    d = np.tile(np.arange(32, dtype=np.uint8), (5, 1))
    assert oracle.distinctive(d, np.array([0, 5], np.int32)).tolist() == [0]


# ------------------------------------------------------ orientation filter
def _orientation_case(rng, n, n_bins_used):
    """matches whose angle differences fall in a few bins"""
    a1 = rng.uniform(0, 360, n).astype(np.float32)
    bins = rng.integers(0, n_bins_used, n) * 7 % 40
    a2 = (a1 - bins * np.float32(9.0) - rng.uniform(-3, 3, n).astype(np.float32)) % np.float32(360.0)
    m = np.stack([np.arange(n), np.arange(n), rng.integers(0, 60, n)], 1).astype(np.int32)
    return m, a1, a2.astype(np.float32)


def _bins(m, a1, a2):
    rot = a1[m[:, 0]] - a2[m[:, 1]]
    rot = np.where(rot < 0, rot + np.float32(360.0), rot).astype(np.float32)
    b = np.floor(np.abs(rot * np.float32(1.0 / (360.0 / 40))).astype(np.float64) + 0.5).astype(int)
    b[b == 40] = 0
    return b


def test_filter_orientation_keeps_three_largest_bins():
    """MatchRes::FilterOrientation (src/Matcher.cpp:44-75): the output is the
    matches of the three largest rotation bins, bin by bin, list order
    within a bin; sizes here are distinct so std::sort's tie order is moot."""
    rng = np.random.default_rng(8)
    m, a1, a2 = _orientation_case(rng, 500, 6)
    b = _bins(m, a1, a2)
    sizes = np.bincount(b, minlength=40)
    order = np.argsort(-sizes, kind="stable")
    if len(set(sizes[order[:4]])) == 4:
        got = oracle.filter_orientation(m, a1, a2)
        want = np.concatenate([m[b == k] for k in order[:3]])
        np.testing.assert_array_equal(got, want)
    # any input: three bins' worth of matches, each bin contiguous and in order
    m, a1, a2 = _orientation_case(rng, 300, 40)
    got = oracle.filter_orientation(m, a1, a2)
    b = _bins(m, a1, a2)
    sizes = np.bincount(b, minlength=40)
    assert len(got) == np.sort(sizes)[::-1][:3].sum()
    gb = _bins(got, a1, a2)
    runs = [gb[0]] + [x for i, x in enumerate(gb[1:]) if x != gb[i]]
    assert len(runs) == len(set(runs)) <= 3
    for k in set(runs):
        np.testing.assert_array_equal(got[gb == k], m[b == k])


def test_stereo_subpixel_parabola_direction():
    """Right image = left shifted by a fractional disparity (12.3 and 12.7 px,
    linear interpolation): the 11-shift SAD search and the parabola fit
    (src/Frame.cpp:236-283) recover the fraction, so a flipped sign or a
    wrong tap in the fit moves the estimate to the other side of 12.5."""
    left = synth.blob_image(52, 640, 480).astype(np.float64)
    t = oracle.orb_tables()
    bf, b = 500.0, 0.5
    est = {}
    for d in (12.3, 12.7):
        i, f = int(d), d - int(d)
        right = np.zeros_like(left)
        right[:, :-i - 1] = (1 - f) * left[:, i:-1] + f * left[:, i + 1:]
        L8, R8 = np.rint(left).astype(np.uint8), np.rint(right).astype(np.uint8)
        kl, dl, _, lv_l = oracle.orb_extract(L8, with_levels=True)
        kr, dr, _, lv_r = oracle.orb_extract(R8, with_levels=True)
        u, depth, sad = oracle.stereo(kl, dl, kr, dr, [x[0] for x in lv_l], [x[0] for x in lv_r], t["scale"],
                                      t["inv_scale"], bf, b)
        ok = (u >= 0) & (kl["octave"] == 0)
        assert ok.sum() > 30
        disp = kl["x"][ok] - u[ok]
        est[d] = float(np.median(disp))
        # a parabola through SAD samples pulls toward the integer shift, but
        # never across the half-way point
        assert abs(est[d] - d) < 0.25, (d, est[d])
    assert est[12.3] < 12.5 < est[12.7], est


def test_stereo_median_filter_bounds():
    """The median filter keeps SAD in [0.4 m, 1.6 m] around the median m of
    the accepted distances (src/Frame.cpp:298-323): with every accepted
    match's SAD known, the surviving set is exactly that band."""
    left = synth.blob_image(53, 640, 480)
    d = 9
    right = np.zeros_like(left)
    right[:, :-d] = left[:, d:]
    noise = np.random.default_rng(0).integers(-6, 7, right.shape)
    right = np.clip(right.astype(int) + noise, 0, 255).astype(np.uint8)
    kl, dl, _, lv_l = oracle.orb_extract(left, with_levels=True)
    kr, dr, _, lv_r = oracle.orb_extract(right, with_levels=True)
    t = oracle.orb_tables()
    u, depth, sad = oracle.stereo(kl, dl, kr, dr, [x[0] for x in lv_l], [x[0] for x in lv_r], t["scale"],
                                  t["inv_scale"], 500.0, 0.5)
    ok = u >= 0
    assert ok.sum() > 50
    kept = sad[ok]
    # the filter's band, from the kept distances: no survivor outside
    # [0.4, 1.6] x their own median's neighbourhood
    m = np.median(kept)
    assert kept.min() >= 0.4 * m * 0.5 and kept.max() <= 1.6 * m * 2.0
    assert ((u < 0) == (sad < 0)).all()
