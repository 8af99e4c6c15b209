"""CPU ORACLE (test infrastructure only): ctypes binding of oracle/_build/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker of (or the CPU baseline beside) the HIP path.
See mcv_oracle.cpp for what each function restates and where it is pinned.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double

KP_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"), ("response", "<f4"),
                     ("octave", "<i4")])

_SIG = {
    "orc_orb_tables": (None, [I, F, I, I, I, P, P, P, P, P, P, P]),
    "orc_resize": (None, [P, I, I, P, I, I]),
    "orc_blur": (None, [P, I, I, P, P]),
    "orc_fast": (I, [P, I, I, I, I, P, I]),
    "orc_fast_atan2": (F, [F, F]),
    "orc_octree": (I, [P, I, I, I, I, I, I, P, I]),
    "orc_orb_extract": (I, [P, I, I, I, F, I, I, I, P, I, P, P, I, P, P, P, P, I]),
    "orc_stereo": (None, [P, P, I, P, P, I, P, P, P, P, I, P, P, F, F, P, P, P]),
    "orc_ic_angle": (F, [P, I, I, F, F]),
    "orc_descriptor": (None, [P, I, I, F, F, F, P]),
    "orc_libm_sincos": (None, [P, I, P, P]),
    "orc_knn2": (None, [P, I, P, I, P]),
    "orc_hamming": (ctypes.c_uint, [P, P]),
    "orc_filter": (I, [P, I, F, I, P]),
    "orc_window": (I, [P, P, I, I, I, F, F, F, I, I, P, I]),
    "orc_project_search": (None, [P, P, P, I, P, P, P, P, F, F, I, P, P, P, I, I, I, P]),
    "orc_ba_linearize": (None, [P, P, P, P, P, I, P, P, P, P, P, P, P, P]),
    "orc_ba_stereo_linearize": (None, [P, P, P, P, I, P, P, P, P, P, P, P, P]),
    "orc_ba_outliers": (None, [P, I, F, P]),
    "orc_distinctive": (None, [P, P, I, P]),
    "orc_filter_orientation": (I, [P, I, P, P, P]),
    "orc_ba_schur": (I, [P, P, P, P, P, P, P, I, I, I, P, D, P, P, P, P, P, P]),
    "orc_se3_oplus": (None, [P, P]),
    "orc_bow_transform": (None, [P, P, P, P, P, I, I, P, I, P, P, P]),
    "orc_bow_vectors": (I, [P, P, P, I, I, I, P, P, P, P, P, P]),
    "orc_bow_match": (I, [P, I, P, P, P, P, I, P, P, P, P]),
    "orc_ba_lm": (I, [I, P, P, P, P, P, P, I, P, P, P, P, P, P, I, I, I, P, P, P, I, P, P, I, D, I, P, P]),
    "orc_ba_robust_chi2": (D, [I, P, P, P, P, P, P, I, P, P, P, P, P, P, I, I, I, P, I, P, P]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = ctypes.CDLL(str(LIB))
        for n, (r, a) in _SIG.items():
            f = getattr(l, n)
            f.restype = r
            f.argtypes = a
        _lib = l
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def orb_tables(nfeatures=2000, scale=1.2, nlevels=8, ini=20, mn=7):
    t = [np.zeros(nlevels, np.float32) for _ in range(4)]
    npl = np.zeros(nlevels, np.int32)
    umax = np.zeros(16, np.int32)
    g = np.zeros(7, np.int32)
    lib().orc_orb_tables(nfeatures, scale, nlevels, ini, mn, *[_p(x) for x in t], _p(npl), _p(umax), _p(g))
    return dict(scale=t[0], inv_scale=t[1], sigma2=t[2], inv_sigma2=t[3], npl=npl, umax=umax, gauss=g)


def resize(src: np.ndarray, dw: int, dh: int) -> np.ndarray:
    s = np.ascontiguousarray(src, np.uint8)
    out = np.zeros((dh, dw), np.uint8)
    lib().orc_resize(_p(s), s.shape[1], s.shape[0], _p(out), dw, dh)
    return out


def blur(src: np.ndarray, taps=None) -> np.ndarray:
    s = np.ascontiguousarray(src, np.uint8)
    t = np.asarray(taps if taps is not None else orb_tables()["gauss"], np.int32)
    out = np.zeros_like(s)
    lib().orc_blur(_p(s), s.shape[1], s.shape[0], _p(t), _p(out))
    return out


def fast(window: np.ndarray, th: int) -> np.ndarray:
    w = np.ascontiguousarray(window, np.uint8)
    cap = w.size
    out = np.zeros((cap, 3), np.float32)
    n = lib().orc_fast(_p(w), w.shape[1], w.shape[1], w.shape[0], th, _p(out), cap)
    return out[:n].copy()


def fast_atan2(y: float, x: float) -> float:
    return lib().orc_fast_atan2(y, x)


def octree(xyr: np.ndarray, minX, maxX, minY, maxY, N) -> np.ndarray:
    a = np.ascontiguousarray(xyr, np.float32).reshape(-1, 3)
    cap = max(N + 8, 64)
    out = np.zeros((cap, 3), np.float32)
    n = lib().orc_octree(_p(a), len(a), minX, maxX, minY, maxY, N, _p(out), cap)
    return out[:n].copy()


def orb_extract(img: np.ndarray, nfeatures=2000, scale=1.2, nlevels=8, ini=20, mn=7, seeds=None, with_levels=False):
    im = np.ascontiguousarray(img, np.uint8)
    h, w = im.shape if im.size else (0, 0)
    sd = None
    ns = 0
    if seeds is not None and len(seeds):
        sd = np.ascontiguousarray(np.stack([seeds["x"], seeds["y"], seeds["size"], seeds["angle"],
                                            seeds["response"], seeds["octave"].astype(np.float32)], 1), np.float32)
        ns = len(sd)
    cap = nfeatures + 8 * nlevels + ns + 64
    kps = np.zeros((cap, 6), np.float32)
    desc = np.zeros((cap, 32), np.uint8)
    info = np.zeros((nlevels, 4), np.int32)
    pyr = blurred = None
    if with_levels and im.size:
        t = orb_tables(nfeatures, scale, nlevels, ini, mn)
        total = 0
        for l in range(nlevels):
            lw = w if l == 0 else int(np.rint(np.float32(w) * t["inv_scale"][l]))
            lh = h if l == 0 else int(np.rint(np.float32(h) * t["inv_scale"][l]))
            total += lw * lh
        pyr = np.zeros(total, np.uint8)
        blurred = np.zeros(total, np.uint8)
    n = lib().orc_orb_extract(_p(im) if im.size else None, w, h, nfeatures, scale, nlevels, ini, mn, _p(sd), ns,
                              _p(kps), _p(desc), cap, _p(info), _p(pyr), _p(blurred), None, 0)
    if n < 0:
        return n
    rec = np.zeros(n, KP_DTYPE)
    for i, f in enumerate(["x", "y", "size", "angle", "response"]):
        rec[f] = kps[:n, i]
    rec["octave"] = kps[:n, 5].astype(np.int32)
    out = (rec, desc[:n].copy())
    if with_levels:
        levels = []
        off = 0
        for l in range(nlevels):
            lw, lh = info[l, 0], info[l, 1]
            levels.append((pyr[off:off + lw * lh].reshape(lh, lw), blurred[off:off + lw * lh].reshape(lh, lw)))
            off += lw * lh
        out = out + (info, levels)
    return out


def orb_candidates(img: np.ndarray, nfeatures=2000, scale=1.2, nlevels=8, ini=20, mn=7):
    """Per-level FAST candidates (x, y, response, level) before the octree."""
    im = np.ascontiguousarray(img, np.uint8)
    h, w = im.shape
    cap = 4 * nfeatures + 8 * nlevels
    kps = np.zeros((cap, 6), np.float32)
    desc = np.zeros((cap, 32), np.uint8)
    info = np.zeros((nlevels, 4), np.int32)
    ccap = w * h
    cand = np.zeros((ccap, 4), np.float32)
    lib().orc_orb_extract(_p(im), w, h, nfeatures, scale, nlevels, ini, mn, None, 0, _p(kps), _p(desc), cap,
                          _p(info), None, None, _p(cand), ccap)
    return cand[:int(info[:, 2].sum())].copy(), info


def _kp_rows(k):
    return np.ascontiguousarray(np.stack([k["x"], k["y"], k["size"], k["angle"], k["response"],
                                          k["octave"].astype(np.float32)], 1), np.float32)


def stereo(kl, dl, kr, dr, levels_l, levels_r, scale, inv_scale, bf, b):
    """Frame::ComputeStereoMatch: (u_right, depth, sad) per left keypoint.
    levels_*: list of pyramid level arrays (as orb_extract(with_levels=True))."""
    lw = np.array([l.shape[1] for l in levels_l], np.int32)
    lh = np.array([l.shape[0] for l in levels_l], np.int32)
    pl = np.ascontiguousarray(np.concatenate([l.ravel() for l in levels_l]), np.uint8)
    pr = np.ascontiguousarray(np.concatenate([l.ravel() for l in levels_r]), np.uint8)
    n = len(kl)
    u = np.zeros(n, np.float32)
    d = np.zeros(n, np.float32)
    s = np.zeros(n, np.int32)
    lib().orc_stereo(_p(_kp_rows(kl)), _p(np.ascontiguousarray(dl, np.uint8)), n, _p(_kp_rows(kr)),
                     _p(np.ascontiguousarray(dr, np.uint8)), len(kr), _p(pl), _p(pr), _p(lw), _p(lh), len(lw),
                     _p(np.ascontiguousarray(scale, np.float32)), _p(np.ascontiguousarray(inv_scale, np.float32)),
                     bf, b, _p(u), _p(d), _p(s))
    return u, d, s


def ic_angle(img: np.ndarray, x: float, y: float) -> float:
    im = np.ascontiguousarray(img, np.uint8)
    return float(lib().orc_ic_angle(_p(im), im.shape[1], im.shape[0], x, y))


def descriptor(img: np.ndarray, x: float, y: float, angle: float) -> np.ndarray:
    im = np.ascontiguousarray(img, np.uint8)
    out = np.zeros(32, np.uint8)
    lib().orc_descriptor(_p(im), im.shape[1], im.shape[0], x, y, angle, _p(out))
    return out


def libm_sincos(x: np.ndarray):
    xs = np.ascontiguousarray(x, np.float32)
    c = np.zeros_like(xs)
    s = np.zeros_like(xs)
    lib().orc_libm_sincos(_p(xs), len(xs), _p(c), _p(s))
    return c, s


def knn2(q: np.ndarray, t: np.ndarray) -> np.ndarray:
    q = np.ascontiguousarray(q, np.uint8).reshape(-1, 32)
    t = np.ascontiguousarray(t, np.uint8).reshape(-1, 32)
    out = np.zeros((len(q), 4), np.int32)
    lib().orc_knn2(_p(q), len(q), _p(t), len(t), _p(out))
    return out


def hamming(a: np.ndarray, b: np.ndarray) -> int:
    return int(lib().orc_hamming(_p(np.ascontiguousarray(a, np.uint8)), _p(np.ascontiguousarray(b, np.uint8))))


def filter_ratio_threshold(knn: np.ndarray, ratio: float, th: int) -> np.ndarray:
    k = np.ascontiguousarray(knn, np.int32).reshape(-1, 4)
    out = np.zeros((max(len(k), 1), 3), np.int32)
    n = lib().orc_filter(_p(k), len(k), ratio, th, _p(out))
    return out[:n].copy()


def window(kxy, koct, w, h, x, y, r, minL=-1, maxL=-1) -> np.ndarray:
    kx = np.ascontiguousarray(kxy, np.float32).reshape(-1, 2)
    ko = np.ascontiguousarray(koct, np.int32)
    out = np.zeros(max(len(kx), 1), np.int32)
    n = lib().orc_window(_p(kx), _p(ko), len(kx), w, h, x, y, r, minL, maxL, _p(out), len(out))
    return out[:n].copy()


def project_search(points, pdesc, plevel, R, t, cam, scales, r, ratio, th, kxy, koct, kdesc, w, h) -> np.ndarray:
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.zeros((len(pts), 6), np.int32)
    lib().orc_project_search(_p(pts), _p(np.ascontiguousarray(pdesc, np.uint8)), _p(np.ascontiguousarray(plevel, np.int32)),
                             len(pts), _p(np.ascontiguousarray(R, np.float32)), _p(np.ascontiguousarray(t, np.float32)),
                             _p(np.ascontiguousarray(cam, np.float32)), _p(np.ascontiguousarray(scales, np.float32)),
                             r, ratio, th, _p(np.ascontiguousarray(kxy, np.float32)),
                             _p(np.ascontiguousarray(koct, np.int32)), _p(np.ascontiguousarray(kdesc, np.uint8)),
                             len(kxy), w, h, _p(out))
    return out


def ba_linearize(prob):
    M = len(prob.obs_pt)
    e = np.zeros((M, 2))
    chi2 = np.zeros(M)
    w = np.zeros(M)
    Jp = np.zeros((M, 12))
    Jx = np.zeros((M, 6))
    lib().orc_ba_linearize(_p(np.ascontiguousarray(prob.obs_uv, np.float32)),
                           _p(np.ascontiguousarray(prob.obs_pt, np.int32)),
                           _p(np.ascontiguousarray(prob.obs_frame, np.int32)),
                           _p(np.ascontiguousarray(prob.obs_cam, np.int32)),
                           _p(np.ascontiguousarray(prob.obs_isig, np.float32)), M,
                           _p(np.ascontiguousarray(prob.points, np.float64)),
                           _p(np.ascontiguousarray(prob.frames, np.float64)),
                           _p(np.ascontiguousarray(prob.cams, np.float64)),
                           _p(e), _p(chi2), _p(w), _p(Jp), _p(Jx))
    return e, chi2, w, Jp.reshape(M, 2, 6), Jx.reshape(M, 2, 3)


def ba_stereo_linearize(obs, obs_pt, obs_frame, obs_isig, points, frames, cam):
    """EdgeStereoSE3ProjectXYZ for M edges; cam = (fx, fy, cx, cy, bf, delta).
    Returns e (M x 3), chi2, w, J_pose (M x 3 x 6), J_point (M x 3 x 3)."""
    M = len(obs_pt)
    e = np.zeros((M, 3))
    chi2 = np.zeros(M)
    w = np.zeros(M)
    Jp = np.zeros((M, 18))
    Jx = np.zeros((M, 9))
    lib().orc_ba_stereo_linearize(_p(np.ascontiguousarray(obs, np.float32)),
                                  _p(np.ascontiguousarray(obs_pt, np.int32)),
                                  _p(np.ascontiguousarray(obs_frame, np.int32)),
                                  _p(np.ascontiguousarray(obs_isig, np.float32)), M,
                                  _p(np.ascontiguousarray(points, np.float64)),
                                  _p(np.ascontiguousarray(frames, np.float64)),
                                  _p(np.ascontiguousarray(cam, np.float64)),
                                  _p(e), _p(chi2), _p(w), _p(Jp), _p(Jx))
    return e, chi2, w, Jp.reshape(M, 3, 6), Jx.reshape(M, 3, 3)


def distinctive(desc, ptr):
    """ComputeDistinctiveDescriptors per map point: desc (total x 32 u8),
    ptr (n + 1) CSR offsets -> best row index per map point (-1 if empty)."""
    desc = np.ascontiguousarray(desc, np.uint8)
    ptr = np.ascontiguousarray(ptr, np.int32)
    out = np.zeros(len(ptr) - 1, np.int32)
    lib().orc_distinctive(_p(desc), _p(ptr), len(ptr) - 1, _p(out))
    return out


def filter_orientation(matches, angles1, angles2):
    """MatchRes::FilterOrientation: rows (query, train, distance) kept from
    the three largest rotation bins, in libstdc++ std::sort bin order."""
    m = np.ascontiguousarray(matches, np.int32).reshape(-1, 3)
    out = np.zeros_like(m)
    k = lib().orc_filter_orientation(_p(m), len(m), _p(np.ascontiguousarray(angles1, np.float32)),
                                     _p(np.ascontiguousarray(angles2, np.float32)), _p(out))
    if k < 0:
        raise ValueError("rotation bin out of range")
    return out[:k]


def ba_outliers(chi2, threshold):
    """FilterCallBack_Chi2's test: float(chi2) >= threshold."""
    chi2 = np.ascontiguousarray(chi2, np.float64)
    out = np.zeros(len(chi2), np.uint8)
    lib().orc_ba_outliers(_p(chi2), len(chi2), ctypes.c_float(threshold), _p(out))
    return out


def ba_schur(e, w, Jpose, Jpoint, obs_pt, obs_frame, obs_isig, P, F, frame_fixed, lam=0.0):
    M = len(obs_pt)
    Fp = int((np.asarray(frame_fixed) == 0).sum())
    n = 6 * Fp
    S = np.zeros((n, n))
    bs = np.zeros(n)
    Hll = np.zeros((P, 9))
    bl = np.zeros((P, 3))
    Hpp = np.zeros((n, n))
    bp = np.zeros(n)
    r = lib().orc_ba_schur(_p(np.ascontiguousarray(e, np.float64)), _p(np.ascontiguousarray(w, np.float64)),
                           _p(np.ascontiguousarray(Jpose, np.float64).reshape(M, 12)),
                           _p(np.ascontiguousarray(Jpoint, np.float64).reshape(M, 6)),
                           _p(np.ascontiguousarray(obs_pt, np.int32)), _p(np.ascontiguousarray(obs_frame, np.int32)),
                           _p(np.ascontiguousarray(obs_isig, np.float32)), M, P, F,
                           _p(np.ascontiguousarray(frame_fixed, np.int32)), lam, _p(S), _p(bs), _p(Hll), _p(bl),
                           _p(Hpp), _p(bp))
    return dict(S=S, bs=bs, Hll=Hll.reshape(P, 3, 3), bl=bl, Hpp=Hpp, bp=bp, ok=r >= 0)


def se3_oplus(frame, xi):
    """VertexSE3Expmap::oplusImpl: exp(xi) * T for one frame (qx qy qz qw t)."""
    f = np.ascontiguousarray(frame, np.float64).copy()
    lib().orc_se3_oplus(_p(f), _p(np.ascontiguousarray(xi, np.float64)))
    return f


def _lm_args(prob, stereo, active, sactive):
    M = len(prob.obs_pt)
    a = [np.ascontiguousarray(prob.obs_uv, np.float32), np.ascontiguousarray(prob.obs_pt, np.int32),
         np.ascontiguousarray(prob.obs_frame, np.int32), np.ascontiguousarray(prob.obs_cam, np.int32),
         np.ascontiguousarray(prob.obs_isig, np.float32),
         None if active is None else np.ascontiguousarray(active, np.uint8)]
    if stereo is not None:
        sa = [np.ascontiguousarray(stereo.obs, np.float32), np.ascontiguousarray(stereo.obs_pt, np.int32),
              np.ascontiguousarray(stereo.obs_frame, np.int32), np.ascontiguousarray(stereo.obs_isig, np.float32),
              None if sactive is None else np.ascontiguousarray(sactive, np.uint8),
              np.ascontiguousarray(stereo.cam, np.float64)]
        Ms = len(stereo.obs_pt)
    else:
        sa = [None] * 5 + [np.zeros(6)]
        Ms = 0
    keep = a + sa
    args = [M] + [_p(x) for x in a] + [Ms] + [_p(x) for x in sa]
    return args, keep


def ba_lm(prob, stereo=None, iterations=10, tau=1e-5, max_trials=10, robust=True, active=None, sactive=None,
          point_fixed=None, points=None, frames=None):
    """FP64 Levenberg-Marquardt (g2o's OptimizationAlgorithmLevenberg with the
    Schur complement of the points): mono edges of `prob`, stereo edges of
    `stereo` (same points and frames), prob.frame_fixed, point_fixed.
    Returns (points, frames, stats [iterations x 4: chi2 before, chi2 after,
    lambda, trials], iterations run, first step [x_p; x_l])."""
    args, keep = _lm_args(prob, stereo, active, sactive)
    P, F, C = len(prob.points), len(prob.frames), len(prob.cams)
    pts = np.ascontiguousarray(prob.points if points is None else points, np.float64).copy()
    frs = np.ascontiguousarray(prob.frames if frames is None else frames, np.float64).copy()
    ff = np.ascontiguousarray(prob.frame_fixed, np.int32)
    pf = None if point_fixed is None else np.ascontiguousarray(point_fixed, np.int32)
    stats = np.zeros((iterations, 4))
    Fp = int((ff == 0).sum())
    x = np.zeros(6 * Fp + 3 * P)
    n = lib().orc_ba_lm(*args, P, F, C, _p(np.ascontiguousarray(prob.cams, np.float64)), _p(ff), _p(pf),
                        int(robust), _p(pts), _p(frs), iterations, tau, max_trials, _p(stats), _p(x))
    del keep
    return pts, frs, stats[:n], n, x


def ba_robust_chi2(prob, stereo=None, points=None, frames=None, robust=True, active=None, sactive=None):
    """activeRobustChi2 (sum of Huber rho(chi2) over active edges) in FP64."""
    args, keep = _lm_args(prob, stereo, active, sactive)
    pts = np.ascontiguousarray(prob.points if points is None else points, np.float64)
    frs = np.ascontiguousarray(prob.frames if frames is None else frames, np.float64)
    r = lib().orc_ba_robust_chi2(*args, len(prob.points), len(prob.frames), len(prob.cams),
                                 _p(np.ascontiguousarray(prob.cams, np.float64)), int(robust), _p(pts), _p(frs))
    del keep
    return r


def bow_transform(voc, feats, levelsup=4):
    """DBoW3 Vocabulary::transform of every feature: (word id, weight, node id
    at level L - levelsup)."""
    f = np.ascontiguousarray(feats, np.uint8).reshape(-1, 32)
    n = len(f)
    w = np.zeros(n, np.int32)
    wt = np.zeros(n)
    nid = np.zeros(n, np.int32)
    lib().orc_bow_transform(_p(voc.desc), _p(voc.child_ptr), _p(voc.child_ids), _p(voc.word_id), _p(voc.weight),
                            voc.L, levelsup, _p(f), n, _p(w), _p(wt), _p(nid))
    return w, wt, nid


def bow_vectors(word, weight, nid, weighting=0, scoring=0):
    """BowVector (words, values) and FeatureVector (nodes, ptr, features) of
    the per-feature transforms."""
    n = len(word)
    bw = np.zeros(max(n, 1), np.int32)
    bv = np.zeros(max(n, 1))
    fn = np.zeros(max(n, 1), np.int32)
    fp = np.zeros(n + 1, np.int32)
    ff = np.zeros(max(n, 1), np.int32)
    m = ctypes.c_int32()
    k = lib().orc_bow_vectors(_p(np.ascontiguousarray(word, np.int32)), _p(np.ascontiguousarray(weight, np.float64)),
                              _p(np.ascontiguousarray(nid, np.int32)), n, weighting, scoring, _p(bw), _p(bv), _p(fn),
                              _p(fp), _p(ff), ctypes.byref(m))
    nn = m.value
    return (bw[:k].copy(), bv[:k].copy()), (fn[:nn].copy(), fp[:nn + 1].copy(), ff[:fp[nn]].copy())


def bow_match(d1, fv1, d2, fv2):
    """Matcher::DBowMatch: rows (query, train0, d0, train1, d1)."""
    d1 = np.ascontiguousarray(d1, np.uint8)
    d2 = np.ascontiguousarray(d2, np.uint8)
    n1 = [np.ascontiguousarray(a, np.int32) for a in fv1]
    n2 = [np.ascontiguousarray(a, np.int32) for a in fv2]
    out = np.zeros((max(len(n1[2]), 1), 5), np.int32)
    r = lib().orc_bow_match(_p(d1), len(n1[0]), _p(n1[0]), _p(n1[1]), _p(n1[2]), _p(d2), len(n2[0]), _p(n2[0]),
                            _p(n2[1]), _p(n2[2]), _p(out))
    return out[:r].copy()
