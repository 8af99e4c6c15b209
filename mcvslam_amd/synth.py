"""Seeded synthetic inputs for the hot path (images, rig scenes, BA problems).

Every generator draws from numpy's PCG64 with an explicit seed, so the same
call gives the same arrays on any machine with the same numpy.  Images follow
the recipe: 400 Gaussian blobs (sigma 2-20 px, amplitude 40-200, random
sign) on a zero canvas, 60 filled axis-aligned rectangles in painter's order
(sides 8-120 px, value 0-255), additive N(0, 3) noise, clamp, round half to
even, uint8.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def blob_image(seed: int, width: int, height: int, n_blobs: int = 400, n_rects: int = 60,
               noise_sigma: float = 3.0) -> np.ndarray:
    rng = _rng(seed)
    canvas = np.zeros((height, width), np.float64)
    cx = rng.uniform(0, width, n_blobs)
    cy = rng.uniform(0, height, n_blobs)
    sig = rng.uniform(2.0, 20.0, n_blobs)
    amp = rng.uniform(40.0, 200.0, n_blobs) * np.where(rng.random(n_blobs) < 0.5, -1.0, 1.0)
    for x0, y0, s, a in zip(cx, cy, sig, amp):
        r = int(math.ceil(4.0 * s))
        xa, xb = max(0, int(x0) - r), min(width, int(x0) + r + 1)
        ya, yb = max(0, int(y0) - r), min(height, int(y0) + r + 1)
        if xa >= xb or ya >= yb:
            continue
        xs = np.arange(xa, xb) - x0
        ys = np.arange(ya, yb) - y0
        g = np.exp(-(ys[:, None] ** 2 + xs[None, :] ** 2) / (2.0 * s * s))
        canvas[ya:yb, xa:xb] += a * g
    for _ in range(n_rects):
        x0 = rng.uniform(0, width)
        y0 = rng.uniform(0, height)
        w = rng.uniform(8, 120)
        h = rng.uniform(8, 120)
        v = rng.uniform(0, 255)
        xa, xb = int(x0), min(width, int(x0 + w))
        ya, yb = int(y0), min(height, int(y0 + h))
        canvas[ya:yb, xa:xb] = v
    canvas += rng.normal(0.0, noise_sigma, canvas.shape)
    return np.rint(np.clip(canvas, 0.0, 255.0)).astype(np.uint8)


def homography(seed: int, width: int, height: int, max_rot_deg: float = 10.0,
               scale_range=(0.9, 1.1), max_shift: float = 20.0) -> np.ndarray:
    """H = T * R * S about the image centre (maps source to warped view)."""
    rng = _rng(seed)
    th = math.radians(rng.uniform(-max_rot_deg, max_rot_deg))
    s = rng.uniform(*scale_range)
    tx, ty = rng.uniform(-max_shift, max_shift, 2)
    c = np.array([[1, 0, width / 2.0], [0, 1, height / 2.0], [0, 0, 1]])
    ci = np.array([[1, 0, -width / 2.0], [0, 1, -height / 2.0], [0, 0, 1]])
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    S = np.diag([s, s, 1.0])
    T = np.array([[1, 0, tx], [0, 1, ty], [0, 0, 1]])
    return T @ c @ R @ S @ ci


def warp(img: np.ndarray, H: np.ndarray, out_w: int | None = None, out_h: int | None = None) -> np.ndarray:
    """Inverse-mapped bilinear warp; samples outside the source are 0."""
    h, w = img.shape
    out_w = out_w or w
    out_h = out_h or h
    Hi = np.linalg.inv(H)
    ys, xs = np.mgrid[0:out_h, 0:out_w].astype(np.float64)
    den = Hi[2, 0] * xs + Hi[2, 1] * ys + Hi[2, 2]
    sx = (Hi[0, 0] * xs + Hi[0, 1] * ys + Hi[0, 2]) / den
    sy = (Hi[1, 0] * xs + Hi[1, 1] * ys + Hi[1, 2]) / den
    x0 = np.floor(sx).astype(np.int64)
    y0 = np.floor(sy).astype(np.int64)
    fx = sx - x0
    fy = sy - y0
    src = img.astype(np.float64)

    def tap(yy, xx):
        ok = (xx >= 0) & (xx < w) & (yy >= 0) & (yy < h)
        v = np.zeros_like(sx)
        v[ok] = src[yy[ok], xx[ok]]
        return v

    out = (tap(y0, x0) * (1 - fx) * (1 - fy) + tap(y0, x0 + 1) * fx * (1 - fy)
           + tap(y0 + 1, x0) * (1 - fx) * fy + tap(y0 + 1, x0 + 1) * fx * fy)
    return np.rint(np.clip(out, 0, 255)).astype(np.uint8)


def image_pair(seed_img: int, seed_warp: int, width: int, height: int):
    a = blob_image(seed_img, width, height)
    b = warp(a, homography(seed_warp, width, height))
    return a, b


# Rig geometry: LEFT = body (Tcb = I), WIDE at +0.5 m, RIGHT at +1.0 m along x
# (config/frame.yaml:1-2 of the reference: Trl, Twl translations).
NARROW_W, NARROW_H = 1280, 720
WIDE_W, WIDE_H = 640, 480
FX_NARROW = (NARROW_W / 2.0) / math.tan(math.radians(15.0))   # 30 deg horizontal FOV
FX_WIDE = (WIDE_W / 2.0) / math.tan(math.radians(55.0))       # 110 deg horizontal FOV


def rig_frame(seed: int):
    """One rig frame: (left 1280x720, right = warp of left, wide 640x480)."""
    left = blob_image(seed, NARROW_W, NARROW_H)
    right = warp(left, homography(seed + 1, NARROW_W, NARROW_H))
    wide = blob_image(seed + 2, WIDE_W, WIDE_H)
    return left, right, wide


def rig_batch(n_frames: int = 64, n_unique: int = 16):
    """The batched-throughput workload: n_unique seeded rig frames
    (seeds 1000 + 7 f) tiled to n_frames, and per-image ORB feature counts
    U{500..4000} drawn with seed 40 + frame (3 per frame, L, R, W).
    Returns (unique frames, frame -> unique index, counts [3 n_frames])."""
    uniq = [rig_frame(1000 + 7 * f) for f in range(min(n_unique, n_frames))]
    index = [f % len(uniq) for f in range(n_frames)]
    nfeat = np.concatenate([np.random.Generator(np.random.PCG64(40 + f)).integers(500, 4001, 3)
                            for f in range(n_frames)]).astype(np.int32)
    return uniq, index, nfeat


@dataclass
class Camera:
    fx: float
    fy: float
    cx: float
    cy: float
    width: int
    height: int
    baseline_x: float  # camera centre along body x (m); Tcw = [I | -b] for a body at identity


def rig_cameras():
    return [
        Camera(FX_NARROW, FX_NARROW, NARROW_W / 2.0, NARROW_H / 2.0, NARROW_W, NARROW_H, 0.0),
        Camera(FX_WIDE, FX_WIDE, WIDE_W / 2.0, WIDE_H / 2.0, WIDE_W, WIDE_H, 0.5),
        Camera(FX_NARROW, FX_NARROW, NARROW_W / 2.0, NARROW_H / 2.0, NARROW_W, NARROW_H, 1.0),
    ]


def random_descriptors(rng: np.random.Generator, n: int) -> np.ndarray:
    return rng.integers(0, 256, size=(n, 32), dtype=np.uint8)


def flip_bits(rng: np.random.Generator, desc: np.ndarray, max_flips: int) -> np.ndarray:
    out = desc.copy()
    for i in range(out.shape[0]):
        k = int(rng.integers(0, max_flips + 1))
        if k:
            bits = rng.choice(256, size=k, replace=False)
            for b in bits:
                out[i, b >> 3] ^= np.uint8(1 << (b & 7))
    return out


@dataclass
class ProjectionScene:
    points: np.ndarray       # (P, 3) f32 world points
    pdesc: np.ndarray        # (P, 32) u8
    plevel: np.ndarray       # (P,) i32
    cams: list               # Camera per view
    R: np.ndarray            # (C, 3, 3) f32 world->camera rotation
    t: np.ndarray            # (C, 3) f32 world->camera translation
    kxy: list                # per camera (K, 2) f32 keypoint positions
    koct: list               # per camera (K,) i32 octaves
    kdesc: list              # per camera (K, 32) u8


def projection_scene(seed: int = 20, n_points: int = 5000, n_kps: int = 2000, frac_true: float = 0.7,
                     max_flips: int = 30, scale_factor: float = 1.2, nlevels: int = 8) -> ProjectionScene:
    rng = _rng(seed)
    pts = np.stack([rng.uniform(-5, 5, n_points), rng.uniform(-5, 5, n_points),
                    rng.uniform(2, 7, n_points)], 1).astype(np.float32)
    pdesc = random_descriptors(rng, n_points)
    plevel = rng.integers(0, nlevels, n_points).astype(np.int32)
    cams = rig_cameras()
    R = np.stack([np.eye(3, dtype=np.float32)] * len(cams))
    t = np.stack([np.array([-c.baseline_x, 0, 0], np.float32) for c in cams])
    kxy, koct, kdesc = [], [], []
    for ci, c in enumerate(cams):
        pc = pts + t[ci]
        u = c.fx * pc[:, 0] / pc[:, 2] + c.cx
        v = c.fy * pc[:, 1] / pc[:, 2] + c.cy
        vis = np.nonzero((pc[:, 2] > 0) & (u >= 0) & (u < c.width) & (v >= 0) & (v < c.height))[0]
        n_true = min(int(round(frac_true * n_kps)), len(vis))
        pick = rng.choice(vis, size=n_true, replace=False) if n_true else np.zeros(0, np.int64)
        xy_t = np.stack([u[pick], v[pick]], 1) + rng.normal(0, 1.0, (n_true, 2))
        d_t = flip_bits(rng, pdesc[pick], max_flips)
        o_t = plevel[pick]
        n_rand = n_kps - n_true
        xy_r = np.stack([rng.uniform(0, c.width, n_rand), rng.uniform(0, c.height, n_rand)], 1)
        d_r = random_descriptors(rng, n_rand)
        o_r = rng.integers(0, nlevels, n_rand)
        xy = np.concatenate([xy_t, xy_r]).astype(np.float32)
        xy[:, 0] = np.clip(xy[:, 0], 0, c.width - 1e-3)
        xy[:, 1] = np.clip(xy[:, 1], 0, c.height - 1e-3)
        perm = rng.permutation(n_kps)
        kxy.append(np.ascontiguousarray(xy[perm]))
        koct.append(np.concatenate([o_t, o_r]).astype(np.int32)[perm])
        kdesc.append(np.ascontiguousarray(np.concatenate([d_t, d_r])[perm]))
    return ProjectionScene(pts, pdesc, plevel, cams, R, t, kxy, koct, kdesc)


# ---------------------------------------------------------------------------
# Bundle-adjustment problems (multi-camera reprojection edges).

def _quat_from_yaw(yaw: float):
    return np.array([0.0, 0.0, math.sin(yaw / 2.0), math.cos(yaw / 2.0)])


def _quat_to_R(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@dataclass
class BAProblem:
    obs_uv: np.ndarray      # (M, 2) f32
    obs_pt: np.ndarray      # (M,) i32
    obs_frame: np.ndarray   # (M,) i32
    obs_cam: np.ndarray     # (M,) i32
    obs_level: np.ndarray   # (M,) i32
    obs_isig: np.ndarray    # (M,) f32 inverse level sigma^2
    points: np.ndarray      # (P, 3) f64
    frames: np.ndarray      # (F, 7) f64 Tbw as (qx qy qz qw tx ty tz)
    cams: np.ndarray        # (C, 12) f64 (fx fy cx cy qx qy qz qw tx ty tz delta)
    frame_fixed: np.ndarray  # (F,) i32


def ba_problem(seed: int = 30, n_kf: int = 10, n_points: int = 20000, target_obs: int = 150000,
               outlier_frac: float = 0.05, huber_delta: float | None = None,
               scale_factor: float = 1.2, nlevels: int = 8, perturb: float = 0.0, noise: float = 1.0) -> BAProblem:
    """Keyframes along +x (0.3 m steps, yaw within +-5 deg), points in a
    10x10x5 m box ahead of the rig, every visible (point, keyframe, camera)
    triple observed, subsampled to about target_obs, N(0,1) px noise and
    outlier_frac uniform outliers.  `perturb` adds N(0, perturb) m to the
    point estimates (the observations stay generated from the true points)."""
    rng = _rng(seed)
    delta = float(np.float32(math.sqrt(5.991))) if huber_delta is None else huber_delta
    cams_geo = rig_cameras()
    cams = []
    for c in cams_geo:
        # Pinhole keeps its intrinsics as float (modules/camera/BaseCamera.hpp:37-43)
        f32 = lambda v: float(np.float32(v))
        cams.append([f32(c.fx), f32(c.fy), f32(c.cx), f32(c.cy), 0, 0, 0, 1, -c.baseline_x, 0, 0, delta])
    cams = np.array(cams, np.float64)
    frames = []
    Rs, ts = [], []
    for k in range(n_kf):
        yaw = math.radians(rng.uniform(-5, 5))
        q = _quat_from_yaw(yaw)
        R = _quat_to_R(q)
        pos = np.array([0.3 * k, 0.0, 0.0])
        t = -R @ pos   # Tbw = [R | -R pos]
        frames.append(np.concatenate([q, t]))
        Rs.append(R)
        ts.append(t)
    frames = np.array(frames)
    span_x = 0.3 * n_kf
    pts = np.stack([rng.uniform(-5, 5 + span_x, n_points), rng.uniform(-5, 5, n_points),
                    rng.uniform(4, 9, n_points)], 1)
    uv, pid, fid, cid = [], [], [], []
    for k in range(n_kf):
        pb = pts @ Rs[k].T + ts[k]
        for ci, c in enumerate(cams_geo):
            pc = pb - np.array([c.baseline_x, 0, 0])
            z = pc[:, 2]
            with np.errstate(divide="ignore", invalid="ignore"):
                u = c.fx * pc[:, 0] / z + c.cx
                v = c.fy * pc[:, 1] / z + c.cy
            ok = np.nonzero((z > 0.5) & (u >= 0) & (u < c.width) & (v >= 0) & (v < c.height))[0]
            uv.append(np.stack([u[ok], v[ok]], 1))
            pid.append(ok)
            fid.append(np.full(len(ok), k))
            cid.append(np.full(len(ok), ci))
    uv = np.concatenate(uv)
    pid = np.concatenate(pid)
    fid = np.concatenate(fid)
    cid = np.concatenate(cid)
    if len(pid) > target_obs:
        keep = np.sort(rng.choice(len(pid), size=target_obs, replace=False))
        uv, pid, fid, cid = uv[keep], pid[keep], fid[keep], cid[keep]
    M = len(pid)
    uv = uv + rng.normal(0, 1.0, uv.shape) * noise
    n_out = int(round(outlier_frac * M))
    out_idx = rng.choice(M, size=n_out, replace=False)
    for i in out_idx:
        c = cams_geo[cid[i]]
        uv[i] = [rng.uniform(0, c.width), rng.uniform(0, c.height)]
    level = rng.integers(0, nlevels, M).astype(np.int32)
    # use the ORBextractor::init tables (repeated f32 products)
    sc = [np.float32(1.0)]
    for _ in range(1, nlevels):
        sc.append(np.float32(sc[-1] * np.float32(scale_factor)))
    inv_s2 = np.array([np.float32(1.0) / np.float32(s * s) for s in sc], np.float32)
    isig = inv_s2[level]
    # order observations by point (CSR-friendly), keep it deterministic
    order = np.lexsort((cid, fid, pid))
    est = pts + (rng.normal(0, perturb, pts.shape) if perturb > 0 else 0.0)
    fixed = np.zeros(n_kf, np.int32)
    fixed[0] = 1
    return BAProblem(uv[order].astype(np.float32), pid[order].astype(np.int32), fid[order].astype(np.int32),
                     cid[order].astype(np.int32), level[order], isig[order].astype(np.float32), est,
                     frames, cams, fixed)


@dataclass
class StereoProblem:
    obs: np.ndarray         # (M, 3) f32: u, v, u_right
    obs_pt: np.ndarray      # (M,) i32
    obs_frame: np.ndarray   # (M,) i32
    obs_isig: np.ndarray    # (M,) f32
    points: np.ndarray      # (P, 3) f64
    frames: np.ndarray      # (F, 7) f64 camera (= body) pose Tcw
    cam: np.ndarray         # (6,) f64: fx fy cx cy bf delta


def stereo_problem(seed: int = 50, n_kf: int = 6, n_points: int = 4000, target_obs: int = 30000,
                   outlier_frac: float = 0.05) -> StereoProblem:
    """Left-camera observations of ba_problem (the left camera is the body)
    with a right-image coordinate u_r = u - bf / z (baseline 1.0 m, the R
    camera of the rig), N(0, 1) px noise on u_r too, and outlier_frac of the
    u_r replaced by uniform values; Huber delta sqrt(7.815) as thHuber3D."""
    prob = ba_problem(seed=seed, n_kf=n_kf, n_points=n_points, target_obs=target_obs * 3,
                      outlier_frac=outlier_frac)
    rng = _rng(seed + 1)
    sel = np.nonzero(prob.obs_cam == 0)[0][:target_obs]
    fx, fy, cx, cy = prob.cams[0, :4]
    bf = float(np.float32(fx * 1.0))
    pts = prob.points[prob.obs_pt[sel]]
    z = np.empty(len(sel))
    for k, i in enumerate(sel):
        f = prob.frames[prob.obs_frame[i]]
        R = _quat_to_R(f[:4] / np.linalg.norm(f[:4]))
        z[k] = (R @ pts[k] + f[4:])[2]
    ur = prob.obs_uv[sel, 0].astype(np.float64) - bf / z + rng.normal(0, 1.0, len(sel))
    n_out = int(round(outlier_frac * len(sel)))
    ur[rng.choice(len(sel), size=n_out, replace=False)] = rng.uniform(0, NARROW_W, n_out)
    obs = np.concatenate([prob.obs_uv[sel].astype(np.float64), ur[:, None]], 1).astype(np.float32)
    delta = float(np.float32(math.sqrt(7.815)))
    return StereoProblem(obs, prob.obs_pt[sel], prob.obs_frame[sel], prob.obs_isig[sel], prob.points,
                         prob.frames, np.array([fx, fy, cx, cy, bf, delta], np.float64))


@dataclass
class LocalBAProblem:
    mono: BAProblem          # mono edges (LEFT without u_right, WIDE), estimates in .points / .frames
    stereo: StereoProblem    # stereo edges of LEFT observations with u_right (same points and frames)
    point_fixed: np.ndarray  # (P,) i32
    true_points: np.ndarray
    true_frames: np.ndarray


def local_ba_problem(seed: int = 60, n_kf: int = 10, n_points: int = 20000, target_obs: int = 150000,
                     stereo_frac: float = 0.3, outlier_frac: float = 0.05, noise: float = 1.0,
                     perturb_point: float = 0.05, perturb_pose: float = 0.01, fixed_point_frac: float = 0.1,
                     scale_factor: float = 1.2, nlevels: int = 8) -> LocalBAProblem:
    """The problem PoseEstimation::BoundleAdjustment builds
    (include/PoseEstimation.hpp:35-119): keyframe 0 fixed, LEFT observations
    with a right-image coordinate as stereo edges (addEdgeStereo, Huber
    sqrt(7.815)), the other LEFT and all WIDE/RIGHT observations as mono edges
    (addEdge / addEdgeToBody), a fraction of the points fixed
    (ThisMapPointBeFix) at their true positions.  The estimates start from
    the truth perturbed by N(0, perturb_point) m per point coordinate and
    exp(N(0, perturb_pose)) per free keyframe."""
    from .autograd import apply_increments
    prob = ba_problem(seed=seed, n_kf=n_kf, n_points=n_points, target_obs=target_obs, outlier_frac=outlier_frac,
                      scale_factor=scale_factor, nlevels=nlevels, noise=noise)
    rng = _rng(seed + 7)
    true_pts = prob.points.copy()
    true_frs = prob.frames.copy()
    left = np.nonzero(prob.obs_cam == 0)[0]
    sel = np.sort(left[rng.random(len(left)) < stereo_frac])
    fx, fy, cx, cy = prob.cams[0, :4]
    bf = float(np.float32(fx * 1.0))
    z = np.empty(len(sel))
    for k, i in enumerate(sel):
        f = true_frs[prob.obs_frame[i]]
        R = _quat_to_R(f[:4] / np.linalg.norm(f[:4]))
        z[k] = (R @ true_pts[prob.obs_pt[i]] + f[4:])[2]
    ur = prob.obs_uv[sel, 0].astype(np.float64) - bf / z + rng.normal(0, 1.0, len(sel)) * noise
    n_out = int(round(outlier_frac * len(sel)))
    if n_out:
        ur[rng.choice(len(sel), size=n_out, replace=False)] = rng.uniform(0, NARROW_W, n_out)
    obs = np.concatenate([prob.obs_uv[sel].astype(np.float64), ur[:, None]], 1).astype(np.float32)
    delta3 = float(np.float32(math.sqrt(7.815)))
    keep = np.ones(len(prob.obs_pt), bool)
    keep[sel] = False
    mono = BAProblem(prob.obs_uv[keep], prob.obs_pt[keep], prob.obs_frame[keep], prob.obs_cam[keep],
                     prob.obs_level[keep], prob.obs_isig[keep], true_pts.copy(), true_frs.copy(), prob.cams,
                     prob.frame_fixed)
    P = len(true_pts)
    pfix = (rng.random(P) < fixed_point_frac).astype(np.int32)
    dpt = rng.normal(0, perturb_point, true_pts.shape) * (1 - pfix)[:, None]
    dpose = rng.normal(0, perturb_pose, (n_kf, 6)) * (1 - prob.frame_fixed)[:, None]
    pts, frs = apply_increments(true_pts, true_frs, dpose, dpt)
    mono.points, mono.frames = pts, frs
    stereo = StereoProblem(obs, prob.obs_pt[sel], prob.obs_frame[sel], prob.obs_isig[sel], pts, frs,
                           np.array([fx, fy, cx, cy, bf, delta3], np.float64))
    return LocalBAProblem(mono, stereo, pfix, true_pts, true_frs)


@dataclass
class Vocabulary:
    """A DBoW3-style vocabulary tree (node 0 the root, children in CSR, in
    breadth-first numbering as DBoW3's HKmeans builds it)."""
    k: int
    L: int
    desc: np.ndarray       # (n, 32) u8 node descriptors (row 0 unused)
    child_ptr: np.ndarray  # (n + 1,) i32
    child_ids: np.ndarray  # (n - 1,) i32
    word_id: np.ndarray    # (n,) i32, -1 for inner nodes
    weight: np.ndarray     # (n,) f64 (idf-like; 0 = stopped word)
    weighting: int = 0     # DBoW3 WeightingType: 0 TF_IDF, 1 TF, 2 IDF, 3 BINARY
    scoring: int = 0       # ScoringType: 0 L1, 1 L2, 2 chi-square, 3 KL, 4 Bhattacharyya, 5 dot product

    @property
    def leaves(self):
        return np.nonzero(self.word_id >= 0)[0]


def bow_vocabulary(seed: int = 80, k: int = 8, L: int = 5, flip: int = 24, stop_frac: float = 0.02,
                   weighting: int = 0, scoring: int = 0) -> Vocabulary:
    """k-ary tree of depth L: level-1 descriptors random, every child its
    parent's descriptor with `flip` random bits flipped (a cluster hierarchy);
    leaves get word ids in node order and weights U(0.5, 4) (a fraction 0)."""
    rng = _rng(seed)
    n = sum(k ** l for l in range(L + 1))
    desc = np.zeros((n, 32), np.uint8)
    child_ptr = np.zeros(n + 1, np.int32)
    child_ids = []
    level_start = [0]
    nxt = 1
    for l in range(L):
        first = level_start[-1]
        count = k ** l
        level_start.append(nxt)
        for i in range(first, first + count):
            child_ptr[i] = len(child_ids)
            for c in range(k):
                child_ids.append(nxt)
                if l == 0:
                    desc[nxt] = rng.integers(0, 256, 32, dtype=np.uint8)
                else:
                    bits = rng.choice(256, flip, replace=False)
                    mask = np.zeros(256, np.uint8)
                    mask[bits] = 1
                    desc[nxt] = desc[i] ^ np.packbits(mask, bitorder="little")
                nxt += 1
    for i in range(level_start[-1], n + 1):
        child_ptr[i] = len(child_ids)
    word_id = -np.ones(n, np.int32)
    leaves = np.arange(level_start[-1], n)
    word_id[leaves] = np.arange(len(leaves))
    weight = np.zeros(n)
    weight[leaves] = rng.uniform(0.5, 4.0, len(leaves))
    weight[leaves[rng.random(len(leaves)) < stop_frac]] = 0.0
    return Vocabulary(k, L, desc, child_ptr, np.array(child_ids, np.int32), word_id, weight, weighting, scoring)


def bow_features(voc: Vocabulary, seed: int, n: int, max_flips: int = 30):
    """n descriptors: noisy copies of random leaves (0..max_flips bits)."""
    rng = _rng(seed)
    leaves = voc.leaves
    src = voc.desc[leaves[rng.integers(0, len(leaves), n)]].copy()
    return flip_bits(rng, src, max_flips)
