"""Python mirror of the reference's hot-path interfaces over the C ABI.

    ORB                  MCVSLAM::ORB / ORB_SLAM3::ORBextractor
                         (modules/local_feature/ORB/ORBExtractor.cpp:10-38,
                          orb3_extractor/ORBextractor.h:46-100)
    Matcher.KnnMatch     Matcher::KnnMatch (include/Matcher.hpp:84-86)
    MatchResKnn.FilterRatio(r).FilterThreshold(th)
                         src/Matcher.cpp:100-111, :23-35
    window_search        Object::ProjectBunchMapPoints + GetFeaturesInArea
                         (src/Object.cpp:208-236, :263-308)
    RigBatch             Frame::Frame extraction + camera-pair matching
    BALinearizer         batched EdgeSE3XYZ2Camera linearization
                         (include/OptimizeTypes.hpp:70-117)

Everything runs in the HIP library; nothing here computes a result itself.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from ._native import KeyPoint, MCVError, check, lib

KP_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"), ("response", "<f4"),
                     ("octave", "<i4")])


def _p(a):
    return ctypes.c_void_p(0) if a is None else ctypes.c_void_p(a.ctypes.data)


class Context:
    """One device context (stream, plans, buffers); one per thread."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        code = lib().mcv_create(ctypes.byref(h), device)
        if code != 0:
            raise MCVError(code, f"mcv_create(device={device}) failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mcv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle: int | None):
        """hipStream_t handle (0 / None: the null stream)."""
        check(self.h, lib().mcv_set_stream(self.h, ctypes.c_void_p(stream_handle or 0)))

    def use_own_stream(self):
        check(self.h, lib().mcv_use_own_stream(self.h))

    def use_torch_stream(self):
        import torch
        self.set_stream(torch.cuda.current_stream().cuda_stream)

    def synchronize(self):
        check(self.h, lib().mcv_synchronize(self.h))

    N_KNN_VARIANTS = 12

    def set_match_variant(self, variant: int):
        """Launch variant of the top-2 kernel for Matcher calls
        (0 .. N_KNN_VARIANTS - 1; 8 and 9 compute the distances on the int8
        matrix cores, 10 and 11 on the fp4 ones; all return the same rows)."""
        check(self.h, lib().mcv_match_set_variant(self.h, int(variant)))

    def to_host(self, d_ptr: int, dtype, shape) -> np.ndarray:
        out = np.zeros(shape, dtype)
        check(self.h, lib().mcv_copy_to_host(self.h, _p(out), ctypes.c_void_p(d_ptr), out.nbytes))
        return out


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class ORB:
    """ORB extractor with the reference's parameters and tables."""

    def __init__(self, nfeatures: int = 2000, scale_factor: float = 1.2, nlevels: int = 8, ini_th_fast: int = 20,
                 min_th_fast: int = 7, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.nfeatures, self.scaleFactor, self.nlevels = nfeatures, float(scale_factor), nlevels
        self.iniThFAST, self.minThFAST = ini_th_fast, min_th_fast
        check(self.ctx.h, lib().mcv_orb_configure(self.ctx.h, nfeatures, self.scaleFactor, nlevels, ini_th_fast,
                                                  min_th_fast))
        t = [np.zeros(nlevels, np.float32) for _ in range(4)]
        npl = np.zeros(nlevels, np.int32)
        check(self.ctx.h, lib().mcv_orb_tables(self.ctx.h, *[_p(x) for x in t], _p(npl)))
        self.mvScaleFactor, self.mvInvScaleFactor, self.mvLevelSigma2, self.mvInvLevelSigma2 = t
        self.mnFeaturesPerLevel = npl

    @classmethod
    def from_yaml(cls, path: str, ctx: Context | None = None) -> "ORB":
        """ORB::Parse (ORBExtractor.cpp:10-18): nkeypoints, scale_factor,
        nlevels, ORBextractor.iniThFAST, ORBextractor.minThFAST."""
        import yaml
        with open(path) as f:
            cfg = yaml.safe_load(f)
        return cls(int(cfg["nkeypoints"]), float(cfg["scale_factor"]), int(cfg["nlevels"]),
                   int(cfg["ORBextractor.iniThFAST"]), int(cfg["ORBextractor.minThFAST"]), ctx)

    def capacity(self, width: int, height: int, n_seeds: int = 0) -> int:
        c = ctypes.c_int32()
        check(self.ctx.h, lib().mcv_orb_capacity(self.ctx.h, width, height, n_seeds, ctypes.byref(c)))
        return c.value

    def extract(self, image: np.ndarray, seeds: np.ndarray | None = None):
        """ORB::Extract: returns (keypoints as KP_DTYPE records, N x 32 u8)."""
        img = np.ascontiguousarray(image, dtype=np.uint8)
        if img.ndim != 2 or img.size == 0:
            raise MCVError(_native.MCV_ERR_EMPTY, "empty image")
        h, w = img.shape
        sd = None
        ns = 0
        if seeds is not None and len(seeds):
            sd = np.ascontiguousarray(seeds, dtype=KP_DTYPE)
            ns = len(sd)
        cap = self.capacity(w, h, ns)
        kps = np.zeros(cap, KP_DTYPE)
        desc = np.zeros((cap, 32), np.uint8)
        cnt = ctypes.c_int32()
        check(self.ctx.h, lib().mcv_orb_extract(self.ctx.h, _p(img), w, h, w, _p(sd), ns, _p(kps), _p(desc), cap,
                                                ctypes.byref(cnt)))
        n = cnt.value
        return kps[:n].copy(), desc[:n].copy()

    __call__ = extract

    def pyramid(self, level: int) -> np.ndarray:
        """mvImagePyramid[level] of the last extract()."""
        w, h = ctypes.c_int32(), ctypes.c_int32()
        check(self.ctx.h, lib().mcv_orb_pyramid(self.ctx.h, level, None, 0, ctypes.byref(w), ctypes.byref(h)))
        out = np.zeros((h.value, w.value), np.uint8)
        check(self.ctx.h, lib().mcv_orb_pyramid(self.ctx.h, level, _p(out), out.size, ctypes.byref(w), ctypes.byref(h)))
        return out

    def extract_batch_dev(self, images, graph: bool = False, seeds=None, check_overflow: bool = True):
        """Batched extraction of torch.cuda uint8 images (H x W, contiguous).
        seeds: None or, per image, None / KP_DTYPE records appended to their
        octave (ORBextractor.cc:845-847).  check_overflow waits for the run and
        reruns it after growing the octree pool when that overflowed.
        Returns device tensors (keypoints n x cap x 6 as float32 view of
        KP records, descriptors n x cap x 32, counts n)."""
        import torch
        n = len(images)
        ws = np.array([im.shape[1] for im in images], np.int32)
        hs = np.array([im.shape[0] for im in images], np.int32)
        sd_arrays = [None] * n
        ns = np.zeros(n, np.int32)
        if seeds is not None:
            for i, sd in enumerate(seeds):
                if sd is not None and len(sd):
                    sd_arrays[i] = np.ascontiguousarray(sd, dtype=KP_DTYPE)
                    ns[i] = len(sd_arrays[i])
        cap = max(self.capacity(int(w), int(h), int(k)) for w, h, k in zip(ws, hs, ns))
        dev = images[0].device
        kps = torch.zeros((n, cap, 6), dtype=torch.float32, device=dev)
        desc = torch.zeros((n, cap, 32), dtype=torch.uint8, device=dev)
        counts = torch.zeros(n, dtype=torch.int32, device=dev)
        ptrs = (ctypes.c_void_p * n)(*[im.data_ptr() for im in images])
        sptrs = (ctypes.c_void_p * n)(*[a.ctypes.data if a is not None else None for a in sd_arrays])
        flags = (_native.MCV_FLAG_GRAPH if graph else 0) | (_native.MCV_FLAG_CHECK if check_overflow else 0)
        self.ctx.use_torch_stream()
        check(self.ctx.h, lib().mcv_orb_extract_batch_dev(
            self.ctx.h, n, ptrs, _p(ws), _p(hs), _p(ws), sptrs if seeds is not None else None,
            _p(ns) if seeds is not None else None, ctypes.c_void_p(kps.data_ptr()),
            ctypes.c_void_p(desc.data_ptr()), cap, ctypes.c_void_p(counts.data_ptr()), flags))
        return kps, desc, counts

    def batch_check(self):
        """Raises when the last batched extraction overflowed (mcv_orb_batch_check)."""
        check(self.ctx.h, lib().mcv_orb_batch_check(self.ctx.h))

    def batch_pyramid(self, image: int, level: int, blurred: bool = False):
        """mvImagePyramid[level] of image `image` of the last batched
        extraction (blurred: the GaussianBlur'd copy the descriptors sample),
        as a zero-copy torch view of the resident plane."""
        ptr, pitch, w, h = ctypes.c_void_p(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(self.ctx.h, lib().mcv_orb_batch_pyramid_dev(self.ctx.h, image, level, int(blurred), ctypes.byref(ptr),
                                                          ctypes.byref(pitch), ctypes.byref(w), ctypes.byref(h)))
        return device_plane(ptr.value, h.value, w.value, pitch.value)


class _DevicePlane:
    def __init__(self, ptr: int, h: int, w: int, pitch: int):
        self.__cuda_array_interface__ = {"shape": (h, w), "typestr": "|u1", "data": (ptr, False),
                                         "strides": (pitch, 1), "version": 3}


def device_plane(ptr: int, h: int, w: int, pitch: int):
    """Zero-copy torch view (uint8, h x w, row pitch in bytes) of device memory."""
    import torch
    return torch.as_tensor(_DevicePlane(ptr, h, w, pitch), device="cuda")


def kp_records(kps6) -> np.ndarray:
    """Reinterpret rows of 6 float32 (device output) as KP_DTYPE records."""
    a = np.ascontiguousarray(np.asarray(kps6, dtype=np.float32))
    return a.view(KP_DTYPE).reshape(a.shape[:-1])


class MatchRes(np.ndarray):
    """Rows of (queryIdx, trainIdx, distance)."""


class MatchResKnn:
    """Rows of (trainIdx0, distance0, trainIdx1, distance1) per query."""

    def __init__(self, rows: np.ndarray, ctx: Context):
        self.rows = rows
        self.ctx = ctx
        self._ratio = None

    def __len__(self):
        return len(self.rows)

    def FilterRatio(self, ratio: float = 0.6) -> "_RatioFiltered":
        return _RatioFiltered(self, float(ratio))


class _RatioFiltered:
    """FilterRatio result; materialised by FilterThreshold (or .matches())."""

    def __init__(self, knn: MatchResKnn, ratio: float):
        self.knn, self.ratio = knn, ratio

    def FilterThreshold(self, th: int = 46) -> np.ndarray:
        return Matcher.filter(self.knn.rows, self.ratio, int(th), self.knn.ctx)

    def matches(self) -> np.ndarray:
        # no threshold: every distance passes (<= 256), order unchanged
        return Matcher.filter(self.knn.rows, self.ratio, 1 << 30, self.knn.ctx)


class Matcher:
    @staticmethod
    def KnnMatch(desp1, desp2, k: int = 2, n_threads: int = 0, ctx: Context | None = None) -> MatchResKnn:
        ctx = ctx or default_context()
        q = np.ascontiguousarray(desp1, dtype=np.uint8).reshape(-1, 32)
        t = np.ascontiguousarray(desp2, dtype=np.uint8).reshape(-1, 32)
        out = np.zeros((len(q), 4), np.int32)
        check(ctx.h, lib().mcv_match_knn2(ctx.h, _p(q), len(q), _p(t), len(t), _p(out)))
        return MatchResKnn(out, ctx)

    @staticmethod
    def filter(knn_rows: np.ndarray, ratio: float, th: int, ctx: Context | None = None) -> np.ndarray:
        ctx = ctx or default_context()
        k = np.ascontiguousarray(knn_rows, dtype=np.int32).reshape(-1, 4)
        out = np.zeros((max(len(k), 1), 3), np.int32)
        cnt = ctypes.c_int32()
        check(ctx.h, lib().mcv_match_filter(ctx.h, _p(k), len(k), ctypes.c_float(ratio), th, _p(out),
                                            ctypes.byref(cnt)))
        return out[:cnt.value].copy()


def window_search(points, pdesc, plevel, R, t, cam, scales, radius, ratio, th, kxy, kdesc, width, height,
                  ctx: Context | None = None) -> np.ndarray:
    """Per map point (kp0, d0, kp1, d1, n_candidates, match)."""
    ctx = ctx or default_context()
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    pd = np.ascontiguousarray(pdesc, np.uint8).reshape(-1, 32)
    pl = np.ascontiguousarray(plevel, np.int32)
    Rm = np.ascontiguousarray(R, np.float32).reshape(9)
    tv = np.ascontiguousarray(t, np.float32).reshape(3)
    cm = np.ascontiguousarray(cam, np.float32).reshape(4)
    sc = np.ascontiguousarray(scales, np.float32)
    kx = np.ascontiguousarray(kxy, np.float32).reshape(-1, 2)
    kd = np.ascontiguousarray(kdesc, np.uint8).reshape(-1, 32)
    out = np.zeros((len(pts), 6), np.int32)
    check(ctx.h, lib().mcv_window_search(ctx.h, _p(pts), _p(pd), _p(pl), len(pts), _p(Rm), _p(tv), _p(cm), _p(sc),
                                         len(sc), ctypes.c_float(radius), ctypes.c_float(ratio), th, _p(kx),
                                         _p(kd), len(kx), width, height, _p(out)))
    return out


def window_search_dev(points, pdesc, plevel, Rt, cams, sizes, scales, radius, ratio, th, kxy, kdesc,
                      ctx: Context | None = None):
    """mcv_window_search_dev over several cameras: torch.cuda inputs (points
    P x 3, pdesc P x 32, plevel P; per camera kxy K x 2 and kdesc K x 32),
    host Rt (n_cams x 12), cams (n_cams x 4), sizes (n_cams x 2).  Returns a
    device tensor [n_cams][P][6]."""
    import torch
    ctx = ctx or default_context()
    n_cams = len(kxy)
    P = int(points.shape[0])
    out = torch.empty((n_cams, P, 6), dtype=torch.int32, device=points.device)
    rt = np.ascontiguousarray(Rt, np.float32).reshape(n_cams, 12)
    cm = np.ascontiguousarray(cams, np.float32).reshape(n_cams, 4)
    sz = np.ascontiguousarray(sizes, np.int32).reshape(n_cams, 2)
    sc = np.ascontiguousarray(scales, np.float32)
    nk = np.array([int(k.shape[0]) for k in kxy], np.int32)
    kx = (ctypes.c_void_p * n_cams)(*[k.data_ptr() for k in kxy])
    kd = (ctypes.c_void_p * n_cams)(*[k.data_ptr() for k in kdesc])
    ctx.use_torch_stream()
    check(ctx.h, lib().mcv_window_search_dev(ctx.h, n_cams, ctypes.c_void_p(points.data_ptr()),
                                             ctypes.c_void_p(pdesc.data_ptr()), ctypes.c_void_p(plevel.data_ptr()),
                                             P, _p(rt), _p(cm), _p(sz), _p(sc), len(sc), ctypes.c_float(radius),
                                             ctypes.c_float(ratio), th, kx, kd, _p(nk),
                                             ctypes.c_void_p(out.data_ptr())))
    return out


def window_search_uv_dev(uv, qdesc, radius, kxy, koct, kdesc, width, height, min_level=-1, max_level=-1,
                         ratio: float = 0.6, th: int = 46, ctx: Context | None = None):
    """Tracker::Wnd_Track's query (src/Tracker.cpp:339-353) for every
    position of uv (torch.cuda n x 2) with descriptor qdesc (n x 32) against
    keypoints kxy / koct / kdesc (torch.cuda).  Returns a device tensor n x 6
    (kp0, d0, kp1, d1, n_candidates, match)."""
    import torch
    ctx = ctx or default_context()
    n = int(uv.shape[0])
    out = torch.empty((n, 6), dtype=torch.int32, device=uv.device)
    ctx.use_torch_stream()
    check(ctx.h, lib().mcv_window_search_uv_dev(ctx.h, ctypes.c_void_p(uv.data_ptr()),
                                                ctypes.c_void_p(qdesc.data_ptr()), n, ctypes.c_float(radius),
                                                int(min_level), int(max_level), ctypes.c_float(ratio), int(th),
                                                ctypes.c_void_p(kxy.data_ptr()), ctypes.c_void_p(koct.data_ptr()),
                                                ctypes.c_void_p(kdesc.data_ptr()), int(kxy.shape[0]), int(width),
                                                int(height), ctypes.c_void_p(out.data_ptr())))
    return out


class RigBatch:
    """n rig frames (L, R, W) extracted and matched (L-R, W-L, W-R)."""

    def __init__(self, orb: ORB, n_frames: int, sizes=((1280, 720), (1280, 720), (640, 480)), ratio: float = 0.7,
                 th: int = 50, nfeatures=None):
        self.orb = orb
        self.ctx = orb.ctx
        self.n_frames = n_frames
        w = np.array([s[0] for s in sizes], np.int32)
        h = np.array([s[1] for s in sizes], np.int32)
        self.sizes = sizes
        hd = ctypes.c_void_p()
        nf = None if nfeatures is None else np.ascontiguousarray(nfeatures, np.int32).reshape(3 * n_frames)
        self.nfeatures = nf
        check(self.ctx.h, lib().mcv_rig_create(self.ctx.h, n_frames, _p(w), _p(h), _p(nf), ctypes.c_float(ratio), th,
                                               ctypes.byref(hd)))
        self.h = hd
        ptrs = [ctypes.c_void_p() for _ in range(6)]
        cap = ctypes.c_int32()
        check(self.ctx.h, lib().mcv_rig_outputs(self.h, *[ctypes.byref(p) for p in ptrs], ctypes.byref(cap)))
        self.d_kps, self.d_desc, self.d_counts, self.d_matches, self.d_mcounts, self.d_knn = [p.value for p in ptrs]
        self.cap = cap.value

    def close(self):
        if getattr(self, "h", None):
            lib().mcv_rig_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, image_ptrs, graph: bool = True):
        n = 3 * self.n_frames
        if len(image_ptrs) != n:
            raise ValueError(f"expected {n} image pointers")
        arr = (ctypes.c_void_p * n)(*image_ptrs)
        check(self.ctx.h, lib().mcv_rig_run(self.h, arr, _native.MCV_FLAG_GRAPH if graph else 0))

    N_BLUR_VARIANTS = 4
    N_KP_VARIANTS = 4
    N_HEAP_VARIANTS = 2

    def set_variants(self, knn: int, blur: int, kp: int = 2, heap: int = 1):
        """Fix the launch variants (top-2 kernel; blur: bit 0 the walk, bit 1
        the captured graph runs it on a branch beside the octree heap replay;
        keypoints per wave: 0 one, 1 eight, 2 sixteen, 3 thirty-two; octree
        heap: 0 register top, 1 LDS, the default) instead of tuning."""
        check(self.ctx.h, lib().mcv_rig_set_variants(self.h, int(knn), int(blur), int(kp), int(heap)))

    def variants(self):
        """(knn, blur, kp, heap): the launch variants in use."""
        v = [ctypes.c_int32() for _ in range(4)]
        check(self.ctx.h, lib().mcv_rig_get_variants(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def autotune(self, image_ptrs, reps: int = 3) -> str:
        """Time every launch variant of the tunable kernels on these images
        and keep the fastest (outputs are identical); returns the report."""
        n = 3 * self.n_frames
        arr = (ctypes.c_void_p * n)(*image_ptrs)
        buf = ctypes.create_string_buffer(4096)
        check(self.ctx.h, lib().mcv_rig_autotune(self.h, arr, reps, buf, 4096))
        return buf.value.decode()

    def run_timed(self, image_ptrs):
        """Eager run with an event after every launch: [(kernel, ms), ...]."""
        n = 3 * self.n_frames
        arr = (ctypes.c_void_p * n)(*image_ptrs)
        ms = np.zeros(64, np.float32)
        names = (ctypes.c_char_p * 64)()
        k = ctypes.c_int32()
        check(self.ctx.h, lib().mcv_rig_run_timed(self.h, arr, _p(ms), names, 64, ctypes.byref(k)))
        return [(names[i].decode(), float(ms[i])) for i in range(min(k.value, 64))]

    def stereo(self, bf: float, baseline: float):
        """L-R stereo of every frame of the last batched run; host copies of
        (u_right, depth, sad), each [n_frames][cap]."""
        check(self.ctx.h, lib().mcv_rig_stereo(self.h, ctypes.c_float(bf), ctypes.c_float(baseline)))
        ptrs = [ctypes.c_void_p() for _ in range(3)]
        check(self.ctx.h, lib().mcv_rig_stereo_outputs(self.h, *[ctypes.byref(p) for p in ptrs]))
        shape = (self.n_frames, self.cap)
        return (self.ctx.to_host(ptrs[0].value, np.float32, shape), self.ctx.to_host(ptrs[1].value, np.float32, shape),
                self.ctx.to_host(ptrs[2].value, np.int32, shape))

    def octree_stats(self):
        """(FAST candidates, node-table entries, nodes emitted) of the last run."""
        v = [ctypes.c_int64() for _ in range(3)]
        check(self.ctx.h, lib().mcv_rig_octree_stats(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def pyramid(self, image: int, level: int, blurred: bool = False):
        """mvImagePyramid[level] of image `image` (frame-major L, R, W) of the
        last batched run (blurred: the descriptors' blurred copy), as a
        zero-copy torch view of the resident plane."""
        ptr, pitch, w, h = ctypes.c_void_p(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(self.ctx.h, lib().mcv_rig_pyramid_dev(self.h, image, level, int(blurred), ctypes.byref(ptr),
                                                    ctypes.byref(pitch),
                                                    ctypes.byref(w), ctypes.byref(h)))
        return device_plane(ptr.value, h.value, w.value, pitch.value)

    def read(self):
        """Host copies: (kps records [3n][cap], desc [3n][cap][32], counts [3n],
        matches [3n][cap][3], match counts [3n], knn [3n][cap][4])."""
        check(self.ctx.h, lib().mcv_rig_check(self.h))
        n, c = 3 * self.n_frames, self.cap
        kps = self.ctx.to_host(self.d_kps, np.float32, (n, c, 6))
        desc = self.ctx.to_host(self.d_desc, np.uint8, (n, c, 32))
        counts = self.ctx.to_host(self.d_counts, np.int32, (n,))
        matches = self.ctx.to_host(self.d_matches, np.int32, (n, c, 3))
        mcounts = self.ctx.to_host(self.d_mcounts, np.int32, (n,))
        knn = self.ctx.to_host(self.d_knn, np.int32, (n, c, 4))
        return kp_records(kps), desc, counts, matches, mcounts, knn


def filter_orientation(matches, angles1, angles2, ctx: Context | None = None):
    """MatchRes::FilterOrientation: rows (query, train, distance) of the three
    largest rotation bins."""
    ctx = ctx or default_context()
    m = np.ascontiguousarray(matches, np.int32).reshape(-1, 3)
    a1 = np.ascontiguousarray(angles1, np.float32)
    a2 = np.ascontiguousarray(angles2, np.float32)
    out = np.zeros_like(m)
    k = ctypes.c_int32()
    check(ctx.h, lib().mcv_match_filter_orientation(ctx.h, _p(m), len(m), _p(a1), len(a1), _p(a2), len(a2),
                                                    _p(out), ctypes.byref(k)))
    return out[:k.value]


def distinctive_descriptors(desc, ptr, ctx: Context | None = None):
    """MapPoint::ComputeDistinctiveDescriptors for many map points: desc
    (total x 32 u8), ptr (n + 1) CSR offsets -> (best index per map point,
    chosen descriptors n x 32)."""
    ctx = ctx or default_context()
    desc = np.ascontiguousarray(desc, np.uint8)
    ptr = np.ascontiguousarray(ptr, np.int32)
    n = len(ptr) - 1
    best = np.zeros(n, np.int32)
    out = np.zeros((n, 32), np.uint8)
    check(ctx.h, lib().mcv_distinctive_descriptors(ctx.h, _p(desc), _p(ptr), n, _p(best), _p(out)))
    return best, out


class BowVocabulary:
    """DBoW3 vocabulary tree on the device (modules/DBow3/src/Vocabulary.h)
    with Vocabulary::transform and Matcher::DBowMatch."""

    def __init__(self, voc, ctx: Context | None = None):
        """voc: an object with desc (n x 32 u8), child_ptr, child_ids,
        word_id, weight, L, weighting, scoring (synth.Vocabulary)."""
        self.ctx = ctx or default_context()
        self._keep = [np.ascontiguousarray(voc.child_ptr, np.int32), np.ascontiguousarray(voc.child_ids, np.int32),
                      np.ascontiguousarray(voc.desc, np.uint8), np.ascontiguousarray(voc.word_id, np.int32),
                      np.ascontiguousarray(voc.weight, np.float64)]
        self.n = len(self._keep[3])
        h = ctypes.c_void_p()
        check(self.ctx.h, lib().mcv_bow_vocab_create(self.ctx.h, self.n, *[_p(x) for x in self._keep], int(voc.L),
                                                     int(voc.weighting), int(voc.scoring), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mcv_bow_vocab_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def transform(self, desc, levelsup: int = 4):
        """(word ids, weights, node ids) per descriptor, BowVector (words,
        values) and FeatureVector (nodes, ptr, features)."""
        d = np.ascontiguousarray(desc, np.uint8).reshape(-1, 32)
        n = len(d)
        w, wt, nid = np.zeros(n, np.int32), np.zeros(n), np.zeros(n, np.int32)
        bw, bv = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1))
        fn, fp, ff = np.zeros(max(n, 1), np.int32), np.zeros(n + 1, np.int32), np.zeros(max(n, 1), np.int32)
        nw, nn = ctypes.c_int32(), ctypes.c_int32()
        check(self.ctx.h, lib().mcv_bow_transform(self.h, _p(d), n, levelsup, _p(w), _p(wt), _p(nid), _p(bw), _p(bv),
                                                  ctypes.byref(nw), _p(fn), _p(fp), _p(ff), ctypes.byref(nn)))
        k, m = nw.value, nn.value
        return (w, wt, nid), (bw[:k].copy(), bv[:k].copy()), (fn[:m].copy(), fp[:m + 1].copy(), ff[:fp[m]].copy())

    def transform_dev(self, d_desc, n: int, levelsup: int = 4):
        """Device descriptors (torch.cuda u8 n x 32) -> device (word, weight, nid)."""
        import torch
        dev = d_desc.device
        w = torch.empty(n, dtype=torch.int32, device=dev)
        wt = torch.empty(n, dtype=torch.float64, device=dev)
        nid = torch.empty(n, dtype=torch.int32, device=dev)
        self.ctx.use_torch_stream()
        check(self.ctx.h, lib().mcv_bow_transform_dev(self.h, ctypes.c_void_p(d_desc.data_ptr()), n, levelsup,
                                                      ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(wt.data_ptr()),
                                                      ctypes.c_void_p(nid.data_ptr())))
        return w, wt, nid


def dbow_match(d1, fv1, d2, fv2, ctx: Context | None = None) -> np.ndarray:
    """Matcher::DBowMatch: rows (query, train0, d0, train1, d1)."""
    ctx = ctx or default_context()
    a = np.ascontiguousarray(d1, np.uint8).reshape(-1, 32)
    b = np.ascontiguousarray(d2, np.uint8).reshape(-1, 32)
    f1 = [np.ascontiguousarray(x, np.int32) for x in fv1]
    f2 = [np.ascontiguousarray(x, np.int32) for x in fv2]
    out = np.zeros((max(len(f1[2]), 1), 5), np.int32)
    cnt = ctypes.c_int32()
    check(ctx.h, lib().mcv_bow_match(ctx.h, _p(a), len(a), len(f1[0]), _p(f1[0]), _p(f1[1]), _p(f1[2]), _p(b), len(b),
                                     len(f2[0]), _p(f2[0]), _p(f2[1]), _p(f2[2]), _p(out), ctypes.byref(cnt)))
    return out[:cnt.value].copy()


class BALinearizer:
    """Batched linearization of EdgeSE3XYZ2Camera edges (see mcvslam_gpu.h)."""

    N_PLANES = 22

    def __init__(self, n_obs: int, n_points: int, n_frames: int, n_cams: int, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.M, self.P, self.F, self.C = n_obs, n_points, n_frames, n_cams
        h = ctypes.c_void_p()
        check(self.ctx.h, lib().mcv_ba_create(self.ctx.h, n_obs, n_points, n_frames, n_cams, ctypes.byref(h)))
        self.h = h

    @classmethod
    def from_problem(cls, prob, ctx: Context | None = None) -> "BALinearizer":
        b = cls(len(prob.obs_pt), len(prob.points), len(prob.frames), len(prob.cams), ctx)
        b.set_observations(prob.obs_uv, prob.obs_pt, prob.obs_frame, prob.obs_cam, prob.obs_isig)
        b.set_state(prob.points, prob.frames, prob.cams)
        if getattr(prob, "frame_fixed", None) is not None:
            b.set_fixed(prob.frame_fixed)
        return b

    def close(self):
        if getattr(self, "h", None):
            lib().mcv_ba_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_observations(self, uv, pt, frame, cam, isig):
        a = [np.ascontiguousarray(uv, np.float32), np.ascontiguousarray(pt, np.int32),
             np.ascontiguousarray(frame, np.int32), np.ascontiguousarray(cam, np.int32),
             np.ascontiguousarray(isig, np.float32)]
        check(self.ctx.h, lib().mcv_ba_set_observations(self.h, *[_p(x) for x in a]))

    def set_state(self, points, frames, cams):
        a = [np.ascontiguousarray(points, np.float64), np.ascontiguousarray(frames, np.float64),
             np.ascontiguousarray(cams, np.float64)]
        check(self.ctx.h, lib().mcv_ba_set_state(self.h, *[_p(x) for x in a]))

    FULL, RESIDUAL, POSE_ONLY = 0, 1, 2

    def linearize(self, residual_only: bool = False, mode: int | None = None):
        """mode: FULL (every plane), RESIDUAL (e, chi2, w) or POSE_ONLY (all
        but J_point, the fixed-point edge)."""
        m = (self.RESIDUAL if residual_only else self.FULL) if mode is None else int(mode)
        check(self.ctx.h, lib().mcv_ba_linearize(self.h, m))

    N_VARIANTS = 8

    def set_variant(self, variant: int):
        """Launch shape of the batched linearization (0 .. 7)."""
        check(self.ctx.h, lib().mcv_ba_set_variant(self.h, variant))

    def autotune(self, reps: int = 5) -> str:
        """Time every launch shape on this problem, keep the fastest (values
        per observation are the same); returns the report."""
        buf = ctypes.create_string_buffer(2048)
        check(self.ctx.h, lib().mcv_ba_autotune(self.h, reps, buf, 2048))
        return buf.value.decode()

    def set_stereo(self, obs, point, frame, inv_sigma2, cam):
        """Stereo edges (EdgeStereoSE3ProjectXYZ): obs Ms x 3 (u, v, u_right),
        cam = (fx, fy, cx, cy, bf, delta)."""
        a = [np.ascontiguousarray(obs, np.float32), np.ascontiguousarray(point, np.int32),
             np.ascontiguousarray(frame, np.int32), np.ascontiguousarray(inv_sigma2, np.float32)]
        c = np.ascontiguousarray(cam, np.float64)
        self.Ms = len(a[1])
        check(self.ctx.h, lib().mcv_ba_set_stereo(self.h, self.Ms, *[_p(x) for x in a], _p(c)))

    def linearize_stereo(self, residual_only: bool = False, mode: int | None = None):
        m = (self.RESIDUAL if residual_only else self.FULL) if mode is None else int(mode)
        check(self.ctx.h, lib().mcv_ba_linearize_stereo(self.h, m))

    def stereo_results(self):
        """(e Ms x 3, chi2, w, J_pose Ms x 3 x 6, J_point Ms x 3 x 3)"""
        pl = np.zeros((32, self.Ms), np.float32)
        check(self.ctx.h, lib().mcv_ba_read_stereo_planes(self.h, 0, 32, _p(pl)))
        return (pl[0:3].T, pl[3], pl[4], pl[5:23].T.reshape(-1, 3, 6), pl[23:32].T.reshape(-1, 3, 3))

    def outliers(self, thr_mono: float = 5.991, thr_stereo: float = 7.815, stereo: bool = False):
        """chi2 >= threshold flags of the last linearization (mono, stereo)."""
        fm = np.zeros(self.M, np.uint8)
        fs = np.zeros(getattr(self, "Ms", 0), np.uint8) if stereo else None
        check(self.ctx.h, lib().mcv_ba_outliers(self.h, ctypes.c_float(thr_mono), ctypes.c_float(thr_stereo),
                                                _p(fm), _p(fs) if stereo else None))
        return fm, fs

    def linearize_plain(self, residual_only: bool = False):
        """One launch per (frame, camera) edge group; same outputs."""
        check(self.ctx.h, lib().mcv_ba_linearize_plain(self.h, 1 if residual_only else 0))

    def generation(self) -> int:
        """Version of the planes (changes with every state, observation or
        linearization update)."""
        g = ctypes.c_int64()
        check(self.ctx.h, lib().mcv_ba_generation(self.h, ctypes.byref(g)))
        return g.value

    def planes_ptr(self):
        p, s = ctypes.c_void_p(), ctypes.c_int64()
        check(self.ctx.h, lib().mcv_ba_planes(self.h, ctypes.byref(p), ctypes.byref(s)))
        return p.value, s.value

    def read(self, first: int = 0, count: int = N_PLANES) -> np.ndarray:
        out = np.zeros((count, self.M), np.float32)
        check(self.ctx.h, lib().mcv_ba_read_planes(self.h, first, count, _p(out)))
        return out

    def results(self):
        """(e M x 2, chi2 M, w M, J_pose M x 2 x 6, J_point M x 2 x 3)"""
        pl = self.read()
        e = np.stack([pl[0], pl[1]], 1)
        return e, pl[2], pl[3], pl[4:16].T.reshape(-1, 2, 6), pl[16:22].T.reshape(-1, 2, 3)

    def set_fixed(self, frame_fixed):
        ff = np.ascontiguousarray(frame_fixed, np.int32)
        check(self.ctx.h, lib().mcv_ba_set_fixed(self.h, _p(ff)))

    def set_point_fixed(self, point_fixed=None):
        pf = None if point_fixed is None else np.ascontiguousarray(point_fixed, np.int32)
        check(self.ctx.h, lib().mcv_ba_set_point_fixed(self.h, _p(pf)))

    def set_active(self, active_mono=None, active_stereo=None):
        """Edge levels: False / 0 = level 1 (left out of the optimization)."""
        am = None if active_mono is None else np.ascontiguousarray(active_mono, np.uint8)
        ast = None if active_stereo is None else np.ascontiguousarray(active_stereo, np.uint8)
        check(self.ctx.h, lib().mcv_ba_set_active(self.h, _p(am), _p(ast)))

    def set_robust(self, enabled: bool):
        check(self.ctx.h, lib().mcv_ba_set_robust(self.h, 1 if enabled else 0))

    def robust_chi2(self) -> float:
        """Sum of the Huber rho(chi2) over the active edges of the last
        linearization (activeRobustChi2)."""
        v = ctypes.c_double()
        check(self.ctx.h, lib().mcv_ba_robust_chi2(self.h, ctypes.byref(v)))
        return v.value

    def optimize(self, iterations: int, tau: float = 1e-5, max_trials: int = 10):
        """g2o's Levenberg-Marquardt for `iterations` iterations over the GPU
        data; returns (stats [n x 4: chi2 before, chi2 after, lambda, trials],
        iterations run)."""
        stats = np.zeros((max(iterations, 1), 4))
        done = ctypes.c_int32()
        check(self.ctx.h, lib().mcv_ba_optimize(self.h, int(iterations), ctypes.c_double(tau), int(max_trials),
                                                _p(stats), ctypes.byref(done)))
        return stats[:done.value], done.value

    def get_state(self):
        """(points P x 3, frames F x 7) FP64 estimates."""
        pts = np.zeros((self.P, 3))
        frs = np.zeros((self.F, 7))
        check(self.ctx.h, lib().mcv_ba_get_state(self.h, _p(pts), _p(frs)))
        return pts, frs

    def schur(self, lam: float = 0.0):
        """Reduced camera system of the last linearization: (S, b_schur) as
        FP64 host arrays (computed in FP32 on the GPU)."""
        check(self.ctx.h, lib().mcv_ba_schur(self.h, ctypes.c_float(lam)))
        n = ctypes.c_int32()
        check(self.ctx.h, lib().mcv_ba_read_schur(self.h, None, None, ctypes.byref(n)))
        S = np.zeros((n.value, n.value))
        b = np.zeros(n.value)
        check(self.ctx.h, lib().mcv_ba_read_schur(self.h, _p(S), _p(b), ctypes.byref(n)))
        return S, b

    @staticmethod
    def solve_llt(S: np.ndarray, b: np.ndarray) -> np.ndarray:
        """FP64 Cholesky solve on the host."""
        S = np.ascontiguousarray(S, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        code = lib().mcv_solve_llt(len(b), _p(S), _p(b), _p(x))
        if code != 0:
            raise MCVError(code, "reduced camera system is not positive definite")
        return x

    def copy_planes_dev(self, first: int, count: int, d_ptr: int):
        check(self.ctx.h, lib().mcv_ba_copy_planes_dev(self.h, first, count, ctypes.c_void_p(d_ptr)))

    def adjoint_dev(self, g_ptr: int, gpose_ptr: int, gpoint_ptr: int):
        check(self.ctx.h, lib().mcv_ba_adjoint_dev(self.h, ctypes.c_void_p(g_ptr), ctypes.c_void_p(gpose_ptr),
                                                   ctypes.c_void_p(gpoint_ptr)))
