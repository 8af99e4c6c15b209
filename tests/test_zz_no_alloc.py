"""mcv_match_filter_noalloc_dev (the filter with its scratch in a fixed
per-GPU array, no allocation) against mcv_match_filter, the entry point it
varies, on two calls in a row (the second smaller), through the C ABI."""
import ctypes

import numpy as np
import pytest

import oracle
from mcvslam_amd import Matcher
from mcvslam_amd._native import MCV_ERR_CAPACITY, check, lib

pytestmark = pytest.mark.gpu


def _knn_rows(rng, nq, nt):
    q = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    t = rng.integers(0, 256, (nt, 32), dtype=np.uint8)
    for i in range(0, nq, 2):      # half the queries get a close train
        t[rng.integers(0, nt)] = q[i] ^ (rng.integers(0, 2, 32, dtype=np.uint8) << rng.integers(0, 8))
    return oracle.knn2(q, t)


def _noalloc(ctx, knn, ratio, th):
    import torch
    nq = len(knn)
    d_knn = torch.from_numpy(np.ascontiguousarray(knn, np.int32)).cuda()
    d_out = torch.full((max(nq, 1), 3), -5, dtype=torch.int32, device="cuda")
    d_cnt = torch.full((1,), -5, dtype=torch.int32, device="cuda")
    ctx.use_torch_stream()
    code = lib().mcv_match_filter_noalloc_dev(ctx.h, ctypes.c_void_p(d_knn.data_ptr()), nq, ctypes.c_float(ratio),
                                              th, ctypes.c_void_p(d_out.data_ptr()), ctypes.c_void_p(d_cnt.data_ptr()))
    check(ctx.h, code)
    torch.cuda.synchronize()
    n = int(d_cnt.cpu().numpy()[0])
    return d_out.cpu().numpy()[:n]


def test_noalloc_filter_matches_allocating_entry_point(gpu_ctx):
    rng = np.random.default_rng(12)
    for nq, nt, ratio, th in [(30000, 4000, 0.7, 50), (1700, 900, 0.6, 46)]:
        knn = _knn_rows(rng, nq, nt)
        want = Matcher.filter(knn, ratio, th, ctx=gpu_ctx)      # mcv_match_filter
        got = _noalloc(gpu_ctx, knn, ratio, th)
        assert len(want) > 0
        np.testing.assert_array_equal(got, want, err_msg=f"nq={nq}")


def test_noalloc_filter_limits(gpu_ctx):
    import torch
    rng = np.random.default_rng(13)
    knn = _knn_rows(rng, 65536, 500)
    np.testing.assert_array_equal(_noalloc(gpu_ctx, knn, 0.8, 60), Matcher.filter(knn, 0.8, 60, ctx=gpu_ctx))
    assert len(_noalloc(gpu_ctx, knn[:0], 0.8, 60)) == 0
    d = torch.zeros((65537, 4), dtype=torch.int32, device="cuda")
    o = torch.zeros((65537, 3), dtype=torch.int32, device="cuda")
    c = torch.zeros(1, dtype=torch.int32, device="cuda")
    code = lib().mcv_match_filter_noalloc_dev(gpu_ctx.h, ctypes.c_void_p(d.data_ptr()), 65537, ctypes.c_float(0.7),
                                              50, ctypes.c_void_p(o.data_ptr()), ctypes.c_void_p(c.data_ptr()))
    assert code == MCV_ERR_CAPACITY
