"""Brute-force Hamming top-2 and the ratio/threshold filters on the GPU
against the oracle (src/Matcher.cpp:23-35, :100-111, :245-302)."""
import numpy as np
import pytest

import oracle
from mcvslam_amd import Context, Matcher

pytestmark = pytest.mark.gpu


def _desc(rng, n):
    return rng.integers(0, 256, (n, 32), dtype=np.uint8)


@pytest.mark.parametrize("nq,nt", [(1, 0), (5, 1), (7, 2), (64, 65), (300, 1000), (2000, 2000)])
def test_knn2_random(gpu_ctx, nq, nt):
    rng = np.random.default_rng(nq * 1000 + nt)
    q, t = _desc(rng, nq), _desc(rng, nt)
    got = Matcher.KnnMatch(q, t, ctx=gpu_ctx).rows
    np.testing.assert_array_equal(got, oracle.knn2(q, t))


@pytest.mark.parametrize("variant", range(Context.N_KNN_VARIANTS))
@pytest.mark.parametrize("nq,nt", [(4, 0), (3, 1), (65, 5), (130, 257), (700, 1531)])
def test_knn2_every_variant(gpu_ctx, variant, nq, nt):
    """All launch variants the autotuner chooses from return the oracle's rows
    (ragged query and train counts against the variants' block shapes)."""
    rng = np.random.default_rng(10 * nq + nt + variant)
    q, t = _desc(rng, nq), _desc(rng, nt)
    t[: nt // 3] = t[nt // 3: 2 * (nt // 3)]          # duplicates: scan-order ties
    try:
        gpu_ctx.set_match_variant(variant)
        got = Matcher.KnnMatch(q, t, ctx=gpu_ctx).rows
    finally:
        gpu_ctx.set_match_variant(0)
    np.testing.assert_array_equal(got, oracle.knn2(q, t))


def test_knn2_fp4_variant_large_train_sets(gpu_ctx):
    """The fp4 variant keys train indices in 16 bits: a train set of 65536 or
    more goes to the int8 kernel, and the best match past index 65535 comes
    back whole (planted exact copies there)."""
    rng = np.random.default_rng(77)
    nq, nt = 40, 70000
    q, t = _desc(rng, nq), _desc(rng, nt)
    t[69000:69000 + nq // 2] = q[: nq // 2]
    try:
        gpu_ctx.set_match_variant(10)
        got = Matcher.KnnMatch(q, t, ctx=gpu_ctx).rows
    finally:
        gpu_ctx.set_match_variant(0)
    want = oracle.knn2(q, t)
    np.testing.assert_array_equal(got, want)
    assert (want[: nq // 2, 0] == np.arange(69000, 69000 + nq // 2)).all()


def test_knn2_ties_keep_scan_order(gpu_ctx):
    """Duplicated trains: the first occurrence wins (strict '<' in order)."""
    rng = np.random.default_rng(3)
    base = _desc(rng, 50)
    t = np.concatenate([base, base, base[::-1]])
    q = np.concatenate([base[:20], _desc(rng, 30)])
    got = Matcher.KnnMatch(q, t, ctx=gpu_ctx).rows
    want = oracle.knn2(q, t)
    np.testing.assert_array_equal(got, want)
    assert (got[:20, 1] == 0).all() and (got[:20, 3] == 0).all()


def test_knn2_near_duplicates(gpu_ctx):
    """Trains one or two bit flips from the query: many equal distances."""
    rng = np.random.default_rng(4)
    q = _desc(rng, 256)
    t = np.repeat(q, 4, axis=0).copy()
    for i in range(len(t)):
        b = rng.integers(0, 256, rng.integers(0, 3))
        for bit in b:
            t[i, bit >> 3] ^= np.uint8(1 << (bit & 7))
    rng.shuffle(t)
    np.testing.assert_array_equal(Matcher.KnnMatch(q, t, ctx=gpu_ctx).rows, oracle.knn2(q, t))


@pytest.mark.parametrize("ratio,th", [(0.7, 50), (0.6, 46), (0.7, 34), (1.0, 256), (0.0, 0)])
def test_filter(gpu_ctx, ratio, th):
    rng = np.random.default_rng(int(ratio * 100) + th)
    q, t = _desc(rng, 1500), _desc(rng, 1200)
    # half the queries get a close train so the filters have work to do
    for i in range(0, 1500, 2):
        t[rng.integers(0, 1200)] = q[i] ^ rng.integers(0, 2, 32, dtype=np.uint8)
    knn = oracle.knn2(q, t)
    got = Matcher.filter(knn, ratio, th, ctx=gpu_ctx)
    np.testing.assert_array_equal(got, oracle.filter_ratio_threshold(knn, ratio, th))


def test_filter_edge_rows(gpu_ctx):
    """Sentinels: one candidate (d1 = 999) passes the ratio, none (999/999)
    and 0/0 do not; the threshold pass reorders by swap-with-last."""
    knn = np.array([[0, 999, 0, 999], [3, 10, 0, 999], [1, 0, 2, 0], [4, 60, 5, 200], [2, 5, 7, 9],
                    [6, 12, 8, 40], [9, 51, 1, 300], [5, 1, 3, 90]], np.int32)
    for ratio, th in [(0.7, 50), (0.6, 46), (2.0, 30)]:
        got = Matcher.filter(knn, ratio, th, ctx=gpu_ctx)
        np.testing.assert_array_equal(got, oracle.filter_ratio_threshold(knn, ratio, th))


def test_reference_chain(gpu_ctx):
    rng = np.random.default_rng(9)
    q, t = _desc(rng, 400), _desc(rng, 400)
    t[:200] = q[:200]
    res = Matcher.KnnMatch(q, t, ctx=gpu_ctx).FilterRatio(0.7).FilterThreshold(50)
    np.testing.assert_array_equal(res, oracle.filter_ratio_threshold(oracle.knn2(q, t), 0.7, 50))


def test_distinctive_descriptors(gpu_ctx):
    """MapPoint::ComputeDistinctiveDescriptors batched over map points with
    0..256 observations each, against the oracle (median index, first min)."""
    from mcvslam_amd import distinctive_descriptors
    rng = np.random.default_rng(21)
    sizes = [0, 1, 2, 3, 5, 8, 31, 64, 65, 127, 256] + list(rng.integers(1, 40, 200))
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    desc = np.empty((ptr[-1], 32), np.uint8)
    for m, n in enumerate(sizes):
        base = _desc(rng, 1)[0]
        for k in range(n):
            d = base.copy()
            for b in rng.integers(0, 256, rng.integers(0, 50)):
                d[b >> 3] ^= np.uint8(1 << (b & 7))
            desc[ptr[m] + k] = d
    desc[ptr[5]:ptr[6]] = desc[ptr[5]]                 # all-equal list: ties
    best, out = distinctive_descriptors(desc, ptr, ctx=gpu_ctx)
    want = oracle.distinctive(desc, ptr)
    np.testing.assert_array_equal(best, want)
    for m in range(len(sizes)):
        if want[m] >= 0:
            np.testing.assert_array_equal(out[m], desc[ptr[m] + want[m]])


@pytest.mark.parametrize("seed,n,used", [(1, 1, 40), (2, 37, 40), (3, 500, 40), (4, 2000, 6), (5, 4096, 40),
                                         (6, 300, 3), (7, 120, 12)])
def test_filter_orientation(gpu_ctx, seed, n, used):
    """MatchRes::FilterOrientation against the oracle, whose bins are ordered
    by the host's real std::sort: many equal bin sizes exercise the
    replicated tie order."""
    from mcvslam_amd import filter_orientation
    rng = np.random.default_rng(seed)
    a1 = rng.uniform(0, 360, n).astype(np.float32)
    bins = rng.integers(0, used, n)
    a2 = ((a1 - bins * np.float32(9.0) - rng.uniform(-4, 4, n).astype(np.float32)) % np.float32(360.0))
    a2 = a2.astype(np.float32)
    m = np.stack([np.arange(n), rng.permutation(n), rng.integers(0, 60, n)], 1).astype(np.int32)
    a2p = np.empty_like(a2)
    a2p[m[:, 1]] = a2
    got = filter_orientation(m, a1, a2p, ctx=gpu_ctx)
    np.testing.assert_array_equal(got, oracle.filter_orientation(m, a1, a2p))


@pytest.mark.parametrize("nq,nt", [(1, 0), (9, 1), (2000, 2000), (4000, 3500)])
def test_knn2_dev(gpu_ctx, nq, nt):
    """mcv_match_knn2_dev: device descriptors in, device rows out."""
    import ctypes

    import torch
    from mcvslam_amd._native import check, lib
    rng = np.random.default_rng(nq + 7 * nt)
    q, t = _desc(rng, nq), _desc(rng, nt)
    dq = torch.from_numpy(q).cuda()
    dt = torch.from_numpy(t).cuda() if nt else torch.zeros((1, 32), dtype=torch.uint8, device="cuda")
    out = torch.full((nq, 4), -7, dtype=torch.int32, device="cuda")
    gpu_ctx.use_torch_stream()
    check(gpu_ctx.h, lib().mcv_match_knn2_dev(gpu_ctx.h, ctypes.c_void_p(dq.data_ptr()), nq,
                                              ctypes.c_void_p(dt.data_ptr()), nt, ctypes.c_void_p(out.data_ptr())))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.knn2(q, t))
