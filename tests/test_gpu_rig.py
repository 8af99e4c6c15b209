"""Rig-frame batch (L, R, W extraction + L-R, W-L, W-R matching): the
hipGraph batch, the plain per-frame path and the oracle agree bit for bit."""
import numpy as np
import pytest

import oracle
from mcvslam_amd import RigBatch, synth

pytestmark = pytest.mark.gpu

PAIRS = [(0, 1), (2, 0), (2, 1)]


def assert_same_outputs(a, b):
    """Two rig runs agree on every written slot (slots past an image's
    count or a pair's match count hold whatever the buffer held before)."""
    kps_a, desc_a, counts_a, m_a, mc_a, knn_a = a
    kps_b, desc_b, counts_b, m_b, mc_b, knn_b = b
    np.testing.assert_array_equal(counts_a, counts_b)
    np.testing.assert_array_equal(mc_a, mc_b)
    for i, n in enumerate(counts_a):
        np.testing.assert_array_equal(kps_a[i, :n], kps_b[i, :n])
        np.testing.assert_array_equal(desc_a[i, :n], desc_b[i, :n])
    for p, mc in enumerate(mc_a):
        nq = counts_a[3 * (p // 3) + PAIRS[p % 3][0]]
        np.testing.assert_array_equal(knn_a[p, :nq], knn_b[p, :nq])
        np.testing.assert_array_equal(m_a[p, :mc], m_b[p, :mc])


def test_rig_two_frames(orb_default):
    import torch
    frames = [synth.rig_frame(100 + 10 * f) for f in range(2)]
    imgs = [torch.from_numpy(im).cuda() for fr in frames for im in fr]
    rig = RigBatch(orb_default, 2, ratio=0.7, th=50)
    results = []
    for graph in (True, False, True):
        rig.run([im.data_ptr() for im in imgs], graph=graph)
        results.append(rig.read())
    for r in results[1:]:
        assert_same_outputs(results[0], r)
    kps, desc, counts, matches, mcounts, knn = results[0]
    ref = [oracle.orb_extract(im) for fr in frames for im in fr]
    for i, (rk, rd) in enumerate(ref):
        n = counts[i]
        assert n == len(rk)
        for f in ("x", "y", "size", "angle", "response", "octave"):
            np.testing.assert_array_equal(kps[i, :n][f], rk[f])
        np.testing.assert_array_equal(desc[i, :n], rd)
    for f in range(2):
        for k, (qa, ta) in enumerate(PAIRS):
            rq, rt = ref[3 * f + qa][1], ref[3 * f + ta][1]
            want_knn = oracle.knn2(rq, rt)
            p = 3 * f + k
            np.testing.assert_array_equal(knn[p, :len(rq)], want_knn)
            want = oracle.filter_ratio_threshold(want_knn, 0.7, 50)
            assert mcounts[p] == len(want)
            np.testing.assert_array_equal(matches[p, :mcounts[p]], want)


def test_rig_uneven_feature_counts(orb_default):
    """Per-image feature counts (500-4000) in one batch: each image matches
    the oracle run with its own count."""
    import torch
    frames = [synth.rig_frame(300 + 10 * f) for f in range(2)]
    imgs = [torch.from_numpy(im).cuda() for fr in frames for im in fr]
    nf = np.array([500, 4000, 1200, 3000, 800, 2500], np.int32)
    rig = RigBatch(orb_default, 2, ratio=0.7, th=50, nfeatures=nf)
    outs = []
    for graph in (True, False):
        rig.run([im.data_ptr() for im in imgs], graph=graph)
        outs.append(rig.read())
    assert_same_outputs(outs[0], outs[1])
    kps, desc, counts, matches, mcounts, knn = outs[0]
    ims = [im for fr in frames for im in fr]
    for i, im in enumerate(ims):
        rk, rd = oracle.orb_extract(im, nfeatures=int(nf[i]))
        assert counts[i] == len(rk)
        for f in ("x", "y", "size", "angle", "response", "octave"):
            np.testing.assert_array_equal(kps[i, :counts[i]][f], rk[f])
        np.testing.assert_array_equal(desc[i, :counts[i]], rd)


@pytest.mark.parametrize("knn", [0, 10, 11])
def test_rig_images_without_keypoints(orb_default, knn):
    """Blank images in a batch: a pair without queries (W blank) and pairs
    without trains (R blank) run through the top-2 launch (the fp4 variants:
    the grid of real query tiles in train-count order) and match the oracle
    rows (index 0, distance 999 where a query has no train)."""
    import torch
    frames = [list(synth.rig_frame(500 + 10 * f)) for f in range(3)]
    frames[1][2] = np.full_like(frames[1][2], 128)   # W of frame 1
    frames[2][1] = np.full_like(frames[2][1], 77)    # R of frame 2
    imgs = [torch.from_numpy(np.ascontiguousarray(im)).cuda() for fr in frames for im in fr]
    rig = RigBatch(orb_default, 3, ratio=0.7, th=50)
    rig.set_variants(knn, 2, 2, 1)
    outs = []
    for graph in (True, False):
        rig.run([im.data_ptr() for im in imgs], graph=graph)
        outs.append(rig.read())
    assert_same_outputs(outs[0], outs[1])
    kps, desc, counts, matches, mcounts, knn_rows = outs[0]
    ref = [oracle.orb_extract(im) for fr in frames for im in fr]
    assert counts[5] == 0 and counts[7] == 0
    for i, (rk, rd) in enumerate(ref):
        assert counts[i] == len(rk)
        np.testing.assert_array_equal(desc[i, :counts[i]], rd)
    for f in range(3):
        for k, (qa, ta) in enumerate(PAIRS):
            rq, rt = ref[3 * f + qa][1], ref[3 * f + ta][1]
            want_knn = oracle.knn2(rq, rt)
            p = 3 * f + k
            np.testing.assert_array_equal(knn_rows[p, :len(rq)], want_knn)
            want = oracle.filter_ratio_threshold(want_knn, 0.7, 50)
            assert mcounts[p] == len(want)
            np.testing.assert_array_equal(matches[p, :mcounts[p]], want)


def test_rig_autotune_keeps_outputs(orb_default):
    """The autotuner only changes launch shapes: outputs after tuning (graph
    re-captured with the chosen variants) equal the untuned ones."""
    import torch
    frames = [synth.rig_frame(500 + 10 * f) for f in range(2)]
    imgs = [torch.from_numpy(im).cuda() for fr in frames for im in fr]
    ptrs = [im.data_ptr() for im in imgs]
    rig = RigBatch(orb_default, 2, ratio=0.7, th=50)
    rig.run(ptrs, graph=True)
    before = rig.read()
    report = rig.autotune(ptrs, reps=1)
    assert "chosen:" in report and report.count("k_knn2") >= 5
    assert report.count("k_blur") == 2 + 1                      # the two walks, then the choice
    assert report.count("graph, blur") == 2                     # before FAST / beside the heap
    assert report.count("k_keypoints") == RigBatch.N_KP_VARIANTS + 1
    assert report.count("k_octree_heap") == RigBatch.N_HEAP_VARIANTS + 1
    rig.run(ptrs, graph=True)
    assert_same_outputs(before, rig.read())


@pytest.mark.parametrize("knn,blur,kp,heap", [(1, 1, 1, 0), (6, 0, 2, 1), (7, 1, 0, 0), (0, 0, 3, 1), (2, 0, 2, 0),
                                             (8, 2, 2, 1), (9, 3, 1, 1), (8, 3, 2, 0), (10, 2, 2, 1),
                                             (11, 3, 2, 1)])
def test_rig_fixed_variants(orb_default, knn, blur, kp, heap):
    """Fixed launch variants (top-2 kernel, blur walk and its place in the
    captured graph, keypoints per wave, octree heap replay) give the outputs
    of the first ones, in the graph and in the plain path."""
    import torch
    frames = [synth.rig_frame(700 + 10 * f) for f in range(2)]
    imgs = [torch.from_numpy(im).cuda() for fr in frames for im in fr]
    ptrs = [im.data_ptr() for im in imgs]
    rig = RigBatch(orb_default, 2, ratio=0.7, th=50)
    rig.set_variants(0, 0, 0, 1)
    rig.run(ptrs, graph=True)
    before = rig.read()
    rig.set_variants(knn, blur, kp, heap)
    assert rig.variants() == (knn, blur, kp, heap)
    rig.run(ptrs, graph=True)
    assert_same_outputs(before, rig.read())
    rig.run(ptrs, graph=False)
    assert_same_outputs(before, rig.read())
    for _ in range(2):                      # graph replays after eager work
        rig.run(ptrs, graph=True)
        assert_same_outputs(before, rig.read())
    with pytest.raises(Exception):
        rig.set_variants(0, RigBatch.N_BLUR_VARIANTS)
    with pytest.raises(Exception):
        rig.set_variants(0, 0, RigBatch.N_KP_VARIANTS)
    with pytest.raises(Exception):
        rig.set_variants(0, 0, 2, RigBatch.N_HEAP_VARIANTS)


def test_rigs_in_flight(orb_default):
    """Several rigs on contexts and streams of their own, steps enqueued back
    to back without waiting (the split step forks onto each context's side
    stream), variants switched back and forth between steps (the rig keeps
    one set of graphs per variant set): every rig's outputs equal one rig's
    run on the default context."""
    import torch
    from mcvslam_amd import ORB, Context
    frames = [synth.rig_frame(900 + 10 * f) for f in range(2)]
    imgs = [torch.from_numpy(im).cuda() for fr in frames for im in fr]
    ptrs = [im.data_ptr() for im in imgs]
    ref = RigBatch(orb_default, 2, ratio=0.7, th=50)
    ref.run(ptrs, graph=True)
    want = ref.read()
    torch.cuda.synchronize()
    streams, ctxs, rigs = [], [], []
    for _ in range(3):
        st = torch.cuda.Stream()
        cx = Context(0)
        cx.set_stream(st.cuda_stream)
        streams.append(st)
        ctxs.append(cx)
        rigs.append(RigBatch(ORB(2000, 1.2, 8, 20, 7, ctx=cx), 2, ratio=0.7, th=50))
    variant_sets = [(8, 2, 2, 1), (1, 0, 2, 1), (8, 3, 2, 0), (8, 2, 2, 1)]
    try:
        for vs in variant_sets:
            for rg in rigs:
                rg.set_variants(*vs)
            for k in range(3 * len(rigs)):
                rigs[k % len(rigs)].run(ptrs, graph=True)
            torch.cuda.synchronize()
            for rg in rigs:
                assert_same_outputs(want, rg.read())
    finally:
        for rg in rigs:
            rg.close()
        for cx in ctxs:
            cx.close()
