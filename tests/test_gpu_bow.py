"""DBoW3 on the GPU against the oracle, bit for bit: Vocabulary::transform
(word, weight, node at level L - levelsup) for every descriptor, the
BowVector and FeatureVector, and Matcher::DBowMatch rows
(modules/DBow3/src/Vocabulary.cpp:572-672, src/Matcher.cpp:146-193)."""
import numpy as np
import pytest

import oracle
from mcvslam_amd import BowVocabulary, dbow_match, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,L,levelsup,weighting,scoring", [(8, 5, 4, 0, 0), (10, 4, 2, 1, 1), (5, 6, 4, 2, 5),
                                                             (20, 3, 0, 3, 2)])
def test_transform_vs_oracle(gpu_ctx, k, L, levelsup, weighting, scoring):
    voc = synth.bow_vocabulary(seed=90 + k, k=k, L=L, weighting=weighting, scoring=scoring)
    feats = synth.bow_features(voc, seed=91, n=2500, max_flips=60)
    feats[:3] = voc.desc[voc.leaves[:3]]        # exact leaf copies
    bv = BowVocabulary(voc, ctx=gpu_ctx)
    (w, wt, nid), (bw, bvv), (fn, fp, ff) = bv.transform(feats, levelsup)
    rw, rwt, rnid = oracle.bow_transform(voc, feats, levelsup)
    np.testing.assert_array_equal(w, rw)
    np.testing.assert_array_equal(wt, rwt)
    np.testing.assert_array_equal(nid, rnid)
    (obw, obv), (ofn, ofp, off) = oracle.bow_vectors(rw, rwt, rnid, weighting, scoring)
    np.testing.assert_array_equal(bw, obw)
    np.testing.assert_array_equal(bvv, obv)
    np.testing.assert_array_equal(fn, ofn)
    np.testing.assert_array_equal(fp, ofp)
    np.testing.assert_array_equal(ff, off)
    bv.close()


def test_transform_dev_matches_host(gpu_ctx):
    import torch
    voc = synth.bow_vocabulary(seed=95, k=8, L=5)
    feats = synth.bow_features(voc, seed=96, n=1000)
    bv = BowVocabulary(voc, ctx=gpu_ctx)
    w, wt, nid = bv.transform_dev(torch.from_numpy(feats).cuda(), len(feats), 4)
    torch.cuda.synchronize()
    rw, rwt, rnid = oracle.bow_transform(voc, feats, 4)
    np.testing.assert_array_equal(w.cpu().numpy(), rw)
    np.testing.assert_array_equal(wt.cpu().numpy(), rwt)
    np.testing.assert_array_equal(nid.cpu().numpy(), rnid)


def test_dbow_match_vs_oracle(gpu_ctx):
    voc = synth.bow_vocabulary(seed=97, k=8, L=5)
    d1 = synth.bow_features(voc, seed=98, n=2000, max_flips=40)
    d2 = synth.bow_features(voc, seed=99, n=1800, max_flips=40)
    d2[:100] = d1[:100]                         # zero distances and ties
    bv = BowVocabulary(voc, ctx=gpu_ctx)
    fv1 = bv.transform(d1, 2)[2]
    fv2 = bv.transform(d2, 2)[2]
    got = dbow_match(d1, fv1, d2, fv2, ctx=gpu_ctx)
    want = oracle.bow_match(d1, fv1, d2, fv2)
    np.testing.assert_array_equal(got, want)
    assert len(want) > 500
