"""Stereo matching (Frame::ComputeStereoMatch, src/Frame.cpp:150-328) on the
GPU-resident pyramids of a batched rig run, against the oracle bit for bit."""
import numpy as np
import pytest

import oracle
from mcvslam_amd import RigBatch, synth

pytestmark = pytest.mark.gpu


def _frames():
    out = []
    for f, d in enumerate((12, 31)):
        left = synth.blob_image(600 + f, 1280, 720)
        right = np.zeros_like(left)
        right[:, :-d] = left[:, d:]
        out.append((left, right, synth.blob_image(700 + f, 640, 480)))
    out.append(synth.rig_frame(800))        # warped pair: few stereo matches
    return out


def test_stereo_vs_oracle(orb_default):
    import torch
    frames = _frames()
    imgs = [torch.from_numpy(im).cuda() for fr in frames for im in fr]
    rig = RigBatch(orb_default, len(frames), ratio=0.7, th=50)
    rig.run([im.data_ptr() for im in imgs], graph=True)
    kps, desc, counts, _, _, _ = rig.read()
    bf, b = 2388.46, 1.0
    u, depth, sad = rig.stereo(bf, b)
    t = oracle.orb_tables()
    for f, (left, right, _) in enumerate(frames):
        kl, dl, _, lv_l = oracle.orb_extract(left, with_levels=True)
        kr, dr, _, lv_r = oracle.orb_extract(right, with_levels=True)
        ru, rd, rs = oracle.stereo(kl, dl, kr, dr, [x[0] for x in lv_l], [x[0] for x in lv_r], t["scale"],
                                   t["inv_scale"], bf, b)
        n = counts[3 * f]
        assert n == len(kl)
        np.testing.assert_array_equal(u[f, :n], ru)
        np.testing.assert_array_equal(depth[f, :n], rd)
        np.testing.assert_array_equal(sad[f, :n], rs)
        if f < 2:
            assert (ru >= 0).sum() > 200
