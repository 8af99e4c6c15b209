"""ORB extraction on the GPU against the oracle, bit for bit: keypoints
(x, y, size, angle, response, octave, in order), descriptors and pyramid
(ORBextractor.cc:831-919)."""
import numpy as np
import pytest

import oracle
from mcvslam_amd import ORB, MCVError, synth

pytestmark = pytest.mark.gpu


def _check(orb, img, seeds=None, **kw):
    kps, desc = orb.extract(img, seeds)
    rk, rd = oracle.orb_extract(img, orb.nfeatures, orb.scaleFactor, orb.nlevels, orb.iniThFAST, orb.minThFAST,
                                seeds=seeds)
    assert len(kps) == len(rk), (len(kps), len(rk))
    for f in ("x", "y", "size", "angle", "response", "octave"):
        np.testing.assert_array_equal(kps[f], rk[f], err_msg=f)
    np.testing.assert_array_equal(desc, rd)
    return kps, desc


@pytest.mark.parametrize("w,h,seed", [(640, 480, 1), (1280, 720, 10), (512, 512, 5), (752, 480, 7)])
def test_orb_blob_images(orb_default, w, h, seed):
    _check(orb_default, synth.blob_image(seed, w, h))


def test_orb_warped_pair(orb_default):
    a, b = synth.image_pair(1, 2, 640, 480)
    _check(orb_default, a)
    _check(orb_default, b)


def test_orb_pyramid_levels(orb_default):
    img = synth.blob_image(21, 640, 480)
    orb_default.extract(img)
    _, _, info, levels = oracle.orb_extract(img, with_levels=True)
    for l in range(orb_default.nlevels):
        np.testing.assert_array_equal(orb_default.pyramid(l), levels[l][0], err_msg=f"level {l}")


def test_orb_batch_pyramid_and_blurred_planes(orb_default):
    """The resident pyramid and the blurred copy the descriptors sample
    (GaussianBlur 7x7, sigma 2, REFLECT_101, ORBextractor.cc:874-875), plane
    by plane against the oracle's, for images of three sizes."""
    import torch
    imgs = [synth.blob_image(60 + i, *s) for i, s in enumerate([(640, 480), (1280, 720), (333, 250)])]
    dev = [torch.from_numpy(im).cuda() for im in imgs]
    orb_default.extract_batch_dev(dev, graph=True)
    torch.cuda.synchronize()
    for i, im in enumerate(imgs):
        _, _, info, levels = oracle.orb_extract(im, with_levels=True)
        for l in range(orb_default.nlevels):
            np.testing.assert_array_equal(orb_default.batch_pyramid(i, l).cpu().numpy(), levels[l][0],
                                          err_msg=f"image {i} level {l}")
            # the reference blurs a level only when it has keypoints
            if info[l, 2] > 0:
                np.testing.assert_array_equal(orb_default.batch_pyramid(i, l, blurred=True).cpu().numpy(),
                                              levels[l][1], err_msg=f"image {i} blurred level {l}")


@pytest.mark.parametrize("cfg", [(1000, 1.2, 8, 20, 7), (2000, 1.2, 1, 28, 15), (3000, 1.3, 6, 12, 5),
                                 (500, 1.2, 8, 40, 20), (1000, 1.2, 4, 1, 0), (1000, 1.2, 4, 0, 0),
                                 (1000, 1.2, 4, 200, 130)])
def test_orb_configs(gpu_ctx, cfg):
    orb = ORB(*cfg, ctx=gpu_ctx)
    _check(orb, synth.blob_image(31, 640, 480))


def test_orb_flat_and_noise_images(gpu_ctx):
    orb = ORB(2000, 1.2, 8, 20, 7, ctx=gpu_ctx)
    _check(orb, np.full((480, 640), 128, np.uint8))          # no corners anywhere
    rng = np.random.default_rng(0)
    _check(orb, rng.integers(0, 256, (480, 640), dtype=np.uint8))   # dense corners


def test_orb_seeds(orb_default):
    img = synth.blob_image(3, 640, 480)
    seeds = np.zeros(5, dtype=orb_default_dtype())
    seeds["x"] = [100.5, 200, 300.25, 50, 400]
    seeds["y"] = [100, 150.75, 200, 60, 300]
    seeds["size"] = 31
    seeds["angle"] = [10.0, 359.5, 0.0, 123.25, -1.0]
    seeds["response"] = [5, 6, 7, 8, 9]
    seeds["octave"] = [0, 2, 0, 7, 3]
    _check(orb_default, img, seeds)


def orb_default_dtype():
    from mcvslam_amd import KP_DTYPE
    return KP_DTYPE


def test_orb_empty_image(orb_default):
    with pytest.raises(MCVError):
        orb_default.extract(np.zeros((0, 0), np.uint8))


def test_orb_batch_matches_single(orb_default):
    import torch
    imgs = [synth.blob_image(40 + i, *s) for i, s in enumerate([(640, 480), (1280, 720), (800, 600)])]
    dev = [torch.from_numpy(im).cuda() for im in imgs]
    for graph in (False, True):
        kps, desc, counts = orb_default.extract_batch_dev(dev, graph=graph)
        torch.cuda.synchronize()
        kps, desc, counts = kps.cpu().numpy(), desc.cpu().numpy(), counts.cpu().numpy()
        from mcvslam_amd import kp_records
        for i, im in enumerate(imgs):
            rk, rd = oracle.orb_extract(im)
            n = counts[i]
            assert n == len(rk)
            rec = kp_records(kps[i, :n])
            for f in ("x", "y", "size", "angle", "response", "octave"):
                np.testing.assert_array_equal(rec[f], rk[f], err_msg=f)
            np.testing.assert_array_equal(desc[i, :n], rd)


def _check_batch(orb, imgs, kps, desc, counts, seeds=None):
    from mcvslam_amd import kp_records
    kps, desc, counts = kps.cpu().numpy(), desc.cpu().numpy(), counts.cpu().numpy()
    for i, im in enumerate(imgs):
        sd = None if seeds is None else seeds[i]
        rk, rd = oracle.orb_extract(im, orb.nfeatures, orb.scaleFactor, orb.nlevels, orb.iniThFAST, orb.minThFAST,
                                    seeds=sd)
        n = counts[i]
        assert n == len(rk), (i, n, len(rk))
        rec = kp_records(kps[i, :n])
        for f in ("x", "y", "size", "angle", "response", "octave"):
            np.testing.assert_array_equal(rec[f], rk[f], err_msg=f"image {i} {f}")
        np.testing.assert_array_equal(desc[i, :n], rd, err_msg=f"image {i}")


def test_orb_batch_graph_new_output_buffers(orb_default):
    """Two graph replays of one batch shape into different output tensors,
    the first call's tensors still alive: the graph is recaptured for the new
    buffers, and both results match the oracle."""
    import torch
    imgs = [synth.blob_image(80 + i, *s) for i, s in enumerate([(640, 480), (1280, 720)])]
    dev = [torch.from_numpy(im).cuda() for im in imgs]
    first = orb_default.extract_batch_dev(dev, graph=True)
    second = orb_default.extract_batch_dev(dev, graph=True)
    assert first[1].data_ptr() != second[1].data_ptr()
    torch.cuda.synchronize()
    _check_batch(orb_default, imgs, *first)
    _check_batch(orb_default, imgs, *second)
    orb_default.batch_check()


def test_orb_batch_seeds(orb_default):
    """Seed keypoints per image (the KL-tracked points appended to their
    octave, ORBextractor.cc:845-847) in the batched path, graph and eager,
    twice with new seed values of the same counts."""
    import torch
    from mcvslam_amd import KP_DTYPE
    imgs = [synth.blob_image(90 + i, *s) for i, s in enumerate([(1280, 720), (640, 480), (1280, 720)])]
    dev = [torch.from_numpy(im).cuda() for im in imgs]
    rng = np.random.default_rng(3)

    def make(n, w, h):
        sd = np.zeros(n, KP_DTYPE)
        sd["x"] = rng.uniform(20, w - 20, n).astype(np.float32)
        sd["y"] = rng.uniform(20, h - 20, n).astype(np.float32)
        sd["size"] = 31
        sd["angle"] = rng.uniform(0, 360, n).astype(np.float32)
        sd["response"] = rng.integers(1, 100, n)
        sd["octave"] = 0
        return sd

    for graph in (False, True, True):
        seeds = [make(40, 1280, 720), None, make(7, 1280, 720)]
        seeds[2]["octave"] = [0, 1, 0, 3, 0, 1, 7]
        out = orb_default.extract_batch_dev(dev, graph=graph, seeds=seeds)
        torch.cuda.synchronize()
        _check_batch(orb_default, imgs, *out, seeds=seeds)