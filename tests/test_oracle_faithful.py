"""The oracle's default tracking path (the one the GPU matches bit for bit)
against a reference-faithful restatement of the same method.

The default path follows the GPU's arithmetic: a division-free polynomial
log, the transfer-function remap by reciprocals and Philox words at fixed
positions per (macrocell, channel). The faithful path (oracle_set_faithful)
uses what deviceCode.cu writes: libm's log (deviceCode.cu:563,805,913), the
two divisions of the remap (deviceCode.cu:36-37) and one sequential RNG
stream per ray in prd.rng()'s draw order (deviceCode.cu:296-298,797,830).
Both are unbiased estimators of the same image, so with independent seeds:

* per pixel, the default path's mean lies within the 99% confidence
  interval of the difference of two independent means (z = d / sqrt(se_a^2 +
  se_b^2), |z| <= 2.576) for all but a binomial share of the pixels: at most
  2.5% of them (1% expected);
* summed over the image, the difference is within 4 standard errors;
* against a 4096-frame faithful render, both k-frame images have an RMSE
  within 10% of the model sqrt(s2 (1/k + 1/4096)), s2 the per-frame variance
  averaged over pixels and RGB (estimated from the reference's moments).

A change that biases the default path (a dropped dt factor, the wrong
remap, a threshold from the wrong majorant) fails these checks."""
import math

import numpy as np
import pytest

from mdwtracker_amd import scenes

FRAMES, REF_FRAMES, SIDE = 512, 4096, 64


@pytest.fixture(scope="module")
def scene():
    return scenes.make_scene("ndda_cpu_2ch64", seed=1234)


def _render(po, sc, mode, faithful, seed, frames):
    o = po.OracleRenderer(threads=8)
    scenes.load_scene(o, sc, mode)
    o.set_frame(SIDE, SIDE, 1, True, (0.0, 0.0, 0.0))
    o.set_camera(*scenes.orbit_camera(1.6))  # the volume fills the frame
    o.set_seed(seed)
    o.set_faithful(faithful)
    o.render(frames, threads=8)
    a, m2 = o.read_accum(with_m2=True)
    o.close()
    mean = a[..., :3].astype(np.float64)
    var = np.maximum(m2[..., :3].astype(np.float64) - mean * mean, 0.0) * frames / (frames - 1)
    return mean, var / frames  # mean and the variance of that mean


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_default_path_agrees_with_faithful_path(oracle_mod, scene, mode):
    po = oracle_mod
    a, va = _render(po, scene, mode, False, 101, FRAMES)
    b, vb = _render(po, scene, mode, True, 202, FRAMES)
    se = np.sqrt(va + vb)
    live = se > 0
    z = (a - b)[live] / se[live]
    outside = float(np.mean(np.abs(z) > 2.576))
    zg = float((a - b)[live].sum() / math.sqrt((va + vb)[live].sum()))
    print(f"[faithful] mode {mode}: {live.sum()} values, {100 * outside:.2f}% outside the 99% CI, "
          f"global z {zg:.2f}, image mean {a.mean():.4f} / {b.mean():.4f}")
    assert live.sum() > 0.3 * a.size
    assert outside <= 0.025
    assert abs(zg) < 4.0
    ref, vref = _render(po, scene, mode, True, 303, REF_FRAMES)
    s2 = float(vref.mean() * REF_FRAMES)  # per-frame variance, mean over values
    model = math.sqrt(s2 * (1.0 / FRAMES + 1.0 / REF_FRAMES))
    for name, img in (("default", a), ("faithful", b)):
        rmse = float(np.sqrt(np.mean((img - ref) ** 2)))
        print(f"[faithful] mode {mode} {name}: RMSE vs {REF_FRAMES} frames {rmse:.3e} "
              f"(model {model:.3e})")
        assert 0.9 * model <= rmse <= 1.1 * model, (name, rmse, model)


def test_faithful_path_is_its_own_stream(oracle_mod, scene):
    """The faithful switch changes the random stream (images differ at equal
    seeds) and nothing else: a transparent volume renders the background
    exactly in both paths."""
    po = oracle_mod
    a, _ = _render(po, scene, 2, False, 7, 4)
    b, _ = _render(po, scene, 2, True, 7, 4)
    assert not np.array_equal(a, b)
    for faithful in (False, True):
        o = po.OracleRenderer(threads=4)
        o.push_channel(np.zeros((16, 16, 16), np.uint8))
        o.init(2, (4, 4, 4), True)
        t = np.ones((8, 4), np.float32)
        t[:, 3] = 0.0  # transparent
        o.set_tf(0, t)
        o.set_frame(24, 24, 1, True, (0.2, 0.3, 0.4))
        o.set_faithful(faithful)
        o.render(2, threads=4)
        acc = o.read_accum()
        np.testing.assert_array_equal(acc[..., :3], np.broadcast_to(np.float32([0.2, 0.3, 0.4]),
                                                                    (24, 24, 3)))
        o.close()
