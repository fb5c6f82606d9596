"""CPU checks of the denoiser's host logic: the network input built from the
accumulated moments, the noise-corrected error estimate, the log-log crossing
used for time-to-RMSE, the training views and tiles, and the loss weights.
No GPU calls (the library is loaded, nothing is launched)."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def dn():
    from mdwtracker_amd import denoise
    return denoise


def test_network_input_channels(dn):
    rng = np.random.default_rng(0)
    h, w, k = 5, 7, 16
    samples = rng.uniform(0, 1, (k, h, w, 3))
    accum = np.zeros((h, w, 4), np.float32)
    accum[..., :3] = samples.mean(0)
    m2 = np.zeros_like(accum)
    m2[..., :3] = (samples ** 2).mean(0)
    aov = rng.uniform(0, 1, (h, w, 6)).astype(np.float32)
    x = dn.network_input(accum, m2, aov, k)
    assert x.shape == (h, w, dn.INPUTS) and x.dtype == np.float32
    np.testing.assert_array_equal(x[..., :3], accum[..., :3])
    np.testing.assert_allclose(x[..., 3], aov[..., 0] * dn.DEPTH_SCALE, rtol=1e-6)
    np.testing.assert_array_equal(x[..., 4:9], aov[..., 1:6])
    # standard error of the mean from the unbiased sample variance
    se = np.sqrt(samples.var(0, ddof=1).mean(-1) / k) * dn.STDERR_SCALE
    np.testing.assert_allclose(x[..., 9], se, rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(x[..., 10], math.log2(k) * dn.FRAMES_SCALE)
    x1 = dn.network_input(accum, m2, aov, 1)
    assert np.all(x1[..., 9] == 0) and np.all(x1[..., 10] == 0)
    # ten channels: every channel's hit fraction, then the noise level
    aov10 = rng.uniform(0, 1, (h, w, 12)).astype(np.float32)
    x10 = dn.network_input(accum, m2, aov10, k)
    assert x10.shape == (h, w, dn.input_width(10)) == (h, w, 17)
    np.testing.assert_array_equal(x10[..., 4:15], aov10[..., 1:])
    np.testing.assert_allclose(x10[..., 15], se, rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(x10[..., 16], math.log2(k) * dn.FRAMES_SCALE)


def test_noise_corrected_mse_is_unbiased(dn):
    # img = g + N(0, s1^2), ref = g + N(0, s2^2) independent: the estimate
    # must be s1^2 although the raw squared difference is s1^2 + s2^2
    rng = np.random.default_rng(1)
    g = rng.uniform(0, 1, (256, 256, 3))
    s1, s2 = 2e-3, 5e-3
    img = g + rng.normal(0, s1, g.shape)
    ref = g + rng.normal(0, s2, g.shape)
    est, se = dn.noise_corrected_mse(img, ref, np.full(g.shape, s2 * s2))
    raw = np.mean((img - ref) ** 2)
    assert abs(raw - (s1 * s1 + s2 * s2)) < 5 * se
    assert abs(est - s1 * s1) < 5 * se
    assert se < 0.05 * s1 * s1
    # a biased image: the estimate includes the bias
    est_b, _ = dn.noise_corrected_mse(img + 3e-3, ref, np.full(g.shape, s2 * s2))
    assert abs(est_b - (s1 * s1 + 9e-6)) < 5 * se


def test_frames_to_target(dn):
    levels = [1, 2, 4, 8, 16]
    mses = [1.0 / k for k in levels]
    # exact power law: the log-log interpolation is exact
    assert dn.frames_to_target(levels, mses, 1.0 / 5.0) == pytest.approx(5.0)
    assert dn.frames_to_target(levels, mses, 1.0 / 3.0) == pytest.approx(3.0)
    assert dn.frames_to_target(levels, mses, 2.0) == 1.0
    assert dn.frames_to_target(levels, mses, 1.0 / 64) is None
    # non-monotone tail: the first crossing counts
    assert dn.frames_to_target([1, 2, 4], [1.0, 0.1, 0.5], 0.2) < 2.0


def test_orbit_view(dn):
    rng = np.random.default_rng(2)
    for _ in range(50):
        frm, at, up, fovy = dn.orbit_view(rng)
        assert fovy == 45.0
        assert np.all(np.abs(np.subtract(at, 0.5)) <= 0.15)
        assert 1.5 - 1e-9 <= math.dist(frm, (0.5, 0.5, 0.5)) <= 2.3 + 1e-9
        d = np.subtract(at, frm) / math.dist(frm, at)
        assert abs(float(np.dot(d, up))) < 0.999  # up is never along the view


def test_tiles(dn):
    img = np.arange(8 * 12 * 2).reshape(8, 12, 2)
    t = dn.tiles(img, 4)
    assert len(t) == 6 and all(x.shape == (4, 4, 2) for x in t)
    np.testing.assert_array_equal(t[4], img[4:8, 4:8])
    assert len(dn.tiles(img, 5)) == 2  # partial tiles are left out


def test_level_weights(dn):
    ks = np.array([1, 4, 16, 64])
    w = dn.level_weights(ks)
    assert w.dtype == np.float32 and w.mean() == pytest.approx(1.0)
    np.testing.assert_allclose(w / w[0], ks.astype(np.float64))
