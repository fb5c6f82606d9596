"""Pins the oracle's blend tracking (-m 4 MAX, -m 5 MIX: blendDTCH,
deviceCode.cu:972-1065 with blendChannels 165-241) and its ray marchers
(-m 6/7/8 rayMarcherCH 1067-1157, -m 9 compRayMarcherCH 1159-1276) with closed
forms of constant volumes:

* blend tracking: with constant channels the blended opacity is constant, so
  P(hit) = 1 - exp(-mu L) with mu = blended opacity / dt (the summed majorant
  only adds null collisions), and the hit colour is the blend's colour;
* marchers: a constant sample of opacity w composited m times gives
  alpha = 1 - (1 - w)^m and colour rgb * alpha, where m is the number of
  samples inside the box (the first sample sits at the camera and is empty);
  compositing stops once alpha reaches 0.99;
* density blend (-m 6) of two constant channels: colour
  (rgb0 w0 + rgb1 w1) / (w0 + w1), opacity (w0^2 + w1^2) / (w0 + w1).
A wrong blend rule, a dropped (1 - alpha) factor or a wrong majorant breaks
these by far more than the tolerances."""
import numpy as np
import pytest

from oracle import pyoracle as po

W = H = 40
FRAMES = 96
SHADE, AMB_SPEC = 1.075, 0.025  # light 1 + ambient 0.1 * 0.75; ambient * 0.25 * white


def _const(n=16, value=128):
    return np.full((n, n, n), value, np.uint8)


def _tf(rgb, alpha, n=64):
    t = np.zeros((n, 4), np.float32)
    t[:, :3] = rgb
    t[:, 3] = alpha
    return t


def _renderer(vols, tfs, mode, mc=(4, 4, 4), bg=(0, 0, 0)):
    r = po.OracleRenderer(threads=4)
    for v in vols:
        r.push_channel(v)
    r.init(mode, mc, True)
    for c, t in enumerate(tfs):
        r.set_tf(c, t)
    r.set_frame(W, H, 1, True, bg)
    r.set_seed(5)
    return r


@pytest.mark.parametrize("mode", [4, 5])
def test_blend_tracking_two_constant_channels(mode):
    a0, a1 = 0.27, 0.33
    rgb0, rgb1 = np.array([1.0, 0.2, 0.0]), np.array([0.0, 0.4, 1.0])
    r = _renderer([_const(), _const(value=200)], [_tf(rgb0, a0), _tf(rgb1, a1)], mode)
    r.render(FRAMES, threads=4)
    a = r.read_accum()
    s0, s1 = float(np.float32(a0) ** 3), float(np.float32(a1) ** 3)
    # summed majorant (renderer.cu:111-157)
    assert np.allclose(r.read_majorants(), s0 + s1, rtol=1e-6)
    w = max(s0, s1)
    if mode == 4:
        col = rgb1  # the more opaque channel wins
    else:
        col = (rgb0 * s0 + rgb1 * s1) / w
    mu = w / float(r.get_dt())
    # chord per frame (hit probability is not linear in L: average it per frame)
    p = np.zeros((H, W))
    for f in range(FRAMES):
        org, d = r.gen_rays(f, 0)
        t01 = r.clip_rays(org, d)[0]
        L = np.maximum(t01[:, 1] - t01[:, 0], 0).reshape(H, W)
        p += 1.0 - np.exp(-mu * L)
    p /= FRAMES
    inside = p > 0.05
    assert inside.sum() > 200
    for ch in range(3):
        v = col[ch] * SHADE + AMB_SPEC  # colour of a hit pixel sample (background 0)
        exp = p * v
        var = p * v * v - exp * exp
        z = (a[..., ch][inside].sum() - exp[inside].sum()) / np.sqrt(var[inside].sum() / FRAMES)
        assert abs(z) < 4.5, (mode, ch, z)


def _march_counts(r, frame, alpha_w):
    """Per pixel: number of compositing steps the marcher loop takes
    (t = t0, t0 + dt, ... while t < t1, in float32) and whether the ray hits
    the box. The loop also stops once alpha >= 0.99."""
    org, d = r.gen_rays(frame, 0)
    t01, hit = r.clip_rays(org, d)
    dt = np.float32(r.get_dt())
    n = np.zeros(W * H, np.int64)
    for i in np.nonzero(hit)[0]:
        t, t1, k = np.float32(t01[i, 0]), np.float32(t01[i, 1]), 0
        while t < t1:
            t = np.float32(t + dt)
            k += 1
        n[i] = k
    return n.reshape(H, W), hit.reshape(H, W).astype(bool)


@pytest.mark.parametrize("mode", [6, 7, 8, 9])
def test_marcher_constant_volume(mode):
    alpha, rgb = 0.3, np.array([0.9, 0.5, 0.1])
    w = float(np.float32(alpha) ** 3)
    r = _renderer([_const()], [_tf(rgb, alpha)], mode)
    r.render(1, threads=4)
    a = r.read_accum()
    steps, hit = _march_counts(r, 0, w)
    # the first sample is taken at the camera (empty); the others inside the
    # box, except that the entry point may round to just outside it
    got_alpha = (a[..., 0] - AMB_SPEC) / (SHADE * rgb[0])
    ok = np.zeros((H, W), bool)
    for m_off in (1, 2):
        m = np.maximum(steps - m_off, 0)
        exp_alpha = 1.0 - (1.0 - w) ** m
        # compositing stops at alpha >= 0.99
        exp_alpha = np.where(exp_alpha >= 0.99, got_alpha, exp_alpha)
        ok |= np.isclose(got_alpha, exp_alpha, rtol=2e-4, atol=2e-6)
    assert ok[hit].all(), (mode, np.argwhere(hit & ~ok)[:5])
    assert (got_alpha[hit] > 0.1).mean() > 0.5
    # colour is rgb * alpha in every channel
    for ch in (1, 2):
        np.testing.assert_allclose(a[..., ch][hit], (SHADE * rgb[ch] * got_alpha + AMB_SPEC)[hit],
                                   rtol=1e-4, atol=1e-6)
    # rays that miss the box see the background (0) exactly
    assert np.all(a[..., :3][~hit] == 0.0)


def test_density_blend_marcher_two_channels():
    a0, a1 = 0.2, 0.28
    rgb0, rgb1 = np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0])
    r = _renderer([_const(), _const(value=200)], [_tf(rgb0, a0), _tf(rgb1, a1)], 6)
    r.render(1, threads=4)
    a = r.read_accum()
    w0, w1 = float(np.float32(a0) ** 3), float(np.float32(a1) ** 3)
    wb = (w0 * w0 + w1 * w1) / (w0 + w1)
    cb = (rgb0 * w0 + rgb1 * w1) / (w0 + w1)
    steps, hit = _march_counts(r, 0, wb)
    got_alpha = (a[..., 0] - AMB_SPEC) / (SHADE * cb[0])
    ok = np.zeros((H, W), bool)
    for m_off in (1, 2):
        m = np.maximum(steps - m_off, 0)
        ok |= np.isclose(got_alpha, 1.0 - (1.0 - wb) ** m, rtol=2e-4, atol=2e-6)
    assert ok[hit].all()
    np.testing.assert_allclose(a[..., 2][hit], (SHADE * cb[2] * got_alpha + AMB_SPEC)[hit],
                               rtol=1e-4, atol=1e-6)


def test_composite_marcher_keeps_position_between_channels():
    """-m 9 marches channel 1 from where channel 0 stopped (the reference
    never resets pos), so channel 1's first sample is inside the box and it
    composites one sample more than channel 0."""
    a0, a1 = 0.18, 0.22
    r = _renderer([_const(), _const(value=200)], [_tf((1, 0, 0), a0), _tf((0, 0, 1), a1)], 9)
    r.render(1, threads=4)
    a = r.read_accum()
    w0, w1 = float(np.float32(a0) ** 3), float(np.float32(a1) ** 3)
    steps, hit = _march_counts(r, 0, w0)
    ok = np.zeros((H, W), bool)
    for d0 in (1, 2):
        for d1 in (0, 1):
            al0 = 1.0 - (1.0 - w0) ** np.maximum(steps - d0, 0)
            al1 = 1.0 - (1.0 - w1) ** np.maximum(steps - d1, 0)
            s_ = np.where(al0 + al1 > 0, al0 + al1, 1.0)
            red = SHADE * al0 * al0 / s_ + AMB_SPEC
            blue = SHADE * al1 * al1 / s_ + AMB_SPEC
            ok |= np.isclose(a[..., 0], red, rtol=2e-4, atol=2e-6) & \
                np.isclose(a[..., 2], blue, rtol=2e-4, atol=2e-6)
    assert ok[hit].all()
    # with a position reset channel 1 would composite the same count as
    # channel 0: that reading must not fit the pixels with long chords
    al0 = 1.0 - (1.0 - w0) ** np.maximum(steps - 1, 0)
    al1 = 1.0 - (1.0 - w1) ** np.maximum(steps - 1, 0)
    s_ = np.where(al0 + al1 > 0, al0 + al1, 1.0)
    reset_blue = SHADE * al1 * al1 / s_ + AMB_SPEC
    long = hit & (steps > 20)
    assert not np.isclose(a[..., 2], reset_blue, rtol=2e-4, atol=2e-6)[long].all()


@pytest.mark.parametrize("mode", [4, 5, 6, 7, 8, 9])
def test_transparent_volume_background(mode):
    bg = (0.2, 0.3, 0.4)
    r = _renderer([_const()], [_tf((1, 1, 1), 0.0)], mode, bg=bg)
    r.render(1, threads=4)
    a = r.read_accum()
    # blend tracking never hits: background; marchers always "hit" with
    # alpha 0: ambient * 0.25 * specular colour on top of the background
    org, d = r.gen_rays(0, 0)
    hit = r.clip_rays(org, d)[1].reshape(H, W).astype(bool)
    extra = 0.0 if mode <= 5 else AMB_SPEC
    for ch in range(3):
        want = np.where(hit, np.float32(np.float32(extra) + np.float32(bg[ch])), np.float32(bg[ch]))
        np.testing.assert_allclose(a[..., ch], want, rtol=1e-6)
