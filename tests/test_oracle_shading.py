"""Pins the oracle's lighting path (deviceCode.cu:322-433, 243-281, 1117-1154;
renderer.h:121-142): the portable exp/pow, gradient (Blinn-Phong) shading,
Woodcock shadow rays, the ray marchers' own shadow march, and per-channel
material specular colours.

* exp / pow against float64 numpy;
* gradient shading on a linear ramp: the TF opacity gradient is constant, so
  every hit pixel of one frame has a colour fixed by its view direction alone
  (N, L, V, H and the normalised Blinn-Phong lobe in float64);
* constant volume: zero gradient, so diffuse = the material's coefficient;
* shadows in a constant cube: the expected image is the quadrature over the
  first-hit depth of the unblocked probability exp(-mu * tau_exit) along
  -lightDir (tau in the shadow ray's own parameter, as the reference traces
  the unnormalised direction), for every Woodcock mode;
* marcher shadow: 1 - (1 - w)^n composited along -lightDir from the most
  opaque sample, n counted like the reference's loop;
* material specular colour of the hit channel enters ambient and specular
  terms (0.25 weight).
A dropped (1 - ambient), a normalised shadow direction, a sign flip of N or a
wrong channel's material moves these by far more than the tolerances."""
import numpy as np
import pytest

from oracle import pyoracle as po

W = H = 40
FRAMES = 96
AMB = 0.1


def _const(n=16, value=128):
    return np.full((n, n, n), value, np.uint8)


def _tf(rgb, alpha, n=64):
    t = np.zeros((n, 4), np.float32)
    t[:, :3] = rgb
    t[:, 3] = alpha
    return t


def test_exp_pow_accuracy():
    xs = np.concatenate([np.linspace(-87.0, 87.9, 20001), np.linspace(-1, 1, 2001)]).astype(np.float32)
    got = np.array([po.expf(float(x)) for x in xs], np.float64)
    ref = np.exp(xs.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 3.0
    assert po.expf(0.0) == 1.0 and po.expf(-100.0) == 0.0 and np.isinf(po.expf(89.0))
    rng = np.random.default_rng(3)
    x = rng.uniform(0.002, 1.0, 4000).astype(np.float32)
    y = rng.uniform(1.0, 128.0, 4000).astype(np.float32)
    got = np.array([po.powf(float(a), float(b)) for a, b in zip(x, y)])
    ref = np.power(x.astype(np.float64), y.astype(np.float64))
    m = ref > 1e-30
    # log error (<= 2 ulp) is amplified by |y log x| in the exponent
    tol = 4e-7 * (1.0 + np.abs(y * np.log(x)))
    assert np.all(np.abs(got[m] - ref[m]) / ref[m] <= tol[m])


def _renderer(vols, tfs, mode, mc=(4, 4, 4), bg=(0, 0, 0), ranges=None):
    r = po.OracleRenderer(threads=4)
    for v in vols:
        r.push_channel(v)
    r.init(mode, mc, True)
    for c, t in enumerate(tfs):
        r.set_tf(c, t, 1.0, ranges[c] if ranges else (0.0, 1.0))
    r.set_frame(W, H, 1, True, bg)
    r.set_seed(9)
    return r


def _norm(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


@pytest.mark.parametrize("mode", [0, 2])
def test_gradient_shading_linear_ramp(mode):
    # value rises along x; the TF maps value linearly to cubed opacity, so
    # the opacity gradient points along +x and N = (-1, 0, 0)
    n = 52  # 255 / 51 = 5: the u8 ramp is exactly linear
    x = np.arange(n, dtype=np.float32)
    vol = np.broadcast_to(x * 5, (n, n, n)).astype(np.uint8).copy()
    tex = 256
    a3 = np.linspace(0.0, 0.06, tex, dtype=np.float64)  # opacity after cubing
    t = np.zeros((tex, 4), np.float32)
    t[:, :3] = (0.8, 0.6, 0.4)
    t[:, 3] = np.cbrt(a3).astype(np.float32)
    r = _renderer([vol], [t], mode)
    ld = np.array([-0.6, -0.3, -0.5], np.float32)
    kd, ks, shin = 0.7, 0.4, 16.0
    r.set_light(ld, shadows=False, gradient=True)
    r.set_material(0, kd, ks, shin, 0.01, (0.9, 0.8, 0.7))
    # seen from the low-density (-x) side, so that almost no hit lies within a
    # gradient step of the clamped faces
    r.set_camera((-1.3, 0.9, 1.5), (0.5, 0.5, 0.5), (0, 1, 0), 40.0)
    r.render(1, threads=4)
    a = r.read_accum()
    org, d = r.gen_rays(0, 0)
    Nv = np.array([-1.0, 0.0, 0.0])
    L = _norm(ld.astype(np.float64))
    V = _norm(-d.astype(np.float64))
    Hh = _norm(L[None] + V)
    ndl = max(float(Nv @ L), 0.0)
    ndh = np.maximum(Hh @ Nv, 0.0)
    diff = kd * ndl
    spec = np.where((ndl > 0) & (ndh > 0.001), ks * (shin + 8) / (8 * np.pi) * ndh ** shin, 0.0)
    alb = np.array([0.8, 0.6, 0.4])
    sc = np.array([0.9, 0.8, 0.7])
    mixed = alb * 0.75 + sc * 0.25
    want = diff * alb[None] + spec[:, None] * mixed[None] + AMB * mixed[None]
    hit = (a[..., :3].reshape(-1, 3) != 0).any(axis=1)
    assert hit.mean() > 0.2
    # pixels whose hit lies within one gradient step of a face use one-sided
    # differences of the same linear field: the same gradient
    close = np.isclose(a[..., :3].reshape(-1, 3), want, rtol=2e-4, atol=2e-5).all(axis=1)
    assert close[hit].mean() > 0.99, close[hit].mean()
    assert ndl > 0 and spec[hit].max() > 0.01


def test_constant_volume_gradient_is_zero():
    r = _renderer([_const()], [_tf((0.5, 0.7, 0.9), 0.3)], 2)
    r.set_light((0.2, -1.0, 0.1), shadows=False, gradient=True)
    r.set_material(0, 0.55, 0.4, 8.0, 0.01, (1.0, 1.0, 1.0))
    r.render(1, threads=4)
    a = r.read_accum()
    alb = np.array([0.5, 0.7, 0.9])
    want = 0.55 * alb + AMB * (alb * 0.75 + 0.25)
    px = a[..., :3].reshape(-1, 3)
    hit = (px != 0).any(axis=1)
    assert hit.sum() > 100
    np.testing.assert_allclose(px[hit], np.broadcast_to(want, px[hit].shape), rtol=1e-6)


def _slab_exit(p, d):
    """Ray parameter at which p + tau d leaves [0,1]^3 (p inside)."""
    with np.errstate(divide="ignore"):
        ta = (0.0 - p) / d
        tb = (1.0 - p) / d
    return np.min(np.maximum(ta, tb), axis=-1)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_shadow_rays_constant_cube(mode):
    alpha = 0.3
    rgb = np.array([0.9, 0.6, 0.3])
    r = _renderer([_const()], [_tf(rgb, alpha)], mode)
    # not unit length: the reference traces -lightDir as given
    ld = np.array([0.6, -2.0, 0.4], np.float32)
    r.set_light(ld, shadows=True, gradient=False)
    r.render(FRAMES, threads=4)
    a = r.read_accum()
    mu = float(np.float32(alpha) ** 3) / float(r.get_dt())
    ds = -ld.astype(np.float64)
    q = 96
    u = (np.arange(q) + 0.5) / q
    E = np.zeros((W * H, 3))
    E2 = np.zeros((W * H, 3))
    hitc = rgb  # albedo
    a_unb = (1.0 - AMB) * hitc + AMB * (0.75 * hitc + 0.25)
    b_blk = AMB * (0.75 * hitc + 0.25)
    for f in range(FRAMES):
        org, d = r.gen_rays(f, 0)
        t01, hitb = r.clip_rays(org, d)
        t0, t1 = t01[:, 0].astype(np.float64), t01[:, 1].astype(np.float64)
        Lc = np.where(hitb > 0, t1 - t0, 0.0)
        # first-hit depth s on [0, L] with density mu e^{-mu s}: midpoint rule in
        # the cumulative variable c = 1 - e^{-mu s} in [0, 1 - e^{-mu L}]
        cmax = 1.0 - np.exp(-mu * Lc)
        s = -np.log1p(-u[None, :] * cmax[:, None]) / mu
        x = org[:, None, :].astype(np.float64) + (t0[:, None] + s)[..., None] * d[:, None, :]
        pu = np.exp(-mu * _slab_exit(np.clip(x, 0, 1), ds[None, None, :]))
        w = cmax[:, None] / q
        E += (w * pu).sum(1)[:, None] * a_unb[None] + (w * (1 - pu)).sum(1)[:, None] * b_blk[None]
        E2 += (w * pu).sum(1)[:, None] * a_unb[None] ** 2 + (w * (1 - pu)).sum(1)[:, None] * b_blk[None] ** 2
    E /= FRAMES
    E2 /= FRAMES
    var = E2 - E * E
    got = a[..., :3].reshape(-1, 3)
    inside = E[:, 0] > 0.05
    assert inside.sum() > 300
    for ch in range(3):
        z = (got[inside, ch].sum() - E[inside, ch].sum()) / np.sqrt(var[inside, ch].sum() / FRAMES)
        assert abs(z) < 4.5, (mode, ch, z)
    # without shadows the same scene is brighter by (1 - (1 - AMB) * P(unblocked)) * albedo
    r.set_light(ld, shadows=False, gradient=False)
    r.render(FRAMES, threads=4)
    assert r.read_accum()[..., 0].sum() > got[:, 0].sum() * 1.05


def _march_steps(t_start, t_end, dt):
    t, k = np.float32(t_start), 0
    while t < np.float32(t_end):
        t = np.float32(t + dt)
        k += 1
    return k


def test_marcher_shadow_constant_cube():
    alpha, rgb = 0.25, np.array([0.9, 0.5, 0.2])
    w = float(np.float32(alpha) ** 3)
    r = _renderer([_const()], [_tf(rgb, alpha)], 7)
    ld = np.array([0.2, -1.0, 0.1], np.float32)
    r.set_light(ld, shadows=True, gradient=False)
    r.render(1, threads=4)
    a = r.read_accum()
    org, d = r.gen_rays(0, 0)
    t01, hitb = r.clip_rays(org, d)
    dt = np.float32(r.get_dt())
    ds = -ld.astype(np.float64)
    bad = 0
    checked = 0
    for i in np.nonzero(hitb)[0]:
        t0, t1 = t01[i]
        steps = _march_steps(t0, t1, dt)
        if steps < 6:
            continue
        dn = d[i].astype(np.float64) / np.linalg.norm(d[i].astype(np.float64))
        ok = False
        # the most opaque sample is the first inside the box: loop t of step
        # 1 or 2 (the entry may round to just outside)
        for j in (1, 2):
            hiT = float(t0) + j * float(dt)
            osh = org[i].astype(np.float64) + dn * hiT
            tex = _slab_exit(osh[None], ds[None])[0]
            nsh = _march_steps(1e-5, tex, dt)
            for m_sh in (nsh - 1, nsh - 2):
                ash = 1.0 - (1.0 - w) ** max(m_sh, 0)
                sh = (1 - AMB) * (1 - ash) + AMB
                for m in (steps - 1, steps - 2):
                    al = 1.0 - (1.0 - w) ** m
                    al = min(al, 1.0)
                    want = (1.0 + 0.075) * rgb[0] * al * sh + 0.025
                    if np.isclose(a.reshape(-1, 4)[i, 0], want, rtol=3e-4, atol=3e-6):
                        ok = True
        checked += 1
        bad += not ok
    assert checked > 200 and bad == 0


def test_material_spec_colour_per_channel():
    """Two constant channels with different specular colours: each hit pixel
    sample takes the ambient mix of its own channel's colour."""
    r = _renderer([_const(), _const(value=200)], [_tf((1, 0, 0), 0.3), _tf((0, 0, 1), 0.3)], 2)
    r.set_material(0, spec=(0.0, 1.0, 0.0))
    r.set_material(1, spec=(0.0, 0.0, 0.0))
    r.render(1, threads=4)
    px = r.read_accum()[..., :3].reshape(-1, 3)
    red = np.array([1.0 + AMB * 0.75, AMB * 0.25, 0.0])
    blue = np.array([0.0, 0.0, 1.0 + AMB * 0.75])
    is_red = np.isclose(px, red, rtol=1e-6).all(axis=1)
    is_blue = np.isclose(px, blue, rtol=1e-6).all(axis=1)
    is_bg = (px == 0).all(axis=1)
    assert np.all(is_red | is_blue | is_bg)
    assert is_red.sum() > 50 and is_blue.sum() > 50
