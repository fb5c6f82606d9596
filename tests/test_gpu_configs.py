"""GPU parity on the shapes of the batched and the large multi-channel
workloads, and on the default macrocell grid.

* Ten f32 channels (the 10-channel 512x512x128 workload at 1/8 scale), -m 2
  and -m 3 without lighting, so the plain frame kernels run: accumulated
  images, second moments and work counters bit-exact with the oracle. At full
  size (2048^2, ten 512x512x128 f32 channels) a homogeneous medium meets its
  closed form: hit probability 1 - exp(-sum mu L) and the channel mix
  mu_c / sum mu.
* Four f16 channels (the batched 4 x 512^3 workload at 1/8 scale): 16 views
  with sizes and spp drawn like the batched workload's (sizes scaled by 1/8,
  one view above 256 px), in both launch forms: every view bit-identical to
  its single-view render, two views bit-identical to the oracle (launch
  forms 0, 1 and 2: per-view launches, one launch, one persistent launch
  replayed from a HIP graph); the
  autotuner's full macrocell x workgroup x tile sweep on that batch.
* mc = 0 (CalculateMCGridDims(4), renderer.cpp:641-647,1342-1355) on a
  ragged volume: grid dims, range grid, majorants and image equal the
  oracle's."""
import os

import numpy as np
import pytest

from mdwtracker_amd import scenes

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def capi():
    from mdwtracker_amd import capi
    return capi


@pytest.fixture(scope="module")
def po():
    from oracle import pyoracle
    return pyoracle


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def large8():
    return scenes.make_scene("large_10ch_f32", seed=1234, scale=8)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_ten_channel_f32_plain_kernels_bit_exact(capi, po, large8, mode):
    sc = large8
    print(f"[10ch f32] mode={mode} dims={sc.volumes[0].shape} mc={sc.mc} image={sc.width}", flush=True)
    g, o = capi.MdwtRenderer(0), po.OracleRenderer()
    g.set_moments(True)
    scenes.load_scene(g, sc, mode)
    scenes.load_scene(o, sc, mode)
    np.testing.assert_array_equal(_bits(g.read_majorants()), _bits(o.read_majorants()))
    cg = g.count_frame()
    g.render(3)
    o.render(3, threads=THREADS)
    a_g, m_g = g.read_accum(with_m2=True)
    a_o, m_o = o.read_accum(with_m2=True)
    np.testing.assert_array_equal(_bits(a_g), _bits(a_o))
    np.testing.assert_array_equal(_bits(m_g), _bits(m_o))
    o1 = po.OracleRenderer()
    scenes.load_scene(o1, sc, mode)
    o1.render(1, threads=THREADS)
    assert cg == o1.read_counters()
    assert cg["samples"] > 1000 and a_g[..., :3].max() > 0.05  # not an empty image
    g.close()


def _chords(cam, W, H, jit=0.5):
    """Chord through the unit box of the ray of each pixel at sub-pixel offset
    (jit, jit), the renderer's diagonal jitter (camera ray at (px + jit) / W,
    (py + jit) / H)."""
    org, llc, hor, ver = cam.astype(np.float64)
    ys, xs = np.mgrid[0:H, 0:W]
    p = llc + ((xs + jit) / W)[..., None] * hor + ((ys + jit) / H)[..., None] * ver
    d = p - org
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    with np.errstate(divide="ignore"):
        ta, tb = (0.0 - org) / d, (1.0 - org) / d
    tn = np.max(np.minimum(ta, tb), axis=-1)
    tf = np.min(np.maximum(ta, tb), axis=-1)
    return np.where(tf > tn, tf - tn, 0.0)


def _hit_probability(cam, W, H, M, n=16):
    """Per pixel E_jit[1 - exp(-M L(jit))] over the uniform jitter (midpoint
    rule, n points): near the box's silhouette the chord changes within a
    pixel, and the hit probability of the centre ray alone overstates the
    mean (1 - exp(-M L) is concave in L)."""
    ph = np.zeros((H, W))
    for k in range(n):
        ph += 1.0 - np.exp(-M * _chords(cam, W, H, (k + 0.5) / n))
    return ph / n


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_ten_channel_full_size_homogeneous(capi, mode):
    """Ten constant f32 channels of 512x512x128 at 2048^2: per pixel the hit
    probability is 1 - exp(-M L) with M the summed extinction, and a hit is
    channel c with probability mu_c / M. Every channel's colour is
    (1, c/9, 0), so the mean red over 1.1 estimates the hit probability and
    the mean green the channel mix (shading: h + ambient (0.75 h + 0.25))."""
    nx, ny, nz, C, frames = 512, 512, 128, 10, 12
    alphas = np.linspace(0.18, 0.30, C).astype(np.float32)
    g = capi.MdwtRenderer(0)
    vol = np.full((nz, ny, nx), 0.5, np.float32)
    for _ in range(C):
        g.push_channel(vol)
    del vol
    g.init(mode, (32, 32, 8), True)
    for c in range(C):
        t = np.zeros((32, 4), np.float32)
        t[:, 0] = 1.0
        t[:, 1] = c / 9.0
        t[:, 3] = alphas[c]
        g.set_tf(c, t)
    W = H = 2048
    g.set_frame(W, H, 1, True, (0.0, 0.0, 0.0))
    print(f"[homogeneous 10ch] mode {mode}: 2048^2, {frames} frames", flush=True)
    g.render(frames)
    a = g.read_accum()
    dt = g.get_dt()
    mu = (alphas.astype(np.float64) ** 3) / dt  # sigma = alpha^3 (opacity 1) per dt
    M = mu.sum()
    ph = _hit_probability(g.get_camera(), W, H, M)
    inside = (ph > 0.05) & (ph < 0.95)
    gc = np.arange(C) / 9.0
    pc = mu / M
    ered = ph * 1.1
    egreen = ph * float(np.sum(pc * (1.075 * gc + 0.025)))
    e2green = ph * float(np.sum(pc * (1.075 * gc + 0.025) ** 2))
    for name, val, mean, var in (("red", a[..., 0], ered, ph * 1.21 - ered ** 2),
                                 ("green", a[..., 1], egreen, e2green - egreen ** 2)):
        z = (val[inside].sum() - mean[inside].sum()) / np.sqrt(var[inside].sum() / frames)
        print(f"[homogeneous 10ch] mode {mode} {name}: {inside.sum()} pixels, z = {z:.2f}", flush=True)
        assert abs(z) < 6.0, (name, z)
    g.close()


# ---------------------------------------------------------------- batch ----
@pytest.fixture(scope="module")
def batch8():
    return scenes.make_scene("batch_4ch512_f16", seed=1234, scale=8)


def _batch_views(n, seed, scale=8):
    """n views like the batched workload's (orbit radius 2.2; width and height
    uniform in 256..1024, spp uniform in 1..16), sizes divided by `scale`;
    the last view is kept above 256 px."""
    rng = np.random.default_rng(seed)
    cams = scenes.orbit_cameras(n, rng, radius=2.2)
    out = []
    for i in range(n):
        w = int(rng.integers(256, 1025)) // scale
        h = int(rng.integers(256, 1025)) // scale
        spp = int(rng.integers(1, 17))
        if i == n - 1:
            w, h, spp = 300, 260, 2
        frm = tuple(float(c) for c in cams[i])
        up = (0.0, 1.0, 0.0) if abs(frm[1] - 0.5) < 0.95 * 2.2 else (1.0, 0.0, 0.0)
        out.append({"from": frm, "at": (0.5, 0.5, 0.5), "up": up, "fovy": 45.0,
                    "width": w, "height": h, "spp": spp, "seed": int(rng.integers(0, 2**31))})
    return out


def _single_view(capi, sc, mode, v, frames):
    g = capi.MdwtRenderer(0)
    g.set_moments(True)
    scenes.load_scene(g, sc, mode)
    g.set_frame(v["width"], v["height"], v["spp"], True, (0.0, 0.0, 0.0))
    g.set_camera(v["from"], v["at"], v["up"], v["fovy"])
    g.set_seed(v["seed"])
    cnt = g.count_frame()
    g.render(frames)
    a = g.read_accum(with_m2=True)
    g.close()
    return a, cnt


@pytest.mark.parametrize("launch", [0, 1, 2])
def test_f16_batch_views_equal_single_views_and_oracle(capi, po, batch8, launch):
    sc = batch8
    views = _batch_views(16, seed=9)
    print(f"[f16 batch] launch={launch} dims={sc.volumes[0].shape} "
          f"views={[(v['width'], v['height'], v['spp']) for v in views]}", flush=True)
    g = capi.MdwtRenderer(0)
    g.set_moments(True)
    scenes.load_scene(g, sc, 2)
    g.batch_set_launch(launch)
    g.batch_set_views(views)
    cb = g.batch_count()
    g.batch_render(1)
    g.batch_render(2)
    tot = {"samples": 0, "visits": 0, "rejections": 0, "rays": 0}
    for i, v in enumerate(views):
        (a_s, m_s), cnt = _single_view(capi, sc, 2, v, 3)
        a_b, m_b = g.batch_read_view(i, with_m2=True)
        np.testing.assert_array_equal(_bits(a_b), _bits(a_s))
        np.testing.assert_array_equal(_bits(m_b), _bits(m_s))
        for k in tot:
            tot[k] += cnt[k]
    assert cb == tot
    # two views (the large one and a small one) against the oracle
    for i in (len(views) - 1, 0):
        v = views[i]
        o = po.OracleRenderer()
        scenes.load_scene(o, sc, 2)
        o.set_frame(v["width"], v["height"], v["spp"], True, (0.0, 0.0, 0.0))
        o.set_camera(v["from"], v["at"], v["up"], v["fovy"])
        o.set_seed(v["seed"])
        o.render(3, threads=THREADS)
        a_o, m_o = o.read_accum(with_m2=True)
        a_b, m_b = g.batch_read_view(i, with_m2=True)
        np.testing.assert_array_equal(_bits(a_b), _bits(a_o))
        np.testing.assert_array_equal(_bits(m_b), _bits(m_o))
    g.close()


def test_f16_batch_full_autotune_sweep(capi, batch8):
    sc = batch8
    g = capi.MdwtRenderer(0)
    scenes.load_scene(g, sc, 2)
    g.batch_set_views(_batch_views(16, seed=10))
    print("[f16 batch] autotune s (4, 8, 16, 32) x threads (64, 128, 256) x tiles (8, 16, 32)",
          flush=True)
    table, best = g.autotune((4, 8, 16, 32), (64, 128, 256), (8, 16, 32), frames=1)
    print(f"[f16 batch] best={best}", flush=True)
    assert table.shape == (36, 6)
    assert np.all(table[:, 3] > 0) and np.all(table[:, 4] > 0) and np.all(table[:, 5] > 0)
    combos = {(int(r[0]), int(r[1]), int(r[2])) for r in table}
    assert len(combos) == 36
    row = table[(table[:, 0] == best[0]) & (table[:, 1] == best[1]) & (table[:, 2] == best[2])]
    assert row[0, 3] == table[:, 3].min()
    assert g.info()["mc"] == (64 // best[0],) * 3
    g.batch_render(1)
    assert np.all(np.isfinite(g.batch_read_view(0)))
    g.close()


# ------------------------------------------------------- default grid ------
@pytest.mark.parametrize("mode", [0, 2, 3])
def test_default_macrocell_grid_matches_oracle(capi, po, mode):
    """mc = (0, 0, 0) takes CalculateMCGridDims(4): dims / 4 per axis, here on
    a 23 x 17 x 30 (x, y, z) three-channel volume whose sides 4 does not
    divide."""
    rng = np.random.default_rng(31)
    shape = (30, 17, 23)
    vols = [scenes.quantize_u8(scenes.gaussian_blobs(shape, 3, rng)) for _ in range(3)]
    tfs = [scenes.piecewise_tf(5, rng, c / 3.0, n_texels=64) for c in range(3)]
    g, o = capi.MdwtRenderer(0), po.OracleRenderer()
    for r in (g, o):
        for v in vols:
            r.push_channel(v)
        r.init(mode, (0, 0, 0), True)
        for c, t in enumerate(tfs):
            r.set_tf(c, t, 3.0, (0.0, 1.0))
        r.set_frame(61, 47, 1, True, (0.0, 0.0, 0.0))
        r.set_camera([1.8, 1.1, -0.9], [0.5, 0.45, 0.5], [0, 1, 0], 40.0)
        r.set_seed(8)
    print(f"[default mc] mode={mode} gpu {g.info()['mc']} oracle {o.info()['mc']}", flush=True)
    assert g.info()["mc"] == o.info()["mc"] == (23 // 4, 17 // 4, 30 // 4)
    np.testing.assert_array_equal(_bits(g.read_ranges()), _bits(o.read_ranges()))
    np.testing.assert_array_equal(_bits(g.read_majorants()), _bits(o.read_majorants()))
    g.render(2)
    o.render(2, threads=THREADS)
    np.testing.assert_array_equal(_bits(g.read_accum()), _bits(o.read_accum()))
    g.close()
