"""GPU parity: libmdwt.so on the MI355X against the CPU oracle on the same
seeded inputs. Range grids, majorant grids, camera rays, root clips and DDA
cell sequences must match bit for bit; with the shared counter-based RNG and
portable log, whole accumulated images and the work counters must too."""
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


def _pair(capi, po, scene, mode):
    g = capi.MdwtRenderer(0)
    g.set_moments(True)  # second moments compared too
    o = po.OracleRenderer()
    scenes.load_scene(g, scene, mode)
    scenes.load_scene(o, scene, mode)
    return g, o


SMALL = {}


def _scene(name, scale=1):
    key = (name, scale)
    if key not in SMALL:
        SMALL[key] = scenes.make_scene(name, seed=1234, scale=scale)
    return SMALL[key]


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("name,scale", [("ndda_cpu_2ch64", 1), ("sync_4ch256", 1),
                                        ("sync_4ch256", 4), ("large_10ch_f32", 8)])
def test_ranges_and_majorants_bit_exact(capi, po, name, scale):
    sc = _scene(name, scale)
    print(f"[ranges] {name} scale={scale} dims={sc.volumes[0].shape} mc={sc.mc}", flush=True)
    for mode in (2, 3):
        g, o = _pair(capi, po, sc, mode)
        np.testing.assert_array_equal(_bits(g.read_ranges()), _bits(o.read_ranges()))
        np.testing.assert_array_equal(_bits(g.read_majorants()), _bits(o.read_majorants()))
        for m in (0, 1):
            g.set_mode(m)
            o2 = po.OracleRenderer()
            scenes.load_scene(o2, sc, m)
            np.testing.assert_array_equal(_bits(g.read_majorants()), _bits(o2.read_majorants()))
        g.close()


# (z, y, x) shapes: rows of 33 voxels go through the per-cell range kernel
# (not a multiple of 16 bytes), rows of 40 through the pencil kernel with one
# 16-byte chunk per lane, rows of 300 (f32: 75 chunks) with two.
@pytest.mark.parametrize("shape", [(19, 26, 33), (21, 18, 40), (6, 5, 300)])
@pytest.mark.parametrize("dtype", [np.uint16, np.float16, np.float32])
def test_ranges_other_dtypes_and_bounds(capi, po, dtype, shape):
    rng = np.random.default_rng(17)
    print(f"[ranges] dtype={np.dtype(dtype).name} shape={shape}", flush=True)
    if dtype == np.uint16:
        vols = [rng.integers(0, 65535, shape, dtype=np.uint16) for _ in range(2)]
    else:
        vols = [(rng.standard_normal(shape) * 3).astype(dtype) for _ in range(2)]
    tfs = [scenes.piecewise_tf(6, rng, c / 2.0, n_texels=97) for c in range(2)]
    for r in (capi.MdwtRenderer(0), po.OracleRenderer()):
        for v in vols:
            r.push_channel(v)
        r.set_bounds([1.3, 0.8, 2.2])
        r.init(2, (7, 5, 3), True)
        for c in range(2):
            r.set_tf(c, tfs[c], 1.7, (-0.1, 0.9))
        if isinstance(r, po.OracleRenderer):
            o = r
        else:
            g = r
    np.testing.assert_array_equal(_bits(g.read_ranges()), _bits(o.read_ranges()))
    np.testing.assert_array_equal(_bits(g.read_majorants()), _bits(o.read_majorants()))
    np.testing.assert_array_equal(g.get_domains(), o.get_domains())
    assert g.get_dt() == o.get_dt()
    g.set_frame(96, 64, 2, True, (0.1, 0.2, 0.3))
    o.set_frame(96, 64, 2, True, (0.1, 0.2, 0.3))
    g.set_camera([2.0, -1.0, 3.0], [0.6, 0.4, 1.1], [0, 0, 1], 38.0)
    o.set_camera([2.0, -1.0, 3.0], [0.6, 0.4, 1.1], [0, 0, 1], 38.0)
    g.render(3)
    o.render(3, threads=THREADS)
    np.testing.assert_array_equal(_bits(g.read_accum()), _bits(o.read_accum()))


@pytest.mark.parametrize("shape,mc", [((3, 3, 16), (1, 2, 3)), ((5, 7, 48), (6, 3, 2)),
                                      ((4, 6, 1040), (37, 2, 4)), ((9, 8, 2048), (64, 3, 5)),
                                      ((7, 5, 2064), (50, 2, 3))])
def test_ranges_u8_row_widths(capi, po, shape, mc):
    """u8 rows of 1, 3, 65 and 128 16-byte chunks (the pencil kernel's lane
    layouts, one and two chunks per lane) and of 129 chunks (beyond the
    pencil kernel: per-cell kernel), random voxels, uneven macrocells."""
    rng = np.random.default_rng(sum(shape))
    print(f"[ranges] u8 shape={shape} mc={mc}", flush=True)
    vols = [rng.integers(0, 256, shape, dtype=np.uint8) for _ in range(3)]
    tfs = [scenes.piecewise_tf(5, rng, c / 3.0, n_texels=50) for c in range(3)]
    rs = (capi.MdwtRenderer(0), po.OracleRenderer())
    for r in rs:
        for v in vols:
            r.push_channel(v)
        r.init(2, mc, True)
        for c in range(3):
            r.set_tf(c, tfs[c], 2.0, (0.0, 1.0))
    g, o = rs
    np.testing.assert_array_equal(_bits(g.read_ranges()), _bits(o.read_ranges()))
    np.testing.assert_array_equal(_bits(g.read_majorants()), _bits(o.read_majorants()))
    g.close()


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float16, np.float32])
def test_bricked_sampling_bit_exact(capi, po, dtype):
    """The samplers read the volume's 4 x 4 x 4 brick copy: on a volume whose
    sides are no multiple of the brick side (partial bricks on every face),
    every voxel type and mode renders the oracle's image bit for bit."""
    rng = np.random.default_rng(23)
    shape = (13, 18, 27)  # (z, y, x)
    if dtype == np.uint8:
        vols = [scenes.quantize_u8(scenes.gaussian_blobs(shape, 3, rng)) for _ in range(2)]
    elif dtype == np.uint16:
        vols = [rng.integers(0, 65535, shape, dtype=np.uint16) for _ in range(2)]
    else:
        vols = [(rng.standard_normal(shape) * 3).astype(dtype) for _ in range(2)]
    tfs = [scenes.piecewise_tf(6, rng, c / 2.0, n_texels=64) for c in range(2)]
    for mode in (0, 2, 3):
        print(f"[bricks] dtype={np.dtype(dtype).name} mode={mode}", flush=True)
        g, o = capi.MdwtRenderer(0), po.OracleRenderer()
        for r in (g, o):
            for v in vols:
                r.push_channel(v)
            r.init(mode, (5, 4, 3), True)
            for c in range(2):
                r.set_tf(c, tfs[c], 3.0, (0.0, 1.0))
            r.set_frame(80, 56, 1, True, (0.1, 0.2, 0.3))
            r.set_camera([1.8, -1.1, 2.6], [0.5, 0.5, 0.5], [0, 0, 1], 42.0)
        g.render(3)
        o.render(3, threads=THREADS)
        np.testing.assert_array_equal(_bits(g.read_accum()), _bits(o.read_accum()))
        g.close()
        o.close()


def _ragged_scene(n_ch, seed):
    """n_ch u8 blob channels on a 23 x 17 x 30 (x, y, z) grid: no side a power
    of two, macrocells (5, 4, 7) that do not divide it."""
    rng = np.random.default_rng(seed)
    shape = (30, 17, 23)
    vols = [scenes.quantize_u8(scenes.gaussian_blobs(shape, 3, rng)) for _ in range(n_ch)]
    tfs = [scenes.piecewise_tf(5, rng, c / n_ch, n_texels=64) for c in range(n_ch)]
    return vols, tfs


@pytest.mark.parametrize("n_ch", [1, 12])
@pytest.mark.parametrize("mode", list(range(10)))
def test_channel_count_extremes_ragged_bit_exact(capi, po, n_ch, mode):
    """One channel and the maximum of twelve, on a volume and a frame whose
    sizes divide nothing: images, second moments and counters stay bit-exact."""
    vols, tfs = _ragged_scene(n_ch, 40 + n_ch)
    print(f"[ragged] channels={n_ch} mode={mode}", flush=True)
    g, o = capi.MdwtRenderer(0), po.OracleRenderer()
    g.set_moments(True)
    for r in (g, o):
        for v in vols:
            r.push_channel(v)
        r.init(mode, (5, 4, 7), True)
        for c, t in enumerate(tfs):
            r.set_tf(c, t, 4.0, (0.0, 1.0))
        r.set_frame(77, 45, 2, True, (0.1, 0.0, 0.2))
        r.set_camera([1.9, 1.2, -0.8], [0.5, 0.45, 0.5], [0, 1, 0], 40.0)
        r.set_seed(5)
    np.testing.assert_array_equal(_bits(g.read_majorants()), _bits(o.read_majorants()))
    cg = g.count_frame()
    g.render(2)
    o.render(2, threads=THREADS)
    a_g, m_g = g.read_accum(with_m2=True)
    a_o, m_o = o.read_accum(with_m2=True)
    np.testing.assert_array_equal(_bits(a_g), _bits(a_o))
    np.testing.assert_array_equal(_bits(m_g), _bits(m_o))
    o1 = po.OracleRenderer()
    for v in vols:
        o1.push_channel(v)
    o1.init(mode, (5, 4, 7), True)
    for c, t in enumerate(tfs):
        o1.set_tf(c, t, 4.0, (0.0, 1.0))
    o1.set_frame(77, 45, 2, True, (0.1, 0.0, 0.2))
    o1.set_camera([1.9, 1.2, -0.8], [0.5, 0.45, 0.5], [0, 1, 0], 40.0)
    o1.set_seed(5)
    o1.render(1, threads=THREADS)
    assert cg == o1.read_counters() and cg["samples"] > 0
    # a one-pixel frame of the same scene
    for r in (g, o):
        r.set_frame(1, 1, 3, True, (0.0, 0.0, 0.0))
    g.render(2)
    o.render(2, threads=THREADS)
    np.testing.assert_array_equal(_bits(g.read_accum()), _bits(o.read_accum()))
    g.close()


def test_transparent_volume_renders_background_exactly(capi):
    """Zero voxels under a transfer function that is transparent at 0: every
    majorant is 0, no sample is taken, and every pixel is the background."""
    g = capi.MdwtRenderer(0)
    for _ in range(3):
        g.push_channel(np.zeros((9, 10, 11), np.uint8))
    g.init(2, (3, 3, 3), True)
    tf = np.zeros((16, 4), np.float32)
    tf[:, :3] = 1.0
    tf[8:, 3] = 1.0  # opaque only above the middle of the range
    for c in range(3):
        g.set_tf(c, tf)
    bg = (0.25, 0.5, 0.75)
    g.set_frame(33, 21, 1, True, bg)
    assert np.all(g.read_majorants() == 0)
    cnt = g.count_frame()
    g.render(3)
    a = g.read_accum()
    assert cnt["samples"] == 0
    np.testing.assert_array_equal(a[..., :3], np.broadcast_to(np.float32(bg), a[..., :3].shape))
    g.close()


def test_camera_rays_and_clip_bit_exact(capi, po):
    sc = _scene("ndda_cpu_2ch64")
    g, o = _pair(capi, po, sc, 2)
    for (w, h) in [(256, 256), (200, 120)]:
        g.set_frame(w, h)
        o.set_frame(w, h)
        g.set_camera([-1.2, 0.8, 2.5], [0.5, 0.45, 0.55], [0, 1, 0], 50.0)
        o.set_camera([-1.2, 0.8, 2.5], [0.5, 0.45, 0.55], [0, 1, 0], 50.0)
        np.testing.assert_array_equal(_bits(g.get_camera()), _bits(o.get_camera()))
        go, gd = g.gen_rays(5, 2)
        oo, od = o.gen_rays(5, 2)
        np.testing.assert_array_equal(_bits(go), _bits(oo))
        np.testing.assert_array_equal(_bits(gd), _bits(od))
        gt, gh = g.clip_rays(go, gd)
        ot, oh = o.clip_rays(oo, od)
        np.testing.assert_array_equal(gh, oh)
        np.testing.assert_array_equal(_bits(gt), _bits(ot))
        assert 0 < gh.sum() < gh.size


@pytest.mark.parametrize("mode", [2, 0, 3])
def test_two_axis_jitter_bit_exact(capi, po, mode):
    """mdwt_set_jitter_2d: camera rays and accumulated images equal the
    oracle's with the same option, and differ from the one-scalar default."""
    sc = _scene("ndda_cpu_2ch64")
    print(f"[jitter2d] ndda_cpu_2ch64 mode={mode}", flush=True)
    g1 = capi.MdwtRenderer(0)
    scenes.load_scene(g1, sc, mode)
    g1.render(4)
    a_one = g1.read_accum()
    g1.close()
    g, o = _pair(capi, po, sc, mode)
    g.set_jitter_2d(True)
    o.set_jitter_2d(True)
    go, gd = g.gen_rays(5, 2)
    oo, od = o.gen_rays(5, 2)
    np.testing.assert_array_equal(_bits(go), _bits(oo))
    np.testing.assert_array_equal(_bits(gd), _bits(od))
    g.render(4)
    o.render(4, threads=THREADS)
    a_g, m_g = g.read_accum(with_m2=True)
    a_o, m_o = o.read_accum(with_m2=True)
    np.testing.assert_array_equal(_bits(a_g), _bits(a_o))
    np.testing.assert_array_equal(_bits(m_g), _bits(m_o))
    assert not np.array_equal(_bits(a_g), _bits(a_one))
    g.close()


@pytest.mark.parametrize("mode", [2, 3, 0])
def test_conservative_majorants_bit_exact(capi, po, mode):
    """mdwt_set_conservative_majorants: the majorant grid and the accumulated
    images equal the oracle's with the same option, and the grid differs
    from the reference's (that it bounds every lookup is pinned on the
    oracle)."""
    sc = _scene("sync_4ch256", 4)
    g, o = _pair(capi, po, sc, mode)
    print(f"[conservative] sync_4ch256/4 mode={mode}", flush=True)
    m_ref = g.read_majorants().copy()
    g.set_conservative_majorants(True)
    o.set_conservative_majorants(True)
    m_g, m_o = g.read_majorants(), o.read_majorants()
    np.testing.assert_array_equal(_bits(m_g), _bits(m_o))
    assert not np.array_equal(_bits(m_g), _bits(m_ref))
    g.render(4)
    o.render(4, threads=THREADS)
    np.testing.assert_array_equal(_bits(g.read_accum()), _bits(o.read_accum()))
    g.close()


def test_dda_cells_bit_exact_64k_rays(capi, po):
    sc = _scene("sync_4ch256", 4)  # 64^3 volume, mc 4^3 -> use a finer grid below
    rng = np.random.default_rng(23)
    g = capi.MdwtRenderer(0)
    o = po.OracleRenderer()
    for r in (g, o):
        for v in sc.volumes:
            r.push_channel(v)
        r.init(2, (16, 16, 16), True)
        for c, t in enumerate(sc.tfs):
            r.set_tf(c, t)
        r.set_frame(1024, 1024)
    # all centre-lattice camera rays of a 1024^2 frame at stride 4 ...
    oo, od = o.gen_rays(0, 0)
    lat = (np.arange(0, 1024, 4)[:, None] * 1024 + np.arange(0, 1024, 4)[None, :]).ravel()
    org, d = oo[lat], od[lat]
    # ... plus random rays from outside and inside the volume
    n = 16384
    ro = np.concatenate([0.5 + rng.normal(size=(n // 2, 3)) * 1.5, rng.uniform(0, 1, (n // 2, 3))])
    rd = rng.normal(size=(n, 3))
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    org = np.concatenate([org, ro]).astype(np.float32)
    d = np.concatenate([d, rd]).astype(np.float32)
    gc, gt, gn = g.trace_cells(org, d, 64)
    oc, ot, on = o.trace_cells(org, d, 64)
    print(f"[dda] rays={len(org)} mean cells={on.mean():.2f} max={on.max()}", flush=True)
    np.testing.assert_array_equal(gn, on)
    np.testing.assert_array_equal(gc, oc)
    np.testing.assert_array_equal(_bits(gt), _bits(ot))
    assert on.max() <= 48 and (on > 0).mean() > 0.3 and on[on > 0].mean() > 10


@pytest.mark.parametrize("mode", list(range(10)))
def test_images_bit_exact_small(capi, po, mode):
    sc = _scene("ndda_cpu_2ch64")
    g, o = _pair(capi, po, sc, mode)
    print(f"[image] ndda_cpu_2ch64 mode={mode}", flush=True)
    cg = g.count_frame()
    g.render(16)
    o.render(16, threads=THREADS)
    a_g, m_g = g.read_accum(with_m2=True)
    a_o, m_o = o.read_accum(with_m2=True)
    np.testing.assert_array_equal(_bits(a_g), _bits(a_o))
    np.testing.assert_array_equal(_bits(m_g), _bits(m_o))
    np.testing.assert_array_equal(g.read_fb(), o.read_fb())
    # counting pass of frame 0 equals the oracle's counters of frame 0
    o1 = po.OracleRenderer()
    scenes.load_scene(o1, sc, mode)
    o1.render(1, threads=THREADS)
    assert cg == o1.read_counters()
    assert g.info()["frame_id"] == 16 and o.info()["frame_id"] == 16


@pytest.mark.parametrize("mode", list(range(10)))
def test_images_bit_exact_4ch_quarter(capi, po, mode):
    sc = _scene("sync_4ch256", 4)
    g, o = _pair(capi, po, sc, mode)
    g.render(4)
    o.render(4, threads=THREADS)
    np.testing.assert_array_equal(_bits(g.read_accum()), _bits(o.read_accum()))


@pytest.mark.parametrize("mode", [2, 0])
def test_image_bit_exact_full_4ch256_one_frame(capi, po, mode):
    """The benchmark frame itself (default, wave-coherent schedule)."""
    sc = _scene("sync_4ch256")
    g, o = _pair(capi, po, sc, mode)
    print(f"[image] sync_4ch256 full, mode {mode}, 1 frame", flush=True)
    cg = g.count_frame()
    g.render(1)
    o.render(1, threads=THREADS)
    np.testing.assert_array_equal(_bits(g.read_accum()), _bits(o.read_accum()))
    assert cg == o.read_counters()


def test_tile_order_and_spp_and_no_accumulation(capi, po):
    sc = _scene("ndda_cpu_2ch64")
    g, o = _pair(capi, po, sc, 2)
    for r in (g, o):
        r.set_frame(130, 70, 3, False, (0.2, 0.3, 0.4))
    g.set_tile_order(0)
    g.render(2)
    o.render(2, threads=THREADS)
    np.testing.assert_array_equal(_bits(g.read_accum()), _bits(o.read_accum()))
    g.set_tile_order(1)
    g.reset_accumulation()
    g.set_frame(130, 70, 3, True, (0.2, 0.3, 0.4))
    o.set_frame(130, 70, 3, True, (0.2, 0.3, 0.4))
    g.render(3)
    o.render(3, threads=THREADS)
    np.testing.assert_array_equal(_bits(g.read_accum()), _bits(o.read_accum()))


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5])
def test_schedules_agree_bit_exact(capi, mode):
    """The wave-coherent schedule (the default for -m 0 to -m 5) and the
    thread-per-pixel schedule give
    identical images, second moments and counters, with two samples per
    pixel on a frame whose sides are not multiples of the 16-pixel block."""
    sc = _scene("sync_4ch256", 2)
    out = []
    for schedule in (0, 1):
        g = capi.MdwtRenderer(0)
        g.set_moments(True)
        g.set_schedule(schedule)
        scenes.load_scene(g, sc, mode)
        g.set_frame(sc.width + 13, sc.height - 7, 2, True, (0.05, 0.1, 0.15))
        cnt = g.count_frame()
        g.render(3)
        out.append((g.read_accum(with_m2=True), cnt))
        g.close()
    (a0, m0), c0 = out[0]
    for (a1, m1), c1 in out[1:]:
        np.testing.assert_array_equal(_bits(a0), _bits(a1))
        np.testing.assert_array_equal(_bits(m0), _bits(m1))
        assert c0 == c1


def test_render_requires_every_transfer_function(capi):
    g = capi.MdwtRenderer(0)
    v = np.zeros((8, 8, 8), np.uint8)
    g.push_channel(v)
    g.push_channel(v)
    g.init(2)
    g.set_tf(0, np.ones((8, 4), np.float32))
    with pytest.raises(capi.MdwtError) as e:
        g.render(1)
    assert e.value.code == 2
    with pytest.raises(capi.MdwtError):
        g.set_tf(2, np.ones((8, 4), np.float32))  # channel id out of range


@pytest.mark.parametrize("mode", list(range(10)))
def test_lighting_bit_exact(capi, po, mode):
    """Shadow rays, gradient shading and per-channel materials (-m 0-5), the
    marchers' own shadow march (-m 6-8); -m 9 ignores both."""
    sc = _scene("ndda_cpu_2ch64")
    g, o = _pair(capi, po, sc, mode)
    print(f"[lighting] ndda_cpu_2ch64 mode={mode}", flush=True)
    for r in (g, o):
        r.set_light((0.4, -1.0, 0.3), shadows=True, gradient=True)
        r.set_material(1, 0.6, 0.5, 24.0, 0.02, (0.8, 0.9, 1.0))
        r.set_lighting(1.2, 0.15)
    cg = g.count_frame()
    g.render(3)
    o.render(3, threads=THREADS)
    a_g, m_g = g.read_accum(with_m2=True)
    a_o, m_o = o.read_accum(with_m2=True)
    np.testing.assert_array_equal(_bits(a_g), _bits(a_o))
    np.testing.assert_array_equal(_bits(m_g), _bits(m_o))
    o1 = po.OracleRenderer()
    scenes.load_scene(o1, sc, mode)
    o1.set_light((0.4, -1.0, 0.3), shadows=True, gradient=True)
    o1.set_material(1, 0.6, 0.5, 24.0, 0.02, (0.8, 0.9, 1.0))
    o1.set_lighting(1.2, 0.15)
    o1.render(1, threads=THREADS)
    assert cg == o1.read_counters()
