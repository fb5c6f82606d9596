"""GPU checks of batched multi-view rendering, the workgroup shapes, the
Hilbert tiles, the macrocell rebuild and the autotuner. Batches render with
one single-view launch per view (launch form 0), with one launch whose grid
holds every view's blocks (form 1), or with one persistent launch per frame
replayed from a captured HIP graph (form 2).

The RNG is keyed by (seed, pixel, frame, sample), so a batched view must give
the same bits as a single-view render of the same camera, size, spp and seed,
whatever the batch, the item order, the workgroup shape or the graph replay;
work counters add up across views. Single-view renders are pinned against
the oracle elsewhere, so equality here carries that pin to the batch."""
import numpy as np
import pytest

from mdwtracker_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from mdwtracker_amd import capi
    return capi


@pytest.fixture(scope="module")
def scene():
    return scenes.make_scene("ndda_cpu_2ch64", seed=5)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _views(n, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        th, ph = rng.uniform(0, 2 * np.pi), rng.uniform(0.3, 2.8)
        r = 2.2
        frm = (0.5 + r * np.sin(ph) * np.cos(th), 0.5 + r * np.cos(ph), 0.5 + r * np.sin(ph) * np.sin(th))
        out.append({"from": frm, "at": (0.5, 0.5, 0.5), "up": (0.0, 1.0, 0.0), "fovy": 40.0,
                    "width": int(rng.integers(17, 97)), "height": int(rng.integers(13, 81)),
                    "spp": int(rng.integers(1, 4)), "seed": int(rng.integers(0, 2**31))})
    return out


def _single(capi, scene, mode, v, frames):
    g = capi.MdwtRenderer(0)
    g.set_moments(True)
    scenes.load_scene(g, scene, mode)
    g.set_frame(v["width"], v["height"], v["spp"], True, (0.0, 0.0, 0.0))
    g.set_camera(v["from"], v["at"], v["up"], v["fovy"])
    g.set_seed(v["seed"])
    cnt = g.count_frame()
    g.render(frames)
    a = g.read_accum(with_m2=True)
    g.close()
    return a, cnt


@pytest.mark.parametrize("launch", [0, 1, 2])
@pytest.mark.parametrize("mode", [2, 0, 1, 3, 4, 7])
def test_batched_views_equal_single_view_renders(capi, scene, mode, launch):
    views = _views(6, seed=mode)
    print(f"[batch] launch={launch} mode={mode} "
          f"views={[(v['width'], v['height'], v['spp']) for v in views]}", flush=True)
    g = capi.MdwtRenderer(0)
    g.set_moments(True)
    scenes.load_scene(g, scene, mode)
    g.batch_set_launch(launch)
    g.batch_set_views(views)
    cb = g.batch_count()
    # frames split over several calls: ids and accumulation carry across them
    g.batch_render(1)
    g.batch_render(3)
    g.batch_render(2)
    tot = {"samples": 0, "visits": 0, "rejections": 0, "rays": 0}
    for i, v in enumerate(views):
        (a_s, m_s), cnt = _single(capi, scene, mode, v, 6)
        a_b, m_b = g.batch_read_view(i, with_m2=True)
        np.testing.assert_array_equal(_bits(a_b), _bits(a_s))
        np.testing.assert_array_equal(_bits(m_b), _bits(m_s))
        for k in tot:
            tot[k] += cnt[k]
    assert cb == tot
    t = g.batch_view_times()
    assert np.all(t > 0) and np.all(np.isfinite(t))
    # restart the accumulation: the next frame overwrites, frame ids continue
    g.batch_reset()
    g.batch_render(1)
    a0 = g.batch_read_view(0)
    assert np.all(np.isfinite(a0))
    g.close()


@pytest.mark.parametrize("launch", [0, 1, 2])
def test_batch_with_lighting_equals_single(capi, scene, launch):
    views = _views(3, seed=11)
    print(f"[batch] lighting launch={launch}", flush=True)
    g = capi.MdwtRenderer(0)
    scenes.load_scene(g, scene, 2)
    g.set_light((0.3, -1.0, 0.2), True, True)
    g.batch_set_launch(launch)
    g.batch_set_views(views)
    g.batch_render(2)
    for i, v in enumerate(views):
        s = capi.MdwtRenderer(0)
        scenes.load_scene(s, scene, 2)
        s.set_light((0.3, -1.0, 0.2), True, True)
        s.set_frame(v["width"], v["height"], v["spp"], True, (0.0, 0.0, 0.0))
        s.set_camera(v["from"], v["at"], v["up"], v["fovy"])
        s.set_seed(v["seed"])
        s.render(2)
        np.testing.assert_array_equal(_bits(g.batch_read_view(i)), _bits(s.read_accum()))
        s.close()
    g.close()


@pytest.mark.parametrize("threads", [64, 128, 256])
@pytest.mark.parametrize("tile", [8, 32])
def test_block_shapes_and_tiles_do_not_change_images(capi, scene, threads, tile):
    ref = capi.MdwtRenderer(0)
    scenes.load_scene(ref, scene, 2)
    ref.set_frame(150, 110, 2, True, (0.1, 0.1, 0.1))
    ref.render(3)
    want = ref.read_accum()
    ref.close()
    g = capi.MdwtRenderer(0)
    g.set_block(threads)
    scenes.load_scene(g, scene, 2)
    g.set_tile_order(2, tile)
    g.set_frame(150, 110, 2, True, (0.1, 0.1, 0.1))
    g.render(3)
    np.testing.assert_array_equal(_bits(g.read_accum()), _bits(want))
    g.close()


@pytest.mark.parametrize("threads", [64, 128])
def test_one_launch_batch_block_shapes(capi, scene, threads):
    """The one-launch and persistent batches with 8x8 and 16x8 blocks (items
    rebuilt for the block shape; the persistent grid's items re-dealt by cost
    after each call) equal the default 16x16 per-view launches."""
    views = _views(5, seed=21)
    print(f"[batch] one launch and persistent, {threads} threads", flush=True)
    out = []
    for launch, th in ((0, 256), (1, threads), (2, threads)):
        g = capi.MdwtRenderer(0)
        g.set_block(th)
        scenes.load_scene(g, scene, 2)
        g.batch_set_launch(launch)
        g.batch_set_views(views)
        g.batch_render(1)
        g.batch_render(1)
        g.batch_render(1)
        out.append([g.batch_read_view(i) for i in range(len(views))])
        t = g.batch_view_times()
        assert t.shape == (len(views),) and np.all(t > 0)
        g.close()
    for a, b, c in zip(*out):
        np.testing.assert_array_equal(_bits(a), _bits(b))
        np.testing.assert_array_equal(_bits(a), _bits(c))


def test_set_macrocells_matches_fresh_init(capi, scene):
    a = capi.MdwtRenderer(0)
    scenes.load_scene(a, scene, 2)
    a.set_macrocells((4, 4, 4))
    b = capi.MdwtRenderer(0)
    for v in scene.volumes:
        b.push_channel(v)
    b.init(2, (4, 4, 4), True)
    for c, t in enumerate(scene.tfs):
        b.set_tf(c, t, scene.opacity, (0.0, 1.0))
    if scene.camera is not None:
        b.set_camera(*scene.camera)
    b.set_seed(scene.seed)
    np.testing.assert_array_equal(_bits(a.read_ranges()), _bits(b.read_ranges()))
    np.testing.assert_array_equal(_bits(a.read_majorants()), _bits(b.read_majorants()))
    for r in (a, b):
        r.set_frame(64, 64, 1, True, (0.0, 0.0, 0.0))
        r.render(2)
    np.testing.assert_array_equal(_bits(a.read_accum()), _bits(b.read_accum()))
    a.close()
    b.close()


def test_autotune_batch(capi, scene):
    g = capi.MdwtRenderer(0)
    scenes.load_scene(g, scene, 2)
    g.batch_set_views(_views(4, seed=7))
    table, best = g.autotune((8, 16), (64, 256), (8, 16), frames=2)
    print(f"[autotune] best={best}\n{table}", flush=True)
    assert table.shape == (8, 6)
    assert np.all(table[:, 3] > 0) and np.all(table[:, 4] > 0)
    assert best[0] in (8, 16) and best[1] in (64, 256) and best[2] in (8, 16)
    row = table[(table[:, 0] == best[0]) & (table[:, 1] == best[1]) & (table[:, 2] == best[2])]
    assert row[0, 3] == table[:, 3].min()
    assert g.info()["mc"] == (64 // best[0],) * 3
    g.batch_render(2)
    assert np.all(np.isfinite(g.batch_read_view(0)))
    g.close()


def test_batch_api_errors(capi, scene):
    g = capi.MdwtRenderer(0)
    scenes.load_scene(g, scene, 2)
    with pytest.raises(capi.MdwtError):
        g.batch_render(1)  # no views yet
    bad = _views(1)
    bad[0]["width"] = 0
    with pytest.raises(capi.MdwtError):
        g.batch_set_views(bad)
    g.batch_set_views(_views(2))
    with pytest.raises(capi.MdwtError, match="no batched frame"):
        g.batch_view_times()
    g.batch_render(1)
    assert g.batch_view_times().shape == (2,)
    assert capi.lib.mdwt_batch_read_view(g._h, 2, None, None) != 0  # no such view
    with pytest.raises(capi.MdwtError, match="batch launch"):
        g.batch_set_launch(3)
    g.close()


def test_autotune_single_view(capi, scene):
    g = capi.MdwtRenderer(0)
    scenes.load_scene(g, scene, 2)
    g.set_frame(96, 96, 1, True, (0.0, 0.0, 0.0))
    print("[autotune] single view 96x96, s in (8, 16), threads (64, 256), tiles (8, 16)", flush=True)
    table, best = g.autotune((8, 16), (64, 256), (8, 16), frames=2)
    print(f"[autotune] best={best}\n{table}", flush=True)
    assert table.shape == (8, 6)
    assert np.all(table[:, 3] > 0) and np.all(table[:, 4] > 0)
    row = table[(table[:, 0] == best[0]) & (table[:, 1] == best[1]) & (table[:, 2] == best[2])]
    assert row[0, 3] == table[:, 3].min()
    assert g.info()["mc"] == (64 // best[0],) * 3
    g.render(2)
    assert np.all(np.isfinite(g.read_accum()))
    g.close()


@pytest.mark.parametrize("launch", [0, 1, 2])
def test_batch_with_jitter_2d_and_conservative_majorants(capi, scene, launch):
    """The two options (two-axis jitter, conservative majorants) reach every
    launch form: each batched view equals a single-view render of the same
    camera, size, spp and seed with the same options."""
    views = _views(4, seed=41)
    print(f"[batch] options launch={launch}", flush=True)

    def setup(r):
        r.set_jitter_2d(True)
        r.set_conservative_majorants(True)

    g = capi.MdwtRenderer(0)
    scenes.load_scene(g, scene, 2)
    setup(g)
    g.batch_set_launch(launch)
    g.batch_set_views(views)
    g.batch_render(2)
    g.batch_render(1)
    for i, v in enumerate(views):
        s = capi.MdwtRenderer(0)
        scenes.load_scene(s, scene, 2)
        setup(s)
        s.set_frame(v["width"], v["height"], v["spp"], True, (0.0, 0.0, 0.0))
        s.set_camera(v["from"], v["at"], v["up"], v["fovy"])
        s.set_seed(v["seed"])
        s.render(3)
        np.testing.assert_array_equal(_bits(g.batch_read_view(i)), _bits(s.read_accum()))
        s.close()
    g.close()


def test_persistent_graph_recaptures_on_changes(capi, scene):
    """Launch form 2 replays one captured frame; changing the frame's
    parameters (lighting, accumulation off, block shape) makes the next call
    capture again, and every frame still equals the one-launch form's."""
    views = _views(4, seed=31)
    print("[batch] persistent graph, parameter changes between calls", flush=True)
    out = []
    for launch in (1, 2):
        g = capi.MdwtRenderer(0)
        scenes.load_scene(g, scene, 2)
        g.batch_set_launch(launch)
        g.batch_set_views(views)
        g.batch_render(2)
        g.set_light((0.2, -1.0, 0.4), True, False)
        g.batch_reset()
        g.batch_render(3)
        a = [g.batch_read_view(i) for i in range(len(views))]
        g.set_block(64)
        g.batch_reset()
        g.batch_render(2)
        b = [g.batch_read_view(i) for i in range(len(views))]
        out.append((a, b))
        g.close()
    for (a1, b1), (a2, b2) in [(out[0], out[1])]:
        for x, y in zip(a1 + b1, a2 + b2):
            np.testing.assert_array_equal(_bits(x), _bits(y))


@pytest.mark.parametrize("launch", [0, 1, 2])
def test_ten_channel_batch_equals_single_view_renders(capi, launch):
    """-m 2 on ten f32 channels runs the wave-coherent kernel's twelve-channel
    instantiation in every launch form (one Philox block per four channels per
    cell record); each batched view equals its single-view render bit for
    bit, and the single-view ten-channel render is pinned against the oracle
    in test_gpu_configs.py."""
    sc = scenes.make_scene("large_10ch_f32", seed=1234, scale=8)
    views = _views(4, seed=11)
    print(f"[batch 10ch] launch={launch} views={[(v['width'], v['height'], v['spp']) for v in views]}",
          flush=True)
    g = capi.MdwtRenderer(0)
    g.set_moments(True)
    scenes.load_scene(g, sc, 2)
    g.batch_set_launch(launch)
    g.batch_set_views(views)
    g.batch_render(2)
    g.batch_render(1)
    for i, v in enumerate(views):
        (a_s, m_s), _ = _single(capi, sc, 2, v, 3)
        a_b, m_b = g.batch_read_view(i, with_m2=True)
        np.testing.assert_array_equal(_bits(a_b), _bits(a_s))
        np.testing.assert_array_equal(_bits(m_b), _bits(m_s))
        assert a_s[..., :3].max() > 0.0  # the views see the volume
    g.close()
