"""The allocation-free frame-buffer read (mdwt_read_fb_noalloc), through the
C ABI: it must give the same RGBA8 image as mdwt_read_fb, on two calls in a
row, the second on a smaller frame (the static staging area keeps the first
call's contents beyond the second frame's pixels), and refuse frames larger
than its staging area."""
import ctypes as C

import numpy as np
import pytest

from mdwtracker_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from mdwtracker_amd import capi
    return capi


def _noalloc(capi, g, w, h):
    out = np.zeros((h, w), np.uint32)
    rc = capi.lib.mdwt_read_fb_noalloc(g._h, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return rc, out


def test_read_fb_noalloc_matches_read_fb(capi):
    sc = scenes.make_scene("ndda_cpu_2ch64", seed=9)
    g = capi.MdwtRenderer(0)
    scenes.load_scene(g, sc, 2)
    for w, h, frames in ((200, 150, 3), (48, 40, 2)):
        print(f"[noalloc] {w}x{h}, {frames} frames", flush=True)
        g.set_frame(w, h, 1, True, (0.1, 0.2, 0.3))
        g.render(frames)
        want = g.read_fb()
        rc, got = _noalloc(capi, g, w, h)
        assert rc == 0, capi.lib.mdwt_last_error(g._h)
        np.testing.assert_array_equal(got, want)
        assert np.any(want != want.flat[0])  # not a flat image
    # a frame beyond the 2^20-pixel staging area is refused, not truncated
    g.set_frame(1100, 1000, 1, True, (0.0, 0.0, 0.0))
    g.render(1)
    rc, _ = _noalloc(capi, g, 1100, 1000)
    assert rc == 4
    assert b"staging" in capi.lib.mdwt_last_error(g._h)
    g.close()
