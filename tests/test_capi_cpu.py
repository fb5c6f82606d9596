"""CPU-only checks of libmdwt.so: it loads, exports every entry point that
include/mdwt.h declares, validates arguments on the host before touching a
GPU (status codes instead of the reference's exit()/throw), and its tile
order follows the reference's Hilbert curve (hilbert.h, built from the
reference's own source into oracle/_ref when the reference is mounted)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def capi():
    from mdwtracker_amd import capi
    return capi


def test_library_exports_every_declared_symbol(capi):
    declared = set()
    for hdr in sorted((ROOT / "include").glob("*.h")):
        declared |= set(re.findall(r"^\s*(?:int|const char\*)\s+(mdwt_\w+)\s*\(", hdr.read_text(), re.M))
    assert len(declared) >= 50
    lib = ctypes.CDLL(str(capi.LIB_PATH))
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(capi.SIGNATURES), declared ^ set(capi.SIGNATURES)


def test_host_validation_without_gpu(capi):
    r = capi.MdwtRenderer(0)
    with pytest.raises(capi.MdwtError) as e:
        r.init(2)
    assert e.value.code == 2  # no data source: ERR_STATE
    v = np.zeros((4, 4, 4), np.uint8)
    with pytest.raises(ValueError):
        r.push_channel(np.zeros((4, 4, 4), np.int64))
    with pytest.raises(capi.MdwtError) as e:
        r.push_channel(np.zeros((0, 4, 4), np.uint8))
    assert e.value.code == 1
    for _ in range(12):
        r.push_channel(v)
    with pytest.raises(capi.MdwtError) as e:  # the 13th channel (renderer.cpp:1033 accepts it)
        r.push_channel(v)
    assert e.value.code == 4
    r2 = capi.MdwtRenderer(0)
    r2.push_channel(v)
    with pytest.raises(capi.MdwtError):
        r2.push_channel(np.zeros((4, 4, 5), np.uint8))  # dims differ
    with pytest.raises(capi.MdwtError):
        r2.push_channel(np.zeros((4, 4, 4), np.float32))  # dtype differs
    with pytest.raises(capi.MdwtError):
        r2.set_bounds([1.0, float("nan"), 1.0])
    with pytest.raises(capi.MdwtError) as e:
        r2.init(10)  # modes are 0-9
    assert e.value.code == 1
    with pytest.raises(capi.MdwtError):
        r2.init(2, (0, -1, 0))
    with pytest.raises(capi.MdwtError) as e:
        r2.set_tf(0, np.ones((4, 4), np.float32))  # before init
    assert e.value.code == 2
    with pytest.raises(capi.MdwtError):
        r2.set_camera([0, float("nan"), 0], [0, 0, 0], [0, 1, 0], 45.0)
    with pytest.raises(capi.MdwtError):
        r2.set_frame(0, 10)
    with pytest.raises(capi.MdwtError):
        r2.render(1)  # before init
    f = np.zeros((4, 4, 4), np.float32)
    f[1, 1, 1] = np.inf
    r3 = capi.MdwtRenderer(0)
    with pytest.raises(capi.MdwtError):
        r3.push_channel(f)
    with pytest.raises(capi.MdwtError) as e:
        capi.MdwtRenderer(-1)
    assert e.value.code == 1


def test_denoiser_and_aov_entry_points_validate_without_gpu(capi):
    import ctypes as C
    lib = capi.lib
    h = C.c_void_p()
    # sizes must be multiples of 4 up to 8192, channels 1..12
    for w, hh, ch in ((30, 32, 4), (32, 30, 4), (0, 32, 4), (8196, 32, 4), (32, 32, 0),
                      (32, 32, 13)):
        assert lib.mdwt_dn_create(C.byref(h), 0, w, hh, ch, 1) == 1, (w, hh, ch)
        assert not h.value
    assert lib.mdwt_dn_create(None, 0, 32, 32, 4, 1) == 1
    # more than 2^24 pixels: 32-bit pixel x channel indices would wrap
    for w, hh in ((8192, 4096), (8192, 8192), (4100, 4096)):
        assert lib.mdwt_dn_create(C.byref(h), 0, w, hh, 4, 1) == 4, (w, hh)
        assert not h.value
    n = C.c_int()
    buf = np.zeros(16, np.float32)
    fp = buf.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.mdwt_dn_destroy(None) == 1
    assert lib.mdwt_dn_num_params(None, C.byref(n)) == 1
    assert lib.mdwt_dn_get_params(None, fp) == 1
    assert lib.mdwt_dn_set_params(None, fp) == 1
    assert lib.mdwt_dn_infer(None, fp, fp, None) == 1
    assert lib.mdwt_dn_set_fused(None, 1) == 1
    assert lib.mdwt_dn_gradient(None, fp, fp, fp, None) == 1
    assert lib.mdwt_dn_train_step(None, fp, fp, None, 1, 1e-3, None) == 1
    assert lib.mdwt_dn_last_error(None) == b""
    # auxiliary buffers: a negative frame limit, a missing handle
    assert lib.mdwt_set_aov_frames(None, 4) == 1
    r = capi.MdwtRenderer(0)
    with pytest.raises(capi.MdwtError) as e:
        r.set_aov_frames(-1)
    assert e.value.code == 1
    with pytest.raises(capi.MdwtError) as e:
        r.read_aov()  # before init
    r.close()


def test_hilbert_order_bijective_and_adjacent(capi):
    for bits in range(1, 7):
        n = 1 << bits
        pts = [capi.hilbert_d2xy(bits, d) for d in range(n * n)]
        assert len(set(pts)) == n * n
        assert all(0 <= x < n and 0 <= y < n for x, y in pts)
        steps = np.abs(np.diff(np.array(pts), axis=0)).sum(axis=1)
        assert np.all(steps == 1)


def test_hilbert_order_matches_reference_hilbert_h(capi):
    from oracle import pyoracle
    ref = pyoracle.ref_hilbert()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree not mounted when building)")
    c = (ctypes.c_ulonglong * 2)()
    for bits in range(1, 8):
        for d in range(1 << (2 * bits)):
            ref.ref_hilbert_i2c(2, bits, d, c)
            assert capi.hilbert_d2xy(bits, d) == (c[0], c[1]), (bits, d)


def test_frame_limit_moments_and_stream_entry_points_without_gpu(capi):
    """Frames beyond 2^30 pixels (32-bit pixel indices in the kernels) are
    refused before any allocation; second moments are opt-in and can be
    switched before init; the bandwidth probes validate their arguments."""
    import ctypes as C
    lib = capi.lib
    r = capi.MdwtRenderer(0)
    for w, hh in ((65536, 16385), (40000, 40000)):
        with pytest.raises(capi.MdwtError) as e:
            r.set_frame(w, hh)
        assert e.value.code == 4, (w, hh)
    r.set_frame(32768, 32768)  # exactly 2^30 pixels: accepted (nothing allocated before init)
    r.set_moments(True)
    r.set_moments(False)
    assert lib.mdwt_set_moments(None, 1) == 1
    # 16384 x 16384 voxels per slice: 2^24 bricks of 4 x 4 x 4 per slice,
    # beyond the samplers' 24-bit brick products (refused before any read)
    one = np.zeros(1, np.uint8)
    assert lib.mdwt_push_channel(r._h, one.ctypes.data_as(C.c_void_p), 0, 16384, 16384, 1) == 4
    g = C.c_float()
    assert lib.mdwt_stream_read(None, 1 << 20, 1, C.byref(g)) == 1
    assert lib.mdwt_stream_read(r._h, 8, 1, C.byref(g)) == 1   # fewer than 16 bytes
    assert lib.mdwt_stream_read(r._h, 1 << 20, 0, C.byref(g)) == 1
    assert lib.mdwt_stream_read(r._h, 1 << 20, 1, None) == 1
    r.close()
