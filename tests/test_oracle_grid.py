"""Pins the oracle's value-range grid (renderer.cu:835-893 + 487-519 scatter)
and majorant grid (renderer.cu:159-203, 111-157).

* power-of-two dims and macrocell counts on the unit box: the scatter
  reduces exactly to the window closed form min/max over voxels
  [s*m - 2, s*m + s + 1] per axis (s = N/M);
* any dims/bounds: every voxel whose centre lies in a macrocell is inside
  that macrocell's range (conservative) and the range never exceeds the
  window widened by one more voxel (tight);
* majorants: brute force over the texel interval, empty cells are 0,
  mode 3 is the channel sum, opacity scale is linear; the conservative
  option bounds every lookup where the reference's falls half a texel low."""
import numpy as np
import pytest

from oracle import pyoracle as po


def _build(vols, mc, mode=2, remap=None, tfs=None, os_=1.0):
    r = po.OracleRenderer()
    for v in vols:
        r.push_channel(v)
    if remap is not None:
        r.set_bounds(remap)
    r.init(mode, mc, True)
    rng = np.random.default_rng(5)
    for c in range(len(vols)):
        t = tfs[c] if tfs else np.column_stack([np.ones((256, 3)), rng.uniform(0, 1, 256)])
        r.set_tf(c, t.astype(np.float32), os_, (0.0, 1.0))
    return r


def _window_closed_form(vol, s, M, scale):
    n = vol.shape[0]
    out = np.zeros((M, M, M, 2), np.float32)
    for cz in range(M):
        for cy in range(M):
            for cx in range(M):
                sl = [slice(max(s * m - 2, 0), min(s * m + s + 1, n - 1) + 1) for m in (cz, cy, cx)]
                w = vol[tuple(sl)].astype(np.float32) * scale
                out[cz, cy, cx] = (w.min(), w.max())
    return out.reshape(-1, 2)


@pytest.mark.parametrize("n,M", [(32, 8), (32, 4), (16, 16), (64, 2)])
def test_ranges_power_of_two_closed_form(n, M):
    rng = np.random.default_rng(n * 31 + M)
    vols = [rng.integers(0, 256, (n, n, n), dtype=np.uint8) for _ in range(2)]
    r = _build(vols, (M, M, M))
    got = r.read_ranges().reshape(M ** 3, 2, 2)
    for c, v in enumerate(vols):
        exp = _window_closed_form(v, n // M, M, 255.0)
        np.testing.assert_array_equal(got[:, c, :], exp)


def test_ranges_general_bounds_conservative_and_tight():
    rng = np.random.default_rng(3)
    dims = (23, 29, 37)  # (nz, ny, nx)
    vol = rng.standard_normal(dims).astype(np.float32)
    mc = (5, 4, 3)  # (mx, my, mz)
    remap = (1.3, 0.7, 2.1)
    r = _build([vol], mc, remap=remap)
    rg = r.read_ranges().reshape(mc[2], mc[1], mc[0], 2)
    nz, ny, nx = dims
    # voxel centres in world space and the macrocell holding them
    def cell_of(n, m, ext):
        c = (np.arange(n) + 0.5) / n * ext
        return np.minimum((c / ext * m).astype(int), m - 1)
    cx, cy, cz = cell_of(nx, mc[0], remap[0]), cell_of(ny, mc[1], remap[1]), cell_of(nz, mc[2], remap[2])
    for k in range(mc[2]):
        for j in range(mc[1]):
            for i in range(mc[0]):
                sel = vol[np.ix_(cz == k, cy == j, cx == i)]
                lo, hi = rg[k, j, i]
                if sel.size:
                    assert lo <= sel.min() and hi >= sel.max()
                # tight: within the centre window widened by two voxels
                zz = np.flatnonzero(cz == k)
                yy = np.flatnonzero(cy == j)
                xx = np.flatnonzero(cx == i)
                w = vol[max(zz.min() - 3, 0):zz.max() + 4, max(yy.min() - 3, 0):yy.max() + 4,
                        max(xx.min() - 3, 0):xx.max() + 4]
                assert lo >= w.min() and hi <= w.max()


def _tf_alpha_texels(t, os_):
    return (t[:, 3].astype(np.float32) ** np.float32(3)) * np.float32(os_)


def test_majorants_bound_the_texel_interval():
    rng = np.random.default_rng(11)
    n, M = 32, 4
    vols = [rng.integers(0, 256, (n, n, n), dtype=np.uint8) for _ in range(3)]
    tfs = [np.column_stack([np.ones((256, 3)), rng.uniform(0, 1, 256)]).astype(np.float32)
           for _ in range(3)]
    r = _build(vols, (M, M, M), tfs=tfs, os_=1.5)
    rg = r.read_ranges().reshape(-1, 3, 2)
    maj = r.read_majorants().reshape(-1, 3)
    dom = r.get_domains()
    for c in range(3):
        a = _tf_alpha_texels(tfs[c], 1.5)
        for cell in range(M ** 3):
            mn, mx = rg[cell, c]
            lo = (mn - dom[c, 0]) / (dom[c, 1] - dom[c, 0]) * 256
            hi = (mx - dom[c, 0]) / (dom[c, 1] - dom[c, 0]) * 256
            i0 = int(np.clip(np.floor(min(lo, hi)), 0, 255))
            i1 = int(np.clip(np.ceil(max(lo, hi)), 0, 255))
            m = maj[cell, c]
            assert m >= a[i0:i1 + 1].max() * (1 - 1e-6)
            assert m <= a[i0:i1 + 1].max() * (1 + 1e-6)
            # every TF lookup of a value in the cell's range is bounded, up to
            # the half-texel interpolation the reference's interval omits
            for f in np.linspace(mn, mx, 9):
                ta = r.tf(c, f)[3]
                assert ta <= max(m, a[max(i0 - 1, 0)]) * (1 + 1e-6)


def test_conservative_majorants_bound_every_lookup():
    """oracle_set_conservative_majorants: on a ramp volume (narrow value range
    per cell) with opacity spikes, the reference's majorants fall below some
    lookups of a cell's values (the half texel its interval omits); the
    conservative ones bound every lookup. They can also be lower than the
    reference's, whose interval may run a texel past what lookups read.
    Rebuilt in place when the option changes, for modes 2 and 3."""
    n, M = 64, 16
    ramp = np.broadcast_to((np.arange(n, dtype=np.uint8) * 4)[None, None, :], (n, n, n)).copy()
    tf = np.zeros((256, 4), np.float32)
    tf[:, :3] = 1.0
    tf[[23, 87, 150], 3] = 1.0  # just below the cells' range starts 24, 88, 152
    for mode in (2, 3):
        r = _build([ramp], (M, M, M), mode=mode, tfs=[tf], os_=1.0)
        rg = r.read_ranges().reshape(-1, 2)
        m_ref = r.read_majorants().copy()
        r.set_conservative_majorants(True)
        m_con = r.read_majorants()
        low = 0
        for cell in range(M ** 3):
            mn, mx = rg[cell]
            if mx < mn:
                continue
            ta = np.array([r.tf(0, f)[3] for f in np.linspace(mn, mx, 33)])
            assert np.all(ta <= m_con[cell]), (cell, ta.max(), m_con[cell])
            low += int(np.any(ta > m_ref[cell] * (1 + 1e-6)))
        assert low > 0  # the option changes something the reference gets wrong
        r.set_conservative_majorants(False)
        np.testing.assert_array_equal(r.read_majorants(), m_ref)


def test_majorants_empty_cells_sum_and_scale():
    rng = np.random.default_rng(2)
    n, M = 16, 4
    vols = [rng.integers(0, 256, (n, n, n), dtype=np.uint8) for _ in range(2)]
    tfs = [np.column_stack([np.ones((64, 3)), rng.uniform(0, 1, 64)]).astype(np.float32)
           for _ in range(2)]
    r0 = _build(vols, (M, M, M), mode=0, tfs=tfs)
    r3 = _build(vols, (M, M, M), mode=3, tfs=tfs)
    m0 = r0.read_majorants().reshape(-1, 2)
    m3 = r3.read_majorants()
    np.testing.assert_array_equal(m3, (np.float32(0) + m0[:, 0]) + m0[:, 1])
    r2 = _build(vols, (M, M, M), mode=0, tfs=tfs, os_=2.0)
    np.testing.assert_array_equal(r2.read_majorants(), 2 * r0.read_majorants())
    # a transfer function that is zero everywhere gives zero majorants
    z = [np.zeros((64, 4), np.float32), np.zeros((64, 4), np.float32)]
    rz = _build(vols, (M, M, M), mode=2, tfs=z)
    assert not rz.read_majorants().any()
