"""Pins the oracle's geometry: the DDA (dda.h:82-203) against an exact rational
traversal of the same grid-space rays, its documented zero-direction quirk,
diagonal tie stepping, the root-box clip (deviceCode.cu:1676-1726) and the
camera rays (deviceCode.cu:293-300, 61-76) against the Philox stream."""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import pyoracle as po

f32 = np.float32


def _grid_renderer(mc, remap=(1.0, 1.0, 1.0)):
    r = po.OracleRenderer()
    r.push_channel(np.zeros((8, 8, 8), np.uint8))
    r.set_bounds(remap)
    r.init(2, mc, True)
    t = np.zeros((16, 4), np.float32)
    t[:, 3] = 1.0
    r.set_tf(0, t)
    return r


def _grid_space(org, d, mc, remap):
    """The closed-hit programs' grid transform in IEEE single precision
    (deviceCode.cu:878-888), evaluated with numpy float32 scalars."""
    w2u = [f32(1.0) / f32(remap[i]) for i in range(3)]
    og = [f32(mc[i]) * w2u[i] * f32(org[i]) for i in range(3)]
    dv = [w2u[i] * f32(mc[i]) * f32(d[i]) for i in range(3)]
    dot = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]
    inv = f32(1.0) / np.sqrt(dot)
    return og, [dv[i] * inv for i in range(3)]


def _exact_cells(og, dg, mc):
    """Exact traversal of the ray og + t dg (t >= 0) through [0, M]^3 in
    rationals: list of (cell id, t0, t1) and the smallest gap between distinct
    boundary crossings (to skip near-ties where fp32 may legitimately flip)."""
    o = [Fr(float(v)) for v in og]
    d = [Fr(float(v)) for v in dg]
    M = [Fr(m) for m in mc]
    tn, tf = Fr(0), None
    for i in range(3):
        if d[i] == 0:
            if not (0 <= o[i] <= M[i]):
                return [], 1.0
            continue
        a, b = (0 - o[i]) / d[i], (M[i] - o[i]) / d[i]
        lo, hi = min(a, b), max(a, b)
        tn = max(tn, lo)
        tf = hi if tf is None else min(tf, hi)
    if tf is None or tn > tf:
        return [], 1.0
    events = []
    for i in range(3):
        if d[i] == 0:
            continue
        for k in range(int(mc[i]) + 1):
            t = (k - o[i]) / d[i]
            if tn < t < tf:
                events.append((t, i))
    events.sort()
    times = sorted(set(t for t, _ in events) | {tn, tf})
    gap = min((float(b - a) for a, b in zip(times, times[1:])), default=1.0)
    p = [o[i] + tn * d[i] for i in range(3)]
    cell = [min(max(int(np.floor(float(p[i]))), 0), mc[i] - 1) for i in range(3)]
    # exact floor for rationals
    for i in range(3):
        fl = p[i].numerator // p[i].denominator
        cell[i] = min(max(fl, 0), mc[i] - 1)
    out, t_prev, j = [], tn, 0
    while True:
        # next distinct event time
        t_next = events[j][0] if j < len(events) else tf
        out.append((cell[0] + cell[1] * mc[0] + cell[2] * mc[0] * mc[1], float(t_prev), float(t_next)))
        if j >= len(events):
            break
        left = False
        while j < len(events) and events[j][0] == t_next:
            ax = events[j][1]
            cell[ax] += 1 if d[ax] > 0 else -1
            if cell[ax] < 0 or cell[ax] >= mc[ax]:
                left = True
            j += 1
        if left:
            break
        t_prev = t_next
    return out, gap


@pytest.mark.parametrize("mc,remap", [((8, 8, 8), (1, 1, 1)), ((5, 3, 7), (1.0, 0.6, 1.7))])
def test_dda_matches_exact_traversal(mc, remap):
    r = _grid_renderer(mc, remap)
    rng = np.random.default_rng(7)
    n = 300
    centre = 0.5 * np.asarray(remap)
    org = centre + rng.normal(size=(n, 3)) * 1.5
    org[: n // 5] = rng.uniform(0, 1, (n // 5, 3)) * np.asarray(remap)  # starts inside
    tgt = rng.uniform(0, 1, (n, 3)) * np.asarray(remap)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    org, d = org.astype(np.float32), d.astype(np.float32)
    cells, t01, counts = r.trace_cells(org, d, 64)
    checked = 0
    for k in range(n):
        og, dg = _grid_space(org[k], d[k], mc, remap)
        ex, gap = _exact_cells(og, dg, mc)
        if gap < 1e-4:
            continue
        checked += 1
        assert counts[k] == len(ex), k
        assert cells[k, : counts[k]].tolist() == [c for c, _, _ in ex], k
        got_t = t01[k, : counts[k]].astype(np.float64)
        ex_t = np.array([[a, b] for _, a, b in ex])
        np.testing.assert_allclose(got_t[:, 0], ex_t[:, 0], rtol=1e-5, atol=1e-5)
        # the last cell's t1 is min(t_closest, 1e20): the next crossing, even
        # past the grid exit; only interior t1 values are exit times
        np.testing.assert_allclose(got_t[:-1, 1], ex_t[:-1, 1], rtol=1e-5, atol=1e-5)
    assert checked > 0.9 * n


def test_dda_zero_direction_rays_are_dropped():
    # dda.h:96-116 sets t_nr twice and never t_fr for a zero direction
    # component, so such rays produce no cells even when they cross the grid
    r = _grid_renderer((4, 4, 4))
    org = np.array([[-1.0, 0.3, 0.6], [0.3, -1.0, 0.6], [0.3, 0.6, -1.0]], np.float32)
    d = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    _, _, counts = r.trace_cells(org, d, 16)
    assert counts.tolist() == [0, 0, 0]
    # with tiny non-zero components the same rays walk a full row
    d2 = np.array([[1, 1e-4, 1e-4], [1e-4, 1, 1e-4], [1e-4, 1e-4, 1]], np.float32)
    cells, _, counts = r.trace_cells(org, d2, 16)
    assert counts.tolist() == [4, 4, 4]
    assert cells[0, :4].tolist() == [0 + 1 * 4 + 2 * 16 + i for i in range(4)]


def test_dda_diagonal_ties_step_all_axes():
    r = _grid_renderer((4, 4, 4))
    org = np.array([[-0.5, -0.5, -0.5]], np.float32)
    d = np.full((1, 3), 1.0 / np.sqrt(3.0), np.float32)
    cells, t01, counts = r.trace_cells(org, d, 16)
    assert counts[0] == 4
    assert cells[0, :4].tolist() == [0, 21, 42, 63]


def test_root_clip_closed_forms():
    r = _grid_renderer((2, 2, 2))
    org = np.array([[0.5, 0.5, 3.0], [0.5, 0.5, 0.5], [0.5, 0.5, 3.0], [3.0, 3.0, 3.0]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, -1], [0, 0, 1], [0, 0, -1]], np.float32)
    t01, hit = r.clip_rays(org, d)
    assert hit.tolist() == [1, 1, 0, 0]
    assert t01[0].tolist() == [2.0, 3.0]
    assert t01[1].tolist() == [0.0, 0.5]


def test_camera_rays_follow_philox_jitter():
    r = _grid_renderer((2, 2, 2))
    r.set_frame(16, 8)
    r.set_seed(99)
    org, d = r.gen_rays(frame=3, sample=1)
    cam = r.get_camera().astype(np.float64)
    for i in [0, 5, 17, 64, 127]:
        px, py = i % 16, i // 16
        w = po.philox4x32([3, 1, 0, 0], [99, i])[0]
        jit = (int(w) >> 8) * 2.0 ** -24
        p = cam[1] + (px + jit) / 16 * cam[2] + (py + jit) / 8 * cam[3]
        e = (p - cam[0]) / np.linalg.norm(p - cam[0])
        np.testing.assert_allclose(d[i], e, atol=2e-6)
        np.testing.assert_allclose(org[i], cam[0], atol=0)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)


def test_camera_rays_two_axis_jitter():
    """With two-axis jitter, x keeps word 0 of the jitter block and y takes
    word 1; switching it off again restores the one-scalar rays."""
    r = _grid_renderer((2, 2, 2))
    r.set_frame(16, 8)
    r.set_seed(99)
    org1, d1 = r.gen_rays(frame=3, sample=1)
    r.set_jitter_2d(True)
    org, d = r.gen_rays(frame=3, sample=1)
    cam = r.get_camera().astype(np.float64)
    for i in [0, 5, 17, 64, 127]:
        px, py = i % 16, i // 16
        w = po.philox4x32([3, 1, 0, 0], [99, i])
        jx = (int(w[0]) >> 8) * 2.0 ** -24
        jy = (int(w[1]) >> 8) * 2.0 ** -24
        p = cam[1] + (px + jx) / 16 * cam[2] + (py + jy) / 8 * cam[3]
        e = (p - cam[0]) / np.linalg.norm(p - cam[0])
        np.testing.assert_allclose(d[i], e, atol=2e-6)
        np.testing.assert_allclose(org[i], cam[0], atol=0)
    assert (np.abs(d - d1).max(axis=1) > 0).mean() > 0.9
    r.set_jitter_2d(False)
    _, d0 = r.gen_rays(frame=3, sample=1)
    np.testing.assert_array_equal(d0, d1)
