"""Learned denoiser beside plain accumulation (include/mdwt_denoise.h).

`Denoiser` wraps the C ABI (a BF16 U-Net trained and run by HIP kernels on
the matrix cores). The network input is built from what the renderer
accumulates: the image, its second moment and the auxiliary buffers of
mdwt_set_aov (`network_input`). Training pairs are rendered by the method
itself: one accumulation run per view is read at several frame counts
(inputs at every noise level) against one independent high-frame render of
the same view (the target); the view is a zoomed crop camera whose pixel
footprint matches the full-size image (`crop_camera`). Plain accumulation
stays the fallback: nothing here changes how frames are rendered.

Errors are measured against a plain high-frame render whose own noise is
removed from the estimate (`noise_corrected_mse`): for an image d made
independently of the reference r (another seed), E[(d - r)^2] = (d - g)^2 +
Var(r) per pixel, with g the converged image and Var(r) the variance of the
reference mean, known from its second moment. No reference counterpart: the
reference accumulates plain samples only.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import scenes
from .capi import MdwtError, MdwtRenderer, lib

_fp = C.POINTER(C.c_float)

INPUTS = 11          # values per pixel, kInputs of the kernels
DEPTH_SCALE = 0.25   # first-hit distances of the orbit cameras are ~1..4 world units
STDERR_SCALE = 4.0   # standard error of the pixel mean, ~0.002..0.5
FRAMES_SCALE = 1.0 / 16.0  # log2 of the frame count, 0..16

# world units per pixel at the centre of the volume for the 1024^2 orbit
# camera at distance 1.8 with a 45 degree field of view
ORBIT_FOOTPRINT = 1.8 * 2.0 * math.tan(math.radians(22.5)) / 1024.0


def _f(a: np.ndarray):
    return a.ctypes.data_as(_fp)


class Denoiser:
    """One denoiser instance for width x height images (multiples of 4)."""

    def __init__(self, width: int, height: int, seed: int = 1, device: int = 0):
        h = C.c_void_p()
        rc = lib.mdwt_dn_create(C.byref(h), device, width, height, seed & 0xFFFFFFFF)
        if rc:
            raise MdwtError(rc, f"mdwt_dn_create({width}x{height})")
        self._h = h
        self.width, self.height = width, height

    def close(self) -> None:
        if self._h:
            lib.mdwt_dn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int) -> None:
        if rc:
            raise MdwtError(rc, lib.mdwt_dn_last_error(self._h).decode())

    def num_params(self) -> int:
        n = C.c_int()
        self._chk(lib.mdwt_dn_num_params(self._h, C.byref(n)))
        return n.value

    def get_params(self) -> np.ndarray:
        p = np.empty(self.num_params(), np.float32)
        self._chk(lib.mdwt_dn_get_params(self._h, _f(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params()
        self._chk(lib.mdwt_dn_set_params(self._h, _f(p)))

    def infer(self, x: np.ndarray):
        """x [h][w][INPUTS] -> (denoised [h][w][3], GPU ms of the network)."""
        x = np.ascontiguousarray(x, np.float32)
        assert x.shape == (self.height, self.width, INPUTS)
        out = np.empty((self.height, self.width, 3), np.float32)
        ms = C.c_float()
        self._chk(lib.mdwt_dn_infer(self._h, _f(x), _f(out), C.byref(ms)))
        return out, ms.value

    def gradient(self, x: np.ndarray, target: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        g = np.empty(self.num_params(), np.float32)
        loss = C.c_float()
        self._chk(lib.mdwt_dn_gradient(self._h, _f(x), _f(t), _f(g), C.byref(loss)))
        return g, loss.value

    def train_step(self, xs: np.ndarray, ts: np.ndarray, lr: float, weights=None) -> float:
        xs = np.ascontiguousarray(xs, np.float32)
        ts = np.ascontiguousarray(ts, np.float32)
        assert xs.shape[1:] == (self.height, self.width, INPUTS)
        assert ts.shape[1:] == (self.height, self.width, 3) and len(ts) == len(xs)
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, np.float32)
            assert w.shape == (len(xs),)
        loss = C.c_float()
        self._chk(lib.mdwt_dn_train_step(self._h, _f(xs), _f(ts), _f(w) if w is not None else None,
                                         xs.shape[0], float(lr), C.byref(loss)))
        return loss.value


def network_input(accum: np.ndarray, m2: np.ndarray, aov: np.ndarray, frames: int) -> np.ndarray:
    """[h][w][INPUTS] input from an accumulation of `frames` frames: its mean
    (accum), second moment (m2) and auxiliary buffers (aov)."""
    x = np.empty(accum.shape[:2] + (INPUTS,), np.float32)
    x[..., :3] = accum[..., :3]
    x[..., 3] = aov[..., 0] * DEPTH_SCALE
    x[..., 4:9] = aov[..., 1:6]
    if frames > 1:
        mean = accum[..., :3].astype(np.float64)
        var = np.maximum(m2[..., :3].astype(np.float64) - mean * mean, 0.0) * frames / (frames - 1)
        x[..., 9] = np.sqrt(var.mean(-1) / frames) * STDERR_SCALE
    else:
        x[..., 9] = 0.0
    x[..., 10] = math.log2(frames) * FRAMES_SCALE
    return x


def read_input(r: MdwtRenderer, frames: int) -> np.ndarray:
    a, m2 = r.read_accum(with_m2=True)
    return network_input(a, m2, r.read_aov(), frames)


def accumulate_inputs(r: MdwtRenderer, levels, seed: int):
    """Yields the network inputs of one accumulation run (seed) read at each
    frame count of `levels` (ascending); the run continues between the
    reads, so the renderer must not be used while the generator is live."""
    r.set_seed(seed)
    r.set_aov(True)
    r.reset_accumulation()
    done = 0
    for k in levels:
        assert k > done
        r.render(k - done)
        done = k
        yield read_input(r, k)


def render_mean(r: MdwtRenderer, frames: int, seed: int):
    """(mean image [h][w][3], variance of that mean per pixel and channel) of
    a plain `frames`-frame accumulation with seed."""
    r.set_seed(seed)
    r.reset_accumulation()
    r.render(frames)
    a, m2 = r.read_accum(with_m2=True)
    mean = a[..., :3].astype(np.float64)
    var = np.maximum(m2[..., :3].astype(np.float64) - mean * mean, 0.0) * frames / (frames - 1)
    return a[..., :3].copy(), var / frames


def noise_corrected_mse(img: np.ndarray, ref: np.ndarray, ref_var: np.ndarray):
    """(estimated MSE of img against the converged image, its standard error)
    given an independent reference mean `ref` with per-value variance ref_var."""
    e = (img[..., :3].astype(np.float64) - ref.astype(np.float64)) ** 2 - ref_var
    return float(e.mean()), float(e.std() / math.sqrt(e.size))


def crop_camera(rng: np.random.Generator, size: int, footprint: float = ORBIT_FOOTPRINT,
                jitter: float = 1.4):
    """(from, at, up, fovy) of a size x size view whose pixel footprint at the
    aim point is footprint times a log-uniform factor in [1/jitter, jitter]:
    a crop of a larger image of the same pixel scale."""
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    at = 0.5 + rng.uniform(-0.3, 0.3, 3)
    frm = np.array([0.5, 0.5, 0.5]) + rng.uniform(1.5, 2.6) * v
    dist = float(np.linalg.norm(frm - at))
    fp = footprint * math.exp(rng.uniform(-math.log(jitter), math.log(jitter)))
    fovy = math.degrees(2.0 * math.atan(fp * size / (2.0 * dist)))
    up = (0.0, 1.0, 0.0) if abs(v[1]) < 0.95 else (1.0, 0.0, 0.0)
    return tuple(float(c) for c in frm), tuple(float(c) for c in at), up, fovy


TRAIN_LEVELS = (1, 2, 4, 16, 64, 256, 1024, 2048, 4096, 8192, 16384)


def make_training_set(n_scenes: int = 4, crops: int = 6, size: int = 256, levels=TRAIN_LEVELS,
                      target_frames: int = 32768, seed: int = 0, scale: int = 1,
                      workload: str = "sync_4ch256", log=None):
    """(inputs [N][size][size][INPUTS], targets [N][size][size][3], frame
    counts [N]) from n_scenes seeded volumes of the workload family (seeds
    drawn from `seed`, so held out from the fixed-seed workload), `crops`
    crop cameras each, every crop read at each level of `levels`."""
    rng = np.random.default_rng(seed)
    xs, ts, ks = [], [], []
    for s in range(n_scenes):
        sc = scenes.make_scene(workload, seed=int(rng.integers(1, 2**31)), scale=scale)
        r = MdwtRenderer(0)
        scenes.load_scene(r, sc, 2)
        r.set_frame(size, size, 1, True, (0.0, 0.0, 0.0))
        fp = ORBIT_FOOTPRINT * scale
        for _ in range(crops):
            r.set_camera(*crop_camera(rng, size, fp))
            s_in, s_t = (int(v) for v in rng.integers(1, 2**31, 2))
            r.set_aov(False)
            t, _ = render_mean(r, target_frames, s_t)
            for k, x in zip(levels, accumulate_inputs(r, levels, s_in)):
                xs.append(x)
                ts.append(t)
                ks.append(k)
        r.close()
        if log:
            log(f"[denoise] training scene {s + 1}/{n_scenes}: {len(xs)} pairs")
    return np.stack(xs), np.stack(ts), np.asarray(ks, np.int64)


def level_weights(ks: np.ndarray) -> np.ndarray:
    """Per-pair loss weights proportional to the frame count: the input's
    own error falls as 1/k, so each pair then weighs its relative error."""
    k = np.asarray(ks, np.float64)
    return (k / k.mean()).astype(np.float32)


def train(dn: Denoiser, xs: np.ndarray, ts: np.ndarray, steps: int, batch: int = 8,
          lr: float = 1e-3, seed: int = 0, weights=None, log=None) -> list:
    """Adam over random batches with random flips; the learning rate decays
    exponentially to lr / 10 over the run."""
    rng = np.random.default_rng(seed)
    losses = []
    for i in range(steps):
        idx = rng.choice(len(xs), size=min(batch, len(xs)), replace=False)
        # flips keep the pairs valid and widen the data
        x, t = xs[idx], ts[idx]
        if rng.random() < 0.5:
            x, t = x[:, :, ::-1], t[:, :, ::-1]
        if rng.random() < 0.5:
            x, t = x[:, ::-1], t[:, ::-1]
        step_lr = lr * (0.1 ** (i / max(1, steps)))
        w = None if weights is None else np.asarray(weights)[idx]
        losses.append(dn.train_step(x, t, step_lr, w))
        if log and (i % 200 == 0 or i == steps - 1):
            log(f"[denoise] step {i}: loss {losses[-1]:.3e} lr {step_lr:.2e}")
    return losses


def frames_to_target(levels, mses, target_mse: float):
    """Frame count at which the series mses (at ascending frame counts
    levels) first reaches target_mse, interpolated log-log between the two
    bracketing levels; None if it never does."""
    prev = None
    for k, m in zip(levels, mses):
        if m <= target_mse:
            if prev is None or prev[1] <= 0 or m <= 0:
                return float(k)
            k0, m0 = prev
            t = (math.log(target_mse) - math.log(m0)) / (math.log(m) - math.log(m0))
            return float(math.exp(math.log(k0) + t * (math.log(k) - math.log(k0))))
        prev = (k, m)
    return None


def rmse(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))
