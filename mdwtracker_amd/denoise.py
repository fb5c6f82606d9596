"""Learned denoiser beside plain accumulation (include/mdwt_denoise.h).

`Denoiser` wraps the C ABI (a BF16 U-Net trained and run by HIP kernels on
the matrix cores). `render_pair` renders one training or evaluation pair with
the method itself: a low-sample input (colour plus the auxiliary buffers of
mdwt_set_aov) and a high-sample target from an independent seed. `train`
runs Adam over such pairs; `evaluate` reports the error of the denoised and
of the plain input against a high-sample reference. No reference
counterpart: the reference accumulates plain samples only.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import scenes
from .capi import MdwtError, MdwtRenderer, lib

_fp = C.POINTER(C.c_float)


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
        """x [h][w][9] -> (denoised [h][w][3], GPU ms of the network)."""
        x = np.ascontiguousarray(x, np.float32)
        assert x.shape == (self.height, self.width, 9)
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

    def train_step(self, xs: np.ndarray, ts: np.ndarray, lr: float) -> float:
        xs = np.ascontiguousarray(xs, np.float32)
        ts = np.ascontiguousarray(ts, np.float32)
        assert xs.shape[1:] == (self.height, self.width, 9) and ts.shape[1:] == (self.height, self.width, 3)
        loss = C.c_float()
        self._chk(lib.mdwt_dn_train_step(self._h, _f(xs), _f(ts), xs.shape[0], float(lr), C.byref(loss)))
        return loss.value


DEPTH_SCALE = 0.25  # first-hit distances of the orbit cameras are ~1..4 world units


def network_input(rgb: np.ndarray, aov: np.ndarray) -> np.ndarray:
    """[h][w][9] input from an accumulated image and its auxiliary buffers."""
    x = np.empty(rgb.shape[:2] + (9,), np.float32)
    x[..., :3] = rgb[..., :3]
    x[..., 3] = aov[..., 0] * DEPTH_SCALE
    x[..., 4:9] = aov[..., 1:6]
    return x


def render_pair(r: MdwtRenderer, in_frames: int, target_frames: int, seed: int):
    """(input [h][w][9], target [h][w][3]) of the renderer's current scene:
    in_frames frames with seed, target_frames frames with an independent seed."""
    r.set_seed(seed)
    r.set_aov(True)
    r.reset_accumulation()
    r.render(in_frames)
    x = network_input(r.read_accum(), r.read_aov())
    r.set_seed(seed ^ 0x5A5A5A5A)
    r.reset_accumulation()
    r.render(target_frames)
    t = r.read_accum()[..., :3].copy()
    return x, t


def make_training_set(n_scenes: int, views: int, size: int, in_frames=(1, 2, 4),
                      target_frames: int = 512, seed: int = 0, n: int = 64):
    """Pairs from seeded 4-channel volumes (fBm + blobs), random transfer
    functions and random orbit cameras."""
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    for s in range(n_scenes):
        sc = scenes.make_scene("sync_4ch256", seed=int(rng.integers(1, 2**31)), scale=256 // n)
        r = MdwtRenderer(0)
        scenes.load_scene(r, sc, 2)
        r.set_frame(size, size, 1, True, (0.0, 0.0, 0.0))
        for v in range(views):
            th, ph = rng.uniform(0, 2 * np.pi), rng.uniform(0.4, 2.7)
            rad = rng.uniform(1.6, 2.6)
            frm = (0.5 + rad * np.sin(ph) * np.cos(th), 0.5 + rad * np.cos(ph),
                   0.5 + rad * np.sin(ph) * np.sin(th))
            r.set_camera(frm, (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), float(rng.uniform(35, 50)))
            k = int(rng.choice(in_frames))
            x, t = render_pair(r, k, target_frames, int(rng.integers(1, 2**31)))
            xs.append(x)
            ts.append(t)
        r.close()
    return np.stack(xs), np.stack(ts)


def train(dn: Denoiser, xs: np.ndarray, ts: np.ndarray, steps: int, batch: int = 8,
          lr: float = 1e-3, seed: int = 0, log=None) -> list:
    rng = np.random.default_rng(seed)
    losses = []
    for i in range(steps):
        idx = rng.choice(len(xs), size=min(batch, len(xs)), replace=False)
        # random flips keep the pairs valid and widen the data
        x, t = xs[idx], ts[idx]
        if rng.random() < 0.5:
            x, t = x[:, :, ::-1], t[:, :, ::-1]
        if rng.random() < 0.5:
            x, t = x[:, ::-1], t[:, ::-1]
        step_lr = lr * (0.1 ** (i / max(1, steps)))
        losses.append(dn.train_step(x, t, step_lr))
        if log and (i % 50 == 0 or i == steps - 1):
            log(f"[denoise] step {i}: loss {losses[-1]:.3e} lr {step_lr:.2e}")
    return losses


def rmse(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.sqrt(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)))
