"""Learned denoiser beside plain accumulation (include/mdwt_denoise.h).

`Denoiser` wraps the C ABI (a BF16 U-Net trained and run by HIP kernels on
the matrix cores). The network input is built from what the renderer
accumulates: the image, its second moment and the auxiliary buffers of
mdwt_set_aov (`network_input`). Training pairs are rendered by the method
itself: one accumulation run per view is read at several frame counts
(inputs at every noise level) against one independent high-frame render of
the same view (the target), at the workload's image size and framing, and
cut into tiles for training (`make_training_set`). Plain accumulation
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

def input_width(channels: int) -> int:
    """Values per pixel of the network input for a volume of `channels` channels."""
    return 7 + channels


INPUTS = input_width(4)  # the 4-channel workloads

# (input channels, output channels, level) of the eleven 3x3 convolutions
# (denoise_api.cpp kLayers; None = the network input; level l runs at
# 1 / 4^l of the pixels); the last one writes the kTaps = 49 filter logits
NETWORK = [(None, 32, 0), (32, 32, 0), (32, 64, 1), (64, 64, 1), (64, 96, 2), (96, 96, 2),
           (160, 64, 1), (64, 64, 1), (96, 32, 0), (32, 32, 0), (32, 49, 0)]
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak


def network_gflop(width: int, height: int, channels: int = 4) -> float:
    """Algorithmic GFLOP of one inference: 2 x 9 x cin x cout per output
    pixel of each convolution, real (unpadded) channel counts."""
    n = width * height
    total = 0
    for cin, cout, level in NETWORK:
        cin = input_width(channels) if cin is None else cin
        total += 2 * 9 * cin * cout * (n >> (2 * level))
    return total * 1e-9
DEPTH_SCALE = 0.25   # first-hit distances of the orbit cameras are ~1..4 world units
STDERR_SCALE = 4.0   # standard error of the pixel mean, ~0.002..0.5
FRAMES_SCALE = 1.0 / 16.0  # log2 of the frame count, 0..16

def _f(a: np.ndarray):
    return a.ctypes.data_as(_fp)


class Denoiser:
    """One denoiser instance for width x height images (multiples of 4) of a
    volume with `channels` channels."""

    def __init__(self, width: int, height: int, channels: int = 4, seed: int = 1, device: int = 0):
        h = C.c_void_p()
        rc = lib.mdwt_dn_create(C.byref(h), device, width, height, channels, seed & 0xFFFFFFFF)
        if rc:
            raise MdwtError(rc, f"mdwt_dn_create({width}x{height}, {channels} channels)")
        self._h = h
        self.width, self.height, self.channels = width, height, channels
        self.inputs = input_width(channels)

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
        """x [h][w][inputs] -> (denoised [h][w][3], GPU ms of the network)."""
        x = np.ascontiguousarray(x, np.float32)
        assert x.shape == (self.height, self.width, self.inputs)
        out = np.empty((self.height, self.width, 3), np.float32)
        ms = C.c_float()
        self._chk(lib.mdwt_dn_infer(self._h, _f(x), _f(out), C.byref(ms)))
        return out, ms.value

    def set_fused(self, on: bool = True) -> None:
        """Inference as fused kernels (default) or as the training forward's
        separate kernels; the outputs are identical."""
        self._chk(lib.mdwt_dn_set_fused(self._h, 1 if on else 0))

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
        assert xs.shape[1:] == (self.height, self.width, self.inputs)
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
    """[h][w][7 + C] input from an accumulation of `frames` frames: its mean
    (accum), second moment (m2) and auxiliary buffers (aov, [h][w][2 + C])."""
    nc = aov.shape[-1] - 2
    x = np.empty(accum.shape[:2] + (input_width(nc),), np.float32)
    x[..., :3] = accum[..., :3]
    x[..., 3] = aov[..., 0] * DEPTH_SCALE
    x[..., 4:5 + nc] = aov[..., 1:]
    if frames > 1:
        mean = accum[..., :3].astype(np.float64)
        var = np.maximum(m2[..., :3].astype(np.float64) - mean * mean, 0.0) * frames / (frames - 1)
        x[..., 5 + nc] = np.sqrt(var.mean(-1) / frames) * STDERR_SCALE
    else:
        x[..., 5 + nc] = 0.0
    x[..., 6 + nc] = math.log2(frames) * FRAMES_SCALE
    return x


def read_input(r: MdwtRenderer, frames: int) -> np.ndarray:
    a, m2 = r.read_accum(with_m2=True)
    return network_input(a, m2, r.read_aov(), frames)


AOV_FRAMES = 64  # the auxiliary buffers converge long before the colour


def accumulate_inputs(r: MdwtRenderer, levels, seed: int, aov_frames: int = AOV_FRAMES):
    """Yields the network inputs of one accumulation run (seed) read at each
    frame count of `levels` (ascending); the run continues between the
    reads, so the renderer must not be used while the generator is live.
    The auxiliary buffers accumulate over the first aov_frames frames only."""
    r.set_moments(True)
    r.set_seed(seed)
    r.set_aov(True)
    r.set_aov_frames(aov_frames)
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
    r.set_moments(True)
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


def orbit_view(rng: np.random.Generator, centre=(0.5, 0.5, 0.5)):
    """(from, at, up, fovy) of a random orbit camera framed like the
    workload's: a direction uniform on the sphere, distance 1.5..2.3, the aim
    point within 0.15 of the centre, 45 degree field of view."""
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    at = np.asarray(centre) + rng.uniform(-0.15, 0.15, 3)
    frm = np.asarray(centre) + rng.uniform(1.5, 2.3) * v
    up = (0.0, 1.0, 0.0) if abs(v[1]) < 0.95 else (1.0, 0.0, 0.0)
    return tuple(float(c) for c in frm), tuple(float(c) for c in at), up, 45.0


def tiles(img: np.ndarray, size: int):
    """Non-overlapping size x size tiles of img [h][w][...] (row-major)."""
    h, w = img.shape[:2]
    return [img[y:y + size, x:x + size] for y in range(0, h - size + 1, size)
            for x in range(0, w - size + 1, size)]


TRAIN_LEVELS = (1, 2, 4, 16, 64, 256, 1024, 2048, 4096, 8192)


def make_training_set(n_scenes: int = 3, views: int = 2, size: int = 256, levels=TRAIN_LEVELS,
                      target_frames: int = 16384, seed: int = 0, scale: int = 1,
                      workload: str = "sync_4ch256", min_signal: float = 0.01, log=None):
    """(inputs [N][size][size][7 + C], targets [N][size][size][3], frame
    counts [N]) from n_scenes seeded volumes of the workload family (seeds
    drawn from `seed`, so held out from the fixed-seed workload) seen from
    `views` random orbit cameras each at the workload's image size: one
    target render and one input accumulation read at each of `levels` per
    view, cut into size x size tiles; tiles whose target mean is below
    min_signal (background) are dropped."""
    rng = np.random.default_rng(seed)
    xs, ts, ks = [], [], []
    for s in range(n_scenes):
        sc = scenes.make_scene(workload, seed=int(rng.integers(1, 2**31)), scale=scale)
        r = MdwtRenderer(0)
        scenes.load_scene(r, sc, 2)
        for vi in range(views):
            if log:
                log(f"[denoise] scene {s + 1}/{n_scenes} view {vi + 1}/{views}")
            r.set_camera(*orbit_view(rng))
            s_in, s_t = (int(v) for v in rng.integers(1, 2**31, 2))
            r.set_aov(False)
            t, _ = render_mean(r, target_frames, s_t)
            t_tiles = tiles(t, size)
            keep = [i for i, tt in enumerate(t_tiles) if float(tt.mean()) >= min_signal]
            for k, x in zip(levels, accumulate_inputs(r, levels, s_in)):
                x_tiles = tiles(x, size)
                for i in keep:
                    xs.append(x_tiles[i].copy())
                    ts.append(t_tiles[i].copy())
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
    exponentially to lr / 10 over the run. The next batch is gathered on a
    helper thread while the GPU trains on the current one (the C calls
    release the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(seed)
    plan = []
    for _ in range(steps):
        idx = rng.choice(len(xs), size=min(batch, len(xs)), replace=False)
        plan.append((idx, rng.random() < 0.5, rng.random() < 0.5))

    def gather(i):
        idx, fx, fy = plan[i]
        # flips keep the pairs valid and widen the data
        x, t = xs[idx], ts[idx]
        if fx:
            x, t = x[:, :, ::-1], t[:, :, ::-1]
        if fy:
            x, t = x[:, ::-1], t[:, ::-1]
        w = None if weights is None else np.asarray(weights)[idx]
        return np.ascontiguousarray(x), np.ascontiguousarray(t), w

    losses = []
    with ThreadPoolExecutor(1) as pool:
        nxt = pool.submit(gather, 0) if steps else None
        for i in range(steps):
            x, t, w = nxt.result()
            if i + 1 < steps:
                nxt = pool.submit(gather, i + 1)
            step_lr = lr * (0.1 ** (i / max(1, steps)))
            losses.append(dn.train_step(x, t, step_lr, w))
            if log and (i % 500 == 0 or i == steps - 1):
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


def _main() -> int:
    """Time the denoiser's kernels on synthetic data: training steps at the
    training tile size and inference at the workload's image size."""
    import argparse
    import time

    ap = argparse.ArgumentParser(description="time denoiser training and inference (synthetic data)")
    ap.add_argument("--size", type=int, default=256, help="training tile side")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--infer", type=int, default=1024, help="inference image side")
    ap.add_argument("--infer-runs", type=int, default=5)
    args = ap.parse_args()
    rng = np.random.default_rng(0)
    xs = rng.uniform(0, 1, (2 * args.batch, args.size, args.size, INPUTS)).astype(np.float32)
    ts = rng.uniform(0, 1, (2 * args.batch, args.size, args.size, 3)).astype(np.float32)
    dn = Denoiser(args.size, args.size, seed=1)
    train(dn, xs, ts, 3, batch=args.batch)
    t = time.perf_counter()
    train(dn, xs, ts, args.steps, batch=args.batch)
    step_ms = (time.perf_counter() - t) * 1000.0 / max(1, args.steps)
    print(f"train step {step_ms:.1f} ms (batch {args.batch}, {args.size}^2)", flush=True)
    dn.close()
    big = Denoiser(args.infer, args.infer, seed=1)
    x = rng.uniform(0, 1, (args.infer, args.infer, INPUTS)).astype(np.float32)
    ms = [big.infer(x)[1] for _ in range(args.infer_runs)]
    med = float(np.median(ms))
    gf = network_gflop(args.infer, args.infer)
    print(f"inference {args.infer}^2: {med:.3f} ms (median of {len(ms)}), {gf:.1f} GFLOP, "
          f"{gf / med:.1f} TFLOP/s = {100 * gf / med / BF16_PEAK_TFLOPS:.1f}% of the BF16 peak",
          flush=True)
    big.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
