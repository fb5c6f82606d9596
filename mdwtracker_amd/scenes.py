"""Seeded synthetic multi-channel volumes, transfer functions and cameras for
the benchmark workloads. There are no datasets offline, so every workload is
generated from a seed:

* ``blobs``: sums of Gaussian blobs (centres U[0,1]^3, sigma U[0.05, 0.2],
  amplitude U[0.3, 1.0]) evaluated at voxel centres (i + 0.5)/n;
* ``fbm``: 5-octave value noise (random lattice values, smoothstep-weighted
  trilinear interpolation, frequency x2 and amplitude x0.5 per octave);
* quantisation to uint8 is round-half-even(255 * clamp(v / max v, 0, 1));
* transfer functions: k control points (x sorted U[0,1], alpha U[0,1])
  rasterised piecewise-linearly at 256 texel centres, one hue per channel.

Workload names follow the five benchmark configurations: ``ndda_cpu_2ch64``
(2 x 64^3 u8, blobs, 256^2), ``sync_4ch256`` (4 x 256^3 u8, fBm + 16 blobs,
1024^2), ``batch_4ch512_f16``, ``denoise_4ch256`` and ``large_10ch_f32``.
"""
from __future__ import annotations

import colorsys
from dataclasses import dataclass, field

import numpy as np


def _centres(n: int) -> np.ndarray:
    return (np.arange(n, dtype=np.float64) + 0.5) / n


def gaussian_blobs(shape, n_blobs: int, rng: np.random.Generator,
                   sigma=(0.05, 0.2), amp=(0.3, 1.0)) -> np.ndarray:
    """Sum of n_blobs Gaussians on a (nz, ny, nx) grid over the unit cube."""
    nz, ny, nx = shape
    zs, ys, xs = _centres(nz), _centres(ny), _centres(nx)
    out = np.zeros(shape, np.float32)
    for _ in range(n_blobs):
        c = rng.uniform(0.0, 1.0, 3)
        s = rng.uniform(*sigma)
        a = rng.uniform(*amp)
        gx = np.exp(-((xs - c[0]) ** 2) / (2 * s * s)).astype(np.float32)
        gy = np.exp(-((ys - c[1]) ** 2) / (2 * s * s)).astype(np.float32)
        gz = (a * np.exp(-((zs - c[2]) ** 2) / (2 * s * s))).astype(np.float32)
        out += gz[:, None, None] * (gy[:, None] * gx[None, :])[None, :, :]
    return out


def _lerp_axis(lat: np.ndarray, n: int, freq: int, axis: int) -> np.ndarray:
    u = _centres(n) * freq
    i = np.minimum(np.floor(u).astype(np.int64), freq - 1)
    t = u - i
    w = (t * t * (3.0 - 2.0 * t)).astype(np.float32)
    a = np.take(lat, i, axis=axis)
    b = np.take(lat, i + 1, axis=axis)
    shape = [1, 1, 1]
    shape[axis] = n
    w = w.reshape(shape)
    return a + w * (b - a)


def value_noise_fbm(shape, rng: np.random.Generator, octaves: int = 5, base_freq: int = 4,
                    z_stretch: float = 1.0) -> np.ndarray:
    """fBm of value noise in [0, 1] on a (nz, ny, nx) grid. z_stretch > 1
    elongates features along z (z frequency divided by z_stretch)."""
    nz, ny, nx = shape
    out = np.zeros(shape, np.float32)
    norm = 0.0
    for o in range(octaves):
        f = base_freq * (2 ** o)
        fz = max(1, int(round(f / z_stretch)))
        lat = rng.uniform(0.0, 1.0, (fz + 1, f + 1, f + 1)).astype(np.float32)
        v = _lerp_axis(lat, nx, f, 2)
        v = _lerp_axis(v, ny, f, 1)
        v = _lerp_axis(v, nz, fz, 0)
        a = 0.5 ** o
        out += np.float32(a) * v
        norm += a
    return out / np.float32(norm)


def normalise(v: np.ndarray) -> np.ndarray:
    m = float(v.max())
    return np.clip(v / (m if m > 0 else 1.0), 0.0, 1.0).astype(np.float32)


def quantize_u8(v: np.ndarray) -> np.ndarray:
    """round-half-even(255 * clamp(v / max v, 0, 1))"""
    return np.rint(255.0 * normalise(v)).astype(np.uint8)


def hue_rgb(h: float, s: float = 0.8, v: float = 1.0) -> tuple:
    return colorsys.hsv_to_rgb(h % 1.0, s, v)


def piecewise_tf(n_points: int, rng: np.random.Generator, hue: float,
                 n_texels: int = 256, x_lo: float = 0.0, clear_first: bool = False) -> np.ndarray:
    """(n_texels, 4) float32 RGBA: alpha piecewise-linear through n_points
    seeded control points (x sorted U[x_lo, 1], alpha U[0, 1], constant beyond
    the ends), constant hue. clear_first makes the lowest control point (and
    every value below it) transparent, as scientific-visualisation transfer
    functions usually hide the background."""
    xs = np.sort(rng.uniform(x_lo, 1.0, n_points))
    al = rng.uniform(0.0, 1.0, n_points)
    if clear_first:
        al[0] = 0.0
    u = _centres(n_texels)
    a = np.interp(u, xs, al)
    rgb = hue_rgb(hue)
    t = np.empty((n_texels, 4), np.float32)
    t[:, 0], t[:, 1], t[:, 2] = rgb
    t[:, 3] = a
    return t


def orbit_cameras(n: int, rng: np.random.Generator, radius: float = 2.5,
                  centre=(0.5, 0.5, 0.5)) -> np.ndarray:
    """n camera positions uniform on a sphere around centre: (n, 3)."""
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return (np.asarray(centre) + radius * v).astype(np.float32)


def orbit_camera(radius: float, direction=(0.4, 0.3, 1.0), centre=(0.5, 0.5, 0.5),
                 fovy: float = 45.0) -> tuple:
    """(from, at, up, fovy) on the sphere of `radius` around `centre`."""
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    frm = tuple(float(c + radius * x) for c, x in zip(centre, d))
    return frm, tuple(float(c) for c in centre), (0.0, 1.0, 0.0), fovy


@dataclass
class Scene:
    name: str
    volumes: list
    tfs: list
    mc: tuple
    width: int
    height: int
    spp: int = 1
    frames: int = 1
    modes: tuple = (0,)
    seed: int = 1234
    extra: dict = field(default_factory=dict)
    opacity: float = 1.0
    camera: tuple = None  # (from, at, up, fovy); None = the reference default camera

    @property
    def channels(self) -> int:
        return len(self.volumes)

    def bytes_per_voxel(self) -> int:
        return self.volumes[0].dtype.itemsize


def blob_volume_u8(n: int, n_blobs: int, rng) -> np.ndarray:
    return quantize_u8(gaussian_blobs((n, n, n), n_blobs, rng))


def noise_blob_volume(shape, n_blobs: int, rng, z_stretch: float = 1.0) -> np.ndarray:
    v = value_noise_fbm(shape, rng, z_stretch=z_stretch) + gaussian_blobs(shape, n_blobs, rng)
    return normalise(v)


def make_scene(name: str, seed: int = 1234, scale: int = 1) -> Scene:
    """Build one named workload. ``scale`` > 1 divides every volume and image
    side by ``scale`` (for quick tests of the same scene family)."""
    rng = np.random.default_rng(seed)
    # Opacity: Woodcock steps are scaled by dt = half a voxel (Resetdt), so a
    # voxel's optical depth is 2 * alpha^3 * opacity; the scales below keep
    # every workload semi-transparent (the reference's experiment scripts
    # reach the same regime with -cb -dt 0.5 and low-alpha transfer functions).
    framing = orbit_camera(1.8)
    if name == "ndda_cpu_2ch64":
        n = 64 // scale
        vols = [blob_volume_u8(n, 8, rng) for _ in range(2)]
        hue = float(rng.uniform())
        tfs = [piecewise_tf(4, rng, hue, clear_first=True) for _ in range(2)]
        return Scene(name, vols, tfs, (n // 8,) * 3, 256 // scale, 256 // scale, 1, 64, (0,), seed,
                     opacity=0.05, camera=None)
    if name == "sync_4ch256":
        n = 256 // scale
        vols = [quantize_u8(noise_blob_volume((n, n, n), 16, rng)) for _ in range(4)]
        tfs = [piecewise_tf(8, rng, c / 4.0, x_lo=0.2, clear_first=True) for c in range(4)]
        return Scene(name, vols, tfs, (n // 16,) * 3, 1024 // scale, 1024 // scale, 1, 256,
                     (0, 2), seed, opacity=0.01 * scale, camera=framing)
    if name == "batch_4ch512_f16":
        n = 512 // scale
        vols = [noise_blob_volume((n, n, n), 16, rng).astype(np.float16) for _ in range(4)]
        tfs = [piecewise_tf(8, rng, c / 4.0, x_lo=0.2, clear_first=True) for c in range(4)]
        return Scene(name, vols, tfs, (n // 16,) * 3, 1024 // scale, 1024 // scale, 1, 1, (2,),
                     seed, {"views": 64}, opacity=0.005 * scale, camera=framing)
    if name == "denoise_4ch256":
        n = 256 // scale
        vols = [quantize_u8(noise_blob_volume((n, n, n), 16, rng)) for _ in range(4)]
        tfs = [piecewise_tf(8, rng, c / 4.0, x_lo=0.2, clear_first=True) for c in range(4)]
        return Scene(name, vols, tfs, (n // 16,) * 3, 1024 // scale, 1024 // scale, 1, 4096, (2,),
                     seed, opacity=0.01 * scale, camera=framing)
    if name == "large_10ch_f32":
        nx, ny, nz = 512 // scale, 512 // scale, 128 // scale
        vols = [noise_blob_volume((nz, ny, nx), 32, rng, z_stretch=4.0) for _ in range(10)]
        tfs = [piecewise_tf(8, rng, c / 10.0, x_lo=0.2, clear_first=True) for c in range(10)]
        return Scene(name, vols, tfs, (nx // 16, ny // 16, max(1, nz // 16)), 2048 // scale,
                     2048 // scale, 1, 1, (2, 3), seed, opacity=0.005 * scale, camera=framing)
    raise KeyError(name)


def load_scene(r, scene: Scene, mode: int) -> None:
    """Push a scene into a renderer object with the MdwtRenderer interface
    (push_channel / init / set_tf / set_camera / set_frame / set_seed)."""
    for v in scene.volumes:
        r.push_channel(v)
    r.init(mode, scene.mc, True)
    for c, t in enumerate(scene.tfs):
        r.set_tf(c, t, scene.opacity, (0.0, 1.0))
    r.set_frame(scene.width, scene.height, scene.spp, True, (0.0, 0.0, 0.0))
    if scene.camera is not None:
        r.set_camera(*scene.camera)
    r.set_seed(scene.seed)
