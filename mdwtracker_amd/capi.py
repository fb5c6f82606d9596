"""ctypes binding of libmdwt.so (include/mdwt.h) and a thin object wrapper.

The library is the product: there is no Python or CPU fallback. Importing
this module without a built libmdwt.so raises ImportError (build it with
``python -m mdwtracker_amd.build`` or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libmdwt.so"

U8, U16, F16, F32 = 0, 1, 2, 3
MODE_NDDA, MODE_SYNCDDA, MODE_CUMULATIVE = 0, 2, 3
MAX_CHANNELS = 12
STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_STATE", 3: "ERR_HIP", 4: "ERR_LIMIT", 5: "ERR_OOM"}

_DTYPES = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16, np.dtype(np.float16): F16,
           np.dtype(np.float32): F32}

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_h = C.c_void_p

# name -> argtypes; every function returns int status unless listed in _RESTYPES
SIGNATURES = {
    "mdwt_create": [C.POINTER(C.c_void_p), C.c_int],
    "mdwt_destroy": [_h],
    "mdwt_last_error": [_h],
    "mdwt_push_channel": [_h, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int],
    "mdwt_set_bounds": [_h, _fp],
    "mdwt_init": [_h, C.c_int, _ip, C.c_int],
    "mdwt_set_mode": [_h, C.c_int],
    "mdwt_set_tf": [_h, C.c_int, _fp, C.c_int, C.c_float, C.c_float, C.c_float],
    "mdwt_set_xf_colormap": [_h, C.c_int, _fp, C.c_int],
    "mdwt_set_xf_opacity_scale": [_h, C.c_int, C.c_float],
    "mdwt_set_xf_range": [_h, C.c_int, C.c_float, C.c_float],
    "mdwt_set_camera": [_h, _fp, _fp, _fp, C.c_float],
    "mdwt_set_dt": [_h, C.c_float],
    "mdwt_reset_dt": [_h],
    "mdwt_set_frame": [_h, C.c_int, C.c_int, C.c_int, C.c_int, _fp],
    "mdwt_set_lighting": [_h, C.c_float, C.c_float],
    "mdwt_set_light": [_h, _fp, C.c_int, C.c_int],
    "mdwt_set_material": [_h, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, _fp],
    "mdwt_material_preset": [C.c_char_p, _fp],
    "mdwt_set_seed": [_h, C.c_uint32],
    "mdwt_render": [_h, C.c_int, _fp],
    "mdwt_reset_accumulation": [_h],
    "mdwt_read_accum": [_h, _fp, _fp],
    "mdwt_read_fb": [_h, _up],
    "mdwt_read_ranges": [_h, _fp],
    "mdwt_read_majorants": [_h, _fp],
    "mdwt_get_info": [_h, _ip],
    "mdwt_get_dt": [_h, _fp],
    "mdwt_get_camera": [_h, _fp],
    "mdwt_get_domains": [_h, _fp],
    "mdwt_trace_cells": [_h, _fp, _fp, C.c_int, C.c_int, _ip, _fp, _ip],
    "mdwt_gen_rays": [_h, C.c_uint32, C.c_uint32, _fp, _fp],
    "mdwt_clip_rays": [_h, _fp, _fp, C.c_int, _fp, _ip],
    "mdwt_count_frame": [_h, _u64p],
    "mdwt_last_kernel_ms": [_h, _fp],
    "mdwt_set_tile_order": [_h, C.c_int, C.c_int],
    "mdwt_set_schedule": [_h, C.c_int],
    "mdwt_set_jitter_2d": [_h, C.c_int],
    "mdwt_set_conservative_majorants": [_h, C.c_int],
    "mdwt_stream_copy": [_h, C.c_uint64, C.c_int, _fp],
    "mdwt_stream_read": [_h, C.c_uint64, C.c_int, _fp],
    "mdwt_set_moments": [_h, C.c_int],
    "mdwt_read_fb_noalloc": [_h, _up],
    "mdwt_hilbert_d2xy": [C.c_int, C.c_uint32, _up, _up],
    "mdwt_set_block": [_h, C.c_int],
    "mdwt_set_macrocells": [_h, _ip],
    "mdwt_batch_set_views": [_h, C.c_void_p, C.c_int],
    "mdwt_batch_render": [_h, C.c_int, _fp],
    "mdwt_batch_count": [_h, _u64p],
    "mdwt_batch_reset": [_h],
    "mdwt_batch_set_launch": [_h, C.c_int],
    "mdwt_batch_read_view": [_h, C.c_int, _fp, _fp],
    "mdwt_batch_view_times": [_h, _fp],
    "mdwt_autotune": [_h, _ip, C.c_int, _ip, C.c_int, _ip, C.c_int, C.c_int, _fp, _ip],
    "mdwt_set_aov": [_h, C.c_int],
    "mdwt_set_aov_frames": [_h, C.c_int],
    "mdwt_read_aov": [_h, _fp],
    "mdwt_dn_create": [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32],
    "mdwt_dn_destroy": [_h],
    "mdwt_dn_last_error": [_h],
    "mdwt_dn_num_params": [_h, _ip],
    "mdwt_dn_get_params": [_h, _fp],
    "mdwt_dn_set_params": [_h, _fp],
    "mdwt_dn_infer": [_h, _fp, _fp, _fp],
    "mdwt_dn_set_fused": [_h, C.c_int],
    "mdwt_dn_gradient": [_h, _fp, _fp, _fp, _fp],
    "mdwt_dn_train_step": [_h, _fp, _fp, _fp, C.c_int, C.c_float, _fp],
}


class View(C.Structure):
    """mdwt_view: one camera of a batched launch."""
    _fields_ = [("frm", C.c_float * 3), ("at", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float), ("width", C.c_int), ("height", C.c_int), ("spp", C.c_int),
                ("seed", C.c_uint32)]
_RESTYPES = {"mdwt_last_error": C.c_char_p, "mdwt_dn_last_error": C.c_char_p}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m mdwtracker_amd.build` "
            "(hipcc, gfx950). There is no CPU fallback.")
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    return lib


lib = _load()


class MdwtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def _f(a) -> C.POINTER(C.c_float):
    return a.ctypes.data_as(_fp)


def _i(a) -> C.POINTER(C.c_int32):
    return a.ctypes.data_as(_ip)


def _vec3(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(3))


def material_preset(name: str) -> dict:
    """ApplyMaterialPreset's values (renderer.cpp:1250-1301) for `name`."""
    out = np.zeros(7, np.float32)
    rc = lib.mdwt_material_preset(name.encode(), _f(out))
    if rc:
        raise MdwtError(rc, "material_preset")
    return {"diffuse": float(out[0]), "specular": float(out[1]), "shininess": float(out[2]),
            "gradient_step": float(out[3]), "spec": tuple(float(x) for x in out[4:7])}


def hilbert_d2xy(bits: int, d: int) -> tuple[int, int]:
    x, y = C.c_uint32(), C.c_uint32()
    rc = lib.mdwt_hilbert_d2xy(bits, d, C.byref(x), C.byref(y))
    if rc:
        raise MdwtError(rc, "hilbert_d2xy")
    return x.value, y.value


class MdwtRenderer:
    """One libmdwt handle (one HIP device, one stream). Methods map 1:1 onto
    the C ABI and raise MdwtError on a non-zero status."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib.mdwt_create(C.byref(h), device)
        if rc:
            raise MdwtError(rc, lib.mdwt_last_error(None).decode())
        self._h = h
        self.channels = 0
        self.width, self.height = 1024, 1024

    def close(self) -> None:
        if self._h:
            lib.mdwt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc: int) -> None:
        if rc:
            raise MdwtError(rc, lib.mdwt_last_error(self._h).decode())

    # -- Renderer set-up --------------------------------------------------
    def push_channel(self, vox: np.ndarray) -> None:
        """vox: array of shape (nz, ny, nx) (x fastest), dtype u8/u16/f16/f32."""
        vox = np.ascontiguousarray(vox)
        if vox.ndim != 3 or vox.dtype not in _DTYPES:
            raise ValueError("voxels must be a 3-D uint8/uint16/float16/float32 array")
        nz, ny, nx = vox.shape
        self._chk(lib.mdwt_push_channel(self._h, vox.ctypes.data, _DTYPES[vox.dtype], nx, ny, nz))
        self.channels += 1

    def set_bounds(self, remap) -> None:
        self._chk(lib.mdwt_set_bounds(self._h, _f(_vec3(remap))))

    def init(self, mode: int = 0, mc=(0, 0, 0), auto_camera: bool = True) -> None:
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mc, dtype=np.int32), (3,)))
        self._chk(lib.mdwt_init(self._h, mode, _i(m), int(auto_camera)))

    def set_mode(self, mode: int) -> None:
        self._chk(lib.mdwt_set_mode(self._h, mode))

    def set_tf(self, ch: int, rgba, opacity_scale: float = 1.0, xf_range=(0.0, 1.0)) -> None:
        t = np.ascontiguousarray(np.asarray(rgba, dtype=np.float32).reshape(-1, 4))
        self._chk(lib.mdwt_set_tf(self._h, ch, _f(t), t.shape[0], opacity_scale,
                                  float(xf_range[0]), float(xf_range[1])))

    def set_xf_colormap(self, ch: int, rgba) -> None:
        t = np.ascontiguousarray(np.asarray(rgba, dtype=np.float32).reshape(-1, 4))
        self._chk(lib.mdwt_set_xf_colormap(self._h, ch, _f(t), t.shape[0]))

    def set_xf_opacity_scale(self, ch: int, s: float) -> None:
        self._chk(lib.mdwt_set_xf_opacity_scale(self._h, ch, s))

    def set_xf_range(self, ch: int, lo: float, hi: float) -> None:
        self._chk(lib.mdwt_set_xf_range(self._h, ch, lo, hi))

    def set_camera(self, frm, at, up, fovy: float) -> None:
        self._chk(lib.mdwt_set_camera(self._h, _f(_vec3(frm)), _f(_vec3(at)), _f(_vec3(up)), fovy))

    def set_dt(self, dt: float) -> None:
        self._chk(lib.mdwt_set_dt(self._h, dt))

    def reset_dt(self) -> None:
        self._chk(lib.mdwt_reset_dt(self._h))

    def set_frame(self, width: int, height: int, spp: int = 1, accumulate: bool = True,
                  bg=(0.0, 0.0, 0.0)) -> None:
        self._chk(lib.mdwt_set_frame(self._h, width, height, spp, int(accumulate), _f(_vec3(bg))))
        self.width, self.height = width, height

    def set_lighting(self, light_intensity: float, ambient: float) -> None:
        self._chk(lib.mdwt_set_lighting(self._h, light_intensity, ambient))

    def set_light(self, direction, shadows: bool = False, gradient: bool = False) -> None:
        self._chk(lib.mdwt_set_light(self._h, _f(_vec3(direction)), int(shadows), int(gradient)))

    def set_material(self, ch: int, diffuse: float = 0.8, specular: float = 0.3,
                     shininess: float = 8.0, gradient_step: float = 0.01,
                     spec=(1.0, 1.0, 1.0)) -> None:
        self._chk(lib.mdwt_set_material(self._h, ch, diffuse, specular, shininess, gradient_step,
                                        _f(_vec3(spec))))

    def set_seed(self, seed: int) -> None:
        self._chk(lib.mdwt_set_seed(self._h, seed & 0xFFFFFFFF))

    def set_tile_order(self, order: int, tile: int = 16) -> None:
        self._chk(lib.mdwt_set_tile_order(self._h, order, tile))

    def set_block(self, threads: int) -> None:
        """Workgroup of the frame kernel: 64 (8x8), 128 (16x8) or 256 (16x16) threads."""
        self._chk(lib.mdwt_set_block(self._h, threads))

    def set_macrocells(self, mc) -> None:
        m = np.ascontiguousarray(np.asarray(mc, np.int32).reshape(3))
        self._chk(lib.mdwt_set_macrocells(self._h, _i(m)))

    # -- batched views --------------------------------------------------------
    def batch_set_views(self, views) -> None:
        """views: dicts with from, at, up, fovy, width, height, spp, seed."""
        arr = (View * len(views))()
        for i, v in enumerate(views):
            arr[i].frm[:] = [float(x) for x in v["from"]]
            arr[i].at[:] = [float(x) for x in v["at"]]
            arr[i].up[:] = [float(x) for x in v.get("up", (0.0, 1.0, 0.0))]
            arr[i].fovy = float(v.get("fovy", 45.0))
            arr[i].width, arr[i].height = int(v["width"]), int(v["height"])
            arr[i].spp = int(v.get("spp", 1))
            arr[i].seed = int(v.get("seed", 0)) & 0xFFFFFFFF
        self._views = [(int(v["width"]), int(v["height"])) for v in views]
        self._chk(lib.mdwt_batch_set_views(self._h, C.cast(arr, C.c_void_p), len(views)))

    def batch_render(self, n_frames: int = 1) -> float:
        ms = C.c_float()
        self._chk(lib.mdwt_batch_render(self._h, n_frames, C.byref(ms)))
        return ms.value

    def batch_count(self) -> dict:
        o = np.zeros(4, np.uint64)
        self._chk(lib.mdwt_batch_count(self._h, o.ctypes.data_as(_u64p)))
        return {"samples": int(o[0]), "visits": int(o[1]), "rejections": int(o[2]),
                "rays": int(o[3])}

    def batch_reset(self) -> None:
        self._chk(lib.mdwt_batch_reset(self._h))

    def batch_set_launch(self, launch: int) -> None:
        """0: one launch per view over four streams; 1: one launch for all views;
        2: one persistent launch for all views, replayed from a captured HIP graph."""
        self._chk(lib.mdwt_batch_set_launch(self._h, int(launch)))

    def batch_read_view(self, i: int, with_m2: bool = False):
        w, h = self._views[i]
        a = np.empty((h, w, 4), np.float32)
        m = np.empty_like(a) if with_m2 else None
        self._chk(lib.mdwt_batch_read_view(self._h, i, _f(a), _f(m) if with_m2 else None))
        return (a, m) if with_m2 else a

    def batch_view_times(self) -> np.ndarray:
        us = np.zeros(len(self._views), np.float32)
        self._chk(lib.mdwt_batch_view_times(self._h, _f(us)))
        return us

    def autotune(self, mc_sizes, threads=(64, 128, 256), tiles=(8, 16, 32), frames: int = 4):
        """Times every (macrocell size, workgroup, tile) combination; returns
        (rows of (s, threads, tile, ms, samples/s, visits/s), best (s, threads, tile))."""
        a = np.ascontiguousarray(mc_sizes, np.int32)
        b = np.ascontiguousarray(threads, np.int32)
        c = np.ascontiguousarray(tiles, np.int32)
        table = np.zeros((len(a) * len(b) * len(c), 6), np.float32)
        best = np.zeros(3, np.int32)
        self._chk(lib.mdwt_autotune(self._h, _i(a), len(a), _i(b), len(b), _i(c), len(c), frames,
                                    _f(table), _i(best)))
        return table, tuple(int(x) for x in best)

    def set_aov(self, on: bool = True) -> None:
        """Accumulate the denoiser's auxiliary inputs with every frame."""
        self._chk(lib.mdwt_set_aov(self._h, int(on)))

    def set_aov_frames(self, frames: int) -> None:
        """Accumulate the auxiliary buffers over the first `frames` frames only (0: all)."""
        self._chk(lib.mdwt_set_aov_frames(self._h, frames))

    def read_aov(self) -> np.ndarray:
        """[h][w][2 + C]: first-hit distance, transmittance, hit fraction of each channel."""
        a = np.empty((self.height, self.width, 2 + self.info()["channels"]), np.float32)
        self._chk(lib.mdwt_read_aov(self._h, _f(a)))
        return a

    def set_schedule(self, schedule: int) -> None:
        """0: one thread per pixel, nested loops; 1: wave-coherent tracking
        (the default; -m 0 to -m 5 without lighting; the ray marchers and
        lighting run schedule 0). Images do not depend on it."""
        self._chk(lib.mdwt_set_schedule(self._h, schedule))

    def set_jitter_2d(self, on: bool = True) -> None:
        """y jitter from its own random word (a box filter over the pixel)
        instead of the reference's one scalar for both axes (the default)."""
        self._chk(lib.mdwt_set_jitter_2d(self._h, int(bool(on))))

    def set_conservative_majorants(self, on: bool = True) -> None:
        """Majorants that bound every opacity a TF lookup of the cell's value
        range interpolates (the default keeps the reference's, which can be
        half a texel low); rebuilds them and restarts the accumulation."""
        self._chk(lib.mdwt_set_conservative_majorants(self._h, int(bool(on))))

    # -- frames -------------------------------------------------------------
    def render(self, n_frames: int = 1) -> float:
        ms = C.c_float()
        self._chk(lib.mdwt_render(self._h, n_frames, C.byref(ms)))
        return ms.value

    def reset_accumulation(self) -> None:
        self._chk(lib.mdwt_reset_accumulation(self._h))

    def read_accum(self, with_m2: bool = False):
        a = np.empty((self.height, self.width, 4), np.float32)
        m = np.empty_like(a) if with_m2 else None
        self._chk(lib.mdwt_read_accum(self._h, _f(a), _f(m) if with_m2 else None))
        return (a, m) if with_m2 else a

    def read_fb(self) -> np.ndarray:
        a = np.empty((self.height, self.width), np.uint32)
        self._chk(lib.mdwt_read_fb(self._h, a.ctypes.data_as(_up)))
        return a

    # -- inspection ---------------------------------------------------------
    def info(self) -> dict:
        o = np.zeros(10, np.int32)
        self._chk(lib.mdwt_get_info(self._h, _i(o)))
        return {"mc": tuple(int(v) for v in o[0:3]), "dims": tuple(int(v) for v in o[3:6]),
                "channels": int(o[6]), "mode": int(o[7]), "frame_id": int(o[8]),
                "accum_id": int(o[9])}

    def read_ranges(self) -> np.ndarray:
        i = self.info()
        n = int(np.prod(i["mc"])) * i["channels"]
        a = np.empty((n, 2), np.float32)
        self._chk(lib.mdwt_read_ranges(self._h, _f(a)))
        return a

    def read_majorants(self) -> np.ndarray:
        i = self.info()
        n = int(np.prod(i["mc"])) * (i["channels"] if i["mode"] <= 2 else 1)
        a = np.empty(n, np.float32)
        self._chk(lib.mdwt_read_majorants(self._h, _f(a)))
        return a

    def get_dt(self) -> float:
        v = C.c_float()
        self._chk(lib.mdwt_get_dt(self._h, C.byref(v)))
        return v.value

    def get_camera(self) -> np.ndarray:
        a = np.empty(12, np.float32)
        self._chk(lib.mdwt_get_camera(self._h, _f(a)))
        return a.reshape(4, 3)

    def get_domains(self) -> np.ndarray:
        a = np.empty(2 * max(self.channels, 1), np.float32)
        self._chk(lib.mdwt_get_domains(self._h, _f(a)))
        return a.reshape(-1, 2)[: self.channels]

    def trace_cells(self, org, dirs, max_cells: int):
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = org.shape[0]
        cells = np.full((n, max_cells), -1, np.int32)
        t01 = np.zeros((n, max_cells, 2), np.float32)
        counts = np.zeros(n, np.int32)
        self._chk(lib.mdwt_trace_cells(self._h, _f(org), _f(dirs), n, max_cells, _i(cells),
                                       _f(t01), _i(counts)))
        return cells, t01, counts

    def gen_rays(self, frame: int, sample: int):
        o = np.empty((self.height * self.width, 3), np.float32)
        d = np.empty_like(o)
        self._chk(lib.mdwt_gen_rays(self._h, frame, sample, _f(o), _f(d)))
        return o, d

    def clip_rays(self, org, dirs):
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = org.shape[0]
        t01 = np.zeros((n, 2), np.float32)
        hit = np.zeros(n, np.int32)
        self._chk(lib.mdwt_clip_rays(self._h, _f(org), _f(dirs), n, _f(t01), _i(hit)))
        return t01, hit

    def count_frame(self) -> dict:
        o = np.zeros(4, np.uint64)
        self._chk(lib.mdwt_count_frame(self._h, o.ctypes.data_as(_u64p)))
        return {"samples": int(o[0]), "visits": int(o[1]), "rejections": int(o[2]),
                "rays": int(o[3])}

    def last_kernel_ms(self) -> float:
        v = C.c_float()
        self._chk(lib.mdwt_last_kernel_ms(self._h, C.byref(v)))
        return v.value

    def stream_copy(self, nbytes: int, iters: int = 10) -> float:
        v = C.c_float()
        self._chk(lib.mdwt_stream_copy(self._h, nbytes, iters, C.byref(v)))
        return v.value

    def stream_read(self, nbytes: int, iters: int = 10) -> float:
        """GB/s of a read-only stream over nbytes (mdwt_stream_read)."""
        v = C.c_float()
        self._chk(lib.mdwt_stream_read(self._h, nbytes, iters, C.byref(v)))
        return v.value

    def set_moments(self, on: bool = True) -> None:
        """Second-moment buffers (needed by read_accum(with_m2=True))."""
        self._chk(lib.mdwt_set_moments(self._h, 1 if on else 0))
