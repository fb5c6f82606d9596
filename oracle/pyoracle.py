"""ctypes wrapper of the CPU oracle (libmdwt_oracle.so). TEST INFRASTRUCTURE:
imported only by tests/, __graft_entry__.smoke() and bench.py's CPU baseline
leg. Mirrors mdwtracker_amd.capi.MdwtRenderer method for method so that a test
can drive both with the same calls."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmdwt_oracle.so"
REF_HILBERT = HERE / "_ref" / "libref_hilbert.so"

_DTYPES = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float16): 2,
           np.dtype(np.float32): 3}
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_h = C.c_void_p

_SIG = {
    "oracle_create": ([], C.c_void_p),
    "oracle_destroy": ([_h], None),
    "oracle_last_error": ([_h], C.c_char_p),
    "oracle_push_channel": ([_h, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "oracle_set_bounds": ([_h, _fp], C.c_int),
    "oracle_init": ([_h, C.c_int, _ip, C.c_int], C.c_int),
    "oracle_set_tf": ([_h, C.c_int, _fp, C.c_int, C.c_float, C.c_float, C.c_float], C.c_int),
    "oracle_set_camera": ([_h, _fp, _fp, _fp, C.c_float], C.c_int),
    "oracle_set_dt": ([_h, C.c_float], C.c_int),
    "oracle_reset_dt": ([_h], C.c_int),
    "oracle_set_frame": ([_h, C.c_int, C.c_int, C.c_int, C.c_int, _fp], C.c_int),
    "oracle_set_seed": ([_h, C.c_uint32], C.c_int),
    "oracle_set_lighting": ([_h, C.c_float, C.c_float], C.c_int),
    "oracle_set_light": ([_h, _fp, C.c_int, C.c_int], C.c_int),
    "oracle_set_material": ([_h, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, _fp], C.c_int),
    "oracle_set_threads": ([_h, C.c_int], C.c_int),
    "oracle_set_faithful": ([_h, C.c_int], C.c_int),
    "oracle_set_jitter_2d": ([_h, C.c_int], C.c_int),
    "oracle_set_conservative_majorants": ([_h, C.c_int], C.c_int),
    "oracle_render": ([_h, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "oracle_reset_accumulation": ([_h], C.c_int),
    "oracle_read_accum": ([_h, _fp, _fp], C.c_int),
    "oracle_read_fb": ([_h, _up], C.c_int),
    "oracle_read_ranges": ([_h, _fp], C.c_int),
    "oracle_read_majorants": ([_h, _fp], C.c_int),
    "oracle_read_counters": ([_h, _u64p], C.c_int),
    "oracle_get_info": ([_h, _ip], C.c_int),
    "oracle_get_dt": ([_h], C.c_float),
    "oracle_get_camera": ([_h, _fp], C.c_int),
    "oracle_get_domains": ([_h, _fp], C.c_int),
    "oracle_trace_cells": ([_h, _fp, _fp, C.c_int, C.c_int, _ip, _fp, _ip], C.c_int),
    "oracle_gen_rays": ([_h, C.c_uint32, C.c_uint32, _fp, _fp], C.c_int),
    "oracle_clip_rays": ([_h, _fp, _fp, C.c_int, _fp, _ip], C.c_int),
    "oracle_philox4x32": ([_up, _up, C.c_int, _up], None),
    "oracle_philox_rounds": ([], C.c_int),
    "oracle_logf": ([C.c_float], C.c_float),
    "oracle_logf_poly": ([C.c_float], C.c_float),
    "oracle_expf": ([C.c_float], C.c_float),
    "oracle_powf": ([C.c_float, C.c_float], C.c_float),
    "oracle_sample": ([_h, C.c_int, _fp], C.c_float),
    "oracle_tf": ([_h, C.c_int, C.c_float, _fp], None),
}


def build() -> None:
    """Compile the oracle (g++) in place."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    for name, (args, res) in _SIG.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()
PHILOX_ROUNDS = int(lib.oracle_philox_rounds())


def _f(a):
    return a.ctypes.data_as(_fp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _vec3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(3))


class OracleError(RuntimeError):
    pass


def philox4x32(ctr, key, rounds=None) -> np.ndarray:
    """One Philox4x32 block; rounds defaults to the renderer's (7)."""
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    n = PHILOX_ROUNDS if rounds is None else int(rounds)
    lib.oracle_philox4x32(c.ctypes.data_as(_up), k.ctypes.data_as(_up), n, out.ctypes.data_as(_up))
    return out


def logf(x: float) -> float:
    return lib.oracle_logf(float(x))


def logf_poly(x: float) -> float:
    """The division-free portable log of every free-flight step."""
    return lib.oracle_logf_poly(float(x))


def expf(x: float) -> float:
    return lib.oracle_expf(float(x))


def powf(x: float, y: float) -> float:
    return lib.oracle_powf(float(x), float(y))


class OracleRenderer:
    def __init__(self, threads: int = 0):
        self._h = C.c_void_p(lib.oracle_create())
        lib.oracle_set_threads(self._h, threads if threads > 0 else min(16, os.cpu_count() or 1))
        self.channels = 0
        self.width, self.height = 1024, 1024

    def close(self):
        if self._h:
            lib.oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            raise OracleError(lib.oracle_last_error(self._h).decode())

    def push_channel(self, vox):
        vox = np.ascontiguousarray(vox)
        nz, ny, nx = vox.shape
        self._chk(lib.oracle_push_channel(self._h, vox.ctypes.data, _DTYPES[vox.dtype], nx, ny, nz))
        self.channels += 1

    def set_bounds(self, remap):
        self._chk(lib.oracle_set_bounds(self._h, _f(_vec3(remap))))

    def init(self, mode=0, mc=(0, 0, 0), auto_camera=True):
        m = np.ascontiguousarray(np.broadcast_to(np.asarray(mc, dtype=np.int32), (3,)))
        self._chk(lib.oracle_init(self._h, mode, _i(m), int(auto_camera)))

    def set_tf(self, ch, rgba, opacity_scale=1.0, xf_range=(0.0, 1.0)):
        t = np.ascontiguousarray(np.asarray(rgba, dtype=np.float32).reshape(-1, 4))
        self._chk(lib.oracle_set_tf(self._h, ch, _f(t), t.shape[0], opacity_scale,
                                    float(xf_range[0]), float(xf_range[1])))

    def set_camera(self, frm, at, up, fovy):
        self._chk(lib.oracle_set_camera(self._h, _f(_vec3(frm)), _f(_vec3(at)), _f(_vec3(up)), fovy))

    def set_dt(self, dt):
        self._chk(lib.oracle_set_dt(self._h, dt))

    def reset_dt(self):
        self._chk(lib.oracle_reset_dt(self._h))

    def set_frame(self, width, height, spp=1, accumulate=True, bg=(0.0, 0.0, 0.0)):
        self._chk(lib.oracle_set_frame(self._h, width, height, spp, int(accumulate), _f(_vec3(bg))))
        self.width, self.height = width, height

    def set_seed(self, seed):
        self._chk(lib.oracle_set_seed(self._h, seed & 0xFFFFFFFF))

    def set_faithful(self, on=True):
        """Reference-faithful tracking: libm log, the TF remap's divisions,
        one sequential RNG stream per ray (see oracle_set_faithful)."""
        self._chk(lib.oracle_set_faithful(self._h, int(bool(on))))

    def set_jitter_2d(self, on=True):
        """y jitter from word 1 of the jitter block (see oracle_set_jitter_2d)."""
        self._chk(lib.oracle_set_jitter_2d(self._h, int(bool(on))))

    def set_conservative_majorants(self, on=True):
        """Majorants that bound every interpolated TF opacity of the cell's
        value range (see oracle_set_conservative_majorants)."""
        self._chk(lib.oracle_set_conservative_majorants(self._h, int(bool(on))))

    def set_lighting(self, light_intensity, ambient):
        self._chk(lib.oracle_set_lighting(self._h, light_intensity, ambient))

    def set_light(self, direction, shadows=False, gradient=False):
        self._chk(lib.oracle_set_light(self._h, _f(_vec3(direction)), int(shadows), int(gradient)))

    def set_material(self, ch, diffuse=0.8, specular=0.3, shininess=8.0, gradient_step=0.01,
                     spec=(1.0, 1.0, 1.0)):
        self._chk(lib.oracle_set_material(self._h, ch, diffuse, specular, shininess, gradient_step,
                                          _f(_vec3(spec))))

    def render(self, n_frames=1, rows=None, threads=1):
        y0, y1 = (0, 0) if rows is None else rows
        self._chk(lib.oracle_render(self._h, n_frames, y0, y1, threads))

    def reset_accumulation(self):
        self._chk(lib.oracle_reset_accumulation(self._h))

    def read_accum(self, with_m2=False):
        a = np.empty((self.height, self.width, 4), np.float32)
        m = np.empty_like(a)
        self._chk(lib.oracle_read_accum(self._h, _f(a), _f(m)))
        return (a, m) if with_m2 else a

    def read_fb(self):
        a = np.empty((self.height, self.width), np.uint32)
        self._chk(lib.oracle_read_fb(self._h, a.ctypes.data_as(_up)))
        return a

    def info(self):
        o = np.zeros(10, np.int32)
        self._chk(lib.oracle_get_info(self._h, _i(o)))
        return {"mc": tuple(int(v) for v in o[0:3]), "dims": tuple(int(v) for v in o[3:6]),
                "channels": int(o[6]), "mode": int(o[7]), "frame_id": int(o[8]),
                "accum_id": int(o[9])}

    def read_ranges(self):
        i = self.info()
        a = np.empty((int(np.prod(i["mc"])) * i["channels"], 2), np.float32)
        self._chk(lib.oracle_read_ranges(self._h, _f(a)))
        return a

    def read_majorants(self):
        i = self.info()
        n = int(np.prod(i["mc"])) * (i["channels"] if i["mode"] <= 2 else 1)
        a = np.empty(n, np.float32)
        self._chk(lib.oracle_read_majorants(self._h, _f(a)))
        return a

    def read_counters(self):
        o = np.zeros(4, np.uint64)
        self._chk(lib.oracle_read_counters(self._h, o.ctypes.data_as(_u64p)))
        return {"samples": int(o[0]), "visits": int(o[1]), "rejections": int(o[2]),
                "rays": int(o[3])}

    def get_dt(self):
        return lib.oracle_get_dt(self._h)

    def get_camera(self):
        a = np.empty(12, np.float32)
        self._chk(lib.oracle_get_camera(self._h, _f(a)))
        return a.reshape(4, 3)

    def get_domains(self):
        a = np.empty(2 * max(self.channels, 1), np.float32)
        self._chk(lib.oracle_get_domains(self._h, _f(a)))
        return a.reshape(-1, 2)[: self.channels]

    def trace_cells(self, org, dirs, max_cells):
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = org.shape[0]
        cells = np.full((n, max_cells), -1, np.int32)
        t01 = np.zeros((n, max_cells, 2), np.float32)
        counts = np.zeros(n, np.int32)
        self._chk(lib.oracle_trace_cells(self._h, _f(org), _f(dirs), n, max_cells, _i(cells),
                                         _f(t01), _i(counts)))
        return cells, t01, counts

    def gen_rays(self, frame, sample):
        o = np.empty((self.height * self.width, 3), np.float32)
        d = np.empty_like(o)
        self._chk(lib.oracle_gen_rays(self._h, frame, sample, _f(o), _f(d)))
        return o, d

    def clip_rays(self, org, dirs):
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        n = org.shape[0]
        t01 = np.zeros((n, 2), np.float32)
        hit = np.zeros(n, np.int32)
        self._chk(lib.oracle_clip_rays(self._h, _f(org), _f(dirs), n, _f(t01), _i(hit)))
        return t01, hit

    def sample(self, ch, p):
        return lib.oracle_sample(self._h, ch, _f(_vec3(p)))

    def tf(self, ch, value):
        out = np.zeros(4, np.float32)
        lib.oracle_tf(self._h, ch, float(value), _f(out))
        return out


def ref_hilbert():
    """The reference's hilbert.h compiled from its own source (oracle/_ref),
    or None when it has not been built."""
    if not REF_HILBERT.exists():
        return None
    lib_ = C.CDLL(str(REF_HILBERT))
    lib_.ref_hilbert_i2c.argtypes = [C.c_uint, C.c_uint, C.c_ulonglong, C.POINTER(C.c_ulonglong)]
    lib_.ref_hilbert_i2c.restype = None
    lib_.ref_hilbert_c2i.argtypes = [C.c_uint, C.c_uint, C.POINTER(C.c_ulonglong)]
    lib_.ref_hilbert_c2i.restype = C.c_ulonglong
    return lib_
