"""The reference's host API on top of libmdwt.

``Renderer`` keeps the method names, argument meaning and error behaviour of
``dtracker::Renderer`` (renderer.h:52-222, RAW path of renderer.cpp);
``RawR`` reads ``<name>_<X>x<Y>x<Z>_<dtype>.raw`` files like ``raw::RawR``
(rawFile.cpp:31-128); ``Camera`` holds the ``camera::Camera`` orientation
(camera.h:83-110). All compute runs in libmdwt.so.
"""
from __future__ import annotations

import os
import re
import time
from dataclasses import dataclass, field

import numpy as np

from .capi import MAX_CHANNELS, MdwtError, MdwtRenderer, material_preset

# rawFile.cpp:63-82 dtype names -> numpy dtype
_RAW_DTYPES = {"int8": np.int8, "int16": np.int16, "int32": np.int32, "uint8": np.uint8,
               "uint16": np.uint16, "uint32": np.uint32, "float32": np.float32,
               "double64": np.float64, "float16": np.float16}


class RawR:
    """RAW volume file (rawFile.h:93-165). dims and dtype come from the file
    name: ``..._<X>x<Y>x<Z>_<dtype>.raw``; unspecified dtype means uint8
    (rawFile.h:165). ``data`` keeps the stored values; uint8/uint16/float16/
    float32 are pushed as they are (the library applies the reference's
    conversion, rawFile.cpp:100-123), the other types are converted on the
    host with the same formulas and pushed as float32."""

    def __init__(self, path: str | os.PathLike | None = None, data: np.ndarray | None = None):
        self.fileName = str(path) if path is not None else "<memory>"
        self.bounds = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], np.float32)  # rawFile.h:162
        if data is not None:
            self.data = np.ascontiguousarray(data)
            self.dims = tuple(int(d) for d in self.data.shape[::-1])
            return
        dims, dtype = None, "uint8"
        for tok in os.path.basename(self.fileName).split("_"):
            tok = tok.split(".")[0]
            m = re.fullmatch(r"(\d+)x(\d+)x(\d+)", tok)
            if m:
                dims = tuple(int(v) for v in m.groups())
            elif tok in _RAW_DTYPES:
                dtype = tok
        if dims is None:
            raise ValueError(f"{self.fileName}: no <X>x<Y>x<Z> token in the file name")
        self.dims = dims
        n = dims[0] * dims[1] * dims[2]
        self.data = np.fromfile(self.fileName, dtype=_RAW_DTYPES[dtype], count=n).reshape(
            dims[2], dims[1], dims[0])

    def getDims(self):
        return self.dims

    def getBytesPerVoxel(self) -> int:
        return self.data.dtype.itemsize

    def reshapeBounds(self, remap=None):
        """rawFile.cpp:217-237: scale the [0,1]^3 box by the dims (or remap)."""
        s = np.asarray(self.dims if remap is None else remap, np.float32)
        self.bounds = self.bounds * s

    def pushable(self) -> np.ndarray:
        d = self.data
        if d.dtype in (np.uint8, np.uint16, np.float16, np.float32):
            return d
        f = d.astype(np.float32)
        if d.dtype == np.int8:
            return f * np.float32(255) - np.float32(128)
        if d.dtype == np.int16:
            return f * np.float32(65535) - np.float32(32768)
        if d.dtype == np.int32:
            return f * np.float32(4294967295) - np.float32(2147483648)
        if d.dtype == np.uint32:
            return f * np.float32(4294967295)
        return d.astype(np.float32)  # double64 -> float


@dataclass
class Camera:
    """camera::Camera set through setOrientation (camera.h:83-110)."""
    origin: tuple = (0.0, -1.0, 0.0)
    interest: tuple = (0.0, -1.0, -1.0)
    up: tuple = (0.0, 1.0, 0.0)
    fovyInDegrees: float = 60.0
    given: bool = False

    def setOrientation(self, origin, interest, up, fovyInDegrees):
        self.origin, self.interest, self.up = tuple(origin), tuple(interest), tuple(up)
        self.fovyInDegrees = float(fovyInDegrees)
        self.given = True


@dataclass
class TFData:
    """renderer.h:39-50"""
    colorMap: np.ndarray | None = None
    xfDomain: tuple = (0.0, 1.0)
    opacityScale: float = 1.0


@dataclass
class MaterialData:
    """renderer.h:134-141 per-channel material."""
    diffuse: float = 0.8
    specular: float = 0.3
    shininess: float = 8.0
    gradientStep: float = 0.01
    specularColor: tuple = (1.0, 1.0, 1.0)


class Renderer:
    """dtracker::Renderer (RAW path) backed by libmdwt."""

    def __init__(self, device: int = 0):
        self._r = MdwtRenderer(device)
        self.rawPtrs: list[RawR] = []
        self.camera = Camera()
        self.macrocellDims = [0, 0, 0]   # renderer.h:146
        self.fbSize = (1024, 1024)       # renderer.h:127
        self.spp = 1                     # renderer.h:111
        self.enableAccumulation = True   # renderer.h:123
        self.bgColor = (0.0, 0.0, 0.0)   # renderer.h:128
        self.lightIntensity = 1.0        # renderer.h:125
        self.ambient = 0.1               # renderer.h:126
        self.enableShadows = False       # renderer.h:121
        self.enableGradientShading = False  # renderer.h:122
        self.lightDir = (0.0001, -1.0, 0.0001)  # renderer.h:124
        self.materials: list[MaterialData] = []  # renderer.h:133-142
        self._pushed_light = None
        self.dt = 0.05                   # renderer.h:119
        self.mode = 0
        self.tfdatas: list[TFData] = []
        self.accumID = 0
        self.frameID = 0
        self.avgTime = 0.0
        self.minTime = float("inf")
        self.totalTime = 0.0
        self._inited = False

    # renderer.cpp:1031-1044 (the reference accepts a 13th channel; we do not)
    def PushMesh(self, raw: RawR) -> bool:
        if len(self.rawPtrs) >= MAX_CHANNELS:
            return False
        self.rawPtrs.append(raw)
        self.materials.append(MaterialData())
        return True

    # renderer.cpp:324-942
    def Init(self, mode: int = 0, autoSetCamera: bool = True) -> None:
        if not self.rawPtrs:
            raise MdwtError(2, "No data source specified!")
        self.mode = mode
        for raw in self.rawPtrs:
            self._r.push_channel(raw.pushable())
        hi = self.rawPtrs[-1].bounds[1]
        if not np.allclose(hi, 1.0):
            self._r.set_bounds(hi)
        self._r.init(mode, tuple(self.macrocellDims), autoSetCamera and not self.camera.given)
        self.macrocellDims = list(self._r.info()["mc"])
        self.tfdatas = [TFData() for _ in self.rawPtrs]
        self._r.set_frame(self.fbSize[0], self.fbSize[1], self.spp, self.enableAccumulation,
                          self.bgColor)
        self._r.set_lighting(self.lightIntensity, self.ambient)
        self._push_light()
        for c in range(len(self.materials)):
            self._push_material(c)
        if self.camera.given:
            self.UpdateCamera()
        self.dt = self._r.get_dt()
        self._inited = True
        self.ResetAccumulation()

    # renderer.cpp:1094-1193: bad ids are ignored, every change rebuilds majorants
    def SetXFColormap(self, newCM, tfID: int = 0) -> None:
        if tfID >= len(self.tfdatas):
            return
        cm = np.asarray(newCM, np.float32).reshape(-1, 4)
        self.tfdatas[tfID].colorMap = cm
        self._r.set_xf_colormap(tfID, cm)
        self.ResetAccumulation()

    def SetXFOpacityScale(self, opacityScale: float, tfID: int = 0) -> None:
        if tfID >= len(self.tfdatas):
            return
        self.tfdatas[tfID].opacityScale = float(opacityScale)
        self._r.set_xf_opacity_scale(tfID, float(opacityScale))
        self.ResetAccumulation()

    def SetXFRange(self, xfDomain, tfID: int = 0) -> None:
        if tfID >= len(self.tfdatas):
            return
        self.tfdatas[tfID].xfDomain = (float(xfDomain[0]), float(xfDomain[1]))
        self._r.set_xf_range(tfID, float(xfDomain[0]), float(xfDomain[1]))
        self.ResetAccumulation()

    # renderer.cpp:1058-1092
    def UpdateCamera(self) -> None:
        c = self.camera
        self._r.set_camera(c.origin, c.interest, c.up, c.fovyInDegrees)
        self.ResetAccumulation()

    # renderer.cpp:1460-1466, 1444-1453
    def Setdt(self, dt: float) -> None:
        self._r.set_dt(dt)
        self.dt = self._r.get_dt()
        self.ResetAccumulation()

    def Resetdt(self) -> None:
        self._r.reset_dt()
        self.dt = self._r.get_dt()
        self.ResetAccumulation()

    def SetLightIntensity(self, lightIntensity: float) -> None:
        self.lightIntensity = float(lightIntensity)
        self._r.set_lighting(self.lightIntensity, self.ambient)
        self.ResetAccumulation()

    def SetAmbient(self, ambient: float) -> None:
        self.ambient = float(ambient)
        self._r.set_lighting(self.lightIntensity, self.ambient)
        self.ResetAccumulation()

    # renderer.cpp:1468-1474
    def SetLightDirection(self, lightDir) -> None:
        self.lightDir = tuple(float(x) for x in lightDir)
        if self._inited:
            self._push_light()
        self.ResetAccumulation()

    # renderer.cpp:1195-1248 (unknown channel ids are ignored, as there)
    def SetMaterialDiffuse(self, v: float, channelID: int = 0) -> None:
        self._set_material(channelID, diffuse=float(v))

    def SetMaterialSpecular(self, v: float, channelID: int = 0) -> None:
        self._set_material(channelID, specular=float(v))

    def SetMaterialShininess(self, v: float, channelID: int = 0) -> None:
        self._set_material(channelID, shininess=float(v))

    def SetMaterialGradientStep(self, v: float, channelID: int = 0) -> None:
        self._set_material(channelID, gradientStep=float(v))

    def SetMaterialSpecularColor(self, v, channelID: int = 0) -> None:
        self._set_material(channelID, specularColor=tuple(float(x) for x in v))

    # renderer.cpp:1250-1301 (the table lives in the library)
    def ApplyMaterialPreset(self, presetName: str, channelID: int = 0) -> None:
        p = material_preset(presetName)
        self._set_material(channelID, diffuse=p["diffuse"], specular=p["specular"],
                           shininess=p["shininess"], gradientStep=p["gradient_step"],
                           specularColor=p["spec"])

    def _set_material(self, ch: int, **kw) -> None:
        if ch >= len(self.materials):
            return
        for k, v in kw.items():
            setattr(self.materials[ch], k, v)
        if self._inited:
            self._push_material(ch)
        self.ResetAccumulation()

    def _push_material(self, ch: int) -> None:
        m = self.materials[ch]
        self._r.set_material(ch, m.diffuse, m.specular, m.shininess, m.gradientStep,
                             m.specularColor)

    def _push_light(self) -> None:
        cur = (tuple(self.lightDir), bool(self.enableShadows), bool(self.enableGradientShading))
        self._r.set_light(cur[0], cur[1], cur[2])
        self._pushed_light = cur

    # renderer.cpp:1046-1056
    def Resize(self, newSize) -> None:
        self.fbSize = (int(newSize[0]), int(newSize[1]))
        self.Update()

    # renderer.cpp:975-1008: push spp / accumulation / background
    def Update(self) -> None:
        if self._inited:
            self._r.set_frame(self.fbSize[0], self.fbSize[1], self.spp, self.enableAccumulation,
                              self.bgColor)
            cur = (tuple(self.lightDir), bool(self.enableShadows), bool(self.enableGradientShading))
            if cur != self._pushed_light:
                self._push_light()

    # renderer.cpp:944-973: one frame, timed (avg over the accumulation, min)
    def Render(self, heatMapMode: int = 0) -> None:
        t0 = time.perf_counter()
        self._r.render(1)
        dt = time.perf_counter() - t0
        self.totalTime += dt
        self.avgTime = self.totalTime / (self.accumID + 1)
        self.minTime = min(self.minTime, dt)
        self.accumID += 1
        self.frameID += 1

    # renderer.cpp:1488-1495
    def ResetAccumulation(self) -> None:
        if self._inited:
            self._r.reset_accumulation()
        self.accumID = 0
        self.totalTime = 0.0
        self.minTime = float("inf")
        self.avgTime = 0.0

    def Terminate(self) -> None:
        self._r.close()

    # owlBufferGetPointer(frameBuffer) / accumBuffer
    def GetFrameBuffer(self) -> np.ndarray:
        return self._r.read_fb()

    def GetAccumBuffer(self) -> np.ndarray:
        return self._r.read_accum()
