"""Build recipe of libmdwt.so (hipcc, gfx950), in-tree.

Every HIP source is compiled with -ffp-contract=off and correctly rounded
fp32 division/sqrt: the kernels' fp32 arithmetic must round exactly like the
host-side reference semantics (see mdwt_device.hpp). Packed fp32 VALU ops are
disabled (-packed-fp32-ops): they are not needed here and some of their forms
are not trusted on the target boxes.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libmdwt.so"
OBJ = PKG / "_obj"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MDWT_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"--offload-arch={ARCH}",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
    "-Wall",
]

SOURCES = ["kernels_build.hip", "kernels_render.hip", "kernels_batch.hip", "kernels_wave.hip",
           "kernels_denoise.hip", "mdwt_api.cpp", "denoise_api.cpp"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = (res.stdout or "") + (res.stderr or "")
    # hipcc's host pass does not know the packed-fp32 feature; that is expected
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n" + "\n".join(lines))
    if verbose and lines:
        print("\n".join(lines), flush=True)


def build(verbose: bool = False, extra_flags: list[str] | None = None,
          out: Path | None = None) -> Path:
    """Compile every HIP/C++ source and link libmdwt.so (or `out`). Always
    recompiles. extra_flags (e.g. tuning -D defines) go to every compile."""
    out = Path(out) if out is not None else LIB
    obj_dir = OBJ if out == LIB else out.parent / "_obj"
    obj_dir.mkdir(parents=True, exist_ok=True)
    extra = list(extra_flags or [])
    jobs = []
    for src in SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append([HIPCC, *COMMON, *extra, *lang, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    objs = [str(obj_dir / (Path(s).stem + ".o")) for s in SOURCES]
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *objs], verbose)
    os.replace(tmp, out)
    if out == LIB:
        # offline driver: plain C++ on the C ABI (no HIP code), rpath to this folder
        _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}",
              str(CSRC / "mdwt_offline.cpp"), f"-L{PKG}", "-lmdwt", "-Wl,-rpath,$ORIGIN",
              "-o", str(PKG / "mdwt_offline")], verbose)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
