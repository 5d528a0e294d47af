"""Build the native traversal library `arcadia_amd/libarcadia_hip.so` for gfx950.

Host sources (world builder, generators, ray sets) are compiled with g++, the HIP sources
with hipcc, all with -ffp-contract=off so every binary32 operation is rounded on its own
(bit-exact traversal). The HIP code is also built without packed FP32 instructions.

Usage: python -m arcadia_amd.build  (or arcadia_amd.build.build())
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libarcadia_hip.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HOST_SOURCES = ["world_build.cpp", "world_edit.cpp", "patch_plan.cpp", "worldgen.cpp", "rays.cpp", "snapshot.cpp"]
HIP_HOST_SOURCES = ["api.cpp"]
HIP_SOURCES = ["trace_kernels.hip", "shadow.hip", "rigid.hip", "gbuffer.hip", "patch.hip", "sharc.hip", "pathtrace.hip", "nrc.hip", "restir.hip"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _hipcc() -> str:
    p = ROCM / "bin" / "hipcc"
    return str(p) if p.exists() else (shutil.which("hipcc") or "hipcc")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")


def build(verbose: bool = False) -> Path:
    """Compile every source and link the shared library; returns its path."""
    OBJ.mkdir(exist_ok=True)
    objs: list[str] = []
    for src in HOST_SOURCES:
        o = OBJ / (src + ".o")
        _run(["g++", *COMMON, "-pthread", "-c", str(CSRC / src), "-o", str(o)])
        objs.append(str(o))
    for src in HIP_HOST_SOURCES:
        o = OBJ / (src + ".o")
        _run(["g++", *COMMON, "-pthread", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", "-c", str(CSRC / src), "-o", str(o)])
        objs.append(str(o))
    for src in HIP_SOURCES:
        o = OBJ / (src + ".o")
        _run([_hipcc(), *COMMON, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(str(o))
    tmp = LIB.with_suffix(".so.tmp")
    _run([_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
