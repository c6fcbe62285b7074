"""Build recipe for the native libraries (gfx950 HIP backend + C++ host layer).

Outputs, in-tree under crystalgraphics_amd/lib/:
  libcg_device.so        HIP kernels + the C-ABI of include/cg_device.h
  libcrystal_graphics.so crystal::graphics host layer (scene prep, API)
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every float operation is one IEEE operation, in the
# order written, on both the device and the host side.  Packed FP32 VALU
# instructions are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_device() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libcg_device.so"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", str(out), str(PKG / "csrc/device/cg_device.hip")])
    return out


def build_host() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libcrystal_graphics.so"
    srcs = [str(PKG / "csrc/host" / f) for f in ("scene_prep.cpp", "scene_gen.cpp", "graphics.cpp", "serve.cpp",
                                                              "gltf.cpp", "jpeg.cpp")]
    _run(["g++", *HOST_FLAGS, "-shared", "-o", str(out), *srcs,
          f"-L{LIB}", "-lcg_device", "-Wl,-rpath,$ORIGIN", "-lpthread", "-lz"])
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


def build_all() -> None:
    build_device()
    build_host()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built:", *sorted(p.name for p in LIB.glob("*.so")))
