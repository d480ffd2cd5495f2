"""Builds world_amd/libworld_amd.so for gfx950 (hipcc for the kernels, g++ for
host-only C++ so the <random> initializers match the CPU oracle's libstdc++
code), and the C++ façade test program under tests/cpp.

Run as ``python -m world_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = ROOT / "_build" / "world_amd"
LIB = PKG / "libworld_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Reproducible FP32: no FMA contraction, IEEE division and square root, no
# packed-FP32 instructions (-packed-fp32-ops removes the target feature).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
             "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]

HOST_SOURCES = ["sim_config.cpp", "lattice.cpp", "capi_util.cpp"]
HIP_SOURCES = ["engine.hip"]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library() -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        out = OBJ / (src + ".o")
        _run([CXX, *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(out)])
        objs.append(str(out))
    for src in HIP_SOURCES:
        out = OBJ / (src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(out)])
        objs.append(str(out))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    return LIB


def build_cpp_tests() -> Path:
    """The reference's GoogleTest cases re-expressed against the C++ façade."""
    src = ROOT / "tests" / "cpp" / "particle_dynamics_test.cpp"
    out = ROOT / "tests" / "cpp" / "particle_dynamics_test"
    _run([CXX, "-O2", "-std=c++17", "-I", str(ROOT / "include"), str(src), "-o", str(out),
          "-L", str(PKG), "-lworld_amd", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../world_amd"])
    return out


DRIVERS = ("dataset", "chaos", "model_problem", "stability_and_accuracy", "broadcast")


def build_drivers() -> list[Path]:
    """The batched counterparts of the reference's ensemble / dataset /
    model-problem drivers, linked against the library (world_amd/bin/)."""
    src_dir = PKG / "drivers"
    out_dir = PKG / "bin"
    out_dir.mkdir(exist_ok=True)
    outs = []
    for name in DRIVERS:
        out = out_dir / name
        _run([CXX, "-O2", "-std=c++17", "-I", str(ROOT / "include"), "-I", str(ROOT / "include" / "world"),
              "-I", str(src_dir), str(src_dir / f"{name}.cpp"), "-o", str(out), "-L", str(PKG),
              "-lworld_amd", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."])
        outs.append(out)
    return outs


def main() -> int:
    build_library()
    if (ROOT / "tests" / "cpp" / "particle_dynamics_test.cpp").exists():
        build_cpp_tests()
    build_drivers()
    return 0


if __name__ == "__main__":
    sys.exit(main())
