"""Build of the gfx950 shared library (HIP kernels + C ABI) and the acouav CLI.

Every build compiles from source (hipcc cross-compiles gfx950 without a GPU).
Outputs land in acogpu_amd/_build/ (git-ignored, shipped with the tree).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libacogpu_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", os.path.join(ROCM, "bin", "hipcc"))

# -ffp-contract=off: fp32 expressions evaluate one IEEE operation at a time
# (bit-exact replay against the CPU oracle).  -packed-fp32-ops: no packed fp32
# instructions are generated.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-command-line-argument",
]

LIB_SOURCES = ["aco_kernels.hip", "aco_capi.hip", "aco_batch.hip", "aco_prior.hip"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> str:
    os.makedirs(OUT, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


APP_SOURCES = ["colony/Colony.cpp", "colony/AntSystem.cpp", "colony/RankBasedAntSystem.cpp",
               "app/Commandline.cpp", "app/TSPReader.cpp", "app/Writer.cpp", "app/TSP.cpp", "app/main.cpp"]
# program names of the reference's build variants (release/makefile:52-59)
APP_VARIANTS = ["acouav_gpu_aging_broker", "acouav_gpu_aging_nobroker", "acouav_gpu_noaging_broker",
                "acouav_gpu_noaging_nobroker"]


def build_app() -> str:
    """acouav: the C++ Colony classes and command line over the C ABI."""
    exe = os.path.join(OUT, "acouav")
    srcs = [os.path.join(CSRC, s) for s in APP_SOURCES]
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(CSRC, "colony"), "-I", os.path.join(CSRC, "app"), "-I", os.path.join(ROCM, "include"),
          *srcs, "-L", OUT, "-lacogpu_amd", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-o", exe + ".tmp"])
    os.replace(exe + ".tmp", exe)
    for v in APP_VARIANTS:
        link = os.path.join(OUT, v)
        if os.path.lexists(link):
            os.remove(link)
        os.symlink("acouav", link)
    return exe


def build_all() -> None:
    build_lib()
    build_app()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
