"""Build the engine's native libraries in-tree.

libgkmhip.so   the HIP kernels and their C-ABI (hipcc, gfx950): the pairwise,
               explain and ISM kernels and the l-mer index (gkm_kernels.hip),
               the exact l-mer tables and DeltaSVM (gkm_tables.hip)
libgkm_emu.so  a CPU build of the same lane code, used by the CPU test suite
               to check the bit-sliced algorithm without a GPU
libgkm_smo.so  the SMO training solver (host C, kernel columns by callback)

Run: python gkmsvm-lite_amd/gkmsvm_amd/build_ext.py
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GKM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{CSRC}", f"-I{INCLUDE}",
]

HIP_SOURCES = ["gkm_kernels.hip", "gkm_tables.hip", "gkm_peaks.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ...")


def build_hip() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libgkmhip.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *(os.path.join(CSRC, f) for f in HIP_SOURCES)])
    os.replace(tmp, out)
    return out


def build_emu() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libgkm_emu.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{CSRC}", "-o", tmp,
          os.path.join(CSRC, "gkm_emu.cpp")])
    os.replace(tmp, out)
    return out


def build_smo() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libgkm_smo.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-o", tmp,
          os.path.join(CSRC, "gkm_smo.c"), "-lm"])
    os.replace(tmp, out)
    return out


def build_all() -> list[str]:
    return [build_hip(), build_emu(), build_smo()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
