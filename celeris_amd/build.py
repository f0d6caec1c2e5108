"""Build recipes: the HIP library (product) and the oracle (test infrastructure).

    python -m celeris_amd.build            # library + oracle (+ reference build if present)

The library is compiled for gfx950 only.  Packed-FP32 instructions are disabled
(``-packed-fp32-ops``): every float op is a plain VALU op.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "celeris_amd", "csrc")
LIBDIR = os.path.join(ROOT, "celeris_amd", "_lib")
LIB = os.path.join(LIBDIR, "libceleris_rs.so")
ORACLE = os.path.join(ROOT, "oracle")
REFERENCE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-value",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
SOURCES = ["rs_kernels.hip", "rs_api.cpp", "rs_analysis.cpp", "rs_tolerance.cpp",
           "propagate_shim.cpp"]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [ln for ln in res.stdout.splitlines()
             if "packed-fp32-ops" not in ln and "hip-link" not in ln]
    if res.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")
    return lines


def build_library() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


DRIVER_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "_build", "dropin_driver")


def build_dropin_driver() -> str:
    """The C++ drop-in test caller: plain g++ against include/celeris/cuda/propagate.hpp,
    linked to the built library (found next to the binary's tree through its rpath)."""
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), DRIVER_SRC,
          "-L" + LIBDIR, "-lceleris_rs", "-Wl,-rpath,$ORIGIN/../../celeris_amd/_lib", "-pthread",
          "-o", DRIVER])
    return DRIVER


PROBE_SRC = os.path.join(CSRC, "issue_probe.hip")
PROBE = os.path.join(LIBDIR, "issue_probe")


def build_issue_probe() -> str:
    """The VALU issue-cost probe (shader-clock cycles per wave64 FP32 FMA, v_sin_f32 and
    FP64 FMA per SIMD), the measured side of the pair kernels' roofline."""
    os.makedirs(LIBDIR, exist_ok=True)
    flags = [f for f in HIP_FLAGS if f not in ("-fPIC", "-shared")]
    _run([HIPCC, *flags, PROBE_SRC, "-o", PROBE])
    return PROBE


REREAD_SRC = os.path.join(CSRC, "reread_probe.hip")
REREAD = os.path.join(LIBDIR, "reread_probe")


def build_reread_probe() -> str:
    """The work-layout probe: the plain FP32 pair sum with each work-group summing every
    pillar (the reference's layout) against pillar chunks split over work-groups."""
    os.makedirs(LIBDIR, exist_ok=True)
    flags = [f for f in HIP_FLAGS if f not in ("-fPIC", "-shared")]
    _run([HIPCC, *flags, REREAD_SRC, "-o", REREAD])
    return REREAD


def build_oracle() -> str:
    _run(["make", "-s", "-C", ORACLE])
    return os.path.join(ORACLE, "_build", "liboracle_rs.so")


def build_reference() -> bool:
    """Compile the reference's CPU path into oracle/_ref (only where its sources are)."""
    if not os.path.isdir(os.path.join(REFERENCE, "src", "celeris")):
        return False
    if shutil.which("g++") is None:
        return False
    _run(["make", "-s", "-C", ORACLE, "ref"])
    return True


def build_all() -> None:
    build_library()
    build_issue_probe()
    build_reread_probe()
    build_dropin_driver()
    build_oracle()
    build_reference()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
