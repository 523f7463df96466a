"""Builds the gfx950 shared library ``libstardust_amd.so`` in-tree with hipcc,
then the C++ ``Sim`` shim over it (``libstardust_sim.so``, g++) and the C++
regression-check binary ``tests/cpp/sim_tests``.

Usage: ``python -m stardust_amd.build`` (or ``stardust_amd.build.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libstardust_amd.so")
SIMLIB = os.path.join(HERE, "libstardust_sim.so")
SIM_TESTS = os.path.join(HERE, "..", "tests", "cpp", "sim_tests")

SOURCES = ["sd_fft.hip", "sd_sort.hip", "sd_particles.hip", "sd_api.hip", "sd_mlp.hip", "sd_batch.hip"]
# The per-particle and API kernels restate the reference's arithmetic with
# explicit fmaf; contraction stays off there so results are reproducible.
NO_CONTRACT = {"sd_sort.hip", "sd_particles.hip", "sd_api.hip", "sd_batch.hip"}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _flags(src: str) -> list[str]:
    f = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(HERE, "..", "include"), "-I" + CSRC]
    if src in NO_CONTRACT:
        f.append("-ffp-contract=off")
    return f


def _compile(src: str) -> str:
    out = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [hipcc(), "-c", "-x", "hip", os.path.join(CSRC, src), "-o", out] + _flags(src)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp()
    if verbose:
        print(f"[stardust_amd] built {LIB}, {SIMLIB}")
    return LIB


def _gxx(args: list[str]) -> None:
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra"] + args, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{r.stdout}\n{r.stderr}")


def build_cpp() -> None:
    """Host-only C++: the Sim shim links the HIP library by file name and finds
    it beside itself ($ORIGIN); the test binary finds both in stardust_amd/."""
    _gxx(["-shared", "-fPIC", os.path.join(HERE, "cpp", "Sim.cpp"), "-o", SIMLIB + ".tmp",
          "-L" + HERE, "-lstardust_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(SIMLIB + ".tmp", SIMLIB)
    src = os.path.join(HERE, "..", "tests", "cpp", "sim_tests.cpp")
    if os.path.exists(src):
        _gxx([src, "-o", SIM_TESTS + ".tmp", "-L" + HERE, "-lstardust_sim", "-lstardust_amd",
              "-Wl,-rpath,$ORIGIN/../../stardust_amd"])
        os.replace(SIM_TESTS + ".tmp", SIM_TESTS)


if __name__ == "__main__":
    build()
    sys.exit(0)
