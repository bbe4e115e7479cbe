"""Builds libdemon_gpu.so (the HIP/C-ABI product) in-tree for gfx950.

Usage: python demon_simulation_code_amd/build.py [--verbose]
(run by path: importing the package would load the library being rebuilt)

The library is compiled with hipcc directly (no torch extension machinery):
every .hip file under csrc/ is one translation unit of a single shared
object.  Flags:
  --offload-arch=gfx950                 MI355X only
  -ffp-contract=off                     no implicit FMA: the exact pair path
                                        keeps the reference's unfused rounding;
                                        the prefilter uses explicit fma()
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                        no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
LIB_NAME = "libdemon_gpu.so"
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


HOST = PKG / "host"
HOST_LIB = "libdemon_host.so"
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]


def cxx() -> str:
    return os.environ.get("CXX") or shutil.which("g++") or "g++"


def build_host(verbose: bool = False) -> list[Path]:
    """The host C++ mirror (libdemon_host.so), the `demon` driver and the
    `demon_fft` analysis tool, linked against libdemon_gpu.so with an $ORIGIN rpath."""
    inc = [f"-I{INCLUDE}", f"-I{HOST}"]
    rpath = "-Wl,-rpath,$ORIGIN"
    lib = PKG / HOST_LIB
    steps = [
        [cxx(), *CXX_FLAGS, *inc, "-shared", "-o", str(lib), str(HOST / "demon_host.cpp"),
         str(HOST / "demon_fits.cpp"), f"-L{PKG}", "-ldemon_gpu", rpath],
        [cxx(), *CXX_FLAGS, *inc, "-o", str(PKG / "demon"), str(HOST / "driver.cpp"),
         f"-L{PKG}", "-ldemon_host", "-ldemon_gpu", rpath],
        [cxx(), *CXX_FLAGS, *inc, "-o", str(PKG / "demon_fft"), str(HOST / "fft_analysis.cpp"),
         f"-L{PKG}", "-ldemon_host", "-ldemon_gpu", rpath],
    ]
    test_src = PKG.parent / "tests" / "host" / "test_host.cpp"
    if test_src.exists():  # the host mirror's test driver (tests/test_host_mirror.py)
        steps.append([cxx(), *CXX_FLAGS, *inc, "-o", str(test_src.with_suffix("")), str(test_src),
                      f"-L{PKG}", "-ldemon_host", "-ldemon_gpu", f"-Wl,-rpath,{PKG}"])
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed:\n{r.stdout}")
    return [lib, PKG / "demon", PKG / "demon_fft"]


def build(verbose: bool = False, jobs: int = 4) -> Path:
    """Compiles every csrc/*.hip, links PKG/libdemon_gpu.so, then the host
    mirror and driver.  Always rebuilds."""
    out = build_hip(verbose, jobs)
    build_host(verbose)
    return out


def build_hip(verbose: bool = False, jobs: int = 4) -> Path:
    out = PKG / LIB_NAME
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in sources():
        obj = objdir / (src.stem + ".o")
        objs.append(obj)
        cmd = [cc, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _drain(procs, verbose)
    _drain(procs, verbose)
    tmp = out.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}")
    os.replace(tmp, out)
    return out


def _drain(procs, verbose):
    while procs:
        src, p = procs.pop(0)
        text = p.communicate()[0].decode(errors="replace")
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{text}")
        if verbose and text.strip():
            print(text, flush=True)


if __name__ == "__main__":
    path = build(verbose="--verbose" in sys.argv)
    print(path)
