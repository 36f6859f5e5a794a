"""Builds the native libraries in-tree.

* ``platoanalyze_amd/lib/libplato_mi355x.so`` -- the gfx950 HIP kernels and the C ABI
  declared in ``include/plato_mi355x.h`` (hipcc, --offload-arch=gfx950).
* ``oracle/_build/liboracle.so`` -- the CPU restatement used only by the tests and the
  CPU baseline (g++ -O3 -fopenmp).

Run ``python -m platoanalyze_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "platoanalyze_amd"
CSRC = PKG / "csrc"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libplato_mi355x.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for all device code (see the kernels' notes).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-march=x86-64-v2"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")
    if verbose and (r.stdout or r.stderr):
        sys.stderr.write(r.stdout + r.stderr)


def _jobs(cmds: list[list[str]], verbose: bool) -> None:
    procs = []
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        procs.append((c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    err = None
    for c, p in procs:
        out, e = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + e)
            err = err or f"command failed ({p.returncode}): {' '.join(c)}"
    if err:
        raise RuntimeError(err)


def build_native(verbose: bool = False) -> Path:
    obj = PKG / "_build"
    obj.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    inc = ["-I", str(ROOT / "include"), "-I", str(CSRC)]
    hip_srcs = ["pa_kernels.hip", "pa_solver.hip", "pa_mf.hip", "pa_mg.hip", "pa_box.hip", "pa_capi.hip", "pa_unet.hip",
                "pa_deflate.hip"]
    cmds = [[HIPCC, *HIP_FLAGS, *inc, "-x", "hip", "-c", str(CSRC / s), "-o", str(obj / (s + ".o"))]
            for s in hip_srcs]
    cmds.append(["g++", *HOST_FLAGS, *inc, "-c", str(CSRC / "pa_host.cpp"), "-o", str(obj / "pa_host.o")])
    _jobs(cmds, verbose)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp),
          *[str(obj / (s + ".o")) for s in hip_srcs], str(obj / "pa_host.o"), "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


# CPU-baseline builds of the oracle for newer instruction sets (AVX2+FMA, AVX-512); the
# timing leg loads the best one the host supports (oracle.select_isa), the tests the
# baseline x86-64-v2 build.
ORACLE_ISAS = ("x86-64-v3", "x86-64-v4")


def oracle_isa_lib(isa: str) -> Path:
    return ORACLE_LIB.parent / f"liboracle_{isa}.so"


def build_oracle(verbose: bool = False) -> Path:
    out = ORACLE_LIB.parent
    out.mkdir(exist_ok=True)
    targets = [("x86-64-v2", ORACLE_LIB)] + [(isa, oracle_isa_lib(isa)) for isa in ORACLE_ISAS]
    _jobs([["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", f"-march={isa}",
            "-o", str(lib.with_suffix(".so.tmp")), str(ORACLE_DIR / "oracle.cpp")] for isa, lib in targets], verbose)
    for _, lib in targets:
        os.replace(lib.with_suffix(".so.tmp"), lib)
    return ORACLE_LIB


def build_all(verbose: bool = False) -> None:
    build_native(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
