"""Builds the native library in-tree: svfsiplus_amd/libsvgpu.so (gfx950).

    python -m svfsiplus_amd.build            # library only
    python -m svfsiplus_amd.build --oracle   # also the CPU oracle (test infrastructure)

hipcc cross-compiles for gfx950 without a GPU present. Every HIP source is
built with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libsvgpu.so"

HIP_SOURCES = ["svg_runtime.hip", "svg_fsils.hip", "svg_assembly.hip", "svg_lpc.hip", "svg_setup.hip",
               "svg_coarse.hip"]
CPP_SOURCES = ["host_setup.cpp"]
HOST_DIR = PKG / "host"
HOST_SOURCES = ["svg_fsils.cpp"]            # C++ host API (into the library)
HOST_PROGRAMS = {"svg_newton_step": "newton_step.cpp"}  # C++ drivers (next to the library)
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

ARCH = os.environ.get("SVG_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, cwd=str(ROOT), capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")


def build_library(verbose: bool = False) -> Path:
    hipcc = _hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
              f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    hip_flags = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    objs = []
    for src in HIP_SOURCES:
        obj = objdir / (src + ".o")
        cmd = [hipcc, *common, *hip_flags, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    for src in CPP_SOURCES:
        obj = objdir / (src + ".o")
        cmd = ["g++", *common, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    host_flags = ["-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", f"-I{HOST_DIR}"]
    for src in HOST_SOURCES:
        obj = objdir / ("host_" + src + ".o")
        cmd = ["g++", *common, *host_flags, "-c", str(HOST_DIR / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs]
    _run(cmd)
    os.replace(tmp, LIB)
    for exe, src in HOST_PROGRAMS.items():
        out = PKG / exe
        cmd = ["g++", *[c for c in common if c != "-fPIC"], *host_flags, str(HOST_DIR / src), "-o", str(out) + ".tmp",
               f"-L{PKG}", "-lsvgpu", f"-L{ROCM / 'lib'}", "-lamdhip64", "-Wl,-rpath,$ORIGIN",
               f"-Wl,-rpath,{ROCM / 'lib'}"]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        os.replace(str(out) + ".tmp", out)
    return LIB


def build_oracle() -> Path:
    oracle = ROOT / "oracle"
    _run(["make", "-s", "-B", "-C", str(oracle)])  # always rebuilt (a few seconds)
    return oracle / "liboracle.so"


def build_reference_harness() -> Path | None:
    """Compiles oracle/_ref/libsvref.so from the reference sources (only
    where /root/reference exists; a machine without the sources uses the
    prebuilt file)."""
    if not Path("/root/reference/Code/Source/solver/fluid.cpp").exists():
        return None
    _run(["make", "-s", "-j8", "-f", str(ROOT / "oracle" / "ref.mk")])
    return ROOT / "oracle" / "_ref" / "libsvref.so"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build_library(a.verbose))
    if a.oracle:
        print(build_oracle())
        print(build_reference_harness())


if __name__ == "__main__":
    main()
