"""Build libppl.so (HIP, gfx950) in-tree.

    python -m prismplan_amd.build            # build (always recompiles)
    python -m prismplan_amd.build --check    # print the library path if present

The library lands in prismplan_amd/_lib/libppl.so. Packed FP32 instructions
are disabled for device code (-packed-fp32-ops); hipcc's host pass warns that
it does not know the feature, which is harmless.
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
LIB_DIR = PKG / "_lib"
LIB = LIB_DIR / "libppl.so"
SOURCES = ["kernels.hip", "capi.hip", "rrtc.hip"]
ARCH = os.environ.get("PPL_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def flags() -> list:
    return ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-mcode-object-version=5", "-I", str(ROOT / "include"), "-I", str(CSRC),
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def build_prrtc(verbose: bool = True) -> Path:
    """The pybind11 module (host C++ over the C ABI), linked to libppl.so."""
    import sysconfig
    import pybind11
    out = LIB_DIR / ("_prrtc" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = out.with_suffix(".tmp")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I", str(ROOT / "include"), "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
           str(CSRC / "prrtc_module.cpp"), "-L", str(LIB_DIR), "-lppl", "-Wl,-rpath,$ORIGIN",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(tmp)]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build(verbose: bool = True) -> Path:
    LIB_DIR.mkdir(exist_ok=True)
    objs = []
    cc = hipcc()
    for src in SOURCES:
        obj = LIB_DIR / (Path(src).stem + ".o")
        cmd = [cc, *flags(), "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print("[build]", " ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-mcode-object-version=5", "-shared", "-fPIC", "-o", str(tmp), *objs]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        Path(o).unlink(missing_ok=True)
    build_prrtc(verbose)
    return LIB


if __name__ == "__main__":
    if "--check" in sys.argv:
        print(LIB if LIB.exists() else "missing")
        sys.exit(0 if LIB.exists() else 1)
    print(build())
