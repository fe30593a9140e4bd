"""Build the device-API test library (TEST INFRASTRUCTURE).

    python tests/devapi/build_devtest.py  ->  tests/devapi/_build/libppl_devtest.so

The kernel in k_devapi.hip uses only include/ppl/ppl_device.hpp; it is
compiled with the same flags as the product library.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
OUT = HERE / "_build" / "libppl_devtest.so"


def build(verbose: bool = True) -> Path:
    from prismplan_amd.build import ARCH, flags, hipcc
    OUT.parent.mkdir(exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *flags(), "-shared", str(HERE / "k_devapi.hip"), "-o", str(tmp)]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
