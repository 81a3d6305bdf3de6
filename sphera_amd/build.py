"""Build recipes for the device library (gfx950) and the test-side CPU checkers.

    python -m sphera_amd.build          # device library + oracle (+ reference build
                                        # when /root/reference is mounted)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "sphera_amd" / "csrc"
OUT = ROOT / "sphera_amd" / "_build"
LIB = OUT / "libsphera_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the FP64 kernels reproduce the reference's separately
# rounded products and sums.  Packed FP32 instructions are disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result", "-Wno-unused-value"]


def _run(cmd, cwd=ROOT):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout[-4000:] + res.stderr[-8000:])
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd))}")
    return res


def build_device(extra=()):
    OUT.mkdir(parents=True, exist_ok=True)
    tmp = OUT / "libsphera_hip.so.tmp"
    _run([HIPCC, *HIP_FLAGS, *extra, "-I", str(ROOT / "include"), str(CSRC / "sph_api.hip"),
          "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    _run(["make", "-s", "-f", "oracle/Makefile"])
    return ROOT / "oracle" / "_build" / "libsph_oracle.so"


def build_reference():
    """flang builds of the reference (3D and 2D) from /root/reference (only
    where it is mounted)."""
    if not Path("/root/reference/src").is_dir():
        return None
    if not shutil.which("/opt/rocm/lib/llvm/bin/flang"):
        return None
    _run(["make", "-s", "-f", "oracle/Makefile.ref", "-j8"])
    _run(["make", "-s", "-f", "oracle/Makefile.ref", "-j8", "DIM=2"])
    return ROOT / "oracle" / "_ref" / "harness3d.x"


def build_all():
    build_device()
    build_oracle()
    build_reference()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
