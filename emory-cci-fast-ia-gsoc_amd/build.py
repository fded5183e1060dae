"""Build script for libnscale_hip.so (gfx950) and the parity oracle.

Compiles every HIP source under csrc/ with hipcc for gfx950 and links the
shared library in-tree (nscale_hip/libnscale_hip.so) so that it travels with
the repository snapshot.  The oracle (test infrastructure) is built by
oracle/Makefile.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
BUILD_DIR = REPO / "_build"
LIB_PATH = PKG_DIR / "nscale_hip" / "libnscale_hip.so"

ARCH = os.environ.get("NSX_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed fp32 ops stay off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]
SOURCES = ["recon.hip", "mtree.hip", "bitplane.hip", "ccl.hip", "dist.hip", "api.hip", "batch.hip", "server.hip", "nsx_gen.cpp"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")


def _compile(src: str) -> Path:
    path = CSRC / src
    obj = BUILD_DIR / (path.stem + ".o")
    if src.endswith(".cpp"):
        cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-c", str(path), "-o", str(obj)]
    else:
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(path), "-o", str(obj)]
    _run(cmd)
    return obj


def build_library(jobs: int = 6) -> Path:
    """Compile csrc/ for gfx950 and link nscale_hip/libnscale_hip.so."""
    BUILD_DIR.mkdir(exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


DROPIN_CHECK = REPO / "tests" / "cpp" / "dropin_check"


def build_dropin_check() -> Path:
    """Compile tests/cpp/dropin_check (the C++ drop-in header's check program)
    against the in-tree library; found at run time through an $ORIGIN rpath."""
    src = REPO / "tests" / "cpp" / "dropin_check.cpp"
    rpath = "$ORIGIN/" + os.path.relpath(LIB_PATH.parent, DROPIN_CHECK.parent)
    _run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{REPO / 'include'}", f"-I{src.parent}", str(src),
          f"-L{LIB_PATH.parent}", "-lnscale_hip", f"-Wl,-rpath,{rpath}", "-o", str(DROPIN_CHECK)])
    return DROPIN_CHECK


def build_oracle() -> None:
    """Build oracle/liboracle.so (and oracle/_ref when the reference is present)."""
    if shutil.which("make") is None:
        raise RuntimeError("make not found")
    _run(["make", "-C", str(REPO / "oracle"), "all"])


if __name__ == "__main__":
    print(build_library())
    print(build_dropin_check())
    build_oracle()
