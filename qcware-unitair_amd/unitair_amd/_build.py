"""Build libunitair_amd.so (HIP, gfx950) in-tree.

Every ``csrc/*.hip`` file is compiled to an object with hipcc, in parallel,
then linked into ``unitair_amd/libunitair_amd.so``.  Run as
``python -m unitair_amd._build`` from ``qcware-unitair_amd/``, or call
:func:`build`.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
PROJECT_DIR = PKG_DIR.parent
CSRC_DIR = PROJECT_DIR / "csrc"
BUILD_DIR = PROJECT_DIR / "_build"
LIB_NAME = "libunitair_amd.so"
LIB_PATH = PKG_DIR / LIB_NAME

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMPILE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay disabled in every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mcode-object-version=5",
    # keep per-case register arrays in registers: otherwise SimplifyCFG merges
    # the arms of runtime switches over cube bits into dynamically indexed
    # stack memory (scratch)
    "-mllvm", "-simplifycfg-sink-common=false",
]


def _sources():
    return sorted(CSRC_DIR.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = BUILD_DIR / (src.stem + ".o")
    cmd = [HIPCC, *COMPILE_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(
            f"hipcc failed for {src.name}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def _stale(src: Path) -> bool:
    """The object is missing or older than its source or any csrc header."""
    obj = BUILD_DIR / (src.stem + ".o")
    if not obj.exists():
        return True
    newest = max([src.stat().st_mtime] + [h.stat().st_mtime for h in CSRC_DIR.glob("*.h")] +
                 [(PROJECT_DIR.parent / "include" / "unitair_amd.h").stat().st_mtime])
    return obj.stat().st_mtime < newest


def build(verbose: bool = False, changed_only: bool = False) -> Path:
    """Compile every HIP source (with changed_only, only those whose object is
    out of date) and link the shared library; returns its path."""
    BUILD_DIR.mkdir(exist_ok=True)
    srcs = _sources()
    objs_all = [BUILD_DIR / (src.stem + ".o") for src in srcs]
    if changed_only:
        srcs = [src for src in srcs if _stale(src)]
    workers = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        list(ex.map(_compile, srcs))
    objs = objs_all
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC",
           *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}", file=sys.stderr)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True, changed_only="--changed" in sys.argv)
