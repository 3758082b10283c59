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


def _compile(src: Path, flags=None, out_dir: Path = None) -> Path:
    obj = (out_dir or BUILD_DIR) / (src.stem + ".o")
    cmd = [HIPCC, *(flags or COMPILE_FLAGS), "-c", str(src), "-o", str(obj)]
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


def build(verbose: bool = False, changed_only: bool = False, asan: bool = False) -> Path:
    """Compile every HIP source (with changed_only, only those whose object is
    out of date) and link the shared library; returns its path.  asan: also
    the host-instrumented AddressSanitizer library (build_asan), compiled in
    the same worker pool."""
    BUILD_DIR.mkdir(exist_ok=True)
    srcs = _sources()
    objs_all = [BUILD_DIR / (src.stem + ".o") for src in srcs]
    jobs = [(src, None, BUILD_DIR) for src in srcs if not changed_only or _stale(src)]
    if asan:
        ASAN_BUILD_DIR.mkdir(exist_ok=True)
        jobs += [(src, ASAN_FLAGS, ASAN_BUILD_DIR) for src in srcs]
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        list(ex.map(lambda j: _compile(*j), jobs))
    _link(objs_all, LIB_PATH)
    if asan:
        _link([ASAN_BUILD_DIR / (src.stem + ".o") for src in srcs], ASAN_LIB_PATH,
              ["-Xarch_host", "-fsanitize=address", "-shared-libasan"])
    if verbose:
        print(f"built {LIB_PATH}" + (f" and {ASAN_LIB_PATH}" if asan else ""), file=sys.stderr)
    return LIB_PATH


def _link(objs, path: Path, extra=()):
    tmp = path.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *extra,
           *map(str, objs), "-lhiprtc", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, path)


ASAN_BUILD_DIR = PROJECT_DIR / "_build_asan"
ASAN_LIB_PATH = ASAN_BUILD_DIR / "libunitair_amd_asan.so"
# AddressSanitizer on the host code only (the C-ABI argument checks and
# launch set-up); the device code is built as in the release library
ASAN_FLAGS = COMPILE_FLAGS + ["-Xarch_host", "-fsanitize=address",
                              "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]


def asan_runtime() -> Path:
    """The clang AddressSanitizer runtime matching HIPCC (LD_PRELOAD it)."""
    llvm = Path(HIPCC).resolve().parent.parent / "lib" / "llvm" / "lib" / "clang"
    hits = sorted(llvm.glob("*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not hits:
        raise FileNotFoundError("libclang_rt.asan-x86_64.so not found under " + str(llvm))
    return hits[-1]


def build_asan(verbose: bool = False) -> Path:
    """Only the host-instrumented AddressSanitizer library, in _build_asan/
    (for the C-ABI validation tests, tests/test_abi_asan.py)."""
    ASAN_BUILD_DIR.mkdir(exist_ok=True)
    srcs = _sources()
    workers = max(1, min(len(srcs), 8, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(lambda src: _compile(src, ASAN_FLAGS, ASAN_BUILD_DIR), srcs))
    _link([ASAN_BUILD_DIR / (src.stem + ".o") for src in srcs], ASAN_LIB_PATH,
          ["-Xarch_host", "-fsanitize=address", "-shared-libasan"])
    if verbose:
        print(f"built {ASAN_LIB_PATH}", file=sys.stderr)
    return ASAN_LIB_PATH


if __name__ == "__main__":
    if "--asan" in sys.argv:
        build_asan(verbose=True)
    else:
        build(verbose=True, changed_only="--changed" in sys.argv)
