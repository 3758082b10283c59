"""The C-ABI argument checks under AddressSanitizer: tests/test_abi.py runs in
a child process against the build of the library whose host code is
instrumented (unitair_amd._build.build_asan), with the matching ASan runtime
preloaded.  Skipped when that build is absent (__graft_entry__.build() makes
it)."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_abi_checks_under_address_sanitizer():
    from unitair_amd import _build
    lib = _build.ASAN_LIB_PATH
    if not lib.exists():
        pytest.skip("no ASan build (python -m unitair_amd._build --asan)")
    env = dict(os.environ, LD_PRELOAD=str(_build.asan_runtime()),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UNITAIR_AMD_LIB=str(lib))
    res = subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests" / "test_abi.py"),
                          "-q", "-p", "no:cacheprovider", "-m", "not gpu"],
                         env=env, cwd=str(ROOT), capture_output=True, text=True, timeout=900)
    out = res.stdout[-4000:] + res.stderr[-4000:]
    assert res.returncode == 0, out
    assert "AddressSanitizer" not in res.stderr, out
    assert " passed" in res.stdout, out
