"""The C-ABI library's host code under AddressSanitizer, on the CPU.

Builds the host-instrumented library (rqutils_amd/asan_build.py) and runs
tests/asan_driver.py against it in a child process with clang's ASan runtime
preloaded: every entry point's argument validation, error messages and
host-side size arithmetic, and the failure path of the first HIP call on a
machine without a GPU."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def asan_lib():
    if not (os.path.exists("/opt/rocm/bin/hipcc") or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    from rqutils_amd import asan_build
    try:
        rt = asan_build.asan_runtime()
    except RuntimeError as e:
        pytest.skip(str(e))
    return asan_build.build_asan(), rt


def test_c_abi_host_paths_under_asan(asan_lib):
    lib, rt = asan_lib
    env = dict(os.environ, RQSQD_LIB=str(lib), LD_PRELOAD=str(rt),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               HIP_VISIBLE_DEVICES="", PYTHONPATH=str(ROOT))
    res = subprocess.run([sys.executable, str(ROOT / "tests" / "asan_driver.py")], env=env,
                         capture_output=True, text=True, timeout=600, cwd=str(ROOT))
    out = res.stdout + res.stderr
    assert "AddressSanitizer" not in res.stderr, out[-4000:]
    assert res.returncode == 0, out[-4000:]
    last = [ln for ln in res.stdout.splitlines() if ln.startswith("ASAN-DRIVER-OK")]
    assert last and int(last[0].split()[1]) > 60, out[-2000:]
