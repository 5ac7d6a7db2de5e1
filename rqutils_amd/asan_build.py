"""Host-side AddressSanitizer build of the C-ABI library (a CPU check).

Every ``csrc/*.hip`` file is compiled for gfx950 as ``build.py`` does, with
only the host pass instrumented (``-Xarch_host -fsanitize=address``; the
device code is plain and never runs from this build), and linked by clang++
against clang's shared sanitizer runtime into ``_asan/librqsqd_asan.so``.
Python must start with that runtime preloaded (``asan_runtime()``), so the
checks run in a child process: ``tests/test_asan_host.py`` drives every
entry point's argument validation, error formatting and host-side size
arithmetic through it on a machine without a GPU.

    python -m rqutils_amd.asan_build      # builds and prints the two paths
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

from . import build as _build

OUT = _build.PKG / "_asan"
LIB = OUT / "librqsqd_asan.so"

_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]


def asan_runtime() -> Path:
    """clang's shared AddressSanitizer runtime of the ROCm toolchain."""
    cands = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not cands:
        raise RuntimeError("clang's libclang_rt.asan-x86_64.so not found under /opt/rocm/llvm")
    return Path(cands[-1])


def _clangxx() -> str:
    path = "/opt/rocm/llvm/bin/clang++"
    if not os.path.exists(path):
        raise RuntimeError("clang++ of the ROCm toolchain not found")
    return path


def _compile(src: Path) -> Path:
    obj = OUT / (src.stem + ".o")
    flags = [f for f in _build.HIPCC_FLAGS if f != "-O3"]
    cmd = [_build.hipcc(), *flags, "-O1", "-g", *_SAN, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc (host ASan) failed on {src.name}:\n{res.stderr}")
    return obj


def build_asan(force: bool = False) -> Path:
    """Builds ``_asan/librqsqd_asan.so`` unless it is newer than every source."""
    sources = sorted(_build.CSRC.glob("*.hip"))
    deps = sources + sorted(_build.CSRC.glob("*.hpp")) + sorted(
        (_build.PKG.parent / "include").glob("*.h"))
    if (not force and LIB.exists()
            and LIB.stat().st_mtime > max(p.stat().st_mtime for p in deps)):
        return LIB
    OUT.mkdir(exist_ok=True)
    workers = max(1, min(len(sources), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_clangxx(), "-shared", "-fPIC", "-fsanitize=address", "-shared-libsan",
           *map(str, objs), "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
           "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link of {LIB.name} failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_asan(force="--force" in sys.argv))
    print(asan_runtime())
