"""Drives the C-ABI library's host code under AddressSanitizer (no GPU needed).

Run by test_asan_host.py in a child process started with clang's ASan
runtime preloaded and RQSQD_LIB naming the host-instrumented build. Every
entry point declared in include/sqdk.h is called with NULL / zero arguments,
then with host structures that reach its deeper validation (unknown enums,
sizes out of range, workspaces too small, mismatched operators), the
host-side size arithmetic, and calls that pass validation so that the first
HIP call fails on a machine without a GPU (the error path that formats HIP's
message). A status other than OK must leave a non-empty message. Prints one
line per call and ``ASAN-DRIVER-OK <calls>`` at the end; an AddressSanitizer
report aborts the process before that line.
"""
from __future__ import annotations

import ctypes as C
import sys
import threading

from rqutils_amd import _lib

L = _lib.lib()
CALLS = 0


def _zero(t):
    if t in (C.c_double, C.c_float):
        return 0.0
    if t in (C.c_int32, C.c_int64, C.c_size_t, C.c_uint64, C.c_int):
        return 0
    return None        # c_void_p, POINTER(...), function pointers


def call(name, *args, expect=None):
    """Calls ``name``; checks the status against ``expect`` (a set) and that a
    failure left a message."""
    global CALLS
    rc = getattr(L, name)(*args)
    CALLS += 1
    msg = L.sqdk_last_error()
    msg = msg.decode(errors="replace") if msg else ""
    print(f"{name}: rc={rc} {msg[:120] if rc != _lib.OK else ''!r}", flush=True)
    if rc != _lib.OK and not msg:
        raise AssertionError(f"{name} failed with {rc} and no message")
    if expect is not None and rc not in expect:
        raise AssertionError(f"{name}: status {rc}, expected one of {sorted(expect)} ({msg})")
    return rc, msg


BAD = {_lib.EINVAL, _lib.ERANGE}
HIP_OR_BAD = BAD | {_lib.EHIP}


def all_null():
    for name, (_, argtypes) in _lib.SIGNATURES.items():
        if name in ("sqdk_abi_version", "sqdk_last_error"):
            continue
        # an empty input may be a valid no-op (OK); anything else must be refused
        call(name, *[_zero(t) for t in argtypes], expect=BAD | {_lib.OK})


def workspaces():
    nb = C.c_size_t(0)
    call("sqdk_gemm_bf16_workspace", 100, 200, 3, C.byref(nb), expect={_lib.OK})
    assert nb.value == 4 * 3 * 100 * 200
    call("sqdk_gemm_bf16_workspace", -1, 200, 3, C.byref(nb), expect=BAD)
    call("sqdk_dots_workspace", 1000, 4, C.byref(nb), expect={_lib.OK})
    call("sqdk_dots_workspace", 1000, 17, C.byref(nb), expect=BAD)
    call("sqdk_intake_unique_workspace", 1 << 20, 28, C.byref(nb), expect={_lib.OK})
    assert nb.value > 0
    call("sqdk_intake_unique_workspace", 1 << 20, 65, C.byref(nb), expect=BAD)
    call("sqdk_intake_unique_workspace", 1 << 40, 28, C.byref(nb), expect=BAD)
    space = _lib.SqdkSpace(n=1 << 20, num_qubits=28, bucket_bits=16, fine_bits=4,
                           chunk_rows=2048, n_items=512, max_bucket=64)
    tune = _lib.SqdkTune(rows_per_thread=4, group_split=3)
    call("sqdk_matvec_workspace", C.byref(space), C.byref(tune), _lib.C64, C.byref(nb),
         expect={_lib.OK})
    call("sqdk_matvec_workspace", C.byref(space), C.byref(tune), 7, C.byref(nb), expect=BAD)
    ns = C.c_int64(0)
    call("sqdk_cache_slices", C.byref(space), C.byref(ns), expect={_lib.OK})
    assert ns.value == 512 * 2048 // 64
    space.chunk_rows = 1000
    call("sqdk_cache_slices", C.byref(space), C.byref(ns), expect=BAD)
    batch = _lib.SqdkBatch(n=1 << 16, n_inst=4, n_items=64, chunk_rows=4096)
    call("sqdk_batch_cache_slices", C.byref(batch), C.byref(ns), expect={_lib.OK})
    call("sqdk_batch_matvec_workspace", C.byref(batch), C.byref(tune), _lib.C128, C.byref(nb),
         expect={_lib.OK})
    assert nb.value == 2 * (1 << 16) * 16
    opts = _lib.SqdkLocgOpts(mode=_lib.LOCG_FUSED, precision=_lib.C128)
    call("sqdk_batch_locg_workspace", C.byref(batch), C.byref(tune), C.byref(opts), C.byref(nb),
         expect={_lib.OK})
    opts.precision = 9
    call("sqdk_batch_locg_workspace", C.byref(batch), C.byref(tune), C.byref(opts), C.byref(nb),
         expect=BAD)


def _noop(ctx, vin, vout, stream):
    return 0


def locg_validation():
    fn = _lib.MATVEC_FN(_noop)
    op = _lib.SqdkOp(fn=fn)
    n = 1000
    opts = _lib.SqdkLocgOpts(mode=_lib.LOCG_FUSED, precision=_lib.C128,
                             criterion=_lib.CRIT_RELATIVE, maxiter=10, tol=1e-8, check_every=4,
                             refresh_every=16)
    nb = C.c_size_t(0)
    call("sqdk_locg_workspace", C.byref(op), n, C.byref(opts), C.byref(nb), expect={_lib.OK})
    buf = (C.c_uint8 * max(nb.value, 4096))()
    x = (C.c_double * (2 * n))()
    res = _lib.SqdkLocgResult()

    def solve(o, work_bytes=None, size=n, operator=op):
        return call("sqdk_locg_solve", C.byref(operator), size, C.cast(x, C.c_void_p),
                    C.byref(o), C.cast(buf, C.c_void_p),
                    nb.value if work_bytes is None else work_bytes, C.byref(res), None,
                    expect=BAD if work_bytes is not None or o is not opts else HIP_OR_BAD)

    for field, value in (("mode", 5), ("maxiter", -1), ("criterion", 3), ("precond", 2),
                         ("precond", 1)):
        o = _lib.SqdkLocgOpts.from_buffer_copy(opts)
        setattr(o, field, value)
        solve(o)
    o = _lib.SqdkLocgOpts.from_buffer_copy(opts)
    o.debug = C.cast(buf, C.c_void_p)            # diagnostics need the reference mode
    solve(o)
    o = _lib.SqdkLocgOpts.from_buffer_copy(opts)
    o.mode = _lib.LOCG_REFERENCE
    o.debug_vectors = C.cast(buf, C.c_void_p)    # vectors without the diagnostics buffer
    solve(o)
    _, msg = solve(opts, work_bytes=1)           # workspace too small: the %zu message
    assert "needed" in msg, msg
    # a built-in operator whose size differs from the vector's
    space = _lib.SqdkSpace(n=n + 1, num_qubits=12, chunk_rows=1024)
    ham = _lib.SqdkHam(n_groups=1, n_terms=1, num_qubits=12)
    op2 = _lib.SqdkOp(ham=C.pointer(ham), space=C.pointer(space))
    _, msg = solve(opts, work_bytes=0, operator=op2)
    assert "size" in msg, msg
    # everything valid: the solve starts and its first HIP call fails without a GPU
    solve(opts)


def gemm_messages():
    big = (1 << 62) + 12345
    _, msg = call("sqdk_gemm_bf16", 0, big, big, None, 1, 0, None, 1, 0, None, 1, None, 1,
                  None, 0, None, expect=BAD)
    buf = (C.c_uint8 * 256)()
    p = C.cast(buf, C.c_void_p)
    _, msg = call("sqdk_gemm_bf16", 0, big, big, p, 1, 0, p, 1, 0, p, 1, None, 1, None, 0, None,
                  expect=BAD)
    assert str(big) in msg, msg
    call("sqdk_gemm_bf16", 16, 16, 16, p, 16, 0, p, 16, 0, p, 16, None, 9000, None, 0, None,
         expect=BAD)
    call("sqdk_gemm_bf16", 16, 16, 16, p, 16, 0, p, 16, 0, p, 16, None, 4, None, 0, None,
         expect=BAD)


def hip_failures():
    """Calls that pass validation: without a GPU the first HIP call fails and
    the library formats HIP's message (no kernel touches these host buffers
    where a GPU is absent; the driver only runs where none is)."""
    buf = (C.c_float * 64)()
    p = C.cast(buf, C.c_void_p)
    call("sqdk_colsum", 4, 4, p, p, None, expect=HIP_OR_BAD)
    call("sqdk_adam", 4, p, p, p, p, None, 1e-3, 0.9, 0.999, 1e-8, 1, None, expect=HIP_OR_BAD)
    a = (C.c_void_p * 1)(p)
    call("sqdk_dots_noalloc", 4, _lib.C64, 1, a, a, p, None, expect=HIP_OR_BAD)
    call("sqdk_dots_noalloc", 4, _lib.C64, 17, a, a, p, None, expect=BAD)


def threads():
    """Each thread keeps its own last message."""
    errs = []

    def work(k):
        try:
            for _ in range(50):
                rc = L.sqdk_dots_workspace(10, 100 + k, C.byref(C.c_size_t(0)))
                msg = L.sqdk_last_error().decode()
                if rc == _lib.OK or f"npairs {100 + k}" not in msg:
                    errs.append((k, rc, msg))
        except Exception as e:     # noqa: BLE001
            errs.append((k, repr(e)))

    ts = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs[:3]


def main():
    import torch
    if torch.cuda.is_available():
        raise SystemExit("the driver passes host buffers where device pointers belong: "
                         "run it only where no GPU is visible")
    all_null()
    workspaces()
    locg_validation()
    gemm_messages()
    hip_failures()
    threads()
    print(f"ASAN-DRIVER-OK {CALLS}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
