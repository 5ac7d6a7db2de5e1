"""ctypes binding of the C-ABI kernel library ``librqsqd.so`` (include/sqdk.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (both carry the soname ``libamdhip64.so.7``). Importing this
module never compiles anything; a missing library raises ``ImportError`` with
the build command, so no code path can silently run without the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch  # noqa: F401  (binds the process to torch's HIP runtime first)

ABI_VERSION = 20

C128, C64 = 0, 1
LOCG_REFERENCE, LOCG_FUSED = 0, 1
CRIT_REFERENCE, CRIT_RELATIVE = 0, 1

OK, EINVAL, ERANGE, EHIP, ENOSPC, ECALLBACK = 0, -1, -2, -3, -4, -5


class SqdkSpace(C.Structure):
    _fields_ = [
        ("keys", C.c_void_p), ("bstart", C.c_void_p), ("fdir", C.c_void_p),
        ("items", C.c_void_p), ("pc_mask", C.c_void_p),
        ("n", C.c_int64), ("num_qubits", C.c_int32), ("bucket_bits", C.c_int32),
        ("fine_bits", C.c_int32), ("chunk_rows", C.c_int32), ("n_items", C.c_int32),
        ("max_bucket", C.c_int32), ("fbits", C.c_void_p),
    ]


class SqdkHam(C.Structure):
    _fields_ = [
        ("xmask", C.c_void_p), ("grp_off", C.c_void_p), ("zmask", C.c_void_p),
        ("coef", C.c_void_p), ("coef32", C.c_void_p), ("class_hi", C.c_void_p),
        ("class_off", C.c_void_p),
        ("n_groups", C.c_int32), ("n_terms", C.c_int32), ("n_classes", C.c_int32),
        ("diag_t0", C.c_int32), ("diag_t1", C.c_int32), ("num_qubits", C.c_int32),
        ("bucket_bits", C.c_int32), ("reserved", C.c_int32),
    ]


class SqdkGate(C.Structure):
    _fields_ = [("x", C.c_uint64), ("z", C.c_uint64), ("cos", C.c_double), ("sin", C.c_double)]


class SqdkInst(C.Structure):
    _fields_ = [("bstart_base", C.c_int64), ("fdir_base", C.c_int64), ("class_base", C.c_int32),
                ("n_classes", C.c_int32), ("diag_t0", C.c_int32), ("diag_t1", C.c_int32),
                ("bucket_bits", C.c_int32), ("fine_bits", C.c_int32), ("fbits_base", C.c_int64)]


class SqdkCache(C.Structure):
    _fields_ = [("n", C.c_int64), ("n_slices", C.c_int64), ("slot_row", C.c_void_p),
                ("slice_base", C.c_void_p), ("col", C.c_void_p), ("coef", C.c_void_p),
                ("diag", C.c_void_p), ("row_len", C.c_void_p), ("precision", C.c_int32),
                ("slices_per_wave", C.c_int32)]


class SqdkBatch(C.Structure):
    _fields_ = [
        ("keys", C.c_void_p), ("bstart", C.c_void_p), ("fdir", C.c_void_p), ("items", C.c_void_p),
        ("item_inst", C.c_void_p), ("item_base", C.c_void_p), ("row_base", C.c_void_p),
        ("xmask", C.c_void_p), ("grp_off", C.c_void_p), ("zmask", C.c_void_p),
        ("coef", C.c_void_p), ("coef32", C.c_void_p), ("class_hi", C.c_void_p),
        ("class_off", C.c_void_p), ("inst", C.c_void_p),
        ("n", C.c_int64), ("n_inst", C.c_int32), ("n_items", C.c_int32),
        ("chunk_rows", C.c_int32), ("max_bucket", C.c_int32), ("max_fine_bits", C.c_int32),
        ("max_bucket_bits", C.c_int32), ("fbits", C.c_void_p), ("dense_fine_bits", C.c_int32),
        ("reserved", C.c_int32), ("cache", C.POINTER(SqdkCache)),
    ]


class SqdkTune(C.Structure):
    _fields_ = [("rows_per_thread", C.c_int32), ("group_split", C.c_int32),
                ("xcd_remap", C.c_int32), ("stage_v", C.c_int32), ("compact", C.c_int32)]


class SqdkLocgOpts(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("precision", C.c_int32), ("criterion", C.c_int32),
        ("maxiter", C.c_int32), ("tol", C.c_double), ("n_for_tol", C.c_int64),
        ("check_every", C.c_int32), ("refresh_every", C.c_int32), ("use_graph", C.c_int32),
        ("precond", C.c_int32), ("reserved", C.c_int32), ("debug", C.c_void_p),
        ("debug_vectors", C.c_void_p),
    ]


class SqdkLocgResult(C.Structure):
    _fields_ = [
        ("eigval", C.c_double), ("resid_norm", C.c_double), ("ax_norm", C.c_double),
        ("niter", C.c_int32), ("converged", C.c_int32), ("matvecs", C.c_int32),
        ("status", C.c_int32),
    ]


COEF_TERMS, COEF_DIAG, COEF_SIGNS = 0, 1, 2


class SqdkGroups(C.Structure):
    _fields_ = [
        ("keys", C.c_void_p), ("n", C.c_int64), ("n_real", C.c_int64),
        ("n_groups", C.c_int32), ("coef_form", C.c_int32), ("xmask", C.c_void_p),
        ("xsource", C.c_void_p), ("ld", C.c_int64), ("zmask", C.c_void_p),
        ("coef", C.c_void_p), ("kmax", C.c_int32), ("sign_bytes", C.c_int32),
        ("diag", C.c_void_p), ("signs", C.c_void_p),
    ]


MATVEC_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)


class SqdkOp(C.Structure):
    _fields_ = [("ham", C.POINTER(SqdkHam)), ("space", C.POINTER(SqdkSpace)),
                ("tune", C.POINTER(SqdkTune)), ("fn", MATVEC_FN), ("ctx", C.c_void_p),
                ("cache", C.POINTER(SqdkCache))]


# name -> (restype, argtypes); every function declared in include/sqdk.h
SIGNATURES = {
    "sqdk_abi_version": (C.c_int, []),
    "sqdk_last_error": (C.c_char_p, []),
    "sqdk_space_scan": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "sqdk_space_dir": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p,
                                 C.c_void_p]),
    "sqdk_space_fine": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_void_p, C.c_void_p, C.c_void_p]),
    "sqdk_intake_pack": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int64, C.c_void_p,
                                   C.c_void_p]),
    "sqdk_intake_unique_workspace": (C.c_int, [C.c_int64, C.c_int32,
                                                C.POINTER(C.c_size_t)]),
    "sqdk_intake_unique": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_size_t, C.c_void_p]),
    "sqdk_intake_unpack": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                     C.c_void_p]),
    "sqdk_space_bits": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sqdk_space_items": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int64,
                                   C.c_void_p, C.c_void_p]),
    "sqdk_xsource": (C.c_int, [C.POINTER(SqdkSpace), C.c_uint64, C.c_void_p, C.c_void_p]),
    "sqdk_diagonal": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                C.c_int64, C.c_void_p, C.c_void_p]),
    "sqdk_matvec_workspace": (C.c_int, [C.POINTER(SqdkSpace), C.POINTER(SqdkTune), C.c_int32,
                                        C.POINTER(C.c_size_t)]),
    "sqdk_matvec": (C.c_int, [C.POINTER(SqdkHam), C.POINTER(SqdkSpace), C.c_int32,
                              C.c_void_p, C.c_void_p, C.POINTER(SqdkTune), C.c_void_p,
                              C.c_size_t, C.c_void_p]),
    "sqdk_batch_matvec_workspace": (C.c_int, [C.POINTER(SqdkBatch), C.POINTER(SqdkTune),
                                              C.c_int32, C.POINTER(C.c_size_t)]),
    "sqdk_batch_matvec": (C.c_int, [C.POINTER(SqdkBatch), C.c_int32, C.c_void_p, C.c_void_p,
                                    C.POINTER(SqdkTune), C.c_void_p, C.c_size_t, C.c_void_p]),
    "sqdk_batch_locg_workspace": (C.c_int, [C.POINTER(SqdkBatch), C.POINTER(SqdkTune),
                                            C.POINTER(SqdkLocgOpts), C.POINTER(C.c_size_t)]),
    "sqdk_batch_locg_solve": (C.c_int, [C.POINTER(SqdkBatch), C.POINTER(SqdkTune), C.c_void_p,
                                        C.POINTER(SqdkLocgOpts), C.c_void_p, C.c_size_t,
                                        C.POINTER(SqdkLocgResult), C.c_void_p]),
    "sqdk_dots": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
                            C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_size_t,
                            C.c_void_p]),
    "sqdk_dots_workspace": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_size_t)]),
    "sqdk_dots_noalloc": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p]),
    "sqdk_convert": (C.c_int, [C.c_int64, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p,
                               C.c_void_p]),
    "sqdk_locg_workspace": (C.c_int, [C.POINTER(SqdkOp), C.c_int64, C.POINTER(SqdkLocgOpts),
                                      C.POINTER(C.c_size_t)]),
    "sqdk_cache_slices": (C.c_int, [C.POINTER(SqdkSpace), C.POINTER(C.c_int64)]),
    "sqdk_cache_workspace": (C.c_int, [C.POINTER(SqdkHam), C.POINTER(SqdkSpace),
                                       C.POINTER(C.c_size_t)]),
    "sqdk_cache_count": (C.c_int, [C.POINTER(SqdkHam), C.POINTER(SqdkSpace), C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                   C.c_void_p]),
    "sqdk_cache_fill": (C.c_int, [C.POINTER(SqdkHam), C.POINTER(SqdkSpace), C.c_int32,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "sqdk_cache_matvec": (C.c_int, [C.POINTER(SqdkCache), C.c_void_p, C.c_void_p, C.c_void_p]),
    "sqdk_batch_cache_slices": (C.c_int, [C.POINTER(SqdkBatch), C.POINTER(C.c_int64)]),
    "sqdk_batch_cache_workspace": (C.c_int, [C.POINTER(SqdkBatch), C.c_int32,
                                             C.POINTER(C.c_size_t)]),
    "sqdk_batch_cache_count": (C.c_int, [C.POINTER(SqdkBatch), C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t,
                                         C.c_void_p]),
    "sqdk_batch_cache_fill": (C.c_int, [C.POINTER(SqdkBatch), C.c_int32, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t,
                                        C.c_void_p]),
    "sqdk_locg_solve": (C.c_int, [C.POINTER(SqdkOp), C.c_int64, C.c_void_p,
                                  C.POINTER(SqdkLocgOpts), C.c_void_p, C.c_size_t,
                                  C.POINTER(SqdkLocgResult), C.c_void_p]),
    "sqdk_small_eigenpair": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sqdk_gemm_bf16_workspace": (C.c_int, [C.c_int64, C.c_int64, C.c_int32,
                                           C.POINTER(C.c_size_t)]),
    "sqdk_gemm_bf16": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.c_void_p, C.c_int64,
                                 C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                 C.c_int64, C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t,
                                 C.c_void_p]),
    "sqdk_mlp_infer": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int32,
                                 C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sqdk_mlp_encode": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int32,
                                  C.c_int32, C.c_void_p, C.c_void_p]),
    "sqdk_mlp_act": (C.c_int, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p]),
    "sqdk_mlp_loss": (C.c_int, [C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_double,
                                C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                C.c_void_p]),
    "sqdk_colsum": (C.c_int, [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sqdk_adam": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                            C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32,
                            C.c_void_p]),
    "sqdk_rows_to_keys": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]),
    "sqdk_keys_to_rows": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int32, C.c_void_p,
                                    C.c_void_p]),
    "sqdk_groups_matvec": (C.c_int, [C.POINTER(SqdkGroups), C.c_int32, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "sqdk_diag_signs": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_int32,
                                  C.c_void_p, C.c_void_p]),
    "sqdk_compute_diagonal": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int32,
                                        C.c_void_p, C.c_void_p]),
    "sqdk_svsim": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                             C.c_void_p, C.c_void_p]),
}


def library_path() -> Path:
    env = os.environ.get("RQSQD_LIB")
    if env:
        return Path(env)
    return Path(__file__).resolve().parent / "librqsqd.so"


_LIB = None


def lib():
    """The loaded library (loaded once); raises ImportError if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not path.exists():
        raise ImportError(f"{path} is not built; run `python -m rqutils_amd.build` "
                          "(or __graft_entry__.build())")
    handle = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    ver = handle.sqdk_abi_version()
    if ver != ABI_VERSION:
        raise ImportError(f"{path} has ABI version {ver}, expected {ABI_VERSION}; rebuild it")
    _LIB = handle
    return handle


class SqdkError(RuntimeError):
    """A HIP runtime or callback failure reported by the kernel library."""


def check(rc: int, what: str = "") -> None:
    """Map a status code to the reference's host-side exceptions."""
    if rc == OK:
        return
    msg = lib().sqdk_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc in (EINVAL, ERANGE, ENOSPC):
        raise ValueError(text)
    raise SqdkError(text)


def stream_ptr(device=None) -> int:
    """Raw hipStream_t of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int:
    return 0 if t is None else t.data_ptr()
