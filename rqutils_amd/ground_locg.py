r"""Single-vector LOBPCG on the GPU (drop-in for ``rqutils.ground_locg``).

``ground_locg(mat, xinit, args=(), maxiter=1000, tol=None, vspace=None,
log_level=WARNING)`` keeps the reference's call signature
(rqutils/ground_locg.py:128-157). The iteration runs in the HIP library
(``sqdk_locg_solve``): vectors stay in HBM, every scalar stays on the device
and the host polls the convergence flag.

``mat`` may be

* a :class:`~rqutils_amd.operator.PauliOperator` (the built-in bucket-join
  matvec; the fast path),
* a callable ``mat(x, *args) -> A @ x`` on torch device tensors (called from
  the solver through a C callback), or
* a dense matrix (numpy array or torch tensor), applied with ``torch.mv`` on
  the device as ``_ground_locg_matrix`` does (rqutils/ground_locg.py:160-179).

Modes: ``mode="reference"`` (default) reproduces the reference's iteration
and stopping rule step by step; ``mode="fused"`` is the one-matvec-per-
iteration variant with the Gram/Rayleigh-Ritz algebra in FP64 on the device.
The vector precision follows ``xinit`` (complex128/float64 -> FP64 storage,
complex64/float32 -> FP32 storage) unless ``precision`` is given.
"""
from __future__ import annotations

import ctypes as C
import logging
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .operator import PauliOperator

__all__ = ["ground_locg", "solve", "LocgInfo", "eigenpair_2x2", "eigenpair_3x3"]

LOG = logging.getLogger(__name__)

STATUS_BREAKDOWN = 2      # LocgInfo.status: Rayleigh-Ritz breakdown (Cholesky pivot floor)
STATUS_NONFINITE = 3      # LocgInfo.status: NaN or Inf in the sums or small eigenproblems
STATUS_INVARIANT = 4      # LocgInfo.status: exact eigenvector start with more than one state

_MODES = {"reference": _lib.LOCG_REFERENCE, "fused": _lib.LOCG_FUSED}
_CRITERIA = {"reference": _lib.CRIT_REFERENCE, "relative": _lib.CRIT_RELATIVE}
_PRECISIONS = {"c128": _lib.C128, "c64": _lib.C64}
_PRECONDS = {None: 0, "none": 0, "diag": 1}     # sqdk_locg_precond


@dataclass
class LocgInfo:
    """Diagnostics of one solve."""
    eigval: float
    niter: int
    converged: bool
    matvecs: int
    resid_norm: float      # |r| as tracked by the solver (explicit at the end for "fused")
    ax_norm: float
    status: int
    diagnostics: dict | None = None   # per-step records of a debug solve


def diagnostics_dict(buf: torch.Tensor, count: int, vectors=None, real: bool = False) -> dict:
    """The reference's diagnostics (rqutils/ground_locg.py:211-233) from the
    device records: arrays over the first step, the 2x2 step and each loop
    iteration: theta, rho, kappa [3], sas [3, 3] (S^H A S of x, y, r), reltol,
    converged, and the residual norm; with ``vectors`` also x, y, r."""
    rec = buf.cpu().numpy().reshape(-1, 15, 2)[:count]
    cplx = rec[..., 0] + 1j * rec[..., 1]
    out = {"theta": rec[:, 0, 0].copy(), "rho": rec[:, 0, 1].copy(),
           "reltol": rec[:, 1, 0].copy(), "converged": rec[:, 1, 1] != 0,
           "resid_norm": rec[:, 2, 0].copy(), "kappa": cplx[:, 3:6].copy(),
           "sas": cplx[:, 6:15].reshape(-1, 3, 3).copy()}
    if vectors is not None:
        v = vectors[:count].cpu().numpy()
        if real:
            v = v.real
        out["x"], out["y"], out["r"] = v[:, 0], v[:, 1], v[:, 2]
    return out


def _precision_of(dtype: torch.dtype) -> int:
    if dtype in (torch.complex128, torch.float64):
        return _lib.C128
    if dtype in (torch.complex64, torch.float32):
        return _lib.C64
    raise ValueError(f"unsupported vector dtype {dtype}")


def _cdtype(prec: int) -> torch.dtype:
    return torch.complex128 if prec == _lib.C128 else torch.complex64


class _Callback:
    """C matvec callback that runs a Python operator on views of solver buffers."""

    def __init__(self, fn, n, dtype, real, buffers):
        self.fn = fn
        self.n = n
        self.dtype = dtype
        self.real = real
        self.buffers = buffers      # list of (base_ptr, nbytes, tensor) the solver owns
        self.error = None
        self.cfunc = _lib.MATVEC_FN(self._call)

    def _view(self, p):
        es = 16 if self.dtype == torch.complex128 else 8
        for base, nbytes, t in self.buffers:
            off = p - base
            if 0 <= off and off + self.n * es <= nbytes:
                raw = t.view(torch.uint8).reshape(-1)[off:off + self.n * es]
                return raw.view(self.dtype)
        raise RuntimeError(f"callback pointer {p:#x} is outside the solver's buffers")

    def _call(self, ctx, pin, pout, stream):
        try:
            vin = self._view(pin)
            vout = self._view(pout)
            if self.real:
                res = self.fn(vin.real.contiguous())
                vout.copy_(res.to(vout.dtype))
            else:
                vout.copy_(self.fn(vin))
            return 0
        except Exception as exc:  # noqa: BLE001 - reported after the C call returns
            self.error = exc
            return -1


def solve(op, x0: torch.Tensor, *, mode: str = "fused", precision: str | None = None,
          criterion: str = "relative", tol: float = 1e-6, maxiter: int = 1000,
          n_for_tol: int | None = None, check_every: int = 4, refresh_every: int = 16,
          real: bool = False, precond: str | None = None,
          debug: bool | str = False, use_graph: bool = False) -> tuple[torch.Tensor, LocgInfo]:
    """Lowest eigenpair of ``op`` from start vector ``x0`` (device tensor).

    ``op`` is a PauliOperator or a callable on device tensors. Returns the
    normalised eigenvector (complex, solver precision) and :class:`LocgInfo`.
    ``precond="diag"`` (fused mode, PauliOperator only) searches along the
    residual scaled by ``1 / max(d - min(theta, min d), floor)``, ``d`` the
    diagonal from the operator's X = 0 terms and ``floor = max(0.02 (mean d -
    min d), 1e-3 |theta|)``.

    ``use_graph`` (fused mode, PauliOperator): replay each refresh period of
    ``refresh_every`` iterations as one captured HIP graph on a private stream.

    ``debug`` (reference mode): record the reference's per-step diagnostics
    (``LocgInfo.diagnostics``, see ``diagnostics_dict``); ``"vectors"`` also
    keeps x, y and r of every step. Every one of ``maxiter`` iterations then
    runs, as in the reference's diagnostics scan.
    """
    L = _lib.lib()
    if mode not in _MODES:
        raise ValueError(f"mode must be one of {sorted(_MODES)}")
    if criterion not in _CRITERIA:
        raise ValueError(f"criterion must be one of {sorted(_CRITERIA)}")
    prec = _PRECISIONS[precision] if precision is not None else _precision_of(x0.dtype)
    if x0.device.type != "cuda":
        raise ValueError("x0 must be a GPU tensor")
    n = x0.numel()
    x = x0.reshape(-1).to(_cdtype(prec)).clone()
    opts = _lib.SqdkLocgOpts()
    opts.mode = _MODES[mode]
    opts.precision = prec
    opts.criterion = _CRITERIA[criterion]
    opts.maxiter = int(maxiter)
    opts.tol = float(tol)
    opts.n_for_tol = int(n_for_tol if n_for_tol is not None else n)
    opts.check_every = int(check_every)
    opts.refresh_every = int(refresh_every)
    opts.use_graph = 1 if use_graph else 0
    if precond not in _PRECONDS:
        raise ValueError(f"precond must be one of {sorted(_PRECONDS, key=str)}")
    opts.precond = _PRECONDS[precond]
    dbg = dbg_vec = None
    if debug:
        if mode != "reference":
            raise ValueError("debug diagnostics are recorded by mode='reference' only")
        dbg = torch.zeros((maxiter + 2) * 30, dtype=torch.float64, device=x.device)
        opts.debug = dbg.data_ptr()
        if debug == "vectors":
            dbg_vec = torch.zeros((maxiter + 2, 3, n), dtype=x.dtype, device=x.device)
            opts.debug_vectors = dbg_vec.data_ptr()
    if opts.precond and (mode != "fused" or not isinstance(op, PauliOperator)):
        raise ValueError("precond='diag' needs mode='fused' and a PauliOperator")
    cop = _lib.SqdkOp()
    keep = []
    cb = None
    if isinstance(op, PauliOperator):
        if op.space.n != n:
            raise ValueError(f"operator size {op.space.n} != vector size {n}")
        h, s, t = op.structs()
        keep += [h, s, t]
        cop.ham = C.pointer(h)
        cop.space = C.pointer(s)
        cop.tune = C.pointer(t)
        cs = op.cache_struct(prec)
        if cs is not None:
            keep.append(cs)
            cop.cache = C.pointer(cs)
    elif not callable(op):
        raise ValueError("op must be a PauliOperator or a callable")
    nbytes = C.c_size_t(0)
    _lib.check(L.sqdk_locg_workspace(C.byref(cop), n, C.byref(opts), C.byref(nbytes)),
               "sqdk_locg_workspace")
    work = torch.empty(nbytes.value, dtype=torch.uint8, device=x.device)
    if not isinstance(op, PauliOperator):
        cb = _Callback(op, n, _cdtype(prec), real,
                       [(work.data_ptr(), work.numel(), work),
                        (x.data_ptr(), x.numel() * x.element_size(), x)])
        cop.fn = cb.cfunc
    res = _lib.SqdkLocgResult()
    with torch.cuda.device(x.device):
        rc = L.sqdk_locg_solve(C.byref(cop), n, x.data_ptr(), C.byref(opts), work.data_ptr(),
                               nbytes.value, C.byref(res), _lib.stream_ptr(x.device))
    if cb is not None and cb.error is not None:
        raise cb.error
    _lib.check(rc, "sqdk_locg_solve")
    info = LocgInfo(res.eigval, res.niter, bool(res.converged), res.matvecs, res.resid_norm,
                    res.ax_norm, res.status)
    if dbg is not None:
        info.diagnostics = diagnostics_dict(dbg, res.niter + 2, dbg_vec, real)
    if info.status == STATUS_NONFINITE:
        LOG.warning("LOBPCG stopped at iteration %d: NaN or Inf in the operator or vectors",
                    info.niter)
    elif info.status == STATUS_BREAKDOWN:
        LOG.warning("LOBPCG stopped at iteration %d: Rayleigh-Ritz breakdown", info.niter)
    elif info.status == STATUS_INVARIANT:
        LOG.warning("LOBPCG start is an exact eigenvector (zero residual) of a %d-state "
                    "operator: theta = %.12g is an eigenvalue but not necessarily the lowest",
                    n, info.eigval)
    return x, info


def ground_locg(mat, xinit, args: tuple = (), maxiter: int = 1000, tol: float | None = None,
                vspace=None, log_level: int = logging.WARNING, *, mode: str = "reference",
                precision: str | None = None, criterion: str = "reference",
                check_every: int | None = None, refresh_every: int = 16,
                device=None, precond: str | None = None, debug: bool | str = False):
    r"""Single-vector LOBPCG (rqutils/ground_locg.py:128-157).

    Args:
        mat: Matrix :math:`A` (dense array/tensor), a callable
            :math:`x \mapsto Ax` on device tensors, or a ``PauliOperator``.
        xinit: Initial vector, or an integer for a one-hot start (then
            ``vspace=(dim, dtype)`` is required for a callable ``mat``). With a
            ``PauliOperator`` also ``"min_diag"``, ``"gaussian"`` or a trained
            ``StartNet`` (``sqd.start_vector``: a network start carries a
            Gaussian part of norm ``startnet.DEFAULT_MIX``).
        args: Extra arguments passed to a callable ``mat``.
        maxiter: Maximum number of loop iterations.
        tol: Convergence tolerance; defaults to the machine epsilon of the
            vector dtype (ground_locg.py:208-209).
        vspace: ``(dimension, dtype)`` for an integer ``xinit``.
        log_level: Verbosity (DEBUG logs the result).
        precond: ``"diag"`` for the Jacobi-preconditioned fused iteration
            (``mode="fused"`` with a ``PauliOperator``; the reference has no
            preconditioner).
        debug: Record the reference's diagnostics (its internal
            ``_ground_locg_callable(debug=True)``, ground_locg.py:211-233,
            358-376): every one of ``maxiter`` iterations runs and a fourth
            result, a dict of per-step arrays, is returned; ``"vectors"`` adds
            x, y and r of every step.

    Returns:
        ``(eigval, eigvec, niter)``: the Rayleigh-Ritz value as a float, the
        eigenvector as a device tensor (real if the problem is real) and the
        iteration count.
    """
    dev = torch.device(device) if device is not None else None
    # Build the operator.
    dense = None
    if isinstance(mat, PauliOperator):
        op = mat
        dev = op.space.device
        n = op.space.n
        base_dtype = torch.complex128
    elif callable(mat):
        user = mat
        op = (lambda v: user(v, *args)) if args else user
        n = None
        base_dtype = None
    else:
        dense = mat if isinstance(mat, torch.Tensor) else torch.as_tensor(np.asarray(mat))
        if dev is None:
            dev = dense.device if dense.device.type == "cuda" else torch.device("cuda")
        dense = dense.to(dev)
        if dense.ndim != 2 or dense.shape[0] != dense.shape[1]:
            raise ValueError("matrix must be square")
        n = dense.shape[0]
        base_dtype = dense.dtype

    # Start vector.
    from .startnet import StartNet
    if isinstance(xinit, (str, StartNet)):
        if not isinstance(mat, PauliOperator):
            raise ValueError("a named or network start needs a PauliOperator")
        from .sqd import start_vector
        x0 = start_vector(mat, xinit)
        if not mat.is_complex:
            x0 = x0.real.contiguous() if xinit == "min_diag" else x0
    elif isinstance(xinit, (int, np.integer)) or (isinstance(xinit, torch.Tensor) and
                                                 xinit.ndim == 0 and not xinit.is_floating_point()
                                                 and not xinit.is_complex()):
        idx = int(xinit)
        if dense is not None:
            dim, vdtype = n, dense.dtype
        elif isinstance(mat, PauliOperator):
            dim = n
            vdtype = torch.complex128 if mat.is_complex else torch.float64
        else:
            if vspace is None:
                raise ValueError("vspace=(dim, dtype) is required for an integer xinit")
            dim = int(vspace[0])
            vdtype = vspace[1] if isinstance(vspace[1], torch.dtype) else \
                torch.from_numpy(np.zeros(1, dtype=vspace[1])).dtype
        x0 = torch.zeros(dim, dtype=vdtype, device=dev or "cuda")
        x0[idx] = 1
    else:
        x0 = xinit if isinstance(xinit, torch.Tensor) else torch.as_tensor(np.asarray(xinit))
        x0 = x0.to(dev or (x0.device if x0.device.type == "cuda" else "cuda"))
    dev = x0.device
    real = not x0.is_complex() and (dense is None or not dense.is_complex()) and \
        (not isinstance(mat, PauliOperator) or not mat.is_complex)
    if tol is None:
        tol = float(torch.finfo(x0.dtype).eps)
    if dense is not None:
        dmat = dense if not real else dense

        def op(v):  # noqa: F811 - dense matvec on the device
            return torch.mv(dmat.to(v.dtype), v)
    if check_every is None:
        check_every = 4
    vec, info = solve(op, x0, mode=mode, precision=precision, criterion=criterion, tol=tol,
                      maxiter=maxiter, check_every=check_every, refresh_every=refresh_every,
                      real=real, precond=precond, debug=debug)
    if log_level <= logging.DEBUG:
        LOG.debug("LOBPCG: eigval %s after %d iterations (converged=%s, |r|=%g)",
                  info.eigval, info.niter, info.converged, info.resid_norm)
    if real:
        vec = vec.real.contiguous()
    if debug:
        return info.eigval, vec, info.niter, info.diagnostics
    return info.eigval, vec, info.niter


def _small_eigenpair(mat) -> tuple[float, np.ndarray]:
    m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128))
    dim = m.shape[0]
    dm = torch.from_numpy(m.view(np.float64).copy()).cuda()
    out = torch.empty(1 + 2 * dim, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().sqdk_small_eigenpair(dim, dm.data_ptr(), out.data_ptr(),
                                               _lib.stream_ptr()), "sqdk_small_eigenpair")
    o = out.cpu().numpy()
    return float(o[0]), o[1:].view(np.complex128).copy()


def eigenpair_2x2(mat) -> tuple[float, np.ndarray]:
    """Lowest eigenpair of a 2x2 Hermitian matrix by the device closed form
    (rqutils/ground_locg.py:424-435)."""
    return _small_eigenpair(mat)


def eigenpair_3x3(mat) -> tuple[float, np.ndarray]:
    """Lowest eigenpair of a 3x3 Hermitian matrix by the device Cardano/Kopp
    closed form (rqutils/ground_locg.py:438-471)."""
    return _small_eigenpair(mat)
