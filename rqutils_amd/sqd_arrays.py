r"""The operator functions of ``rqutils.sqd`` on the reference's own arrays.

Callers that hold the reference's packed data -- padded bit-packed state rows
``uint8[N, Bp]`` with 0xFF fill rows, X signatures ``uint8[J, Bp]``, Z
signatures ``uint8[J, Kmax, Bp]``, phased coefficients ``[J, Kmax]``, cached
source indices ``int32[J, N]``, sign bits ``uint8[J, N, ceil(Kmax/8)]`` or
diagonals ``[J, N]`` -- call these with the reference's argument lists
(rqutils/sqd.py:414-684):

``uniquify_states``, ``get_xsource``, ``get_diag_signs``,
``compute_diagonal``, ``get_diagonal``, ``apply_xgrp``, ``apply_h``,
``apply_h_s_cached``, ``apply_h_z_cached``, ``apply_h_x_cached``,
``apply_h_xs_cached``, ``apply_h_xz_cached``.

Arrays may be numpy arrays or torch tensors; results are torch tensors on
the GPU. Every product runs in the HIP library (``sqdk_groups_matvec``: one
thread per row, groups in the reference's scan order, FP64 sums); a packed
row of up to 8 bytes is handled as one big-endian integer, whose order is the
reference's row order. The fast path for the same products is
``SQDProblem``/``PauliOperator`` (bucket join or hit table).

Where the reference's own code misbehaves, these compute the intended value:
sums over a group's terms skip zero coefficients instead of stopping at the
first (sqd.py:562-564), ``compute_diagonal`` reads term k's sign bit at
7 - k % 8 (the reference's ``7 - (k & 255)`` turns every term from the ninth
on into +1, sqd.py:544), and the products give fill rows 0 (outside the
subspace; the reference's values there come from the sort-join of equal fill
rows and are never read by its solver, whose vectors are 0 there).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .intake import unique_keys

__all__ = ["uniquify_states", "get_xsource", "get_diag_signs", "compute_diagonal",
           "get_diagonal", "apply_xgrp", "apply_h", "apply_h_s_cached", "apply_h_z_cached",
           "apply_h_x_cached", "apply_h_xs_cached", "apply_h_xz_cached"]


def _device(*arrays):
    for a in arrays:
        if isinstance(a, torch.Tensor) and a.device.type == "cuda":
            return a.device
    return torch.device("cuda", torch.cuda.current_device())


def _on(a, dev, dtype=None) -> torch.Tensor:
    t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(np.asarray(a)))
    t = t.to(dev)
    if dtype is not None:
        t = t.to(dtype)
    return t.contiguous()


def _rows(states, dev) -> torch.Tensor:
    st = _on(states, dev)
    if st.dtype != torch.uint8 or st.dim() != 2:
        raise ValueError("states must be packed uint8 rows [N, Bp]")
    if not 1 <= st.shape[1] <= 8:
        raise ValueError(f"packed rows of {st.shape[1]} bytes: 1 to 8 (up to 63 qubits) "
                         "are supported")
    return st


def _row_keys(rows: torch.Tensor) -> torch.Tensor:
    n, bp = rows.shape
    keys = torch.empty(n, dtype=torch.int64, device=rows.device)
    if n:
        _lib.check(_lib.lib().sqdk_rows_to_keys(rows.data_ptr(), n, bp, keys.data_ptr(),
                                                _lib.stream_ptr(rows.device)),
                   "sqdk_rows_to_keys")
    return keys


def _sig_keys(sigs) -> np.ndarray:
    """Big-endian integers of packed signatures [..., Bp] (host, uint64)."""
    a = np.asarray(sigs.cpu() if isinstance(sigs, torch.Tensor) else sigs, dtype=np.uint8)
    k = np.zeros(a.shape[:-1], np.uint64)
    for b in range(a.shape[-1]):
        k = (k << np.uint64(8)) | a[..., b].astype(np.uint64)
    return k


def _n_real(keys: torch.Tensor, bp: int) -> int:
    """Rows before the first fill row (the pad bit, the top bit of the row
    key, is set only on fill rows, which sort last)."""
    top = torch.tensor(np.uint64(1 << (8 * bp - 1)).view(np.int64), device=keys.device)
    pad = torch.bitwise_and(keys, top) != 0
    return int(keys.numel() - torch.count_nonzero(pad).item())


def _coef(coeffs, dev):
    c = coeffs.cpu().numpy() if isinstance(coeffs, torch.Tensor) else np.asarray(coeffs)
    return torch.from_numpy(np.ascontiguousarray(c.astype(np.complex128)).view(np.float64)).to(dev)


def _result(out: torch.Tensor, vec, coeff_complex: bool):
    """The reference returns a real vector when the vector and coefficients
    are real."""
    real_in = not (vec.is_complex() if isinstance(vec, torch.Tensor) else np.iscomplexobj(vec))
    if real_in and not coeff_complex:
        return out.real.contiguous()
    return out


def uniquify_states(states_p, states_size: int) -> torch.Tensor:
    """Sorted distinct packed rows, ``states_size`` of them: past the unique
    count the rows are filled with 255 (rqutils/sqd.py:414-445). Bit-exact with
    the reference (``lax.sort`` + unique over the packed bytes)."""
    dev = _device(states_p)
    rows = _rows(states_p, dev)
    n, bp = rows.shape
    if states_size < 1:
        raise ValueError("states_size must be at least 1")
    u = unique_keys(_row_keys(rows), 8 * bp) if n else torch.zeros(0, dtype=torch.int64,
                                                                     device=dev)
    out = torch.empty((states_size, bp), dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib().sqdk_keys_to_rows(u.data_ptr() if u.numel() else 0, min(u.numel(),
                                            states_size), states_size, bp, out.data_ptr(),
                                            _lib.stream_ptr(dev)), "sqdk_keys_to_rows")
    return out


def get_xsource(xsignature, states) -> torch.Tensor:
    """Source index of ``S[i] ^ x`` in ``S`` or -1 (rqutils/sqd.py:448-503),
    on packed sorted rows with optional fill rows (these get -1). Bit-exact
    with the reference on the distinct rows."""
    from .sqd import get_xsource as _gx
    return _gx(xsignature, states)


def get_diag_signs(zsignatures, states) -> torch.Tensor:
    """Packed sign bits ``uint8[N, ceil(K/8)]`` of one group's K Z signatures:
    bit 7 - k % 8 of byte k / 8 is the parity of ``z_k & s_i``
    (rqutils/sqd.py:506-528)."""
    dev = _device(states, zsignatures)
    rows = _rows(states, dev)
    keys = _row_keys(rows)
    z = _sig_keys(zsignatures).reshape(-1)
    nb = max(1, -(-z.size // 8))
    zt = torch.from_numpy(z.view(np.int64)).to(dev)
    out = torch.empty((rows.shape[0], nb), dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib().sqdk_diag_signs(zt.data_ptr() if z.size else 0, z.size, keys.data_ptr(),
                                          rows.shape[0], nb, out.data_ptr(), _lib.stream_ptr(dev)),
               "sqdk_diag_signs")
    return out


def compute_diagonal(diag_signs, coeffs) -> torch.Tensor:
    """Diagonal of one group from its sign bits and coefficients
    (rqutils/sqd.py:531-551; term k's bit at 7 - k % 8); complex128."""
    dev = _device(diag_signs)
    s = _on(diag_signs, dev, torch.uint8)
    c = _coef(coeffs, dev)
    k = c.numel() // 2
    out = torch.empty(s.shape[0], dtype=torch.complex128, device=dev)
    _lib.check(_lib.lib().sqdk_compute_diagonal(s.data_ptr(), s.shape[0], s.shape[1],
                                                c.data_ptr(), k, out.data_ptr(),
                                                _lib.stream_ptr(dev)), "sqdk_compute_diagonal")
    return out


def get_diagonal(zsignatures, coeffs, states) -> torch.Tensor:
    """Composed diagonal of one X group (rqutils/sqd.py:554-575), complex128."""
    from .sqd import get_diagonal as _gd
    return _gd(zsignatures, coeffs, states)


def apply_xgrp(xsource, diagonal, vec) -> torch.Tensor:
    """``vec[xsource] * diagonal`` with 0 where the source is -1
    (rqutils/sqd.py:578-587)."""
    dev = _device(xsource, diagonal, vec)
    xs = _on(xsource, dev, torch.int32).reshape(1, -1)
    d = _on(diagonal, dev, torch.complex128).reshape(1, -1)
    n = xs.shape[1]
    keys = torch.zeros(n, dtype=torch.int64, device=dev)
    out = _groups(keys, n, xs, None, _lib.COEF_DIAG, diag=d, vec=vec, dev=dev)
    return _result(out, vec, torch.is_complex(_on(diagonal, dev)))


def _groups(keys, n_real, xsource, xmask, form, *, zmask=None, coef=None, kmax=0, diag=None,
            signs=None, vec=None, dev=None):
    v = _on(vec, dev)
    prec = _lib.C64 if v.dtype == torch.complex64 else _lib.C128
    v = v.to(torch.complex64 if prec == _lib.C64 else torch.complex128).contiguous()
    n = keys.numel()
    if v.numel() != n:
        raise ValueError(f"vector of length {v.numel()} for {n} states")
    g = _lib.SqdkGroups()
    g.keys = keys.data_ptr()
    g.n = n
    g.n_real = n_real
    nj = (xsource.shape[0] if xsource is not None else len(xmask))
    g.n_groups = nj
    g.coef_form = form
    keep = []
    if xsource is not None:
        if xsource.shape[1] != n:
            raise ValueError("source indices must be [J, N]")
        g.xsource = xsource.data_ptr()
        g.ld = n
    else:
        xm = torch.from_numpy(np.ascontiguousarray(xmask).view(np.int64)).to(dev)
        keep.append(xm)
        g.xmask = xm.data_ptr() if nj else 0
        g.ld = n
    if zmask is not None:
        g.zmask = zmask.data_ptr()
    if coef is not None:
        g.coef = coef.data_ptr()
    g.kmax = kmax
    if diag is not None:
        if tuple(diag.shape) != (nj, n):
            raise ValueError("diagonals must be [J, N]")
        g.diag = diag.data_ptr()
    if signs is not None:
        if signs.shape[:2] != (nj, n):
            raise ValueError("sign bits must be [J, N, nbytes]")
        g.signs = signs.data_ptr()
        g.sign_bytes = signs.shape[2]
    out = torch.empty_like(v)
    _lib.check(_lib.lib().sqdk_groups_matvec(C.byref(g), prec, v.data_ptr(), out.data_ptr(),
                                             _lib.stream_ptr(dev)), "sqdk_groups_matvec")
    return out


def _states_keys(states, dev):
    rows = _rows(states, dev)
    keys = _row_keys(rows)
    return keys, _n_real(keys, rows.shape[1])


def _terms(zsignatures, coeffs, dev):
    z = _sig_keys(zsignatures)                      # [J, Kmax]
    if z.ndim != 2:
        raise ValueError("Z signatures must be [J, Kmax, Bp]")
    c = coeffs.cpu().numpy() if isinstance(coeffs, torch.Tensor) else np.asarray(coeffs)
    if c.shape != z.shape:
        raise ValueError("coefficients must be [J, Kmax]")
    zt = torch.from_numpy(np.ascontiguousarray(z).view(np.int64)).to(dev)
    return zt, _coef(c, dev), z.shape[1], bool(np.iscomplexobj(c))


def apply_h(vec, xsignatures, zsignatures, coeffs, states) -> torch.Tensor:
    """Hv from the X and Z signatures and coefficients, source indices found
    at every call (rqutils/sqd.py:590-605, cache_level (0, 0))."""
    dev = _device(vec, states)
    keys, n_real = _states_keys(states, dev)
    zt, ct, kmax, cplx = _terms(zsignatures, coeffs, dev)
    out = _groups(keys, n_real, None, _sig_keys(xsignatures).reshape(-1), _lib.COEF_TERMS,
                  zmask=zt, coef=ct, kmax=kmax, vec=vec, dev=dev)
    return _result(out, vec, cplx)


def apply_h_s_cached(vec, xsignatures, states, diag_signs, coeffs) -> torch.Tensor:
    """Hv from cached sign bits (rqutils/sqd.py:608-622, cache_level (0, 1))."""
    dev = _device(vec, states, diag_signs)
    keys, n_real = _states_keys(states, dev)
    c = coeffs.cpu().numpy() if isinstance(coeffs, torch.Tensor) else np.asarray(coeffs)
    s = _on(diag_signs, dev, torch.uint8)
    out = _groups(keys, n_real, None, _sig_keys(xsignatures).reshape(-1), _lib.COEF_SIGNS,
                  coef=_coef(c, dev), kmax=c.shape[1], signs=s, vec=vec, dev=dev)
    return _result(out, vec, bool(np.iscomplexobj(c)))


def apply_h_z_cached(vec, xsignatures, states, diagonals) -> torch.Tensor:
    """Hv from cached diagonals (rqutils/sqd.py:625-637, cache_level (0, 2))."""
    dev = _device(vec, states, diagonals)
    keys, n_real = _states_keys(states, dev)
    d = _on(diagonals, dev)
    cplx = d.is_complex()
    out = _groups(keys, n_real, None, _sig_keys(xsignatures).reshape(-1), _lib.COEF_DIAG,
                  diag=d.to(torch.complex128).contiguous(), vec=vec, dev=dev)
    return _result(out, vec, cplx)


def apply_h_x_cached(vec, xsources, zsignatures, coeffs, states) -> torch.Tensor:
    """Hv from cached source indices (rqutils/sqd.py:640-654, the reference's
    default cache_level (1, 0))."""
    dev = _device(vec, xsources, states)
    keys, n_real = _states_keys(states, dev)
    zt, ct, kmax, cplx = _terms(zsignatures, coeffs, dev)
    out = _groups(keys, n_real, _on(xsources, dev, torch.int32), None, _lib.COEF_TERMS,
                  zmask=zt, coef=ct, kmax=kmax, vec=vec, dev=dev)
    return _result(out, vec, cplx)


def apply_h_xs_cached(vec, xsources, diag_signs, coeffs) -> torch.Tensor:
    """Hv from cached source indices and sign bits (rqutils/sqd.py:657-670).
    Without the states, every row counts as a subspace row."""
    dev = _device(vec, xsources, diag_signs)
    xs = _on(xsources, dev, torch.int32)
    n = xs.shape[1]
    c = coeffs.cpu().numpy() if isinstance(coeffs, torch.Tensor) else np.asarray(coeffs)
    keys = torch.zeros(n, dtype=torch.int64, device=dev)
    out = _groups(keys, n, xs, None, _lib.COEF_SIGNS, coef=_coef(c, dev), kmax=c.shape[1],
                  signs=_on(diag_signs, dev, torch.uint8), vec=vec, dev=dev)
    return _result(out, vec, bool(np.iscomplexobj(c)))


def apply_h_xz_cached(vec, xsources, diagonals) -> torch.Tensor:
    """Hv from cached source indices and diagonals (rqutils/sqd.py:673-684).
    Without the states, every row counts as a subspace row."""
    dev = _device(vec, xsources, diagonals)
    xs = _on(xsources, dev, torch.int32)
    n = xs.shape[1]
    d = _on(diagonals, dev)
    keys = torch.zeros(n, dtype=torch.int64, device=dev)
    out = _groups(keys, n, xs, None, _lib.COEF_DIAG, diag=d.to(torch.complex128).contiguous(),
                  vec=vec, dev=dev)
    return _result(out, vec, d.is_complex())
