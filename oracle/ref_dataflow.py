"""ctypes front end of the oracle's C restatement of the reference's matvec
dataflow (``ref_dataflow.c``, built by ``oracle/build.py`` into
``libsqdref.so``). Test and CPU-baseline infrastructure only: the product
never imports this module.

The tables come from the oracle's ``from_paulisum`` in the reference's packed
layout (groups ascending by X signature, each group's terms zero-padded to
Kmax as in rqutils/paulis/symplectic.py:98-115), turned into uint64 masks;
the C sums stop at a group's first zero coefficient as
rqutils/sqd.py:562-564 does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import sqd_oracle as so

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsqdref.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from .build import build
            build()
        L = C.CDLL(LIB)
        p = C.c_void_p
        L.ref_xsource.argtypes = [p, C.c_int64, C.c_int, C.c_uint64, p]
        L.ref_diagonal.argtypes = [p, C.c_int64, p, p, C.c_int32, C.c_int32, p]
        L.ref_matvec.argtypes = [p, C.c_int64, C.c_int, p, p, p, p, p, C.c_int32, p, p, p,
                                 C.c_int32]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class RefTables:
    """Group tables of a Hamiltonian over ``num_qubits`` qubits, reference order."""

    def __init__(self, labels, coeffs, num_qubits):
        x, z, c, n = so.from_paulisum(labels, coeffs, force_real=True, add_padding=True)
        assert n == num_qubits
        self.num_qubits = n
        j, kmax = c.shape
        self.xmask = np.ascontiguousarray(so.packed_to_keys(x, n))
        self.zmask = np.ascontiguousarray(so.packed_to_keys(z.reshape(j * kmax, -1), n))
        self.coef = np.ascontiguousarray(c.reshape(-1).astype(np.complex128))
        self.grp_off = np.arange(0, (j + 1) * kmax, kmax, dtype=np.int32)
        self.n_groups = j


def xsource(keys, num_qubits, x):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty(keys.size, np.int32)
    if lib().ref_xsource(_ptr(keys), keys.size, num_qubits + 1, int(x), _ptr(out)) != 0:
        raise MemoryError("ref_xsource")
    return out


def matvec(tables: RefTables, keys, v, groups=None, nthreads: int = 1, xsources=None):
    """sum over ``groups`` (default: all) of C_g * v[xsource_g]; returns
    (out complex128, threads used)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    v = np.ascontiguousarray(v, dtype=np.complex128)
    groups = np.arange(tables.n_groups, dtype=np.int32) if groups is None else \
        np.ascontiguousarray(groups, dtype=np.int32)
    if xsources is not None:
        xsources = np.ascontiguousarray(xsources, dtype=np.int32)
        assert xsources.shape == (groups.size, keys.size)
    out = np.empty(keys.size, np.complex128)
    rc = lib().ref_matvec(_ptr(keys), keys.size, tables.num_qubits + 1, _ptr(tables.xmask),
                          _ptr(tables.grp_off), _ptr(tables.zmask), _ptr(tables.coef),
                          _ptr(groups), groups.size, _ptr(xsources), _ptr(v), _ptr(out),
                          int(nthreads))
    if rc < 0:
        raise MemoryError("ref_matvec")
    return out, rc
