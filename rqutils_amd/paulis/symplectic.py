r"""Symplectic (X/Z-signature) representation of Pauli sums.

Drop-in counterpart of ``rqutils.paulis.symplectic`` (rqutils/paulis/
symplectic.py:42-150). A Pauli string is :math:`Q = (-i)^{x\cdot z} Z^z X^x`
with binary X and Z signatures; a Pauli sum is stored grouped by X signature
with phase-folded coefficients :math:`c\,(-i)^{\mathrm{popcount}(x\,\&\,z)}`.

``PauliSumXZ`` keeps the reference's packed ``uint8`` fields (``x``, ``z``,
``c``, ``num_qubits``) so code written against the reference sees the same
arrays, and adds ``masks()``, the uint64 integer view used by the device
operator (bit q of a mask = qubit q; character 0 of a Pauli string acts on
qubit n-1, the Qiskit order).
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass
from typing import Any

import numpy as np

try:  # optional, as in the reference (rqutils/paulis/symplectic.py:35-39)
    from qiskit.quantum_info import SparsePauliOp
    HAS_QISKIT = True
except ImportError:  # pragma: no cover - qiskit is not installed here
    SparsePauliOp = None
    HAS_QISKIT = False

__all__ = ["PauliSumXZ", "pack_rows", "unpack_rows", "phase_power"]

# (-i)^k, k = 0..3: the phase that folds Y = i X Z into the coefficient.
_MINUS_I_POWERS = np.array([1.0 + 0.0j, -1.0j, -1.0 + 0.0j, 1.0j])


def phase_power(xbits: np.ndarray, zbits: np.ndarray) -> np.ndarray:
    """``popcount(x & z) mod 4`` per row of 0/1 arrays."""
    return np.count_nonzero(np.logical_and(xbits, zbits), axis=-1) & 3


def pack_rows(bits: np.ndarray, pad: bool) -> np.ndarray:
    """Pack 0/1 rows big-endian into bytes, optionally behind one zero bit."""
    bits = np.asarray(bits, dtype=np.uint8)
    if pad:
        widths = [(0, 0)] * (bits.ndim - 1) + [(1, 0)]
        bits = np.pad(bits, widths)
    return np.packbits(bits, axis=-1)


def unpack_rows(packed: np.ndarray, num_qubits: int, pad: bool) -> np.ndarray:
    """Inverse of :func:`pack_rows` (returns 0/1 ``uint8`` rows)."""
    bits = np.unpackbits(np.asarray(packed, dtype=np.uint8), axis=-1)
    start = 1 if pad else 0
    return bits[..., start:start + num_qubits]


def _bits_to_u64(bits: np.ndarray) -> np.ndarray:
    """Integer value of 0/1 rows whose column 0 is the most significant bit."""
    bits = np.asarray(bits, dtype=np.uint8)
    n = bits.shape[-1]
    if n > 64:
        raise ValueError(f"{n} qubits exceed the 64-bit key width")
    lead = (-n) % 8
    widths = [(0, 0)] * (bits.ndim - 1) + [(lead, 0)]
    packed = np.packbits(np.pad(bits, widths), axis=-1)
    nbytes = packed.shape[-1]
    widths = [(0, 0)] * (packed.ndim - 1) + [(8 - nbytes, 0)]
    wide = np.ascontiguousarray(np.pad(packed, widths))
    return wide.view(">u8")[..., 0].astype(np.uint64)


@dataclass(frozen=True)
class PauliSumXZ:
    """Symplectic representation of a sum of Pauli strings.

    Fields as in rqutils/paulis/symplectic.py:42-49: ``x`` uint8[J, Bp] packed
    X signatures (one per group, ascending), ``z`` uint8[J, Kmax, Bp] packed Z
    signatures (zero-padded), ``c`` [J, Kmax] phased coefficients (float64 if
    all real, else complex128; zero-padded), ``num_qubits``.
    """

    x: np.ndarray
    z: np.ndarray
    c: np.ndarray
    num_qubits: int
    padding: bool = False  # packed signatures carry the leading pad bit

    @classmethod
    def from_paulisum(cls, paulisum: Any, force_real: bool = False,
                      add_padding: bool = False) -> "PauliSumXZ":
        """Group a Pauli sum by X signature (rqutils/paulis/symplectic.py:51-125).

        ``paulisum`` is ``(list[str], list[Number])`` or a qiskit
        ``SparsePauliOp``. Tuple input: zero coefficients are dropped, then
        repeated strings merged; strings are ordered lexicographically by
        letter. ``force_real`` keeps the real part (with a warning if an
        imaginary part was nonzero). ``add_padding`` prepends the zero pad bit
        that ``sqd`` uses to mark fill rows.
        """
        if isinstance(paulisum, tuple):
            labels, coeffs = paulisum
            if len(labels) != len(coeffs):
                raise ValueError("Lengths of Pauli and coeff lists do not match")
            if len(set(len(p) for p in labels)) != 1:
                raise ValueError("Pauli strings have non-uniform lengths")
            coeffs = np.asarray(coeffs)
            letters = np.array([list(p.upper()) for p in labels])
            keep = coeffs != 0
            letters, coeffs = letters[keep], coeffs[keep]
            letters, where = np.unique(letters, axis=0, return_inverse=True)
            merged = np.zeros(letters.shape[0], dtype=np.result_type(coeffs.dtype, np.float64))
            np.add.at(merged, np.ravel(where), coeffs)
            coeffs = merged
            xbits = (letters == "X") | (letters == "Y")
            zbits = (letters == "Y") | (letters == "Z")
            num_qubits = letters.shape[1]
        elif HAS_QISKIT and isinstance(paulisum, SparsePauliOp):  # pragma: no cover
            if not np.allclose(paulisum.coeffs.imag, 0.0):
                raise ValueError("Coefficients of Paulis must be real for the Hamiltonian to be"
                                 " Hermitian.")
            paulisum = paulisum.simplify()
            coeffs = paulisum.coeffs
            xbits = paulisum.paulis.x[:, ::-1]
            zbits = paulisum.paulis.z[:, ::-1]
            num_qubits = paulisum.num_qubits
        else:
            raise ValueError("Unsupported input type")

        if force_real:
            if np.any(np.imag(coeffs) != 0.0):
                warnings.warn("Found nonzero imaginary part when force_real=True")
            coeffs = np.real(coeffs)

        # Groups: distinct X signatures in ascending (lexicographic) order.
        xsig, group_of, sizes = np.unique(xbits, axis=0, return_inverse=True,
                                          return_counts=True)
        group_of = np.ravel(group_of)
        n_groups, kmax = xsig.shape[0], int(np.max(sizes))
        zsig = np.zeros((n_groups, kmax, num_qubits), dtype=np.uint8)
        phc = np.zeros((n_groups, kmax), dtype=np.complex128)
        order = np.argsort(group_of, kind="stable")
        starts = np.concatenate([[0], np.cumsum(sizes)])
        for g in range(n_groups):
            members = order[starts[g]:starts[g + 1]]
            zsig[g, : sizes[g]] = zbits[members]
            phc[g, : sizes[g]] = coeffs[members] * _MINUS_I_POWERS[
                phase_power(xsig[g][None, :], zbits[members])]
        if np.all(phc.imag == 0.0):
            phc = phc.real
        return cls(pack_rows(xsig, add_padding), pack_rows(zsig, add_padding), phc,
                   int(num_qubits), bool(add_padding))

    @property
    def arrays(self):
        return self.x, self.z, self.c

    def masks(self, padded: bool | None = None):
        """Integer view: ``(xmask u64[J], zmask u64[J, Kmax], coef c128[J, Kmax], count i32[J])``.

        Zero-padded slots have coefficient 0 and are excluded by ``count``
        (terms per group: the number of leading nonzero coefficients, or all
        slots up to the last nonzero one).
        """
        if padded is None:
            padded = self.padding
        xb = unpack_rows(self.x, self.num_qubits, padded)
        zb = unpack_rows(self.z, self.num_qubits, padded)
        xm = _bits_to_u64(xb)
        zm = _bits_to_u64(zb)
        coef = np.asarray(self.c, dtype=np.complex128)
        nz = coef != 0
        count = np.where(nz.any(axis=1),
                         coef.shape[1] - np.argmax(nz[:, ::-1], axis=1), 0).astype(np.int32)
        return xm, zm, coef, count

    def matmul(self, rhs):
        """Full-space product ``H @ rhs`` over all ``2**num_qubits`` basis states.

        Counterpart of rqutils/paulis/symplectic.py:131-150, evaluated by the
        device matvec with the whole space as the subspace (every source index
        exists, so the search always hits). ``rhs`` is a vector (numpy or a
        torch tensor); the result has the same type and device.
        """
        from ..operator import full_space_matmul
        return full_space_matmul(self, rhs)
