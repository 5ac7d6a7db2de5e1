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
from dataclasses import dataclass, field
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


# 2-bit codes of the Pauli letters (either case) in their lexicographic order;
# 255 marks any other byte.
_PAULI_CODE = np.full(256, 255, dtype=np.uint8)
for _code, _letter in enumerate("IXYZ"):
    _PAULI_CODE[ord(_letter)] = _code
    _PAULI_CODE[ord(_letter.lower())] = _code


def _pauli_codes(labels, width: int):
    """``[n, width]`` letter codes (I, X, Y, Z -> 0..3, either case) of ASCII
    Pauli strings, or None if a string holds any other character."""
    try:
        raw = "".join(labels).encode("ascii")
    except (UnicodeEncodeError, TypeError):
        return None
    if len(raw) != width * len(labels):
        return None
    codes = _PAULI_CODE[np.frombuffer(raw, dtype=np.uint8)]
    if np.any(codes == 255):
        return None
    return codes.reshape(len(labels), width)


def _codes_unique(codes: np.ndarray):
    """``(first, inverse)`` of the distinct rows of 2-bit letter codes in
    lexicographic order: four codes per byte, big-endian, compared as byte
    strings (code order is letter order, and every row is padded alike)."""
    n, width = codes.shape
    if width <= 32:
        # two bits per letter, the first letter highest: one integer per row
        # orders the rows like the byte strings below
        weights = np.uint64(1) << (np.uint64(2) * np.arange(width - 1, -1, -1, dtype=np.uint64))
        keys = (codes.astype(np.uint64) * weights).sum(axis=1, dtype=np.uint64)
        _, first, inverse = np.unique(keys, return_index=True, return_inverse=True)
        return first, np.ravel(inverse)
    cols = -(-width // 4) * 4
    c = np.zeros((n, cols), dtype=np.uint8)
    c[:, :width] = codes
    c = c.reshape(n, cols // 4, 4)
    packed = np.ascontiguousarray((c[..., 0] << 6) | (c[..., 1] << 4) | (c[..., 2] << 2) | c[..., 3])
    keys = packed.view(np.dtype((np.void, cols // 4))).ravel()
    _, first, inverse = np.unique(keys, return_index=True, return_inverse=True)
    return first, np.ravel(inverse)


def _masks_to_packed(masks: np.ndarray, num_qubits: int, pad: bool) -> np.ndarray:
    """The rows :func:`pack_rows` makes of the bit rows whose integer values
    are ``masks`` (column 0 the most significant bit), for rows of at most 64
    bits with the pad bit."""
    nbits = num_qubits + (1 if pad else 0)
    nbytes = -(-nbits // 8)
    value = np.asarray(masks, dtype=np.uint64) << np.uint64(8 * nbytes - nbits)
    be = np.ascontiguousarray(value.astype(">u8")).view(np.uint8)
    return np.ascontiguousarray(be.reshape(value.shape + (8,))[..., 8 - nbytes:])


def _letter_codes(labels) -> np.ndarray:
    """Upper-cased Pauli strings as an ``[n, L]`` array of character codes."""
    width = len(labels[0])
    if width == 0:
        raise ValueError("empty Pauli strings")
    text = np.array([p.upper() for p in labels], dtype=f"<U{width}")
    return text.view(np.uint32).reshape(len(labels), width)


def _unique_rows(rows: np.ndarray):
    """``(first, inverse)`` of the distinct rows in ascending lexicographic
    order (``np.unique(rows, axis=0)``'s order): the index of each distinct
    row's first occurrence and, per row, the rank of its distinct row. Rows
    are compared as big-endian byte strings (or as integers when they fit 64
    bits), which orders them element by element."""
    n, width = rows.shape
    if rows.dtype == np.uint8 and width <= 64 and np.all(rows <= 1):
        keys = _bits_to_u64(rows)
    else:
        big = np.ascontiguousarray(rows.astype(rows.dtype.newbyteorder(">")))
        keys = big.view(np.dtype((np.void, big.itemsize * width))).ravel()
    _, first, inverse = np.unique(keys, return_index=True, return_inverse=True)
    return first, np.ravel(inverse)


def _packed_to_u64(packed: np.ndarray, num_qubits: int, pad: bool) -> np.ndarray:
    """Masks of rows packed by :func:`pack_rows` without unpacking them: the
    packed bytes read as one big-endian integer hold the row's bits followed
    by the zero fill of the last byte."""
    packed = np.asarray(packed, dtype=np.uint8)
    nbits = num_qubits + (1 if pad else 0)
    nbytes = packed.shape[-1]
    if num_qubits > 64:
        raise ValueError(f"{num_qubits} qubits exceed the 64-bit key width")
    if nbytes > 8:
        return _bits_to_u64(unpack_rows(packed, num_qubits, pad))
    wide = np.zeros(packed.shape[:-1] + (8,), dtype=np.uint8)
    wide[..., 8 - nbytes:] = packed
    value = wide.view(">u8")[..., 0].astype(np.uint64) >> np.uint64(8 * nbytes - nbits)
    if num_qubits < 64:
        value &= np.uint64((1 << num_qubits) - 1)
    return value


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
    # (xmask u64[J], zmask u64[J, Kmax]) when the constructor already had them
    _mask_cache: Any = field(default=None, repr=False, compare=False)
    # the nonzero terms grouped (see ``ragged``) when the constructor had them
    _ragged_cache: Any = field(default=None, repr=False, compare=False)

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
            if len(set(map(len, labels))) != 1:
                raise ValueError("Pauli strings have non-uniform lengths")
            coeffs = np.asarray(coeffs)
            width = len(labels[0])
            codes = _pauli_codes(labels, width) if 0 < width <= 63 else None
            if codes is not None:
                return cls._from_codes(codes, coeffs, force_real, add_padding)
            letters = _letter_codes(labels)
            keep = coeffs != 0
            letters, coeffs = letters[keep], coeffs[keep]
            first, where = _unique_rows(letters)
            letters = letters[first]
            merged = np.zeros(letters.shape[0], dtype=np.result_type(coeffs.dtype, np.float64))
            np.add.at(merged, np.ravel(where), coeffs)
            coeffs = merged
            xbits = (letters == ord("X")) | (letters == ord("Y"))
            zbits = (letters == ord("Y")) | (letters == ord("Z"))
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

        # Groups: distinct X signatures in ascending (lexicographic) order; a
        # group's terms keep their order (lexicographic by letter).
        xbits = np.asarray(xbits, dtype=np.uint8)
        zbits = np.asarray(zbits, dtype=np.uint8)
        first_x, group_of = _unique_rows(xbits)
        xsig = xbits[first_x]
        sizes = np.bincount(group_of, minlength=xsig.shape[0])
        n_groups, kmax = xsig.shape[0], int(np.max(sizes))
        zsig = np.zeros((n_groups, kmax, num_qubits), dtype=np.uint8)
        phc = np.zeros((n_groups, kmax), dtype=np.complex128)
        order = np.argsort(group_of, kind="stable")
        starts = np.concatenate([[0], np.cumsum(sizes)])
        slot = np.empty(group_of.size, dtype=np.int64)
        slot[order] = np.arange(group_of.size) - starts[group_of[order]]
        zsig[group_of, slot] = zbits
        phc[group_of, slot] = coeffs * _MINUS_I_POWERS[phase_power(xbits, zbits)]
        if np.all(phc.imag == 0.0):
            phc = phc.real
        return cls(pack_rows(xsig, add_padding), pack_rows(zsig, add_padding), phc,
                   int(num_qubits), bool(add_padding))

    @classmethod
    def _from_codes(cls, codes: np.ndarray, coeffs: np.ndarray, force_real: bool,
                    add_padding: bool) -> "PauliSumXZ":
        """The tuple path of :meth:`from_paulisum` on integer masks, for strings
        of I, X, Y, Z of at most 63 letters: the same arrays, without building
        the [J, Kmax, n] bit array."""
        keep = coeffs != 0
        codes, coeffs = codes[keep], coeffs[keep]
        first, where = _codes_unique(codes)
        codes = codes[first]
        merged = np.zeros(codes.shape[0], dtype=np.result_type(coeffs.dtype, np.float64))
        np.add.at(merged, where, coeffs)
        coeffs = merged
        num_qubits = codes.shape[1]
        xm = _bits_to_u64((codes == 1) | (codes == 2))
        zm = _bits_to_u64((codes == 2) | (codes == 3))
        if force_real:
            if np.any(np.imag(coeffs) != 0.0):
                warnings.warn("Found nonzero imaginary part when force_real=True")
            coeffs = np.real(coeffs)
        # groups: distinct X masks ascending (the lexicographic order of the
        # signatures); a group's terms keep their order
        xsig, first_x, group_of = np.unique(xm, return_index=True, return_inverse=True)
        group_of = np.ravel(group_of)
        sizes = np.bincount(group_of, minlength=xsig.shape[0])
        n_groups, kmax = xsig.shape[0], int(np.max(sizes))
        order = np.argsort(group_of, kind="stable")
        starts = np.concatenate([[0], np.cumsum(sizes)])
        slot = np.empty(group_of.size, dtype=np.int64)
        slot[order] = np.arange(group_of.size) - starts[group_of[order]]
        # The padded [J, Kmax] arrays are mostly zero padding where one group
        # (the X = 0 terms, say) is large: they are scattered from the T
        # terms, and the ragged form is kept for the operator tables.
        zsig = np.zeros((n_groups, kmax), dtype=np.uint64)
        zsig[group_of, slot] = zm
        phase = np.bitwise_count(xm & zm).astype(np.int64) & 3
        phased = coeffs * _MINUS_I_POWERS[phase]
        real = bool(np.all(phased.imag == 0.0))
        phc = np.zeros((n_groups, kmax), dtype=np.float64 if real else np.complex128)
        phc[group_of, slot] = phased.real if real else phased
        packed_z = np.zeros((n_groups, kmax, -(-(num_qubits + add_padding) // 8)), np.uint8)
        packed_z[group_of, slot] = _masks_to_packed(zm, num_qubits, add_padding)
        tz, tc = zm[order], phased[order].astype(np.complex128)
        nonzero = tc != 0
        counts = np.bincount(group_of[order][nonzero], minlength=n_groups)
        offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        xsig = xsig.astype(np.uint64)
        return cls(_masks_to_packed(xsig, num_qubits, add_padding), packed_z, phc,
                   int(num_qubits), bool(add_padding), (xsig, zsig),
                   (xsig, offs, tz[nonzero], tc[nonzero]))

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
        if self._mask_cache is not None and padded == self.padding:
            xm, zm = self._mask_cache
        else:
            xm = _packed_to_u64(self.x, self.num_qubits, padded)
            zm = _packed_to_u64(self.z, self.num_qubits, padded)
        coef = np.asarray(self.c, dtype=np.complex128)
        nz = coef != 0
        count = np.where(nz.any(axis=1),
                         coef.shape[1] - np.argmax(nz[:, ::-1], axis=1), 0).astype(np.int32)
        return xm, zm, coef, count

    def ragged(self):
        """The nonzero terms grouped: ``(xmask u64[J], offs i64[J+1], zmask
        u64[T], coef c128[T])``; group g's terms are ``offs[g]:offs[g + 1]``, in
        slot order (the terms ``masks()`` counts, zero coefficients left out)."""
        if self._ragged_cache is not None:
            return self._ragged_cache
        xm, zm, coef, count = self.masks()
        slots = np.arange(coef.shape[1])
        valid = (slots[None, :] < count[:, None]) & (coef != 0)
        offs = np.concatenate([[0], np.cumsum(np.count_nonzero(valid, axis=1))]).astype(np.int64)
        return xm, offs, zm[valid], coef[valid]

    def matmul(self, rhs):
        """Full-space product ``H @ rhs`` over all ``2**num_qubits`` basis states.

        Counterpart of rqutils/paulis/symplectic.py:131-150, evaluated by the
        device matvec with the whole space as the subspace (every source index
        exists, so the search always hits). ``rhs`` is a vector (numpy or a
        torch tensor); the result has the same type and device.
        """
        from ..operator import full_space_matmul
        return full_space_matmul(self, rhs)
