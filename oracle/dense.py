"""Dense Kronecker-product Pauli sums for small qubit counts (test oracle).

Test infrastructure only. Builds ``H = sum_k c_k P_k`` from 2x2 Pauli matrices
with no use of the symplectic (X/Z mask) representation, so it independently
pins the phase, sign and ordering conventions of the restated operator path.
The basis index is ``sum_q s_q 2^q`` and character 0 of a Pauli string acts on
qubit n-1 (the convention of rqutils/paulis/symplectic.py:86-87 and
examples/sqd.py:61-63).
"""
from __future__ import annotations

import numpy as np

__all__ = ["pauli_matrix", "dense_hamiltonian", "projected", "states_to_index"]

_P = {
    "I": np.eye(2, dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def pauli_matrix(label):
    """Kronecker product of single-qubit Paulis, leftmost factor most significant."""
    out = np.ones((1, 1), dtype=np.complex128)
    for ch in label.upper():
        out = np.kron(out, _P[ch])
    return out


def dense_hamiltonian(paulis, coeffs):
    """``sum_k c_k P_k`` as a dense complex matrix (n <= 12 or so)."""
    n = len(paulis[0])
    h = np.zeros((2 ** n, 2 ** n), dtype=np.complex128)
    for lab, c in zip(paulis, coeffs):
        h += c * pauli_matrix(lab)
    return h


def states_to_index(states):
    """Basis index of binary rows whose column 0 is qubit n-1."""
    states = np.asarray(states).astype(np.int64)
    n = states.shape[1]
    return states @ (1 << np.arange(n - 1, -1, -1, dtype=np.int64))


def projected(h, index):
    """``H`` restricted to the rows/columns listed in ``index``."""
    return h[np.ix_(index, index)]
