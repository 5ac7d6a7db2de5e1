"""NumPy complex128 restatement of the reference's state-vector simulator (test oracle).

Test infrastructure only: the product never imports this module.

- ``to_circuitxz`` follows rqutils/svsim.py:130-194 on the gate-list path
  (the QuantumCircuit path of :138-160 needs qiskit, which is absent here):
  each gate becomes an X mask, a Z mask and ``(cos, sin) = (cos(-a/2),
  sin(-a/2))`` for a rotation by angle ``a``, or ``(0, -1)`` for the
  fixed gates x, y, z (:168-188).
- ``svsim`` follows rqutils/svsim.py:96-122 (``do_svsim``): every gate maps
  the state to ``i sin (-1)^{popcount(i & z)} psi[i ^ x] + cos psi[i]``,
  in circuit order (``lax.scan``); an integer initial state is a one-hot
  vector (:103-104).

The map is ``cos I + i sin Z^z X^x``. For the single-qubit y gate and ry it
uses ``Z X = i Y`` without the ``(-i)^{x.z}`` phase of the Pauli algebra, so
``y`` gives ``Y`` but ``ry(a)`` gives ``cos(a/2) I + sin(a/2) Y``, which is not
unitary; x and z give ``-i X`` and ``-i Z``. The oracle keeps these
conventions (they are the reference's outputs), and ``tests/test_oracle.py``
pins each gate against dense Kronecker matrices written out independently.
"""
from __future__ import annotations

import numpy as np

__all__ = ["GATE_SPECS", "to_circuitxz", "svsim", "gate_matrix_dense"]

# name -> (x bit, z bit, angle or "pi") per qubit, rqutils/svsim.py:168-184
GATE_SPECS = {
    "x": (1, 0, "pi"), "y": (1, 1, "pi"), "z": (0, 1, "pi"),
    "rx": (1, 0, None), "ry": (1, 1, None), "rz": (0, 1, None), "rzz": (0, 1, None),
}


def to_circuitxz(gates, num_qubits=None):
    """Gate list -> ``(x int64[G], z int64[G], cos f64[G], sin f64[G], num_qubits)``.

    A gate is ``(name, qubit or qubits)`` or ``(name, qubits, angle)``.
    ``num_qubits`` defaults to one more than the largest qubit index
    (rqutils/svsim.py:191-192). Unknown names raise ``ValueError`` (:185-186).
    """
    g = len(gates)
    x = np.zeros(g, np.int64)
    z = np.zeros(g, np.int64)
    cos = np.zeros(g, np.float64)
    sin = np.zeros(g, np.float64)
    qmax = 0
    for i, gate in enumerate(gates):
        name = gate[0]
        if name not in GATE_SPECS:
            raise ValueError(f"Unsupported gate name {name}")
        xb, zb, ang = GATE_SPECS[name]
        qubits = np.atleast_1d(np.asarray(gate[1], dtype=np.int64))
        x[i] = int(np.sum(np.int64(xb) << qubits))
        z[i] = int(np.sum(np.int64(zb) << qubits))
        if ang == "pi":
            sin[i] = -1.0
        else:
            a = float(gate[2])
            cos[i] = np.cos(-a * 0.5)
            sin[i] = np.sin(-a * 0.5)
        qmax = max(qmax, int(qubits.max()) + 1)
    return x, z, cos, sin, qmax if num_qubits is None else int(num_qubits)


def svsim(x, z, cos, sin, num_qubits, initial_state=0):
    """Apply the gates in order to the 2^n amplitudes (complex128)."""
    n = int(num_qubits)
    idx = np.arange(2 ** n, dtype=np.int64)
    if np.ndim(initial_state) == 0:
        psi = (idx == int(initial_state)).astype(np.complex128)
    else:
        psi = np.asarray(initial_state, dtype=np.complex128).copy()
    for xg, zg, c, s in zip(np.asarray(x), np.asarray(z), np.asarray(cos), np.asarray(sin)):
        par = np.zeros(idx.size, np.int64)
        zz = idx & np.int64(zg)
        while np.any(zz):
            par ^= zz & 1
            zz >>= 1
        signs = 1.0 - 2.0 * par
        xs = psi if xg == 0 else psi[idx ^ np.int64(xg)]
        out = 1j * s * signs * xs
        if c != 0.0:
            out = out + c * psi
        psi = out
    return psi


_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_I = np.eye(2, dtype=np.complex128)


def _on(n, ops):
    """Kronecker product over qubits n-1..0 of ``ops.get(q, I)``."""
    out = np.ones((1, 1), dtype=np.complex128)
    for q in range(n - 1, -1, -1):
        out = np.kron(out, ops.get(q, _I))
    return out


def gate_matrix_dense(name, qubits, angle, n):
    """Dense 2^n matrix of one gate, written from textbook matrices with the
    reference's conventions (module docstring), independent of the masks."""
    qubits = list(np.atleast_1d(qubits))
    eye = np.eye(2 ** n, dtype=np.complex128)
    if name in ("x", "z"):
        p = _on(n, {q: (_X if name == "x" else _Z) for q in qubits})
        return -1j * p
    if name == "y":
        # i sin prod_q (Z X)_q with sin = -1 and Z X = i Y: -i i^k Y^(x)k
        return (-1j) * (1j ** len(qubits)) * _on(n, {q: _Y for q in qubits})
    c, s = np.cos(angle / 2), np.sin(angle / 2)
    if name == "rx":
        return c * eye - 1j * s * _on(n, {q: _X for q in qubits})
    if name in ("rz", "rzz"):
        return c * eye - 1j * s * _on(n, {q: _Z for q in qubits})
    if name == "ry":
        # cos(a/2) I + i sin(-a/2) (Z X) = cos(a/2) I + sin(a/2) (Y per qubit, with i^k)
        zx = _on(n, {q: _Z @ _X for q in qubits})
        return c * eye - 1j * s * zx
    raise ValueError(name)
