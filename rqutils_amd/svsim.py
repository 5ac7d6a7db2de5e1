r"""State-vector simulation of rotation circuits (drop-in for ``rqutils.svsim``).

``svsim(circuit, initial_state=0, out_sharding=None)`` keeps the reference's
signature and results (rqutils/svsim.py:58-88): the circuit (a ``CircuitXZ``
or a list of gate specifiers ``(name, qubits)`` / ``(name, qubits, angle)``
with names x, y, z, rx, ry, rz, rzz) is applied gate by gate to the
2^num_qubits amplitudes, each gate mapping the state to
``cos psi + i sin Z^z X^x psi`` (rqutils/svsim.py:106-117). The gate
conventions are the reference's: x and z act as ``-i X`` and ``-i Z``, and ry
as ``cos(a/2) I + sin(a/2) Y`` (it omits the ``(-i)^{x.z}`` phase, so it is
not unitary).

The engine differs from the reference's ``lax.scan`` of one gather pass per
gate: the library groups consecutive gates whose X masks fit one LDS tile
and applies each group in a single pass over HBM (``sqdk_svsim``). The state
is a torch complex128 (or complex64) tensor on one GPU; ``out_sharding`` must
be None (one device).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib

__all__ = ["CircuitXZ", "to_circuitxz", "svsim", "apply", "GATE_NAMES"]

# name -> (X bit, Z bit, fixed-gate flag) per qubit (rqutils/svsim.py:168-184)
_SPECS = {
    "x": (1, 0, True), "y": (1, 1, True), "z": (0, 1, True),
    "rx": (1, 0, False), "ry": (1, 1, False), "rz": (0, 1, False), "rzz": (0, 1, False),
}
GATE_NAMES = tuple(_SPECS)


@dataclass
class CircuitXZ:
    """Symplectic (XZ) form of a series of rotation gates (rqutils/svsim.py:41-49)."""
    x: np.ndarray
    z: np.ndarray
    cos: np.ndarray
    sin: np.ndarray
    num_qubits: int


def _qiskit_specs(circuit):
    """Gate specifiers of a circuit given as a sequence of instructions (a
    qiskit ``QuantumCircuit`` or anything with its ``data``, ``find_bit`` and
    ``num_qubits``): each instruction's name, qubit indices and first
    parameter. ``cz`` has no rotation form in the gate table, so it becomes
    rzz(pi/2) rz(-pi/2) rz(-pi/2) on its two qubits, which equals
    e^{i pi/4} CZ; the global phase is dropped, as the reference drops it
    (rqutils/svsim.py:147-152)."""
    specs = []
    for inst in circuit.data:
        name = inst.operation.name
        where = [circuit.find_bit(q).index for q in inst.qubits]
        if name == "cz":
            a, b = where
            specs.append(("rzz", [a, b], np.pi / 2.0))
            specs.append(("rz", [a], -np.pi / 2.0))
            specs.append(("rz", [b], -np.pi / 2.0))
        else:
            specs.append((name, where) + tuple(inst.operation.params[:1]))
    return specs, int(circuit.num_qubits)


def to_circuitxz(circuit) -> CircuitXZ:
    """Gate specifiers to X/Z masks and (cos, sin) of half the negated angle
    (rqutils/svsim.py:130-194); fixed gates x, y, z get (0, -1). The number of
    qubits is one more than the largest index used, or the circuit's own for
    a qiskit-style circuit (``_qiskit_specs``; rqutils/svsim.py:138-160)."""
    if isinstance(circuit, CircuitXZ):
        return circuit
    num_qubits = None
    if hasattr(circuit, "data") and hasattr(circuit, "find_bit"):
        circuit, num_qubits = _qiskit_specs(circuit)
    g = len(circuit)
    x = np.zeros(g, np.int64)
    z = np.zeros(g, np.int64)
    cos = np.zeros(g, np.float64)
    sin = np.zeros(g, np.float64)
    qmax = 0
    for i, gate in enumerate(circuit):
        spec = _SPECS.get(gate[0])
        if spec is None:
            raise ValueError(f"Unsupported gate name {gate[0]}")
        qubits = np.atleast_1d(np.asarray(gate[1], dtype=np.int64))
        if qubits.size == 0 or qubits.min() < 0:
            raise ValueError(f"gate {i} has no valid qubit index")
        x[i] = int(np.sum(np.int64(spec[0]) << qubits))
        z[i] = int(np.sum(np.int64(spec[1]) << qubits))
        if spec[2]:
            sin[i] = -1.0
        else:
            cos[i] = np.cos(-float(gate[2]) * 0.5)
            sin[i] = np.sin(-float(gate[2]) * 0.5)
        qmax = max(qmax, int(qubits.max()) + 1)
    return CircuitXZ(x, z, cos, sin, qmax if num_qubits is None else int(num_qubits))


def _gates(c: CircuitXZ):
    n = len(c.x)
    arr = (_lib.SqdkGate * max(1, n))()
    for i in range(n):
        arr[i].x = int(c.x[i]) & ((1 << 64) - 1)
        arr[i].z = int(c.z[i]) & ((1 << 64) - 1)
        arr[i].cos = float(c.cos[i])
        arr[i].sin = float(c.sin[i])
    return arr, n


def svsim(circuit, initial_state=0, out_sharding=None, *, device=None,
          precision: str = "c128", return_runs: bool = False):
    """Simulate the circuit; returns the final state vector (torch, on the GPU).

    ``initial_state``: a one-hot index (int) or 2^n amplitudes. ``precision``:
    "c128" (the reference's x64 result) or "c64". With ``return_runs`` the
    number of passes over the state is returned too."""
    if out_sharding is not None:
        raise ValueError("out_sharding: this build simulates on one GPU (None only)")
    c = to_circuitxz(circuit)
    n = int(c.num_qubits)
    if not 1 <= n <= 40:
        raise ValueError(f"num_qubits {n} outside 1..40")
    if precision not in ("c128", "c64"):
        raise ValueError("precision must be 'c128' or 'c64'")
    dev = torch.device(device if device is not None else "cuda")
    dtype = torch.complex128 if precision == "c128" else torch.complex64
    size = 1 << n
    if isinstance(initial_state, (int, np.integer)) or np.ndim(initial_state) == 0:
        k = int(initial_state)
        if not 0 <= k < size:
            raise ValueError(f"initial_state index {k} outside 0..2^{n}-1")
        state = torch.zeros(size, dtype=dtype, device=dev)
        state[k] = 1.0
    else:
        src = initial_state if isinstance(initial_state, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(np.asarray(initial_state, np.complex128)))
        if src.numel() != size:
            raise ValueError(f"initial_state has {src.numel()} amplitudes, 2^{n} needed")
        state = src.reshape(-1).to(device=dev, dtype=dtype).clone()
    runs = apply(state, c)
    return (state, runs) if return_runs else state


def apply(state: torch.Tensor, circuit) -> int:
    """Apply the circuit in place to a contiguous complex128 or complex64 state
    of 2^num_qubits amplitudes on the GPU (the loop of do_svsim,
    rqutils/svsim.py:106-122). Returns the number of passes over the state."""
    c = to_circuitxz(circuit)
    n = int(c.num_qubits)
    if state.dtype not in (torch.complex128, torch.complex64) or not state.is_contiguous():
        raise ValueError("state must be a contiguous complex128 or complex64 tensor")
    if state.numel() != (1 << n):
        raise ValueError(f"state has {state.numel()} amplitudes, 2^{n} needed")
    arr, ng = _gates(c)
    runs = C.c_int32(0)
    prec = _lib.C128 if state.dtype == torch.complex128 else _lib.C64
    _lib.check(_lib.lib().sqdk_svsim(n, prec, state.data_ptr(), C.cast(arr, C.c_void_p), ng,
                                     C.byref(runs), _lib.stream_ptr(state.device)),
               "sqdk_svsim")
    return runs.value
