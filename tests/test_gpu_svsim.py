"""GPU parity of the state-vector simulator (sqdk_svsim) against the oracle's
restatement of rqutils/svsim.py, on seeded random circuits of every gate type:
complex128 within 1e-12 relative, complex64 within 1e-5. Cases cover tiles
smaller than the state, gates on high qubits (several runs), runs longer
than one launch holds, gates whose X masks exceed a tile (global passes), and
norm preservation at a larger size."""
import numpy as np
import pytest
import torch

from oracle import svsim_oracle as sv

pytestmark = pytest.mark.gpu

NAMES = ["x", "y", "z", "rx", "ry", "rz", "rzz"]


def _circuit(rng, n, count, names=NAMES, maxq=3):
    gates = []
    for _ in range(count):
        name = names[rng.integers(0, len(names))]
        k = 2 if name == "rzz" else int(rng.integers(1, min(maxq, n) + 1))
        qubits = sorted(rng.choice(n, size=min(k, n), replace=False).tolist())
        if name == "rzz" and n < 2:
            name, qubits = "rz", [0]
        if name.startswith("r"):
            gates.append((name, qubits, float(rng.uniform(-np.pi, np.pi))))
        else:
            gates.append((name, qubits))
    return gates


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _run(gates, n, psi, precision="c128"):
    from rqutils_amd.svsim import svsim, to_circuitxz
    c = to_circuitxz(gates)
    c.num_qubits = n
    out, runs = svsim(c, psi, precision=precision, return_runs=True)
    return out.cpu().numpy(), runs


@pytest.mark.parametrize("n,count", [(1, 20), (3, 40), (6, 60), (10, 80), (13, 120), (16, 60)])
def test_svsim_matches_oracle(gpu, n, count):
    rng = np.random.default_rng(100 + n)
    gates = _circuit(rng, n, count)
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    ref = sv.svsim(*sv.to_circuitxz(gates, n), initial_state=psi)
    got, runs = _run(gates, n, psi)
    assert _rel(got, ref) < 1e-12, (n, runs)
    got32, _ = _run(gates, n, psi.astype(np.complex64), "c64")
    ref32 = sv.svsim(*sv.to_circuitxz(gates, n), initial_state=psi.astype(np.complex64))
    assert _rel(got32.astype(np.complex128), ref32) < 1e-5


def test_svsim_runs_split_and_global_passes(gpu):
    """18 qubits: single-qubit X rotations on every qubit need several tiles
    (complex128 tiles hold 5 qubits above the lowest 6); 300 low-qubit gates
    exceed one launch; an x on qubits 6..17 (12 high qubits) takes a global
    pass."""
    n = 18
    rng = np.random.default_rng(5)
    gates = [("rx", q, 0.1 * (q + 1)) for q in range(n)]
    gates += _circuit(rng, 6, 300)                       # low qubits only
    gates += [("x", list(range(6, 18)))]
    gates += [("ry", [2, 17], 0.7), ("rzz", [0, 17], 1.3), ("y", [16])]
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    ref = sv.svsim(*sv.to_circuitxz(gates, n), initial_state=psi)
    got, runs = _run(gates, n, psi)
    assert _rel(got, ref) < 1e-12
    # 18 rx: 12 high qubits in runs of at most 5 -> 3 runs; the 300 low gates
    # (many of them multi-qubit X gates applied in LDS, with register layout
    # changes around them) split into launches of at most 80 ops; the wide x
    # alone; the last three gates one more run. Far fewer passes than gates.
    assert 6 <= runs <= 16, runs


def test_svsim_one_hot_and_norm(gpu):
    """Integer starts are one-hot; without ry (the reference's non-unitary
    rotation) a 22-qubit circuit keeps the norm to rounding."""
    from rqutils_amd.svsim import svsim
    n = 22
    rng = np.random.default_rng(9)
    gates = _circuit(rng, n, 200, names=["x", "y", "z", "rx", "rz", "rzz"])
    out = svsim(gates + [("rz", n - 1, 0.0)], initial_state=12345).cpu().numpy()
    assert abs(np.linalg.norm(out) - 1.0) < 1e-12
    small = svsim([("x", 1)], initial_state=0).cpu().numpy()
    assert np.allclose(small, [0, 0, -1j, 0])
    # the one-hot start equals the explicit vector
    e = np.zeros(2 ** 10, complex)
    e[777] = 1.0
    g10 = _circuit(np.random.default_rng(2), 10, 50)
    a = svsim(g10 + [("rz", 9, 0.1)], initial_state=777).cpu().numpy()
    b = svsim(g10 + [("rz", 9, 0.1)], initial_state=e).cpu().numpy()
    assert np.array_equal(a, b)


def test_svsim_errors(gpu):
    from rqutils_amd.svsim import CircuitXZ, svsim
    with pytest.raises(ValueError):
        svsim([("cz", [0, 1])])
    with pytest.raises(ValueError):
        svsim([("rx", 0, 0.1)], out_sharding="x")
    with pytest.raises(ValueError):
        svsim([("rx", 2, 0.1)], initial_state=np.ones(4))
    bad = CircuitXZ(np.array([1 << 5]), np.array([0]), np.array([0.0]), np.array([-1.0]), 3)
    with pytest.raises(ValueError):
        svsim(bad)
