"""Write the golden fixtures under tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

No reference code is imported or run (the reference cannot run here: it needs
JAX and Python >= 3.12). The expected outputs come from computations that are
independent of the symplectic (X/Z mask) formulation under test:

* ``pauli_phases.npz``: Kronecker products of the textbook 2x2 Pauli matrices.
* ``small16.npz``: the 16-qubit, 200-term, 4,096-state workload. ``hv`` is the
  projected product computed column by column from the one-qubit actions
  X|b> = |1-b>, Y|b> = i(-1)^b |1-b>, Z|b> = (-1)^b |b>; ``xs`` are
  dictionary lookups of s ^ x; ``e0`` is the lowest eigenvalue of the
  projected matrix by dense ``numpy.linalg.eigh``.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from rqutils_amd.workloads import small16  # noqa: E402  (seeded generator only)

HERE = os.path.dirname(os.path.abspath(__file__))

_P = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
      "Z": np.array([[1, 0], [0, -1]])}


def kron_label(lab):
    m = np.ones((1, 1), complex)
    for ch in lab:
        m = np.kron(m, _P[ch])
    return m


def apply_term_columns(label, coef, keys, n):
    """Image of every basis column |s_j> under coef * P: (target keys, amplitudes)."""
    tgt = keys.copy()
    amp = np.full(keys.shape, complex(coef))
    for pos, ch in enumerate(label):
        q = n - 1 - pos
        bit = ((keys >> np.uint64(q)) & np.uint64(1)).astype(np.int64)
        if ch == "X":
            tgt ^= np.uint64(1) << np.uint64(q)
        elif ch == "Y":
            tgt ^= np.uint64(1) << np.uint64(q)
            amp *= 1j * (1 - 2 * bit)
        elif ch == "Z":
            amp *= (1 - 2 * bit)
    return tgt, amp


def projected_dense(labels, coeffs, keys, n):
    size = keys.size
    h = np.zeros((size, size), complex)
    for lab, c in zip(labels, coeffs):
        tgt, amp = apply_term_columns(lab, c, keys, n)
        row = np.searchsorted(keys, tgt)
        ok = (row < size) & (keys[np.minimum(row, size - 1)] == tgt)
        np.add.at(h, (row[ok], np.nonzero(ok)[0]), amp[ok])
    return h


def main():
    labels = ["X", "Y", "Z", "XY", "YX", "YY", "ZY", "YZ", "XZY", "YIY"]
    # matrices zero-padded to 8 x 8 so they stack into one array
    mats = np.stack([np.pad(kron_label(lab), ((0, 8 - 2 ** len(lab)), (0, 8 - 2 ** len(lab))))
                     for lab in labels])
    np.savez(os.path.join(HERE, "pauli_phases.npz"), labels=np.array(labels), mats=mats)
    w = small16()
    n = w.num_qubits
    h = projected_dense(w.labels, w.coeffs, w.keys, n)
    rng = np.random.default_rng(42)
    v = rng.normal(size=w.n) + 1j * rng.normal(size=w.n)
    hv = h @ v
    e0 = np.linalg.eigvalsh(h)[0]
    table = {int(k): i for i, k in enumerate(w.keys)}
    # three X groups: the first off-diagonal, a middle and the last, by ascending X mask
    xsig = sorted({int(sum(1 << (n - 1 - p) for p, ch in enumerate(l) if ch in "XY"))
                   for l in w.labels})
    picks = [1, len(xsig) // 2, len(xsig) - 1]
    xs = np.array([[table.get(int(k) ^ xsig[i], -1) for k in w.keys] for i in picks], np.int32)
    np.savez_compressed(os.path.join(HERE, "small16.npz"), n=n, labels=np.array(w.labels),
                        coeffs=w.coeffs, keys=w.keys, v=v, hv=hv, e0=e0, seed_x0=7,
                        xs=xs, xs_groups=np.array(picks), xs_masks=np.array(
                            [xsig[i] for i in picks], np.uint64))
    print("e0 =", e0, "hermitian:", np.abs(h - h.conj().T).max())


if __name__ == "__main__":
    main()
