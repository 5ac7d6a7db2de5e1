"""Shared builders for the test suite: small seeded problems in both the
reference's packed layout (for the oracle) and the uint64-key layout."""
import numpy as np

from oracle import dense, sqd_oracle as so
from rqutils_amd.operator import keys_to_states


def random_labels(rng, n, count, wmin=1, wmax=None, full=False):
    wmax = n if wmax is None else wmax
    out = []
    for _ in range(count):
        if full:
            out.append("".join("IXYZ"[i] for i in rng.integers(0, 4, n)))
            continue
        w = int(rng.integers(wmin, wmax + 1))
        chars = ["I"] * n
        for q in rng.choice(n, w, replace=False):
            chars[q] = "XYZ"[rng.integers(0, 3)]
        out.append("".join(chars))
    return out


class OracleProblem:
    """A Hamiltonian + sorted distinct keys, with the oracle's packed view."""

    def __init__(self, labels, coeffs, keys, n):
        self.labels, self.coeffs, self.n = labels, np.asarray(coeffs), n
        self.keys = np.asarray(keys, dtype=np.uint64)
        self.x, self.z, self.c, _ = so.from_paulisum(labels, coeffs, force_real=True,
                                                     add_padding=True)
        self.states = keys_to_states(self.keys, n)
        self.packed = so.pack_states(self.states)
        self._xs = None

    @property
    def xsources(self):
        if self._xs is None:
            self._xs = np.stack([so.get_xsource(self.x[g], self.packed)
                                 for g in range(self.x.shape[0])])
        return self._xs

    def matvec(self, v):
        return so.apply_h_x_cached(v, self.xsources, self.z, self.c, self.packed)

    def dense(self):
        """Projected matrix from dense Kronecker products (independent; n <= 11)."""
        if self.n > 11:
            raise ValueError("dense Kronecker construction is limited to 11 qubits here")
        h = dense.dense_hamiltonian(self.labels, np.real(self.coeffs))
        idx = self.keys.astype(np.int64)
        return dense.projected(h, idx)

    def csr(self):
        """Projected matrix from the oracle's sort-join construction."""
        import scipy.sparse as sps
        rows, cols, vals = so.hproj_coo(self.x, self.z, self.c, self.packed)
        n = self.keys.size
        return sps.csr_matrix((vals, (rows, cols)), shape=(n, n))

    def ground(self):
        """Lowest eigenvalue of the projected matrix (sparse Lanczos, FP64)."""
        import scipy.sparse.linalg as spl
        h = self.csr().astype(np.complex128)
        return float(spl.eigsh(h, k=1, which="SA", tol=1e-13)[0][0])


def small_problem(seed=0, n=10, terms=60, nstates=300, full=False):
    rng = np.random.default_rng(seed)
    labels = random_labels(rng, n, terms, 1, 4, full=full)
    coeffs = rng.normal(size=terms)
    keys = np.sort(rng.choice(2 ** n, size=nstates, replace=False)).astype(np.uint64)
    return OracleProblem(labels, coeffs, keys, n)
