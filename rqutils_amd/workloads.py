"""Seeded synthetic SQD workloads (Pauli-sum Hamiltonians + sampled subspaces).

Each generator returns a :class:`Workload`: Pauli labels and coefficients (the
tuple input of ``sqd``; character 0 acts on qubit n-1) and the sorted distinct
uint64 keys of the subspace. All randomness comes from
``numpy.random.default_rng(seed)``.

* ``small16``: 16 qubits, 200 strings of weight 1-4 on random qubits (letters
  uniform in X/Y/Z), coefficients N(0, 1); 4,096 distinct keys uniform in the
  8-ones sector.
* ``heisenberg28``: 28-qubit chain, XX+YY+ZZ on the 27 nearest-neighbour bonds
  (J = 1) plus 2,000 random strings of weight 2-4 with coefficients
  N(0, 0.1); 2^22 distinct keys of the 14-ones sector drawn with probability
  proportional to exp(-0.5 * domain walls).
* ``random40``: 40 qubits, 20,000 strings of weight 2-4 whose X masks come
  from a pool of 5,000 masks, coefficients N(0, 1)/sqrt(weight); 2^24 keys
  uniform in the 20-ones sector.
* ``batch``: instances of 20-32 qubits, 500-5,000 strings of weight 2-4,
  half-filling subspaces of log-uniform size in [2^12, 2^20].
"""
from __future__ import annotations

from dataclasses import dataclass, field
from math import comb

import numpy as np

__all__ = ["Workload", "small16", "heisenberg28", "random40", "batch_instance", "make",
           "sector_uniform_keys", "domain_wall_keys", "random_pauli_strings"]


@dataclass
class Workload:
    name: str
    num_qubits: int
    labels: list
    coeffs: np.ndarray
    keys: np.ndarray          # uint64, sorted, distinct
    seed: int
    params: dict = field(default_factory=dict)

    @property
    def hamiltonian(self):
        return (self.labels, self.coeffs)

    @property
    def n(self) -> int:
        return int(self.keys.shape[0])


def _label(n: int, xbits: np.ndarray, zbits: np.ndarray) -> str:
    """Pauli label from per-qubit bits (index q = qubit q; char 0 = qubit n-1)."""
    chars = np.full(n, "I")
    chars[xbits & ~zbits] = "X"
    chars[xbits & zbits] = "Y"
    chars[~xbits & zbits] = "Z"
    return "".join(chars[::-1])


def random_pauli_strings(rng, n: int, count: int, wmin: int, wmax: int):
    """``count`` strings of uniform weight in [wmin, wmax] on distinct random
    qubits, letters uniform in X/Y/Z. Returns (labels, weights)."""
    labels, weights = [], []
    for _ in range(count):
        w = int(rng.integers(wmin, wmax + 1))
        sites = rng.choice(n, size=w, replace=False)
        letters = rng.integers(0, 3, size=w)        # 0 X, 1 Y, 2 Z
        xb = np.zeros(n, bool)
        zb = np.zeros(n, bool)
        xb[sites] = letters <= 1
        zb[sites] = letters >= 1
        labels.append(_label(n, xb, zb))
        weights.append(w)
    return labels, np.asarray(weights)


def _popcount(a: np.ndarray) -> np.ndarray:
    return np.bitwise_count(a.astype(np.uint64)).astype(np.int64)


def sector_uniform_keys(rng, n: int, ones: int, count: int) -> np.ndarray:
    """``count`` distinct keys drawn uniformly from the ``ones``-ones sector, sorted."""
    total = comb(n, ones)
    if count > total:
        raise ValueError(f"sector of {total} states cannot hold {count}")
    if total <= 4 * count or n <= 20:
        # Enumerate the sector and choose without replacement.
        allk = np.arange(1 << n, dtype=np.uint64)
        sector = allk[_popcount(allk) == ones]
        return np.sort(rng.choice(sector, size=count, replace=False))
    frac = total / float(2 ** n)
    got = np.zeros(0, np.uint64)
    mask = np.uint64((1 << n) - 1) if n < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    while got.size < count:
        need = count - got.size
        batch = int(min(1 << 24, max(1 << 16, 1.3 * need / frac)))
        cand = rng.integers(0, np.iinfo(np.uint64).max, size=batch, dtype=np.uint64,
                            endpoint=True) & mask
        cand = cand[_popcount(cand) == ones]
        got = np.unique(np.concatenate([got, cand]))
        if got.size > count:
            got = np.sort(rng.choice(got, size=count, replace=False))
    return got


def domain_wall_keys(rng, n: int, ones: int, count: int, beta: float = 0.5) -> np.ndarray:
    """``count`` distinct keys of the ``ones``-ones sector, each draw having
    probability proportional to exp(-beta * domain walls) on the open chain.

    Exact sampling of the accepted-draw law of rejection sampling from the
    uniform sector: a draw picks its wall count k and first bit b with weight
    (#states with k walls starting with b) * exp(-beta k), then uniform run
    lengths (a uniform composition of the ones and of the zeros). Draws are
    repeated, in batches, until ``count`` distinct keys are seen; the first
    ``count`` distinct keys in draw order are kept.
    """
    zeros = n - ones
    classes = []
    for k in range(0, n):
        for b in (0, 1):
            runs = k + 1
            r1 = (runs + 1) // 2 if b == 1 else runs // 2
            r0 = runs - r1
            if r1 < 1 or r0 < 0 or (zeros == 0) != (r0 == 0) or (ones == 0) != (r1 == 0):
                continue
            if r1 > ones or r0 > zeros:
                continue
            c = comb(ones - 1, r1 - 1) * (comb(zeros - 1, r0 - 1) if r0 else 1)
            if c:
                classes.append((k, b, r1, r0, c * np.exp(-beta * k)))
    w = np.array([c[4] for c in classes])
    w = w / w.sum()
    seen = np.zeros(0, np.uint64)
    first = np.zeros(0, np.int64)
    drawn = 0
    while seen.size < count:
        m = max(1 << 18, 2 * (count - seen.size))
        cls = rng.choice(len(classes), size=m, p=w)
        keys = np.zeros(m, np.uint64)
        for ci in np.unique(cls):
            sel = np.nonzero(cls == ci)[0]
            k, b, r1, r0, _ = classes[ci]
            l1 = _compositions(rng, ones, r1, sel.size)
            l0 = _compositions(rng, zeros, r0, sel.size) if r0 else np.zeros((sel.size, 0), int)
            keys[sel] = _runs_to_keys(l1, l0, b)
        allk = np.concatenate([seen, keys])
        order = np.concatenate([first, drawn + np.arange(m)])
        uk, idx = np.unique(allk, return_index=True)
        seen, first = uk, order[idx]
        drawn += m
    keep = np.argsort(first, kind="stable")[:count]
    return np.sort(seen[keep])


def _compositions(rng, total: int, parts: int, size: int) -> np.ndarray:
    """``size`` uniform compositions of ``total`` into ``parts`` positive parts."""
    if parts == 1:
        return np.full((size, 1), total)
    cuts = np.sort(np.argsort(rng.random((size, total - 1)), axis=1)[:, : parts - 1] + 1, axis=1)
    edges = np.concatenate([np.zeros((size, 1), int), cuts, np.full((size, 1), total)], axis=1)
    return np.diff(edges, axis=1)


def _runs_to_keys(l1: np.ndarray, l0: np.ndarray, first_bit: int) -> np.ndarray:
    """Keys from alternating run lengths starting at qubit 0 with ``first_bit``."""
    size = l1.shape[0]
    r1, r0 = l1.shape[1], l0.shape[1]
    runs = r1 + r0
    lens = np.zeros((size, runs), np.int64)
    bits = np.zeros(runs, np.int64)
    i1 = i0 = 0
    for r in range(runs):
        is_one = (r % 2 == 0) == (first_bit == 1)
        if is_one:
            lens[:, r] = l1[:, i1]
            i1 += 1
        else:
            lens[:, r] = l0[:, i0]
            i0 += 1
        bits[r] = 1 if is_one else 0
    start = np.cumsum(lens, axis=1) - lens
    keys = np.zeros(size, np.uint64)
    for r in range(runs):
        if bits[r]:
            block = (np.left_shift(np.uint64(1), lens[:, r].astype(np.uint64)) - np.uint64(1))
            keys |= block << start[:, r].astype(np.uint64)
    return keys


def small16(seed: int = 1) -> Workload:
    rng = np.random.default_rng(seed)
    n = 16
    labels, _ = random_pauli_strings(rng, n, 200, 1, 4)
    coeffs = rng.normal(size=200)
    keys = sector_uniform_keys(rng, n, 8, 4096)
    return Workload("small16", n, labels, coeffs, keys, seed,
                    {"terms": 200, "sector_ones": 8})


def heisenberg28(seed: int = 2, n_states: int | None = 1 << 22, n_random: int = 2000,
                 num_qubits: int = 28) -> Workload:
    """Heisenberg chain plus disorder; ``n_states=None`` (or a count at least
    the sector size) takes the whole half-filling sector."""
    rng = np.random.default_rng(seed)
    n = num_qubits
    labels, coeffs = [], []
    for q in range(n - 1):
        for p in "XYZ":
            chars = ["I"] * n
            chars[n - 1 - q] = p
            chars[n - 2 - q] = p
            labels.append("".join(chars))
            coeffs.append(1.0)
    rl, _ = random_pauli_strings(rng, n, n_random, 2, 4)
    labels += rl
    coeffs = np.concatenate([np.asarray(coeffs), rng.normal(0.0, 0.1, size=n_random)])
    sector = comb(n, n // 2)
    if n_states is None or n_states >= sector:
        if n > 24:
            raise ValueError("whole-sector enumeration is limited to 24 qubits")
        allk = np.arange(1 << n, dtype=np.uint64)
        keys = allk[_popcount(allk) == n // 2]
        n_states = sector
    else:
        keys = domain_wall_keys(rng, n, n // 2, n_states, beta=0.5)
    name = "heisenberg28" if (n, n_states) == (28, 1 << 22) else f"heisenberg{n}_{n_states}"
    return Workload(name, n, labels, coeffs, keys, seed,
                    {"bonds": n - 1, "random_terms": n_random, "sector_ones": n // 2,
                     "beta": 0.5, "states": n_states})


def random40(seed: int = 3, n_states: int = 1 << 24, n_terms: int = 20000,
             pool: int = 5000, num_qubits: int = 40) -> Workload:
    rng = np.random.default_rng(seed)
    n = num_qubits
    # X-mask pool: entry 0 is the identity mask (diagonal terms), the rest
    # have 1-4 X/Y sites on distinct random qubits.
    masks = np.zeros((pool, n), bool)
    for i in range(1, pool):
        wx = int(rng.integers(1, 5))
        masks[i, rng.choice(n, size=wx, replace=False)] = True
    labels, coeffs = [], []
    for _ in range(n_terms):
        xb = masks[int(rng.integers(0, pool))].copy()
        wx = int(xb.sum())
        w = int(rng.integers(max(2, wx), 5))
        zb = np.zeros(n, bool)
        zb[xb] = rng.integers(0, 2, size=wx).astype(bool)   # X or Y on the mask sites
        free = np.nonzero(~xb)[0]
        zb[rng.choice(free, size=w - wx, replace=False)] = True
        labels.append(_label(n, xb, zb))
        coeffs.append(rng.normal() / np.sqrt(w))
    keys = sector_uniform_keys(rng, n, n // 2, n_states)
    return Workload("random40", n, labels, np.asarray(coeffs), keys, seed,
                    {"terms": n_terms, "mask_pool": pool, "sector_ones": n // 2})


def batch_instance(seed: int) -> Workload:
    """One instance of the batched workload (qubits 20-32, 500-5000 terms)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 33))
    terms = int(rng.integers(500, 5001))
    size = int(round(2 ** rng.uniform(12, 20)))
    size = min(size, comb(n, n // 2))
    labels, _ = random_pauli_strings(rng, n, terms, 2, 4)
    coeffs = rng.normal(size=terms)
    keys = sector_uniform_keys(rng, n, n // 2, size)
    return Workload(f"batch{seed}", n, labels, coeffs, keys, seed,
                    {"terms": terms, "sector_ones": n // 2})


def ghz_circuit(n: int) -> list:
    """GHZ preparation (h on qubit 0, then cx from each qubit to the next, as in
    the reference's example, examples/svsim.py:55-58) written with rx, rz and
    rzz: h = rz(pi/2) rx(pi/2) rz(pi/2) and cx(c, t) = h_t cz(c, t) h_t with the
    reference's cz = rzz(pi/2) rz_c(-pi/2) rz_t(-pi/2) (rqutils/svsim.py:148-153),
    each up to a global phase. Ends in (|0...0> + |1...1>)/sqrt(2) up to a
    global phase."""
    h = lambda q: [("rz", q, np.pi / 2), ("rx", q, np.pi / 2), ("rz", q, np.pi / 2)]
    gates = h(0)
    for c in range(n - 1):
        t = c + 1
        gates += h(t)
        gates += [("rzz", [c, t], np.pi / 2), ("rz", c, -np.pi / 2), ("rz", t, -np.pi / 2)]
        gates += h(t)
    return gates


def layered_circuit(n: int, layers: int, seed: int = 0) -> list:
    """Seeded random layers: rx and rz on every qubit, then rzz on neighbours."""
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(layers):
        for q in range(n):
            gates.append(("rx", q, float(rng.uniform(-np.pi, np.pi))))
            gates.append(("rz", q, float(rng.uniform(-np.pi, np.pi))))
        for q in range(n - 1):
            gates.append(("rzz", [q, q + 1], float(rng.uniform(-np.pi, np.pi))))
    return gates


def make(name: str, **kw) -> Workload:
    return {"small16": small16, "heisenberg28": heisenberg28, "random40": random40}[name](**kw)
