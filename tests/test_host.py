"""CPU tests of the host side: the C-ABI library loads and exports every
function include/sqdk.h declares (no compute call without a GPU), the Pauli
representation, key packing, the class tables and the workload generators."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import sqd_oracle as so
from tests.helpers import random_labels

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "sqdk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sqdk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_function():
    import torch  # noqa: F401  (the library binds to torch's HIP runtime)
    from rqutils_amd import _lib
    path = _lib.library_path()
    assert path.exists(), "build the library first: python -m rqutils_amd.build"
    handle = ctypes.CDLL(str(path))
    names = _header_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(handle, name), name
    # the ctypes binding declares exactly the header's functions
    assert sorted(_lib.SIGNATURES) == names


def test_abi_version_and_error_path_without_gpu():
    from rqutils_amd import _lib
    L = _lib.lib()
    assert L.sqdk_abi_version() == _lib.ABI_VERSION
    # argument validation happens before any HIP call
    rc = L.sqdk_dots_workspace(10, 99, None)
    assert rc == _lib.EINVAL
    with pytest.raises(ValueError):
        _lib.check(rc, "dots")
    assert b"NULL" in L.sqdk_last_error() or b"npairs" in L.sqdk_last_error()
    nbytes = ctypes.c_size_t(0)
    assert L.sqdk_dots_workspace(100, 4, ctypes.byref(nbytes)) == 0 and nbytes.value > 0


def test_paulisumxz_matches_oracle_packing():
    from rqutils_amd.paulis import PauliSumXZ
    rng = np.random.default_rng(0)
    for n, full in [(5, True), (8, False), (13, False), (16, True)]:
        labels = random_labels(rng, n, 40, 1, 4, full=full)
        coeffs = rng.normal(size=40)
        for pad in (False, True):
            h = PauliSumXZ.from_paulisum((labels, coeffs), force_real=True, add_padding=pad)
            x, z, c, nq = so.from_paulisum(labels, coeffs, force_real=True, add_padding=pad)
            assert np.array_equal(h.x, x) and np.array_equal(h.z, z)
            assert np.allclose(h.c, c, rtol=0, atol=0) and h.c.dtype == c.dtype
            assert h.num_qubits == nq


def test_paulisumxz_mask_path_matches_bit_path(monkeypatch):
    """Strings of I, X, Y, Z (either case, at most 63 letters) take the
    integer-mask construction; it must give the bit-array construction's
    arrays byte for byte, and the same masks, on inputs with repeated strings,
    zero and complex coefficients, lower case and one-letter strings."""
    import rqutils_amd.paulis.symplectic as sym
    from rqutils_amd.paulis import PauliSumXZ
    rng = np.random.default_rng(11)
    cases = []
    for n in (1, 3, 7, 8, 9, 31, 40, 63):
        labels = random_labels(rng, n, 60, 1, min(n, 4))
        labels = labels + labels[:15]                   # repeated strings are merged
        coeffs = rng.normal(size=len(labels))
        coeffs[::7] = 0.0                               # zero coefficients are dropped
        cases.append((labels, coeffs, True))
        cases.append(([l.lower() if i % 3 == 0 else l for i, l in enumerate(labels)],
                      coeffs + 1j * rng.normal(size=len(labels)), False))
    # merged coefficients that cancel leave a zero term in its group
    cases.append((["XZ", "ZX", "XZ", "YY", "XX"], np.array([1.0, 2.0, -1.0, 0.5, 0.25]), True))
    for labels, coeffs, force_real in cases:
        for pad in (False, True):
            fast = PauliSumXZ.from_paulisum((labels, coeffs), force_real=force_real,
                                            add_padding=pad)
            assert fast._mask_cache is not None
            with monkeypatch.context() as m:
                m.setattr(sym, "_pauli_codes", lambda labels, width: None)
                slow = PauliSumXZ.from_paulisum((labels, coeffs), force_real=force_real,
                                                add_padding=pad)
            assert slow._mask_cache is None
            assert fast.x.dtype == slow.x.dtype and np.array_equal(fast.x, slow.x)
            assert np.array_equal(fast.z, slow.z) and fast.z.shape == slow.z.shape
            assert fast.c.dtype == slow.c.dtype and np.array_equal(fast.c, slow.c)
            for a, b in zip(fast.masks(), slow.masks()):
                assert np.array_equal(a, b)
            # the grouped nonzero terms, kept by the mask construction and
            # derived from masks() otherwise
            assert fast._ragged_cache is not None and slow._ragged_cache is None
            for a, b in zip(fast.ragged(), slow.ragged()):
                assert a.dtype == b.dtype and np.array_equal(a, b)
    # other letters, and strings of 64 letters, keep the bit-array construction
    assert sym._pauli_codes(["XQZ"], 3) is None
    long = PauliSumXZ.from_paulisum((["X" * 64, "Z" * 64], [1.0, 2.0]), add_padding=True)
    assert long._mask_cache is None and long.x.shape == (2, 9)


def test_paulisumxz_errors_and_masks():
    from rqutils_amd.paulis import PauliSumXZ
    with pytest.raises(ValueError):
        PauliSumXZ.from_paulisum((["XX", "Y"], [1.0, 2.0]))
    with pytest.raises(ValueError):
        PauliSumXZ.from_paulisum((["XX"], [1.0, 2.0]))
    with pytest.raises(ValueError):
        PauliSumXZ.from_paulisum({"XX": 1.0})
    with pytest.warns(UserWarning):
        PauliSumXZ.from_paulisum((["XY"], [1.0 + 1.0j]), force_real=True)
    h = PauliSumXZ.from_paulisum((["XIZ", "YZI", "ZZZ"], [1.0, 2.0, 3.0]), add_padding=True)
    xm, zm, coef, count = h.masks()
    assert xm.tolist() == [0, 0b100]          # ZZZ group, then X on qubit 2
    assert count.tolist() == [1, 2]
    assert zm[0, 0] == 0b111
    # YZI: x=100, z=110 -> popcount(x&z)=1 -> coefficient 2 * (-i)
    assert 2.0 * -1j in coef[1].tolist()


def test_keys_round_trip_and_order():
    from rqutils_amd.operator import keys_to_states, states_to_keys
    rng = np.random.default_rng(1)
    for n in (3, 8, 17, 40, 64):
        states = rng.integers(0, 2, (200, n)).astype(np.uint8)
        keys = states_to_keys(states)
        assert np.array_equal(keys_to_states(keys, n), states)
        if n < 64:
            assert np.array_equal(keys, so.packed_to_keys(so.pack_states(states), n))


def _max_bucket_brute(keys, m):
    hi = [int(k) >> m for k in keys]
    best, run = 1, 1
    for a, b in zip(hi, hi[1:]):
        run = run + 1 if a == b else 1
        best = max(best, run)
    return best


def test_widest_bucket_bits_matches_brute_force():
    import torch
    from rqutils_amd.operator import Subspace
    rng = np.random.default_rng(3)
    cases = []
    for nq in (6, 12, 20, 40, 63, 64):
        for n in (1, 5, 40, 300):
            if n > 2 ** min(nq, 20):
                continue
            keys = np.unique(rng.integers(0, 2 ** min(nq, 62), 4 * n, dtype=np.uint64)
                             << np.uint64(max(0, nq - 62)))[:n]
            cases.append((nq, keys))
    # clustered keys, one near 2^64 (the unsigned comparison), consecutive runs
    cases.append((64, np.array([1, 2, 3, 2 ** 63, 2 ** 63 + 5, 2 ** 64 - 1], dtype=np.uint64)))
    cases.append((16, np.arange(1000, 1300, dtype=np.uint64)))
    for nq, keys in cases:
        sp = Subspace.__new__(Subspace)
        sp.keys = torch.from_numpy(keys.view(np.int64).copy())
        sp.n = keys.size
        for cap in (1, 2, 3, 7, 64, 65535):
            got = sp._widest_bits(cap)
            want = 0
            while want < 64 and _max_bucket_brute(keys, want + 1) <= cap:
                want += 1
            if keys.size <= cap:
                want = 64
            assert got == want, (nq, keys.size, cap, got, want)


class _FakeSpace:
    def __init__(self, m, parity_single):
        self.bucket_bits = m
        self._p = parity_single

    def single_parity(self):
        return self._p


def test_class_tables_order_and_pruning():
    from rqutils_amd.operator import PauliOperator
    from rqutils_amd.paulis import PauliSumXZ
    rng = np.random.default_rng(2)
    n = 12
    labels = random_labels(rng, n, 300, 1, 4)
    coeffs = rng.normal(size=300)
    h = PauliSumXZ.from_paulisum((labels, coeffs), force_real=True, add_padding=True)
    xm, zm, coef, count = h.masks()
    for m in (0, 4, 9, 12):
        t = PauliOperator._tables(h, _FakeSpace(m, False), prune=True)
        hi = t.xmask >> np.uint64(m)
        # classes partition the ordered groups by high part, ascending
        assert np.all(np.diff(hi.astype(np.int64)) >= 0)
        for c in range(t.class_hi.size):
            seg = hi[t.class_off[c]:t.class_off[c + 1]]
            assert np.all(seg == t.class_hi[c])
        assert t.class_off[-1] == t.xmask.size
        # every nonzero term is present exactly once
        assert t.zmask.size == int(np.count_nonzero(coef))
        assert t.diag_terms == int(np.count_nonzero(coef[xm == 0]))
        assert 0 not in t.xmask.tolist()
        # term ranges: off-diagonal groups start right after the X = 0 terms
        assert t.grp_off[0] == t.diag_terms and t.grp_off[-1] == t.zmask.size
        assert np.all(np.diff(t.grp_off) >= 1)
        for j, x in enumerate(t.xmask):
            g = int(np.nonzero(xm == x)[0][0])
            want = sorted(zm[g, :count[g]][coef[g, :count[g]] != 0].tolist())
            assert sorted(t.zmask[t.grp_off[j]:t.grp_off[j + 1]].tolist()) == want
    t = PauliOperator._tables(h, _FakeSpace(4, True), prune=True)
    w = np.array([bin(int(x)).count("1") for x in t.xmask])
    assert np.all(w % 2 == 0) and t.pruned_groups > 0


def test_workload_generators_small():
    from rqutils_amd import workloads as W
    w = W.small16()
    assert w.n == 4096 and np.all(np.diff(w.keys.astype(np.int64)) > 0)
    assert set(np.bitwise_count(w.keys).tolist()) == {8}
    assert len(w.labels) == 200 and all(1 <= sum(ch != "I" for ch in l) <= 4 for l in w.labels)
    k = W.domain_wall_keys(np.random.default_rng(0), 16, 8, 3000)
    assert k.size == 3000 and set(np.bitwise_count(k).tolist()) == {8}
    assert np.all(np.diff(k.astype(np.int64)) > 0)
    # low domain-wall states are over-represented against the uniform sector
    bits = ((k[:, None] >> np.arange(16, dtype=np.uint64)) & np.uint64(1)).astype(int)
    walls = np.count_nonzero(np.diff(bits, axis=1), axis=1)
    u = W.sector_uniform_keys(np.random.default_rng(0), 16, 8, 3000)
    ub = ((u[:, None] >> np.arange(16, dtype=np.uint64)) & np.uint64(1)).astype(int)
    assert walls.mean() < np.count_nonzero(np.diff(ub, axis=1), axis=1).mean()
    b = W.batch_instance(5)
    assert 20 <= b.num_qubits <= 32 and 2 ** 12 - 1 <= b.n <= 2 ** 20 + 1


def test_sqd_argument_errors():
    """Host-side errors of the reference's sqd() (rqutils/sqd.py:222-225):
    states_size below the number of rows, and an unknown cache_level."""
    from rqutils_amd.sqd import sqd
    states = np.array([[0, 1], [1, 0]], dtype=np.uint8)
    with pytest.raises(ValueError):
        sqd((["XX", "ZZ"], [1.0, 0.5]), states, states_size=1)
    with pytest.raises(ValueError):
        sqd((["XX", "ZZ"], [1.0, 0.5]), states, cache_level=(2, 0))


class _Bit:
    pass


class _Op:
    def __init__(self, name, params):
        self.name, self.params = name, list(params)


class _Inst:
    def __init__(self, name, qubits, params=()):
        self.operation, self.qubits = _Op(name, params), qubits


class _Circuit:
    """The parts of a qiskit QuantumCircuit that the conversion reads."""

    def __init__(self, n):
        self.num_qubits = n
        self.qubits = [_Bit() for _ in range(n)]
        self.data = []

    def find_bit(self, bit):
        class _Loc:
            index = self.qubits.index(bit)
        return _Loc()

    def add(self, name, qs, *params):
        self.data.append(_Inst(name, [self.qubits[q] for q in qs], params))


def test_svsim_qiskit_style_circuit_conversion():
    """A qiskit-shaped circuit converts to the same symplectic gates as the
    equivalent specifier list, and cz becomes e^{i pi/4} CZ (checked on the
    dense matrices of the converted gates; oracle conventions)."""
    from oracle import svsim_oracle as sv
    from rqutils_amd.svsim import to_circuitxz
    c = _Circuit(4)
    c.add("rx", [0], 0.3)
    c.add("cz", [1, 3])
    c.add("x", [2])
    c.add("rzz", [0, 2], -0.7)
    got = to_circuitxz(c)
    want = to_circuitxz([("rx", [0], 0.3), ("rzz", [1, 3], np.pi / 2), ("rz", [1], -np.pi / 2),
                         ("rz", [3], -np.pi / 2), ("x", [2]), ("rzz", [0, 2], -0.7)])
    assert got.num_qubits == 4
    for f in ("x", "z", "cos", "sin"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    # the cz expansion is CZ up to the global phase e^{i pi/4}
    m = np.eye(16, dtype=complex)
    for name, qs, ang in [("rzz", [1, 3], np.pi / 2), ("rz", [1], -np.pi / 2),
                          ("rz", [3], -np.pi / 2)]:
        m = sv.gate_matrix_dense(name, qs, ang, 4) @ m
    idx = np.arange(16)
    cz = np.diag(np.where(((idx >> 1) & 1) & ((idx >> 3) & 1), -1.0, 1.0)).astype(complex)
    ph = m[0, 0] / abs(m[0, 0])
    assert np.allclose(m, ph * cz, atol=1e-12)
