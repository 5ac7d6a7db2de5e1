"""Edge cases of the drop-in boundary (rqutils/sqd.py:181-248): a one-state
subspace given as repeated rows, fill rows (states_size above the row count,
sqd.py:222-225) and the rejection of a smaller states_size, and the widest
keys (63 and 64 qubits, keys at and above 2^63, which order as unsigned
integers) against the oracle's sort-join product and its Lanczos energy."""
import numpy as np
import pytest
import torch

from oracle import dense
from rqutils_amd.operator import keys_to_states
from tests.helpers import OracleProblem, random_labels

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_sqd_one_state_subspace(gpu):
    from rqutils_amd.sqd import sqd
    rng = np.random.default_rng(31)
    n = 9
    labels = random_labels(rng, n, 60, 1, 4)
    coeffs = rng.normal(size=60)
    row = rng.integers(0, 2, n).astype(np.uint8)
    states = np.repeat(row[None, :], 5, axis=0)
    eigval, eigvec, basis = sqd((labels, coeffs), states)
    H = dense.dense_hamiltonian(labels, coeffs)
    idx = dense.states_to_index(row[None, :])
    assert basis.shape == (1, n) and np.array_equal(basis[0], row)
    assert eigvec.shape == (1,) and abs(abs(eigvec[0]) - 1) < 1e-12
    assert abs(eigval - np.real(H[idx[0], idx[0]])) < 1e-12 * max(1.0, abs(eigval))


def test_sqd_states_size_fill_and_rejection(gpu):
    from rqutils_amd.sqd import sqd
    rng = np.random.default_rng(32)
    n = 10
    labels = ["".join("IXYZ"[i] for i in r) for r in rng.integers(0, 4, (80, n))]
    coeffs = rng.uniform(-1, 1, 80)
    states = rng.integers(0, 2, (700, n)).astype(np.uint8)
    e0, v0, b0 = sqd((labels, coeffs), states)
    e1, v1, b1 = sqd((labels, coeffs), states, states_size=states.shape[0] + 1000)
    uniq = np.unique(states, axis=0)
    H = dense.dense_hamiltonian(labels, coeffs)
    w = np.linalg.eigvalsh(dense.projected(H, dense.states_to_index(uniq)))
    assert np.array_equal(b0, uniq) and np.array_equal(b1, uniq)
    assert v1.shape == (uniq.shape[0],)
    for e in (e0, e1):
        assert abs(e - w[0]) < 1e-9 * abs(w[0])
    with pytest.raises(ValueError):
        sqd((labels, coeffs), states, states_size=states.shape[0] - 1)


@pytest.mark.parametrize("n", [63, 64])
def test_widest_keys_matvec_and_energy(gpu, n):
    from rqutils_amd.operator import PauliOperator, Subspace
    from rqutils_amd.sqd import sqd
    rng = np.random.default_rng(n)
    labels = random_labels(rng, n, 120, 1, 3)
    coeffs = rng.normal(size=120)
    base = rng.integers(0, 2 ** 62, size=300, dtype=np.uint64) | (np.uint64(1) << np.uint64(n - 1))
    keys = np.unique(np.concatenate([base, base ^ np.uint64(0b110),
                                     base ^ (np.uint64(3) << np.uint64(n - 2))]))
    p = OracleProblem(labels, coeffs, keys, n)
    v = rng.normal(size=keys.size) + 1j * rng.normal(size=keys.size)
    ref = p.matvec(v)
    for dense_lookup in (False, None):
        sp = Subspace(keys, n, dense=dense_lookup)
        assert np.array_equal(sp.keys.cpu().numpy().view(np.uint64), keys)
        op = PauliOperator((labels, coeffs), sp)
        got = op.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
        assert _rel(got, ref) < 1e-12, dense_lookup
        # the cached hit table (64-bit bucket keys in the join that fills it)
        op.use_cache = True
        got = op.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
        op.use_cache = False
        assert _rel(got, ref) < 1e-12, ("cached", dense_lookup)
    # the whole path from unsorted rows with duplicates
    states = keys_to_states(keys, n)
    order = rng.permutation(np.concatenate([np.arange(keys.size), np.arange(50)]))
    e_ref = p.ground()
    e_join = sqd((labels, coeffs), states[order], cache_level=(0, 0), return_eigvec=False)
    assert abs(e_join - e_ref) < 1e-9 * abs(e_ref), ("join", e_join, e_ref)
    eigval, eigvec, basis = sqd((labels, coeffs), states[order])
    assert np.array_equal(basis, states)
    assert abs(eigval - e_ref) < 1e-9 * abs(e_ref), ("cached", eigval, e_ref)


@pytest.mark.parametrize("criterion", ["relative", "reference"])
def test_batch_with_a_one_state_instance(gpu, criterion):
    """A one-state instance in a batch converges at its diagonal element and
    leaves the other instances' solves unchanged."""
    from rqutils_amd.batch import SQDBatch
    from tests.helpers import small_problem
    one = small_problem(40, n=9, terms=50, nstates=1)
    probs = [small_problem(41, n=10, terms=60, nstates=300), one,
             small_problem(42, n=11, terms=70, nstates=500)]
    b = SQDBatch([((p.labels, p.coeffs), p.keys, p.n) for p in probs], bucket_cap=512)
    rng = np.random.default_rng(7)
    x0 = torch.from_numpy(rng.normal(size=b.n) + 1j * rng.normal(size=b.n)).cuda()
    res = b.solve(x0, precision="c128", criterion=criterion, tol=1e-8, maxiter=2000)
    for i, p in enumerate(probs):
        info = res.infos[i]
        ev = np.real(p.dense()[0, 0]) if p is one else p.ground()
        assert info.converged and np.isfinite(info.eigval), (i, info)
        # the reference criterion stops at |r| < (|Ax| - theta) N 10 tol (7e-4 here)
        etol = 1e-12 if p is one else (1e-9 if criterion == "relative" else 1e-6)
        assert abs(info.eigval - ev) <= etol * max(1.0, abs(ev)), (i, info.eigval, ev)
        x = b.segment(res.vectors, i).cpu().numpy()
        assert np.all(np.isfinite(x)), i


def test_sqd_empty_sample_is_an_error(gpu):
    from rqutils_amd.sqd import sqd
    with pytest.raises(ValueError, match="subspace size 0"):
        sqd((["XZI", "ZZZ"], [0.5, -1.0]), np.zeros((0, 3), dtype=np.uint8))
