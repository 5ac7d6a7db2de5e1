"""GPU parity of the HIP path against the oracle (small, seeded cases).

Integer work (source indices) must match bit for bit; the complex128 matvec
within 1e-12 relative (different summation order only); complex64 within
2e-6 relative; eigenvalues within 1e-9 relative of dense eigh.
"""
import numpy as np
import pytest
import torch

from oracle import dense, locg_oracle as lo, sqd_oracle as so
from tests.helpers import OracleProblem, random_labels, small_problem

pytestmark = pytest.mark.gpu


def _problem_on_gpu(p):
    from rqutils_amd.sqd import SQDProblem
    return SQDProblem((p.labels, p.coeffs), p.keys, p.n)


def _rand_vec(rng, n):
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_xsource_bit_exact(gpu, seed):
    p = small_problem(seed, n=12, terms=50, nstates=700)
    prob = _problem_on_gpu(p)
    for g in range(p.x.shape[0]):
        ref = so.get_xsource(p.x[g], p.packed)
        xm = int(so.packed_to_keys(p.x[g][None, :], p.n)[0])
        got = prob.space.xsource(xm).cpu().numpy()
        assert np.array_equal(got, ref), f"group {g} x={xm:#x}"


def test_xsource_full_space_identity(gpu):
    # S = the whole space: xsource[i] = i ^ x (rqutils/sqd.py:461-463).
    from rqutils_amd.operator import Subspace
    n = 14
    sp = Subspace(np.arange(2 ** n, dtype=np.uint64), n)
    for x in [0, 1, 0x2A5, 2 ** n - 1]:
        got = sp.xsource(x).cpu().numpy()
        assert np.array_equal(got, np.arange(2 ** n) ^ x)


@pytest.mark.parametrize("seed", [0, 3])
def test_matvec_c128_vs_oracle_and_dense(gpu, seed):
    p = small_problem(seed, n=11, terms=80, nstates=900)
    prob = _problem_on_gpu(p)
    rng = np.random.default_rng(100 + seed)
    v = _rand_vec(rng, p.keys.size)
    ref = p.matvec(v)
    got = prob.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
    assert _rel(got, ref) < 1e-12
    assert _rel(got, p.dense() @ v) < 1e-12


def test_matvec_c64_compensated(gpu):
    p = small_problem(5, n=12, terms=120, nstates=1500)
    prob = _problem_on_gpu(p)
    rng = np.random.default_rng(7)
    v = _rand_vec(rng, p.keys.size).astype(np.complex64)
    ref = p.matvec(v.astype(np.complex128))
    got = prob.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
    assert _rel(got.astype(np.complex128), ref) < 2e-6


def test_matvec_many_buckets_and_classes(gpu):
    # A tiny bucket cap forces many buckets, classes and multi-chunk rows.
    from rqutils_amd.operator import PauliOperator, Subspace
    p = small_problem(11, n=14, terms=150, nstates=6000)
    for cap, rows, stage_v in [(64, 4, 0), (512, 8, 1), (4096, 16, 0), (300, 4, 1)]:
        sp = Subspace(p.keys, p.n, bucket_cap=cap, rows_per_thread=rows)
        op = PauliOperator((p.labels, p.coeffs), sp)
        op.tune.stage_v = stage_v
        rng = np.random.default_rng(cap)
        v = _rand_vec(rng, p.keys.size)
        got = op.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
        assert _rel(got, p.matvec(v)) < 1e-12, (cap, rows, stage_v, sp.bucket_bits, sp.n_items)
        v32 = v.astype(np.complex64)
        got32 = op.matvec(torch.from_numpy(v32).cuda()).cpu().numpy()
        assert _rel(got32.astype(np.complex128), p.matvec(v32.astype(np.complex128))) < 2e-6


def test_matvec_wide_keys_and_fine_bits(gpu):
    # bucket_bits > 32 (64-bit LDS keys) and explicit fine-directory sizes
    from rqutils_amd.operator import PauliOperator, Subspace
    rng = np.random.default_rng(17)
    n = 48
    labels = random_labels(rng, n, 200, 1, 3)
    coeffs = rng.normal(size=200)
    base = np.sort(rng.choice(2 ** 14, 400, replace=False)).astype(np.uint64)
    # a structured set closed under some flips so that hits exist
    keys = np.unique(np.concatenate([base, base ^ np.uint64(0b11), base << np.uint64(30)]))
    p = OracleProblem(labels, coeffs, keys, n)
    v = _rand_vec(rng, keys.size)
    ref = p.matvec(v)
    for fine in (0, 3, 9):
        sp = Subspace(keys, n, fine_bits=fine)
        assert sp.bucket_bits > 32
        op = PauliOperator((labels, coeffs), sp)
        got = op.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
        assert _rel(got, ref) < 1e-12, fine


@pytest.mark.parametrize("bucket_bits", [14, 30, 31, 32, 33])
def test_matvec_search_padding_and_clamps(gpu, bucket_bits):
    """The search's padded key slice: 32-bit LDS keys below 32 bucket bits,
    64-bit from 32 on (a target may then equal the 32-bit padding value);
    directories from one sub-bucket per bucket (every lifting stride of 64
    keys or more clamped) to the default, all giving the same products."""
    from rqutils_amd.operator import PauliOperator, Subspace
    rng = np.random.default_rng(bucket_bits)
    n = 40
    labels = random_labels(rng, n, 150, 1, 3)
    coeffs = rng.normal(size=150)
    if bucket_bits == 14:
        base = rng.choice(2 ** 16, 3000, replace=False).astype(np.uint64)
    else:
        base = rng.integers(0, 2 ** n, size=1500, dtype=np.uint64)
    x1 = np.uint64(int("".join(rng.choice(list("01"), n)), 2))
    keys = np.unique(np.concatenate([base, base ^ np.uint64(0b101), base ^ x1]))
    keys = keys[keys < np.uint64(2 ** n)]
    p = OracleProblem(labels + ["".join("X" if (int(x1) >> (n - 1 - q)) & 1 else "I"
                                        for q in range(n))],
                      np.concatenate([coeffs, [0.7]]), keys, n)
    v = _rand_vec(rng, keys.size)
    ref = p.matvec(v)
    outs = []
    for fine in (0, 1, 3, None):
        sp = Subspace(keys, n, bucket_bits=bucket_bits, fine_bits=fine)
        op = PauliOperator((p.labels, p.coeffs), sp)
        got = op.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
        assert _rel(got, ref) < 1e-12, (fine, sp.max_bucket)
        outs.append(got)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_diagonal_group0(gpu):
    p = small_problem(4, n=10, terms=60, nstates=400)
    prob = _problem_on_gpu(p)
    assert np.all(p.x[0] == 0)
    ref = so.get_diagonal(p.z[0], p.c[0], p.packed)
    got = prob.op.diagonal_of_group0().cpu().numpy()
    assert np.allclose(got, ref, rtol=0, atol=1e-12)


def test_small_eigenpairs_match_eigh(gpu):
    from rqutils_amd.ground_locg import eigenpair_2x2, eigenpair_3x3
    rng = np.random.default_rng(3)
    for _ in range(5):
        a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        a = a + a.conj().T
        w, V = np.linalg.eigh(a)
        lam, vec = eigenpair_3x3(a)
        assert abs(lam - w[0]) < 1e-12 * max(1, abs(w[0]))
        assert abs(abs(np.vdot(V[:, 0], vec)) - 1) < 1e-10
        lam_o, vec_o = lo.eigenpair_3x3(a)
        assert abs(lam - lam_o) < 1e-13 * max(1, abs(lam_o))
        b = a[:2, :2]
        w, V = np.linalg.eigh(b)
        lam, vec = eigenpair_2x2(b)
        assert abs(lam - w[0]) < 1e-12 * max(1, abs(w[0]))
        assert abs(abs(np.vdot(V[:, 0], vec)) - 1) < 1e-10


def test_locg_reference_mode_tracks_oracle(gpu):
    p = small_problem(8, n=10, terms=70, nstates=500)
    prob = _problem_on_gpu(p)
    rng = np.random.default_rng(9)
    x0 = _rand_vec(rng, p.keys.size)
    th_o, _, it_o = lo.ground_locg(p.matvec, x0)
    vec, info = prob.solve(torch.from_numpy(x0).cuda(), mode="reference", precision="c128",
                           criterion="reference", tol=float(np.finfo(np.float64).eps),
                           maxiter=1000)
    ev = np.linalg.eigvalsh(p.dense())[0]
    assert info.converged
    assert abs(info.eigval - ev) < 1e-9 * abs(ev)
    assert abs(info.eigval - th_o) < 1e-9 * abs(ev)
    assert abs(info.niter - it_o) <= max(2, 0.05 * it_o), (info.niter, it_o)
    assert info.matvecs == 4 * info.niter + 4


@pytest.mark.parametrize("precision,tol", [("c128", 1e-10), ("c64", 1e-6)])
def test_locg_fused_converges(gpu, precision, tol):
    p = small_problem(12, n=12, terms=120, nstates=2000)
    prob = _problem_on_gpu(p)
    rng = np.random.default_rng(13)
    x0 = _rand_vec(rng, p.keys.size)
    vec, info = prob.solve(torch.from_numpy(x0).cuda(), mode="fused", precision=precision,
                           criterion="relative", tol=tol, maxiter=2000)
    H = p.csr()
    ev = p.ground()
    x = vec.cpu().numpy().astype(np.complex128)
    x /= np.linalg.norm(x)
    r = np.linalg.norm(H @ x - info.eigval * x) / abs(info.eigval)
    assert info.converged
    assert r <= 1.5 * tol
    assert abs(info.eigval - ev) / abs(ev) < max(10 * tol * tol, 1e-12) + 1e-7 * (precision == "c64")


def test_sqd_end_to_end_vs_dense(gpu):
    # The reference notebook's check (tests/sqd.ipynb), seeded and at 10 qubits
    # so that the dense Kronecker oracle stays fast: 100 full-weight terms.
    from rqutils_amd.sqd import sqd
    rng = np.random.default_rng(2024)
    n = 10
    labels = ["".join("IXYZ"[i] for i in row) for row in rng.integers(0, 4, (100, n))]
    coeffs = rng.uniform(-1, 1, 100)
    states = rng.integers(0, 2, (1500, n)).astype(np.uint8)
    eigval, eigvec, basis = sqd((labels, coeffs), states)
    uniq = np.unique(states, axis=0)
    assert np.array_equal(basis, uniq)
    H = dense.dense_hamiltonian(labels, coeffs)
    idx = dense.states_to_index(uniq)
    w, V = np.linalg.eigh(dense.projected(H, idx))
    assert abs(eigval - w[0]) < 1e-9 * abs(w[0])
    assert abs(abs(np.vdot(V[:, 0], eigvec)) - 1) < 1e-6


def test_hproj_matches_dense(gpu):
    from rqutils_amd.sqd import hproj
    p = small_problem(21, n=9, terms=40, nstates=200)
    states = p.states
    m = hproj((p.labels, p.coeffs), states).toarray()
    assert np.allclose(m, p.dense(), atol=1e-12)


@pytest.mark.parametrize("seed,n,nstates,full", [(22, 12, 3000, False), (23, 11, 1500, True)])
def test_hproj_matches_oracle_csr(gpu, seed, n, nstates, full):
    """hproj (unpacked from the hit table) against the oracle's sort-join CSR,
    with repeated input rows; full-weight strings give complex coefficients."""
    from rqutils_amd.sqd import hproj
    p = small_problem(seed, n=n, terms=120, nstates=nstates, full=full)
    rng = np.random.default_rng(seed)
    states = p.states[rng.integers(0, p.states.shape[0], size=2 * p.states.shape[0])]
    states = np.concatenate([p.states, states])
    m = hproj((p.labels, p.coeffs), states)
    ref = p.csr()
    assert m.shape == ref.shape
    assert abs(m - ref).max() <= 1e-12 * max(1.0, abs(ref).max())
    # unique_states=True with the distinct sorted rows gives the same matrix
    m2 = hproj((p.labels, p.coeffs), p.states, unique_states=True)
    assert abs(m2 - m).max() == 0


def test_full_space_matmul_vs_dense(gpu):
    from rqutils_amd.paulis import PauliSumXZ
    rng = np.random.default_rng(5)
    n = 8
    labels = random_labels(rng, n, 30, full=True)
    coeffs = rng.normal(size=30)
    h = PauliSumXZ.from_paulisum((labels, coeffs))
    v = _rand_vec(rng, 2 ** n)
    got = h.matmul(v)
    assert np.allclose(got, dense.dense_hamiltonian(labels, coeffs) @ v, atol=1e-12)


def test_ground_locg_dense_and_callable(gpu):
    from rqutils_amd.ground_locg import ground_locg
    rng = np.random.default_rng(6)
    a = rng.normal(size=(60, 60))
    a = a + a.T
    w = np.linalg.eigvalsh(a)
    x0 = rng.normal(size=60)
    ev, vec, it = ground_locg(a, x0)
    assert abs(ev - w[0]) < 1e-10 * abs(w[0])
    assert vec.dtype == torch.float64
    at = torch.from_numpy(a).cuda()
    ev2, vec2, it2 = ground_locg(lambda v, s: s * torch.mv(at, v), x0, args=(1.0,))
    assert abs(ev2 - w[0]) < 1e-10 * abs(w[0])
    ev3, _, _ = ground_locg(a, int(np.argmin(np.diag(a))))
    assert abs(ev3 - w[0]) < 1e-9 * abs(w[0])


@pytest.mark.parametrize("split", [2, 3, 8])
def test_matvec_group_split(gpu, split):
    from rqutils_amd.operator import PauliOperator, Subspace
    p = small_problem(31, n=13, terms=140, nstates=3000)
    sp = Subspace(p.keys, p.n, bucket_cap=1024, rows_per_thread=4)
    op = PauliOperator((p.labels, p.coeffs), sp)
    op.set_tune(group_split=split)
    rng = np.random.default_rng(split)
    v = _rand_vec(rng, p.keys.size)
    ref = p.matvec(v)
    got = op.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
    assert _rel(got, ref) < 1e-12
    got32 = op.matvec(torch.from_numpy(v.astype(np.complex64)).cuda()).cpu().numpy()
    assert _rel(got32.astype(np.complex128), p.matvec(v.astype(np.complex64).astype(complex))) < 2e-6
    # the solver allocates the split's partial sums from its own workspace
    vec, info = solve_fused(op, v)
    ev = p.ground()
    assert info.converged and abs(info.eigval - ev) < 1e-9 * abs(ev), info


def solve_fused(op, v):
    from rqutils_amd.ground_locg import solve
    return solve(op, torch.from_numpy(v).cuda(), mode="fused", precision="c128",
                 criterion="relative", tol=1e-8, maxiter=3000)


def test_autotune_picks_a_valid_launch(gpu):
    from rqutils_amd.autotune import autotune
    from rqutils_amd.sqd import SQDProblem
    p = small_problem(41, n=14, terms=120, nstates=5000)
    res = autotune((p.labels, p.coeffs), p.keys, p.n, caps=(256, 1024, 8192), rows=(4, 8),
                   reps=1, use_cache=False)
    assert res.ms > 0 and len(res.table) >= 3
    assert abs(res.ms - min(t[3] for t in res.table)) < 1e-3
    prob = SQDProblem((p.labels, p.coeffs), p.keys, p.n, tune="auto")
    v = _rand_vec(np.random.default_rng(0), p.keys.size)
    got = prob.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
    assert _rel(got, p.matvec(v)) < 1e-12


@pytest.mark.parametrize("m,rows,split", [(5, 4, 1), (8, 8, 1), (11, 4, 3), (14, 16, 1),
                                          (13, 4, 2)])
def test_dense_lookup_matches_search_bit_for_bit(gpu, m, rows, split):
    """The dense lookup (bitmap + rank) finds exactly the keys the search finds,
    in the same class and group order, so both give the same bits; both match
    the oracle."""
    from rqutils_amd.operator import PauliOperator, Subspace
    p = small_problem(50 + m, n=14, terms=160, nstates=5000)
    dense_sp = Subspace(p.keys, p.n, dense=True, bucket_bits=m, rows_per_thread=rows)
    assert dense_sp.fine_bits == m - 5 and dense_sp.fbits is not None
    # bitmaps against the keys: bit j of entry e set iff (e << 5) | j is a key
    bits = dense_sp.fbits.cpu().numpy().view(np.uint32)
    member = ((bits[:, None] >> np.arange(32, dtype=np.uint32)) & 1).astype(bool).reshape(-1)
    assert np.array_equal(np.nonzero(member)[0].astype(np.uint64), p.keys)
    search_sp = Subspace(p.keys, p.n, bucket_bits=m, rows_per_thread=rows)
    rng = np.random.default_rng(m)
    v = _rand_vec(rng, p.keys.size)
    ref = p.matvec(v)
    for dt in (np.complex128, np.complex64):
        outs = []
        for sp in (dense_sp, search_sp):
            op = PauliOperator((p.labels, p.coeffs), sp)
            op.set_tune(group_split=split)
            outs.append(op.matvec(torch.from_numpy(v.astype(dt)).cuda()).cpu().numpy())
        assert np.array_equal(outs[0].view(np.float64 if dt == np.complex128 else np.float32),
                              outs[1].view(np.float64 if dt == np.complex128 else np.float32))
        tol = 1e-12 if dt == np.complex128 else 2e-6
        assert _rel(outs[0].astype(np.complex128), p.matvec(v.astype(dt).astype(complex))) < tol


def test_dense_lookup_full_space_and_solver(gpu):
    from rqutils_amd.operator import PauliOperator, Subspace
    from rqutils_amd.sqd import SQDProblem
    n = 12
    rng = np.random.default_rng(77)
    labels = random_labels(rng, n, 90, 1, 4)
    coeffs = rng.normal(size=90)
    keys = np.arange(2 ** n, dtype=np.uint64)
    p = OracleProblem(labels, coeffs, keys, n)
    prob = SQDProblem((labels, coeffs), keys, n, dense=True, bucket_bits=9)
    assert prob.space.dense
    v = _rand_vec(rng, keys.size)
    got = prob.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
    assert _rel(got, p.matvec(v)) < 1e-12
    vec, info = solve_fused(prob.op, v)
    ev = p.ground()
    assert info.converged and abs(info.eigval - ev) < 1e-9 * abs(ev), info
    with pytest.raises(ValueError):
        op = PauliOperator((labels, coeffs), prob.space)
        op.tune.stage_v = 1
        op.matvec(torch.from_numpy(v).cuda())


def test_fused_solver_stops_on_nan(gpu):
    """A NaN in the operator's output ends the fused solve with status 3 at
    once instead of iterating on it (the reference has no such guard)."""
    from rqutils_amd.ground_locg import STATUS_NONFINITE, solve
    n = 64
    a = torch.randn(n, n, dtype=torch.float64)
    a = (a + a.T).to(torch.complex128).cuda()
    calls = []

    def op(x):
        calls.append(1)
        y = a @ x
        if len(calls) > 3:
            y[5] = float("nan")
        return y

    x0 = torch.randn(n, dtype=torch.complex128).cuda()
    _, info = solve(op, x0, mode="fused", precision="c128", criterion="relative", tol=1e-12,
                    maxiter=500)
    assert info.status == STATUS_NONFINITE and not info.converged
    assert len(calls) < 12


@pytest.mark.parametrize("dense,split", [(False, 1), (True, 1), (False, 2), (True, 3)])
def test_matvec_512_thread_chunks_match_256(gpu, dense, split):
    """2,048-row chunks (512-thread workgroups) change only which rows share a
    staged table, not any row's summation order: same bits as 256 threads."""
    from rqutils_amd.operator import PauliOperator, Subspace
    p = small_problem(61, n=14, terms=150, nstates=9000)
    kw = dict(dense=True, bucket_bits=12) if dense else dict(bucket_bits=11)
    rng = np.random.default_rng(5)
    v = _rand_vec(rng, p.keys.size)
    outs = []
    for nt in (256, 512):
        sp = Subspace(p.keys, p.n, rows_per_thread=4, block_threads=nt, **kw)
        assert sp.chunk_rows == 4 * nt
        op = PauliOperator((p.labels, p.coeffs), sp)
        op.set_tune(group_split=split)
        outs.append(op.matvec(torch.from_numpy(v).cuda()).cpu().numpy())
    assert np.array_equal(outs[0].view(np.float64), outs[1].view(np.float64))
    assert _rel(outs[1], p.matvec(v)) < 1e-12
    with pytest.raises(ValueError):
        Subspace(p.keys, p.n, rows_per_thread=8, block_threads=512)


@pytest.mark.parametrize("dense_lookup,compact", [(False, 0), (True, 0), (False, 1), (True, 1)])
def test_zero_sum_terms_and_empty_group(gpu, dense_lookup, compact):
    """Duplicate strings that cancel: the X=110 group (XXI, YYZ) loses every
    term, so its range is empty and the kernel skips it; the X=100 group keeps
    XII and XIZ around a cancelled XZI (the reference's get_diagonal would stop
    at that zero, quirk Q3; the intended operator keeps XIZ). Checked against
    the dense Kronecker sum, which has no notion of groups. The dense lookup
    needs 32-value sub-buckets, so the strings act on the low 3 of 6 qubits."""
    from rqutils_amd.operator import PauliOperator, Subspace
    base = ["XXI", "XXI", "YYZ", "YYZ", "ZIZ", "XII", "XZI", "XZI", "XIZ", "IXX"]
    coeffs = [1.0, -1.0, 0.3, -0.3, 0.5, 1.0, 0.5, -0.5, 2.0, 0.7]
    n = 6
    labels = ["III" + lab for lab in base]
    keys = np.arange(2 ** n, dtype=np.uint64)
    kw = dict(dense=True, bucket_bits=5) if dense_lookup else dict(bucket_bits=3)
    sp = Subspace(keys, n, rows_per_thread=4, **kw)
    op = PauliOperator((labels, coeffs), sp)
    op.set_tune(compact=compact)
    grp = op.grp_off.cpu().numpy()
    assert np.any(grp[1:] == grp[:-1]), "expected an empty group"
    v = _rand_vec(np.random.default_rng(4), keys.size)
    got = op.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
    assert _rel(got, dense.dense_hamiltonian(labels, coeffs) @ v) < 1e-12


def test_ground_locg_maxiter_and_exact_eigenvector(gpu):
    """maxiter=1 stops after one loop iteration (rqutils/ground_locg.py:366-372);
    an exact eigenvector start has a zero residual, which the norm guards
    (ground_locg.py:276, 388) must turn into a clean stop, not NaN."""
    from rqutils_amd.ground_locg import ground_locg
    rng = np.random.default_rng(8)
    n = 40
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = a + a.conj().T
    w, u = np.linalg.eigh(a)
    at = torch.from_numpy(a).cuda()
    x0 = torch.from_numpy(rng.normal(size=n) + 0j).cuda()
    ev, vec, it = ground_locg(at, x0, maxiter=1)
    assert it == 1 and np.isfinite(ev)
    for mode in ("reference", "fused"):
        ev, vec, it = ground_locg(at, torch.from_numpy(u[:, 0].copy()).cuda(), maxiter=50,
                                  mode=mode)
        vv = vec.cpu().numpy()
        assert np.all(np.isfinite(vv)) and abs(ev - w[0]) < 1e-10 * abs(w[0]), (mode, ev, it)
        assert abs(abs(np.vdot(u[:, 0], vv)) / np.linalg.norm(vv) - 1.0) < 1e-10
