"""Pins the oracle itself (CPU only): brute force, closed forms, invariants and
the committed golden fixtures. Each check is chosen so that a plausible slip
in the restatement (a dropped phase, a wrong sign, a transposed operand, an
off-by-one in the compaction) fails it."""
import os

import numpy as np
import pytest

from oracle import dense, locg_oracle as lo, sqd_oracle as so
from rqutils_amd.operator import keys_to_states
from tests.helpers import OracleProblem, random_labels, small_problem

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---- Pauli grouping and phases --------------------------------------------------

def test_single_pauli_phases_match_dense():
    # Each one-qubit string against its 2x2 matrix: pins the (-i)^{x.z} fold and Y.
    for lab in ["X", "Y", "Z", "XY", "YZ", "ZY", "YY", "XZY"]:
        n = len(lab)
        p = OracleProblem([lab], [1.0], np.arange(2 ** n), n)
        v = np.random.default_rng(0).normal(size=2 ** n) + 0j
        assert np.allclose(p.matvec(v), dense.pauli_matrix(lab) @ v), lab


def test_grouping_merges_and_drops_zero():
    x, z, c, n = so.from_paulisum(["XZ", "XZ", "ZZ", "IX", "YI"], [1.0, 2.0, 0.0, 1.5, -1.0])
    # ZZ dropped (zero), XZ merged (3.0); groups ascending by X: IX (x=01), XZ/YI (x=10)
    xb = np.unpackbits(x, axis=1)[:, :n]
    assert xb.tolist() == [[0, 1], [1, 0]]
    assert c.shape == (2, 2)
    # XZ: x=10, z=01 -> (-i)^0 -> 3.0 ; YI: x=10, z=10 -> (-i)^1 * (-1) -> +1j
    assert set(np.round(c[1], 12).tolist()) == {3.0 + 0j, 1j}
    assert np.round(c[0], 12).tolist() == [1.5 + 0j, 0j]


def test_matvec_matches_dense_kronecker():
    for seed in range(4):
        p = small_problem(seed, n=9, terms=40, nstates=180, full=(seed % 2 == 1))
        rng = np.random.default_rng(seed)
        v = rng.normal(size=p.keys.size) + 1j * rng.normal(size=p.keys.size)
        assert np.allclose(p.matvec(v), p.dense() @ v, atol=1e-12)


@pytest.mark.parametrize("seed,full", [(0, False), (1, True), (2, False)])
def test_fullspace_matmul_matches_dense_kronecker(seed, full):
    """The full-space product (symplectic.py:131-150) against the dense
    Kronecker matrix, on a vector and on a matrix of right-hand sides; the
    sign of each term comes from the output index (a slip to the source index
    would transpose the Y phases, which the complex strings expose)."""
    rng = np.random.default_rng(seed)
    n = 7
    from tests.helpers import random_labels
    labels = random_labels(rng, n, 25, 1, 4, full=full)
    coeffs = rng.normal(size=len(labels))
    x, z, c, _ = so.from_paulisum(labels, coeffs, add_padding=True)
    h = dense.dense_hamiltonian(labels, coeffs)
    v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    assert np.allclose(so.fullspace_matmul(x, z, c, n, v), h @ v, atol=1e-12)
    m = rng.normal(size=(2 ** n, 3))
    assert np.allclose(so.fullspace_matmul(x, z, c, n, m), h @ m, atol=1e-12)
    # a sample of groups sums to the whole when all are taken
    part = sum(so.fullspace_matmul(x, z, c, n, v, groups=[g]) for g in range(x.shape[0]))
    assert np.allclose(part, h @ v, atol=1e-12)


def test_matvec_hermitian():
    p = small_problem(7, n=10, terms=60, nstates=400)
    rng = np.random.default_rng(1)
    u = rng.normal(size=p.keys.size) + 1j * rng.normal(size=p.keys.size)
    v = rng.normal(size=p.keys.size) + 1j * rng.normal(size=p.keys.size)
    assert abs(np.vdot(u, p.matvec(v)) - np.vdot(p.matvec(u), v)) < 1e-10


# ---- source indices -------------------------------------------------------------

@pytest.mark.parametrize("seed", range(4))
def test_xsource_sort_join_equals_lookup(seed):
    p = small_problem(seed, n=11, terms=30, nstates=500)
    for g in range(p.x.shape[0]):
        assert np.array_equal(so.get_xsource(p.x[g], p.packed),
                              so.xsource_by_lookup(p.x[g], p.packed))


def test_xsource_full_space_identity():
    n = 10
    states = dense_states = np.array([[(k >> (n - 1 - j)) & 1 for j in range(n)]
                                      for k in range(2 ** n)], dtype=np.uint8)
    packed = so.pack_states(dense_states)
    for x in [1, 0x155, 2 ** n - 1]:
        xb = so.pack_states(np.array([[(x >> (n - 1 - j)) & 1 for j in range(n)]], np.uint8))[0]
        assert np.array_equal(so.get_xsource(xb, packed), np.arange(2 ** n) ^ x)
    assert states.shape == (2 ** n, n)


def test_key_order_equals_packed_row_order():
    rng = np.random.default_rng(3)
    for n in (7, 8, 15, 16, 23, 40):
        states = rng.integers(0, 2, (300, n)).astype(np.uint8)
        packed = so.pack_states(states)
        keys = so.packed_to_keys(packed, n)
        order_rows = np.lexsort(packed.T[::-1])
        assert np.array_equal(np.sort(keys), keys[order_rows])


def test_uniquify_fill_rows():
    rng = np.random.default_rng(4)
    states = rng.integers(0, 2, (50, 6)).astype(np.uint8)
    packed = so.pack_states(states)
    su, nu = so.uniquify_states(packed, 80)
    assert nu == np.unique(states, axis=0).shape[0]
    assert np.all(su[nu:] == 255)
    assert np.all(su[:nu, 0] >> 7 == 0)


# ---- LOBPCG -----------------------------------------------------------------------

def test_eigenpairs_closed_forms():
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        a = a + a.conj().T
        w, V = np.linalg.eigh(a)
        lam, k = lo.eigenpair_3x3(a)
        assert abs(lam - w[0]) < 1e-12 * max(1, abs(w[0]))
        assert abs(abs(np.vdot(V[:, 0], k)) - 1) < 1e-9
        b = a[:2, :2]
        w2, V2 = np.linalg.eigh(b)
        lam2, k2 = lo.eigenpair_2x2(b)
        assert abs(lam2 - w2[0]) < 1e-12 * max(1, abs(w2[0]))
        assert abs(abs(np.vdot(V2[:, 0], k2)) - 1) < 1e-9
    # diagonal 3x3: lowest eigenvalue exact; the cross-product vector degenerates (quirk Q5)
    lam, _ = lo.eigenpair_3x3(np.diag([3.0, -2.0, 1.0]).astype(complex))
    assert abs(lam + 2.0) < 1e-12


def test_project_out_invariants():
    rng = np.random.default_rng(2)
    n = 200
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    x /= np.linalg.norm(x)
    y = rng.normal(size=n) + 1j * rng.normal(size=n)
    y -= x * np.vdot(x, y)
    y /= np.linalg.norm(y)
    r = rng.normal(size=n) + 1j * rng.normal(size=n)
    p = lo.project_out((x, y), r)
    assert abs(np.vdot(x, p)) < 1e-14 and abs(np.vdot(y, p)) < 1e-14
    assert abs(np.linalg.norm(p) - 1) < 1e-12
    # a residual inside span(x, y): the result is still zero-or-orthonormal
    # (normalisation amplifies the rounding left after the first subtraction)
    q = lo.project_out((x, y), 0.3 * x - 2j * y)
    nq = np.linalg.norm(q)
    assert nq == 0.0 or (abs(nq - 1) < 1e-12 and abs(np.vdot(x, q)) < 1e-12)


def test_ground_locg_oracle_vs_eigh():
    p = small_problem(5, n=9, terms=50, nstates=300)
    h = p.dense()
    rng = np.random.default_rng(5)
    x0 = rng.normal(size=h.shape[0]) + 1j * rng.normal(size=h.shape[0])
    th, x, it = lo.ground_locg(lambda v: h @ v, x0)
    w = np.linalg.eigvalsh(h)
    assert abs(th - w[0]) < 1e-10 * abs(w[0])
    assert 0 < it < 1000
    # maxiter exhaustion returns silently with niter == maxiter
    _, _, it1 = lo.ground_locg(lambda v: h @ v, x0, maxiter=1)
    assert it1 == 1


# ---- golden fixtures ------------------------------------------------------------

def _golden(name):
    path = os.path.join(GOLDEN, name)
    return np.load(path, allow_pickle=False)


def test_golden_small16_against_oracle():
    g = _golden("small16.npz")
    labels = [str(s) for s in g["labels"]]
    p = OracleProblem(labels, g["coeffs"], g["keys"], int(g["n"]))
    # matvec against the fixture's dense product
    assert np.allclose(p.matvec(g["v"]), g["hv"], rtol=0, atol=1e-11)
    # source indices of three groups against the fixture's brute-force lookup
    for k, gi in enumerate(g["xs_groups"]):
        assert np.array_equal(so.get_xsource(p.x[gi], p.packed), g["xs"][k])
    # oracle LOBPCG converges to the fixture's dense eigenvalue
    rng = np.random.default_rng(int(g["seed_x0"]))
    x0 = rng.normal(size=p.keys.size) + 1j * rng.normal(size=p.keys.size)
    th, _, _ = lo.ground_locg(p.matvec, x0)
    assert abs(th - float(g["e0"])) < 1e-10 * abs(float(g["e0"]))


def test_golden_phase_table():
    g = _golden("pauli_phases.npz")
    for lab, mat in zip(g["labels"], g["mats"]):
        n = len(str(lab))
        p = OracleProblem([str(lab)], [1.0], np.arange(2 ** n), n)
        cols = np.stack([p.matvec(np.eye(2 ** n)[:, j].astype(complex)) for j in range(2 ** n)],
                        axis=1)
        assert np.allclose(cols, mat[: 2 ** n, : 2 ** n]), str(lab)


# ---- start-vector network oracle ---------------------------------------------------

from oracle import mlp_oracle as mo  # noqa: E402


def _mlp_params(rng, kin, h):
    return [rng.normal(size=(kin, h)) / np.sqrt(kin), 0.1 * rng.normal(size=h),
            rng.normal(size=(h, h)) / np.sqrt(h), 0.1 * rng.normal(size=h),
            rng.normal(size=(h, 2)) / np.sqrt(h), 0.1 * rng.normal(size=2)]


def _mlp_batch(rng, n, nq, kpad):
    keys = rng.integers(0, 2 ** nq, size=n).astype(np.uint64)
    tgt = np.stack([rng.normal(size=n), rng.integers(0, 2, n).astype(float),
                    rng.uniform(0.1, 2.0, n)], axis=1)
    return mo.encode(keys, nq, kpad), tgt


def test_mlp_oracle_gradients_match_finite_differences():
    # Central differences on every parameter tensor: pins the chain rule of the
    # backward pass (a dropped GELU derivative or a transposed weight fails).
    rng = np.random.default_rng(0)
    params = _mlp_params(rng, 8, 6)
    x, tgt = _mlp_batch(rng, 24, 5, 8)
    _, grads, _, _ = mo.loss_and_grads(params, x, tgt, lam=0.7)
    eps = 1e-6
    for p, g in zip(params, grads):
        flat, gf = p.reshape(-1), g.reshape(-1)
        for j in rng.choice(flat.size, size=min(8, flat.size), replace=False):
            old = flat[j]
            flat[j] = old + eps
            lp = mo.loss_and_grads(params, x, tgt, lam=0.7)[0]
            flat[j] = old - eps
            lm = mo.loss_and_grads(params, x, tgt, lam=0.7)[0]
            flat[j] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - gf[j]) <= 2e-7 * max(1.0, abs(fd)), (p.shape, j, fd, gf[j])


def test_mlp_oracle_loss_gelu_adam_match_torch():
    # Library routines: torch's exact-erf GELU, BCE-with-logits and Adam.
    import torch
    import torch.nn.functional as F
    rng = np.random.default_rng(1)
    z = np.linspace(-9, 9, 301)
    zt = torch.tensor(z, requires_grad=True)
    gt = F.gelu(zt)
    gt.sum().backward()
    assert np.allclose(mo.gelu(z), gt.detach().numpy(), rtol=1e-14, atol=1e-15)
    assert np.allclose(mo.gelu_grad(z), zt.grad.numpy(), rtol=1e-13, atol=1e-15)
    params = _mlp_params(rng, 16, 10)
    x, tgt = _mlp_batch(rng, 40, 9, 16)
    tgt[:3, 0] = 0.0
    loss, _, y, _ = mo.loss_and_grads(params, x, tgt, lam=0.5)
    yt = torch.tensor(y)
    tt = torch.tensor(tgt)
    ref = ((tt[:, 2] * (yt[:, 0] - tt[:, 0]) ** 2).sum()
           + 0.5 * F.binary_cross_entropy_with_logits(yt[:, 1], tt[:, 1], weight=tt[:, 2],
                                                      reduction="sum")) / tt[:, 2].sum()
    assert abs(loss - float(ref)) < 1e-12 * abs(float(ref))
    # Adam, two steps, against torch.optim.Adam in float64
    w = rng.normal(size=50)
    g1, g2 = rng.normal(size=50), rng.normal(size=50)
    wt = torch.tensor(w.copy(), requires_grad=True)
    opt = torch.optim.Adam([wt], lr=0.01, betas=(0.8, 0.95), eps=1e-7)
    m = v = np.zeros(50)
    ww = w.copy()
    for step, g in enumerate((g1, g2), start=1):
        wt.grad = torch.tensor(g)
        opt.step()
        ww, m, v = mo.adam(ww, g, m, v, 0.01, 0.8, 0.95, 1e-7, step)
    assert np.allclose(ww, wt.detach().numpy(), rtol=1e-13, atol=1e-15)


def test_bf16_round_matches_torch():
    import torch
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(size=4000) * 10.0 ** rng.integers(-6, 6, size=4000),
                        [0.0, -0.0, 1.0, -1.0, 65504.0, 1e-40]]).astype(np.float32)
    # exact ties: low 16 bits 0x8000 on odd and even upper halves
    ties = (rng.integers(0x3F00, 0x4100, size=200).astype(np.uint32) << 16) | 0x8000
    x = np.concatenate([x, ties.view(np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16_round(x), ref)


def test_mlp_encode_bits():
    keys = np.array([0, 1, 0b1011, (1 << 39) | 5], dtype=np.uint64)
    x = mo.encode(keys, 40, 48)
    for i, k in enumerate(keys.tolist()):
        bits = format(k, "040b")[::-1]
        assert list(x[i, :40]) == [(-1.0 if b == "1" else 1.0) for b in bits]
        assert not x[i, 40:].any()


# ---- state-vector simulator (rqutils/svsim.py) -----------------------------------

def _random_gate(rng, n):
    name = ["x", "y", "z", "rx", "ry", "rz", "rzz"][rng.integers(0, 7)]
    k = 2 if name == "rzz" else int(rng.integers(1, min(3, n) + 1))
    qubits = sorted(rng.choice(n, size=min(k, n), replace=False).tolist())
    ang = float(rng.uniform(-np.pi, np.pi))
    return (name, qubits, ang) if name.startswith("r") else (name, qubits)


def test_svsim_each_gate_matches_dense():
    """Every gate (one and several qubits) against its dense Kronecker matrix:
    pins qubit order, the output-index sign and the reference's phases."""
    from oracle import svsim_oracle as sv
    rng = np.random.default_rng(3)
    n = 4
    for name in sv.GATE_SPECS:
        for qubits in ([2], [0, 3], [1, 2, 3]):
            if name == "rzz" and len(qubits) != 2:
                continue
            ang = 0.7
            gate = (name, qubits, ang) if name.startswith("r") else (name, qubits)
            psi0 = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
            got = sv.svsim(*sv.to_circuitxz([gate], n), initial_state=psi0)
            want = sv.gate_matrix_dense(name, qubits, ang, n) @ psi0
            assert np.allclose(got, want, atol=1e-13), (name, qubits)


def test_svsim_circuit_matches_product_of_dense_gates():
    from oracle import svsim_oracle as sv
    rng = np.random.default_rng(7)
    n = 5
    gates = [_random_gate(rng, n) for _ in range(40)]
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    want = psi.copy()
    for g in gates:
        want = sv.gate_matrix_dense(g[0], g[1], g[2] if len(g) > 2 else None, n) @ want
    got = sv.svsim(*sv.to_circuitxz(gates, n), initial_state=psi)
    assert np.allclose(got, want, atol=1e-12)


def test_svsim_conventions():
    """rx/rz/rzz/x/y/z keep the norm; the reference's ry is
    cos(a/2) I + sin(a/2) Y (not unitary), x and z carry a -i; integer starts
    are one-hot; num_qubits defaults to the largest qubit + 1; unknown
    names raise."""
    from oracle import svsim_oracle as sv
    x, z, c, s, n = sv.to_circuitxz([("rx", 0, 0.3), ("rz", 2, 1.1)])
    assert n == 3 and list(x) == [1, 0] and list(z) == [0, 4]
    assert np.isclose(c[0], np.cos(-0.15)) and np.isclose(s[0], np.sin(-0.15))
    one = sv.svsim(*sv.to_circuitxz([("x", 1)], 2), initial_state=0)
    assert np.allclose(one, [0, 0, -1j, 0])
    psi = np.array([1.0, 0.0], dtype=complex)
    ry = sv.svsim(*sv.to_circuitxz([("ry", 0, 0.8)], 1), initial_state=psi)
    assert np.allclose(ry, [np.cos(0.4), 1j * np.sin(0.4)])   # cos I + sin Y on |0>
    pm = np.array([1.0, 1.0], dtype=complex) / np.sqrt(2)
    ry_pm = sv.svsim(*sv.to_circuitxz([("ry", 0, 0.8)], 1), initial_state=pm)
    assert abs(np.linalg.norm(ry_pm) - 1.0) < 1e-12            # |+> keeps its norm ...
    pi = np.array([1.0, 1j], dtype=complex) / np.sqrt(2)
    ry_pi = sv.svsim(*sv.to_circuitxz([("ry", 0, 0.8)], 1), initial_state=pi)
    assert abs(np.linalg.norm(ry_pi) - 1.0) > 0.1              # ... an eigenvector of Y does not
    rng = np.random.default_rng(0)
    gates = [g for g in (_random_gate(rng, 6) for _ in range(60)) if g[0] != "ry"]
    v = rng.normal(size=64) + 1j * rng.normal(size=64)
    v /= np.linalg.norm(v)
    assert abs(np.linalg.norm(sv.svsim(*sv.to_circuitxz(gates, 6), initial_state=v)) - 1) < 1e-12
    with pytest.raises(ValueError):
        sv.to_circuitxz([("cz", [0, 1])])


# ---- property tests (hypothesis) --------------------------------------------------

from hypothesis import given, settings, strategies as st  # noqa: E402


@st.composite
def _pauli_problem(draw):
    n = draw(st.integers(1, 5))
    terms = draw(st.integers(1, 8))
    labels = [draw(st.text("IXYZ", min_size=n, max_size=n)) for _ in range(terms)]
    coeffs = [draw(st.floats(-2, 2, allow_nan=False, allow_infinity=False)) for _ in range(terms)]
    keys = sorted(draw(st.sets(st.integers(0, 2 ** n - 1), min_size=1, max_size=2 ** n)))
    return labels, coeffs, np.asarray(keys, np.uint64), n


@settings(max_examples=60, deadline=None)
@given(_pauli_problem())
def test_oracle_matvec_is_projected_dense_sum(problem):
    """For any strings (duplicates, zeros, all-identity included) and any
    subspace, the oracle's sort-join matvec equals P H P from Kronecker
    products, and the projected operator is Hermitian."""
    labels, coeffs, keys, n = problem
    if all(c == 0 for c in coeffs):
        return
    p = OracleProblem(labels, coeffs, keys, n)
    v = np.random.default_rng(len(keys)).normal(size=keys.size) + 0j
    want = dense.projected(dense.dense_hamiltonian(labels, coeffs), keys.astype(np.int64)) @ v
    assert np.allclose(p.matvec(v), want, atol=1e-12)
    m = dense.projected(dense.dense_hamiltonian(labels, coeffs), keys.astype(np.int64))
    assert np.allclose(m, m.conj().T)


@settings(max_examples=60, deadline=None)
@given(st.integers(1, 10).flatmap(lambda n: st.tuples(
    st.just(n), st.sets(st.integers(0, 2 ** n - 1), min_size=1, max_size=2 ** n),
    st.integers(0, 2 ** n - 1))))
def test_oracle_xsource_is_a_dictionary_lookup(case):
    """get_xsource (the sort-join of rqutils/sqd.py:476-500) is, for every
    subspace and X mask, the index of s ^ x in the sorted subspace or -1."""
    from rqutils_amd.operator import keys_to_states
    n, keyset, x = case
    keys = np.asarray(sorted(keyset), np.uint64)
    packed = so.pack_states(keys_to_states(keys, n))
    xs = so.pack_states(keys_to_states(np.asarray([x], np.uint64), n))[0]
    got = so.get_xsource(xs, packed)
    pos = {int(k): i for i, k in enumerate(keys)}
    want = np.array([pos.get(int(k) ^ x, -1) for k in keys])
    assert np.array_equal(got, want)


# ---- default start (vinit) -------------------------------------------------------

def _brute_vinit(labels, coeffs, keys, n):
    """Index of the first smallest diagonal entry <s|H|s> over the real rows,
    from dense Kronecker products (independent of the X/Z code), or 0 when no
    term is diagonal."""
    if not any(set(lab) <= {"I", "Z"} for lab, c in zip(labels, coeffs) if c != 0):
        return 0
    diag = np.real(np.diag(dense.dense_hamiltonian(labels, np.real(coeffs))))
    d = diag[keys.astype(np.int64)]
    return int(np.flatnonzero(d == d.min())[0])


@pytest.mark.parametrize("case", ["random", "ties", "fill_lowest", "no_diag"])
def test_vinit_from_min_diag_matches_brute_force(case):
    """Pins the default start (rqutils/sqd.py:373-398): argmin over the real
    rows with the first of equal entries, fill rows (0xFF) never chosen even
    when their diagonal would be the smallest, row 0 without an X = 0 group."""
    rng = np.random.default_rng({"random": 1, "ties": 2, "fill_lowest": 3, "no_diag": 4}[case])
    n = 8
    if case == "random":
        labels = random_labels(rng, n, 40, 1, 3)
        coeffs = rng.normal(size=40)
    elif case == "ties":
        # small integer Z couplings: the diagonal takes few values, many rows tie
        labels = ["ZZIIIIII", "IIZZIIII", "IIIIZIZI"]
        coeffs = np.array([1.0, -1.0, 1.0])
    elif case == "fill_lowest":
        # sum_q Z_q: the all-ones row (the fill row's bits) has the lowest diagonal
        labels = ["I" * q + "Z" + "I" * (n - 1 - q) for q in range(n)] + ["XXIIIIII"]
        coeffs = np.concatenate([np.ones(n), [0.3]])
    else:
        labels = ["XIIIIIIY", "IXZIIIII", "IIIYXIII"]
        coeffs = np.array([1.0, -0.5, 0.7])
    keys = np.sort(rng.choice(2 ** n - 1, size=120, replace=False)).astype(np.uint64)
    keys = keys[::-1][:100][::-1].copy()            # never the all-ones key
    x, z, c, _ = so.from_paulisum(labels, coeffs, force_real=True, add_padding=True)
    packed = so.pack_states(keys_to_states(keys, n))
    su, nu = so.uniquify_states(packed[rng.permutation(len(keys))], len(keys) + 17)
    assert nu == len(keys) and np.all(su[nu:] == 255)
    v = so.vinit_from_min_diag(x, z, c, su, nu)
    want = _brute_vinit(labels, coeffs, keys, n)
    assert v.shape == (len(keys) + 17,) and np.count_nonzero(v) == 1
    assert int(np.flatnonzero(v)[0]) == want, (case, np.flatnonzero(v), want)
    if case == "ties":
        d = np.real(np.diag(dense.dense_hamiltonian(labels, coeffs)))[keys.astype(np.int64)]
        assert np.count_nonzero(d == d.min()) > 1, "the case must hold a tie"
    if case == "fill_lowest":
        # without the masking the fill rows would win: the guard matters here
        diag = np.real(so.get_diagonal(z[0], c[0], su))
        assert diag[nu:].min() < diag[:nu].min()


# ---- where the reference's closed forms fail (the device guards these) ----------

def test_reference_closed_forms_fail_where_guarded():
    """The reference's own formulas (the oracle's restatement) give a wrong
    vector for the diagonal 3x3 (rounding noise in column 1 of A - lambda I
    picks e0 instead of e1) and NaN for the diagonal 2x2 whose lower entry is
    the smaller: the cases the device closed forms guard."""
    a3 = np.diag([3.0, -2.0, 1.0]).astype(complex)
    lam3, v3 = lo.eigenpair_3x3(a3)
    assert abs(lam3 + 2.0) < 1e-12
    assert not np.all(np.isfinite(v3)) or np.linalg.norm(a3 @ v3 - lam3 * v3) > 1.0
    with np.errstate(all="ignore"):
        _, v2 = lo.eigenpair_2x2(np.diag([-1.0, 1.0]).astype(complex))
    assert not np.all(np.isfinite(v2))


def test_oracle_locg_reference_has_no_zero_guard():
    """Documents the reference behaviour the restart replaces: the oracle's
    reference iteration from an exact eigenvector start divides 0 by 0."""
    a = np.diag([1.0, -1.0, 2.0]).astype(complex)
    x0 = np.array([0.0, 1.0, 0.0], complex)
    with np.errstate(all="ignore"):
        th, _, _ = lo.ground_locg(lambda v: a @ v, x0, maxiter=3)
    assert np.isnan(th)


@pytest.mark.parametrize("seed", [0, 5])
def test_apply_h_rows_equals_full_product(seed):
    """The per-row restatement (for full-size subspaces) against the whole
    cached-source product on a small problem, including a complex Hamiltonian
    and rows whose targets fall before the first and after the last row."""
    p = small_problem(seed, n=11, terms=80, nstates=700, full=bool(seed))
    rng = np.random.default_rng(seed)
    v = rng.normal(size=p.keys.size) + 1j * rng.normal(size=p.keys.size)
    full = p.matvec(v)
    rows = np.concatenate([[0, p.keys.size - 1], rng.choice(p.keys.size, 100, replace=False)])
    got = so.apply_h_rows(rows, v, p.x, p.z, p.c, p.packed)
    assert np.array_equal(got, full[rows])


# ---- the C restatement of the reference's dataflow (CPU baseline) -----------------

@pytest.mark.parametrize("seed,full", [(0, False), (6, True)])
def test_c_restatement_matches_numpy_oracle(seed, full):
    """ref_dataflow.c against the NumPy restatement: source indices bit for
    bit for every group, the product of all groups within 1e-12 on one and on
    four threads, with and without precomputed source indices."""
    from oracle import ref_dataflow as rd
    p = small_problem(seed, n=11, terms=90, nstates=900, full=full)
    t = rd.RefTables(p.labels, p.coeffs, p.n)
    for g in range(p.x.shape[0]):
        assert np.array_equal(rd.xsource(p.keys, p.n, t.xmask[g]), p.xsources[g]), g
    rng = np.random.default_rng(seed)
    v = rng.normal(size=p.keys.size) + 1j * rng.normal(size=p.keys.size)
    want = p.matvec(v)
    for threads in (1, 4):
        got, used = rd.matvec(t, p.keys, v, nthreads=threads)
        assert used == threads
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    got, _ = rd.matvec(t, p.keys, v, xsources=p.xsources, nthreads=2)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    # a group subset is the sum of those groups only
    sub = np.array([0, 3, 5], np.int32)
    got, _ = rd.matvec(t, p.keys, v, groups=sub, nthreads=2)
    part = sum(so.apply_xgrp(p.xsources[g], so.get_diagonal(p.z[g], p.c[g], p.packed), v)
               for g in sub)
    assert np.linalg.norm(got - part) <= 1e-12 * np.linalg.norm(part)


@pytest.mark.parametrize("seed", [2, 3])
def test_sign_bit_cache_composes_the_diagonal(seed):
    """get_diag_signs then compute_diagonal gives get_diagonal for groups of
    more than 8 terms (an independent path through the sign bits); with the
    reference's bit index (quirk Q1) it differs exactly when a term from the
    ninth on has an odd parity, and agrees for groups of at most 8 terms."""
    p = small_problem(seed, n=9, terms=300, nstates=250)
    kmax = p.c.shape[1]
    assert kmax > 8
    seen_diff = False
    for g in range(p.x.shape[0]):
        signs = so.get_diag_signs(p.z[g], p.packed)
        d = so.get_diagonal(p.z[g], p.c[g], p.packed)
        assert np.allclose(so.compute_diagonal(signs, p.c[g]), d, rtol=0, atol=1e-12)
        dq = so.compute_diagonal(signs, p.c[g], reference_bit_index=True)
        nterms = int(np.count_nonzero(p.c[g]))
        if nterms <= 8:
            assert np.allclose(dq, d, rtol=0, atol=1e-12)
        elif not np.allclose(dq, d):
            seen_diff = True
    assert seen_diff
