"""GPU parity of the compacted hit path of the matvec.

Compacted mode queues each wave's hits in LDS and applies them one lane per
hit, with the same per-hit arithmetic and group order as register mode, so the
two must agree bit for bit; both must match the oracle. Shapes cover the dense
and search lookups, 4 and 8 rows per thread, 512-thread chunks, group splits,
v staging, 64-bit LDS keys, hit counts above 64 per wave and group (several
queue passes) and the batched launch.
"""
import numpy as np
import pytest
import torch

from tests.helpers import OracleProblem, random_labels, small_problem

pytestmark = pytest.mark.gpu


def _rand_vec(rng, n):
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _both_modes(labels, coeffs, sp, v, split=1, stage_v=0):
    from rqutils_amd.operator import PauliOperator
    outs = []
    for compact in (0, 1):
        op = PauliOperator((labels, coeffs), sp)
        op.set_tune(group_split=split, stage_v=stage_v, compact=compact)
        outs.append(op.matvec(torch.from_numpy(v).cuda()).cpu().numpy())
    return outs


def _same_bits(a, b):
    view = np.float64 if a.dtype == np.complex128 else np.float32
    return np.array_equal(a.view(view), b.view(view))


@pytest.mark.parametrize("kind,rows,threads,split,stage_v", [
    ("dense", 4, 256, 1, 0), ("dense", 8, 256, 1, 0), ("dense", 4, 512, 1, 0),
    ("dense", 4, 256, 3, 0), ("search", 4, 256, 1, 0), ("search", 8, 256, 2, 0),
    ("search", 4, 512, 1, 0), ("search", 4, 256, 1, 1), ("search", 8, 256, 1, 1),
])
def test_compact_equals_register_mode(gpu, kind, rows, threads, split, stage_v):
    from rqutils_amd.operator import Subspace
    p = small_problem(70 + rows + split, n=14, terms=160, nstates=7000)
    if kind == "dense":
        sp = Subspace(p.keys, p.n, dense=True, bucket_bits=12, rows_per_thread=rows,
                      block_threads=threads)
    else:
        sp = Subspace(p.keys, p.n, bucket_bits=11, rows_per_thread=rows, block_threads=threads)
    rng = np.random.default_rng(rows * 10 + split)
    v = _rand_vec(rng, p.keys.size)
    for dt in (np.complex128, np.complex64):
        vd = v.astype(dt)
        reg, cmp_ = _both_modes(p.labels, p.coeffs, sp, vd, split, stage_v)
        assert _same_bits(reg, cmp_), (kind, rows, threads, split, stage_v, dt)
        tol = 1e-12 if dt == np.complex128 else 2e-6
        assert _rel(cmp_.astype(np.complex128), p.matvec(vd.astype(complex))) < tol


def test_compact_many_hits_per_wave(gpu):
    """Whole 12-qubit space: every off-diagonal group hits every row, so a
    wave queues 256 (4 rows) or 512 (8 rows) hits per group and takes 4 or 8
    passes over its queue."""
    from rqutils_amd.operator import Subspace
    n = 12
    rng = np.random.default_rng(3)
    labels = random_labels(rng, n, 120, 1, 4)
    coeffs = rng.normal(size=120)
    keys = np.arange(2 ** n, dtype=np.uint64)
    p = OracleProblem(labels, coeffs, keys, n)
    v = _rand_vec(rng, keys.size)
    for kw in (dict(dense=True, bucket_bits=10), dict(bucket_bits=10)):
        for rows in (4, 8):
            sp = Subspace(keys, n, rows_per_thread=rows, **kw)
            reg, cmp_ = _both_modes(labels, coeffs, sp, v)
            assert _same_bits(reg, cmp_), (kw, rows)
            assert _rel(cmp_, p.matvec(v)) < 1e-12


def test_compact_wide_keys(gpu):
    """bucket_bits > 32: 64-bit low keys in LDS and in the compacted row area."""
    from rqutils_amd.operator import Subspace
    rng = np.random.default_rng(17)
    n = 48
    labels = random_labels(rng, n, 200, 1, 3)
    coeffs = rng.normal(size=200)
    base = np.sort(rng.choice(2 ** 14, 400, replace=False)).astype(np.uint64)
    keys = np.unique(np.concatenate([base, base ^ np.uint64(0b11), base << np.uint64(30)]))
    p = OracleProblem(labels, coeffs, keys, n)
    v = _rand_vec(rng, keys.size)
    sp = Subspace(keys, n, fine_bits=3)
    assert sp.bucket_bits > 32
    reg, cmp_ = _both_modes(labels, coeffs, sp, v)
    assert _same_bits(reg, cmp_)
    assert _rel(cmp_, p.matvec(v)) < 1e-12


def test_compact_rejects_16_rows(gpu):
    from rqutils_amd.operator import PauliOperator, Subspace
    p = small_problem(2, n=12, terms=40, nstates=1000)
    sp = Subspace(p.keys, p.n, rows_per_thread=16)
    op = PauliOperator((p.labels, p.coeffs), sp)
    with pytest.raises(ValueError):
        op.set_tune(compact=1)
    op.tune.compact = 1          # bypass the host check: the library refuses too
    with pytest.raises(ValueError):
        op.matvec(torch.zeros(p.keys.size, dtype=torch.complex128, device="cuda"))


@pytest.mark.parametrize("split", [1, 2])
def test_batch_compact_equals_single(gpu, split):
    from rqutils_amd.batch import SQDBatch
    from rqutils_amd.operator import PauliOperator, Subspace
    probs = [small_problem(90 + i, n=10 + i % 4, terms=60 + 20 * i,
                           nstates=min(2 ** (10 + i % 4) - 1, 300 + 200 * i)) for i in range(5)]
    b = SQDBatch([((p.labels, p.coeffs), p.keys, p.n) for p in probs], bucket_cap=1024,
                 rows_per_thread=4, group_split=split, compact=1)
    rng = np.random.default_rng(split)
    v = _rand_vec(rng, b.n)
    for dtype in (torch.complex128, torch.complex64):
        vt = torch.from_numpy(v).to(dtype).cuda()
        out = b.matvec(vt)
        for i, p in enumerate(probs):
            sp = Subspace(p.keys, p.n, bucket_cap=1024, rows_per_thread=4)
            op = PauliOperator((p.labels, p.coeffs), sp)
            op.set_tune(group_split=split, compact=0)
            ref = op.matvec(b.segment(vt, i).contiguous())
            assert torch.equal(b.segment(out, i), ref), (i, dtype)


def test_compact_solver_converges(gpu):
    from rqutils_amd.sqd import SQDProblem
    p = small_problem(33, n=13, terms=150, nstates=4000)
    prob = SQDProblem((p.labels, p.coeffs), p.keys, p.n, compact=1)
    assert prob.op.tune.compact == 1
    v = _rand_vec(np.random.default_rng(1), p.keys.size)
    _, info = prob.solve(torch.from_numpy(v).cuda(), mode="fused", precision="c128",
                         criterion="relative", tol=1e-8, maxiter=3000)
    ev = p.ground()
    assert info.converged and abs(info.eigval - ev) < 1e-9 * abs(ev), info
