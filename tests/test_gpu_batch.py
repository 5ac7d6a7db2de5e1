"""GPU parity of the batched path: one launch over many instances must give
each instance exactly what its own matvec gives (same kernel, same order),
and the batched fused LOBPCG must reach each instance's ground energy."""
import numpy as np
import pytest
import torch

from tests.helpers import small_problem

pytestmark = pytest.mark.gpu


def _instances(k=6, seed=0):
    probs = []
    for i in range(k):
        n = 10 + (i % 4)
        p = small_problem(seed + i, n=n, terms=40 + 15 * i, nstates=min(2 ** n - 1, 200 + 150 * i))
        probs.append(p)
    return probs


def _rand(rng, n):
    return rng.normal(size=n) + 1j * rng.normal(size=n)


@pytest.mark.parametrize("cap,rows,split", [(8192, 4, 1), (256, 4, 3), (1024, 8, 2)])
def test_batch_matvec_equals_single(gpu, cap, rows, split):
    from rqutils_amd.batch import SQDBatch
    from rqutils_amd.operator import PauliOperator, Subspace
    probs = _instances()
    b = SQDBatch([((p.labels, p.coeffs), p.keys, p.n) for p in probs], bucket_cap=cap,
                 rows_per_thread=rows, group_split=split)
    rng = np.random.default_rng(cap + split)
    v = _rand(rng, b.n)
    for dtype in (torch.complex128, torch.complex64):
        vt = torch.from_numpy(v).to(dtype).cuda()
        out = b.matvec(vt)
        for i, p in enumerate(probs):
            sp = Subspace(p.keys, p.n, bucket_cap=cap, rows_per_thread=rows)
            op = PauliOperator((p.labels, p.coeffs), sp)
            op.set_tune(group_split=split)
            ref = op.matvec(b.segment(vt, i).contiguous())
            assert torch.equal(b.segment(out, i), ref), (i, dtype)
        # and against the oracle
        if dtype == torch.complex128:
            for i, p in enumerate(probs):
                seg = b.segment(out, i).cpu().numpy()
                want = p.matvec(b.segment(vt, i).cpu().numpy())
                assert np.linalg.norm(seg - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("precision,tol", [("c128", 1e-9), ("c64", 1e-6)])
def test_batch_solve_matches_each_instance(gpu, use_graph, precision, tol):
    from rqutils_amd.batch import SQDBatch
    probs = _instances(k=5, seed=50)
    b = SQDBatch([((p.labels, p.coeffs), p.keys, p.n) for p in probs], bucket_cap=512)
    rng = np.random.default_rng(3)
    x0 = torch.from_numpy(_rand(rng, b.n)).cuda()
    res = b.solve(x0, precision=precision, tol=tol, maxiter=2000, use_graph=use_graph)
    for i, p in enumerate(probs):
        info = res.infos[i]
        ev = p.ground()
        assert info.converged, (i, info)
        assert abs(info.eigval - ev) <= max(10 * tol * tol, 1e-12) * abs(ev) + \
            (2e-7 * abs(ev) if precision == "c64" else 0.0), (i, info.eigval, ev)
        x = b.segment(res.vectors, i).cpu().numpy().astype(np.complex128)
        x /= np.linalg.norm(x)
        h = p.csr()
        r = np.linalg.norm(h @ x - info.eigval * x) / abs(info.eigval)
        assert r <= 1.5 * tol, (i, r)


def test_batch_graph_and_direct_agree(gpu):
    from rqutils_amd.batch import SQDBatch
    probs = _instances(k=4, seed=80)
    b = SQDBatch([((p.labels, p.coeffs), p.keys, p.n) for p in probs], bucket_cap=512)
    x0 = torch.from_numpy(_rand(np.random.default_rng(9), b.n)).cuda()
    r1 = b.solve(x0, precision="c128", tol=1e-8, use_graph=False)
    r2 = b.solve(x0, precision="c128", tol=1e-8, use_graph=True)
    # replays of the captured graph run the same kernels with the same data
    for a, c in zip(r1.infos, r2.infos):
        assert a.niter == c.niter and a.eigval == c.eigval
    assert torch.equal(r1.vectors, r2.vectors)


def _mixed_instances():
    """Two instances whose subspaces suit the dense lookup (>= 2^16 keys in a
    2^17 or 2^18 key space) and three small searched ones."""
    probs = [small_problem(300, n=17, terms=50, nstates=70000),
             small_problem(301, n=12, terms=40, nstates=900),
             small_problem(302, n=18, terms=60, nstates=80000),
             small_problem(303, n=10, terms=30, nstates=500),
             small_problem(304, n=13, terms=45, nstates=2000)]
    return probs


@pytest.mark.parametrize("split", [1, 2])
def test_batch_mixed_dense_and_search_equals_single(gpu, split):
    """A batch whose suitable instances take the dense lookup runs two
    launches per matvec; every instance still gets exactly its own matvec."""
    from rqutils_amd.batch import SQDBatch
    from rqutils_amd.operator import PauliOperator, Subspace
    probs = _mixed_instances()
    b = SQDBatch([((p.labels, p.coeffs), p.keys, p.n) for p in probs], bucket_cap=1024,
                 rows_per_thread=4, group_split=split, dense=None)
    assert b.n_dense == 2 and [sp.dense for sp in b.spaces] == [True, False, True, False, False]
    rng = np.random.default_rng(split)
    v = _rand(rng, b.n)
    for dtype in (torch.complex128, torch.complex64):
        vt = torch.from_numpy(v).to(dtype).cuda()
        out = b.matvec(vt)
        for i, p in enumerate(probs):
            sp = Subspace(p.keys, p.n, bucket_cap=1024, rows_per_thread=4, dense=None)
            op = PauliOperator((p.labels, p.coeffs), sp)
            op.set_tune(group_split=split)
            ref = op.matvec(b.segment(vt, i).contiguous())
            assert torch.equal(b.segment(out, i), ref), (i, dtype)
            if dtype == torch.complex128 and i in (0, 1):
                want = p.matvec(b.segment(vt, i).cpu().numpy())
                seg = b.segment(out, i).cpu().numpy()
                assert np.linalg.norm(seg - want) <= 1e-12 * np.linalg.norm(want)


def test_batch_mixed_solve(gpu):
    from rqutils_amd.batch import SQDBatch
    probs = _mixed_instances()
    b = SQDBatch([((p.labels, p.coeffs), p.keys, p.n) for p in probs], dense=None)
    x0 = torch.from_numpy(_rand(np.random.default_rng(0), b.n)).cuda()
    res = b.solve(x0, precision="c128", tol=1e-9, maxiter=4000, use_graph=True)
    for p, info in zip(probs, res.infos):
        ev = p.ground()
        assert info.converged and abs(info.eigval - ev) < 1e-8 * abs(ev), (p.n, info, ev)
