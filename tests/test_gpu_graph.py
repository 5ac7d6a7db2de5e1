"""Graph-replayed refresh periods of the fused single solve
(sqdk_locg_opts.use_graph): the same kernels in the same order, so the result
equals the directly launched solve bit for bit, including a maxiter that is
not a multiple of the period and a refresh period longer than the host poll."""
import pytest
import torch

from tests.helpers import small_problem

pytestmark = pytest.mark.gpu


def _problem(gpu, seed, cached, prec):
    from rqutils_amd import _lib
    from rqutils_amd.sqd import SQDProblem
    p = small_problem(seed, n=12, terms=90, nstates=1500)
    prob = SQDProblem((p.labels, p.coeffs), p.keys, p.n, device=gpu)
    if cached:
        prob.op.build_cache(_lib.C128 if prec == "c128" else _lib.C64)
    prob.op.use_cache = cached
    return prob, p


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("cached", [False, True])
@pytest.mark.parametrize("precond", [None, "diag"])
@pytest.mark.parametrize("tol,maxiter,refresh", [(1e-9, 500, 16), (1e-300, 37, 16),
                                                 (1e-10, 400, 5)])
def test_graph_solve_equals_direct(gpu, prec, cached, precond, tol, maxiter, refresh):
    print(f"graph prec={prec} cached={cached} precond={precond} tol={tol} maxiter={maxiter} "
          f"refresh={refresh}", flush=True)
    prob, p = _problem(gpu, 41, cached, prec)
    g = torch.Generator().manual_seed(9)
    x0 = torch.randn(p.keys.size, dtype=torch.complex128, generator=g).to(gpu)
    if prec == "c64" and tol > 1e-200:
        tol = max(tol, 1e-6)         # FP32 vectors stall near a 1e-7 relative residual
    out = {}
    for graph in (False, True):
        x, info = prob.solve(x0, mode="fused", precision=prec, criterion="relative", tol=tol,
                             maxiter=maxiter, precond=precond, refresh_every=refresh,
                             use_graph=graph)
        torch.cuda.synchronize()
        out[graph] = (x.clone(), info)
    (xa, ia), (xb, ib) = out[False], out[True]
    assert ia.niter == ib.niter and ia.status == ib.status and ia.converged == ib.converged
    assert ia.eigval == ib.eigval
    assert torch.equal(xa, xb)
    if tol > 1e-200:
        e0 = p.ground()
        assert ia.converged and abs(ia.eigval - e0) < (1e-6 if prec == "c64" else 1e-8) * abs(e0)
    else:
        assert ia.niter == maxiter


def test_graph_solve_then_stream_work(gpu):
    """The private stream is drained before the call returns: work queued on
    the caller's stream afterwards sees the final vector."""
    prob, p = _problem(gpu, 43, True, "c128")
    x0 = torch.ones(p.keys.size, dtype=torch.complex128, device=gpu)
    x, info = prob.solve(x0, mode="fused", precision="c128", tol=1e-10, maxiter=400,
                         use_graph=True)
    y = prob.matvec(x)
    assert info.converged and abs(x.norm().item() - 1.0) < 1e-12
    rq = torch.vdot(x, y).real.item()
    assert abs(rq - info.eigval) < 1e-9 * abs(info.eigval)
