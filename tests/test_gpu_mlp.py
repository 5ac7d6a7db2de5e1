"""Start-vector network on the GPU: the MFMA GEMM, the elementwise kernels and
the full forward/backward against the FP64 oracle, then the gain of the
trained network's start vector over a Gaussian start.

Tolerances: the GEMM accumulates BF16 products in FP32, so each entry is held
to 1e-5 of sum_k |a_ik b_kj| (FP32 summation of up to 4,096 terms); the
network's gradients are held to 2e-3 (relative Frobenius) against the oracle
that rounds to BF16 at the same points as the kernels, and to 5e-2 against
exact FP64 calculus.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo
from rqutils_amd import _lib

pytestmark = pytest.mark.gpu


def _bf16_dev(a, dev):
    """(device int16 tensor of bf16 bits, float64 values of those bits)."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    bits = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return torch.from_numpy(bits.view(np.int16).copy()).to(dev), mo.bf16_round(f)


def _bf16_host(t):
    u = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def _f32(a, dev):
    return torch.tensor(np.asarray(a, dtype=np.float32), device=dev)


def _round8(x):
    return (x + 7) // 8 * 8


# (M, N, K, ksplit); the last two shapes take the 128 x 128 tiles (at least
# 1024 of them), the others the 64 x 64 ones
_GEMM_SHAPES = [(16, 16, 16, 1), (37, 45, 29, 1), (64, 2, 256, 1), (256, 256, 1000, 5),
                (5, 300, 64, 2), (4097, 4100, 72, 1), (1000, 1100, 333, 16)]


@pytest.mark.parametrize("M,N,K,ksplit", _GEMM_SHAPES)
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("aligned", [False, True])
def test_gemm_bf16_matches_numpy(gpu, M, N, K, ksplit, ta, tb, aligned):
    """Leading dimensions padded to an odd size take the element-wise loads;
    multiples of 8 take the 16-byte loads (ragged edges element-wise)."""
    print(f"gemm M={M} N={N} K={K} ksplit={ksplit} ta={ta} tb={tb} aligned={aligned}", flush=True)
    rng = np.random.default_rng(M * 7 + N * 3 + K + 2 * ta + tb)
    opa, opb = rng.normal(size=(M, K)), rng.normal(size=(K, N))
    lda, ldb, ldc = (M if ta else K) + 3, (K if tb else N) + 1, N + 2
    if aligned:
        lda, ldb = _round8(M if ta else K), _round8(K if tb else N)
    a_store = np.zeros((K, lda) if ta else (M, lda))
    b_store = np.zeros((N, ldb) if tb else (K, ldb))
    if ta:
        a_store[:, :M] = opa.T
    else:
        a_store[:, :K] = opa
    if tb:
        b_store[:, :K] = opb.T
    else:
        b_store[:, :N] = opb
    a16, ar = _bf16_dev(a_store, gpu)
    b16, br = _bf16_dev(b_store, gpu)
    opa_r = ar[:, :M].T if ta else ar[:, :K]
    opb_r = br[:, :K].T if tb else br[:, :N]
    bias = rng.normal(size=N).astype(np.float32)
    c = torch.full((M, ldc), 7.0, dtype=torch.float32, device=gpu)
    bias_t = _f32(bias, gpu)
    L = _lib.lib()
    nb = C.c_size_t(0)
    _lib.check(L.sqdk_gemm_bf16_workspace(M, N, ksplit, C.byref(nb)), "ws")
    work = torch.empty(max(1, nb.value), dtype=torch.uint8, device=gpu)
    _lib.check(L.sqdk_gemm_bf16(M, N, K, a16.data_ptr(), lda, ta, b16.data_ptr(), ldb, tb,
                                c.data_ptr(), ldc, bias_t.data_ptr(), ksplit,
                                work.data_ptr() if nb.value else None, nb.value,
                                _lib.stream_ptr(gpu)), "gemm")
    torch.cuda.synchronize()
    got = c.cpu().numpy().astype(np.float64)
    ref = opa_r @ opb_r + bias
    bound = np.abs(opa_r) @ np.abs(opb_r)
    assert np.all(np.abs(got[:, :N] - ref) <= 1e-5 * bound + 1e-6)
    assert np.all(got[:, N:] == 7.0)


def test_gemm_bf16_rejects_bad_arguments(gpu):
    L = _lib.lib()
    x = torch.zeros(256, dtype=torch.int16, device=gpu)
    y = torch.zeros(256, dtype=torch.float32, device=gpu)
    with pytest.raises(ValueError):
        _lib.check(L.sqdk_gemm_bf16(16, 16, 16, x.data_ptr(), 16, 0, x.data_ptr(), 16, 0,
                                    y.data_ptr(), 16, None, 4, None, 0, _lib.stream_ptr(gpu)), "g")
    with pytest.raises(ValueError):
        _lib.check(L.sqdk_gemm_bf16(0, 16, 16, x.data_ptr(), 16, 0, x.data_ptr(), 16, 0,
                                    y.data_ptr(), 16, None, 1, None, 0, _lib.stream_ptr(gpu)), "g")


@pytest.mark.parametrize("n,cols", [(1, 1), (7, 3), (4096, 256), (5000, 2), (1000, 37),
                                    (33, 100), (4099, 17)])
def test_colsum_matches_numpy(gpu, n, cols):
    """Column sums (the bias gradients) against an FP64 sum; a second launch
    gives the same bits (fixed summation order)."""
    rng = np.random.default_rng(n + cols)
    x = rng.normal(size=(n, cols)).astype(np.float32)
    xt = _f32(x, gpu)
    outs = []
    for _ in range(2):
        out = torch.full((cols + 1,), 5.0, dtype=torch.float32, device=gpu)
        _lib.check(_lib.lib().sqdk_colsum(n, cols, xt.data_ptr(), out.data_ptr(),
                                          _lib.stream_ptr(gpu)), "colsum")
        outs.append(out.cpu().numpy())
    ref = x.astype(np.float64).sum(axis=0)
    scale = np.abs(x).astype(np.float64).sum(axis=0)
    assert np.all(np.abs(outs[0][:cols] - ref) <= 1e-6 * scale + 1e-30)
    assert outs[0][cols] == 5.0
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("n_extra", [0, 2])
def test_encode_bit_exact(gpu, n_extra):
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 1 << 40, size=3001, dtype=np.uint64)
    kt = torch.from_numpy(keys.view(np.int64)).to(gpu)
    extra = rng.normal(size=(keys.size, n_extra)).astype(np.float32)
    et = _f32(extra, gpu)
    out = torch.empty((keys.size, 48), dtype=torch.int16, device=gpu)
    _lib.check(_lib.lib().sqdk_mlp_encode(kt.data_ptr(), keys.size, 40,
                                          et.data_ptr() if n_extra else None, n_extra, 48,
                                          out.data_ptr(), _lib.stream_ptr(gpu)), "encode")
    ref = mo.encode(keys, 40, 48, extra=mo.bf16_round(extra) if n_extra else None)
    assert np.array_equal(_bf16_host(out), ref)


def test_activation_kernels(gpu):
    rng = np.random.default_rng(6)
    n = 20011
    z = (3.0 * rng.normal(size=n)).astype(np.float32)
    # a few exact bf16 ties for the cast
    z[:64] = ((rng.integers(0x3F00, 0x4100, size=64).astype(np.uint32) << 16) | 0x8000).view(
        np.float32)
    g = rng.normal(size=n).astype(np.float32)
    zt, gt = _f32(z, gpu), _f32(g, gpu)
    L = _lib.lib()
    s = _lib.stream_ptr(gpu)
    for op, ref in ((0, mo.gelu(z.astype(np.float64))),
                    (1, g.astype(np.float64) * mo.gelu_grad(z.astype(np.float64))),
                    (2, z.astype(np.float64))):
        o16 = torch.empty(n, dtype=torch.int16, device=gpu)
        o32 = torch.empty(n, dtype=torch.float32, device=gpu)
        _lib.check(L.sqdk_mlp_act(op, n, zt.data_ptr(), gt.data_ptr(), o16.data_ptr(),
                                  o32.data_ptr(), s), "act")
        torch.cuda.synchronize()
        got32 = o32.cpu().numpy()
        assert np.allclose(got32, ref, rtol=2e-6, atol=2e-7), op
        # the bf16 output is the round-to-nearest-even of the FP32 result
        assert np.array_equal(_bf16_host(o16), mo.bf16_round(got32)), op


def test_loss_kernel(gpu):
    rng = np.random.default_rng(7)
    n = 5003
    y = (2.0 * rng.normal(size=(n, 2))).astype(np.float32)
    y[:5, 1] = [40.0, -40.0, 0.0, 90.0, -90.0]         # saturated logits
    tgt = np.stack([rng.normal(size=n), rng.integers(0, 2, n), rng.uniform(0, 3, n)],
                   axis=1).astype(np.float32)
    wsum = float(tgt[:, 2].astype(np.float64).sum())
    dy32 = torch.empty((n, 2), dtype=torch.float32, device=gpu)
    dy16 = torch.empty((n, 2), dtype=torch.int16, device=gpu)
    part = torch.empty(7, dtype=torch.float64, device=gpu)
    yt, tt = _f32(y, gpu), _f32(tgt, gpu)     # named: a temporary's memory is reused at once
    _lib.check(_lib.lib().sqdk_mlp_loss(n, yt.data_ptr(), 2, tt.data_ptr(),
                                        wsum, 0.8, dy32.data_ptr(), dy16.data_ptr(),
                                        part.data_ptr(), 7, _lib.stream_ptr(gpu)), "loss")
    torch.cuda.synchronize()
    yd, td = y.astype(np.float64), tgt.astype(np.float64)
    a, p, w = td[:, 0], td[:, 1], td[:, 2]
    bce = np.maximum(yd[:, 1], 0) - yd[:, 1] * p + np.log1p(np.exp(-np.abs(yd[:, 1])))
    loss_ref = float(np.sum(w * ((yd[:, 0] - a) ** 2 + 0.8 * bce)) / wsum)
    sg = 1.0 / (1.0 + np.exp(-yd[:, 1]))
    dy_ref = np.stack([2 * w * (yd[:, 0] - a) / wsum, 0.8 * w * (sg - p) / wsum], axis=1)
    got = dy32.cpu().numpy()
    assert np.allclose(got, dy_ref, rtol=1e-5, atol=1e-6 * np.abs(dy_ref).max())
    assert np.array_equal(_bf16_host(dy16), mo.bf16_round(got))
    assert abs(float(part.sum()) - loss_ref) <= 1e-5 * abs(loss_ref)


def test_adam_kernel(gpu):
    rng = np.random.default_rng(8)
    n = 4099
    w, g, m = (rng.normal(size=n).astype(np.float32) for _ in range(3))
    v = rng.uniform(0, 2, size=n).astype(np.float32)
    wt, gt, mt, vt = (_f32(a, gpu) for a in (w, g, m, v))
    w16 = torch.empty(n, dtype=torch.int16, device=gpu)
    _lib.check(_lib.lib().sqdk_adam(n, wt.data_ptr(), gt.data_ptr(), mt.data_ptr(), vt.data_ptr(),
                                    w16.data_ptr(), 0.01, 0.9, 0.99, 1e-8, 3,
                                    _lib.stream_ptr(gpu)), "adam")
    torch.cuda.synchronize()
    wr, mr, vr = mo.adam(*(a.astype(np.float64) for a in (w, g, m, v)), 0.01, 0.9, 0.99, 1e-8, 3)
    assert np.allclose(mt.cpu().numpy(), mr, rtol=1e-6, atol=1e-7)
    assert np.allclose(vt.cpu().numpy(), vr, rtol=1e-6, atol=1e-7)
    assert np.allclose(wt.cpu().numpy(), wr, rtol=1e-6, atol=1e-6)
    assert np.array_equal(_bf16_host(w16), mo.bf16_round(wt.cpu().numpy()))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n_extra", [0, 1])
def test_startnet_forward_backward_matches_oracle(gpu, n_extra):
    from rqutils_amd.startnet import StartNet
    nq, n = 20, 3000
    net = StartNet(nq, hidden=256, device=gpu, seed=3, lam=0.7, n_extra=n_extra)
    rng = np.random.default_rng(9)
    for i in (1, 3, 5):        # non-zero biases so that every term is exercised
        net.params[i].copy_(_f32(0.1 * rng.normal(size=tuple(net.params[i].shape)), gpu))
    net._sync_bf16()
    keys = rng.integers(0, 1 << nq, size=n, dtype=np.uint64)
    tgt = np.stack([rng.normal(size=n), rng.integers(0, 2, n), rng.uniform(0.1, 3, n)],
                   axis=1).astype(np.float32)
    extra = rng.uniform(0, 3, size=(n, n_extra)).astype(np.float32)
    kt = torch.from_numpy(keys.view(np.int64)).to(gpu)
    tt = _f32(tgt, gpu)
    et = _f32(extra, gpu) if n_extra else None
    loss = float(net.loss_grads(kt, tt, extra=et).item())
    y_dev = net.forward(kt, et).cpu().numpy().astype(np.float64)
    grads_dev = [g.cpu().numpy().astype(np.float64) for g in net.grads]
    p = [t.cpu().numpy().astype(np.float64) for t in net.params]
    pb = [mo.bf16_round(p[0]), p[1], mo.bf16_round(p[2]), p[3], mo.bf16_round(p[4]), p[5]]
    x = mo.encode(keys, nq, net.kpad, extra=mo.bf16_round(extra) if n_extra else None)
    td = tgt.astype(np.float64)
    loss_o, grads_o, y_o, _ = mo.loss_and_grads(pb, x, td, lam=0.7, rounding=mo.bf16_round)
    assert _rel(y_dev, y_o) < 1e-4
    assert abs(loss - loss_o) < 1e-4 * abs(loss_o)
    for k, (gd, go) in enumerate(zip(grads_dev, grads_o)):
        if k == 0:   # W1 padding rows see zero inputs
            assert not gd[nq + n_extra:].any()
        assert _rel(gd, go) < 2e-3, (k, _rel(gd, go))
    _, grads_x, _, _ = mo.loss_and_grads(pb, x, td, lam=0.7)
    for k, (gd, gx) in enumerate(zip(grads_dev, grads_x)):
        assert _rel(gd, gx) < 5e-2, (k, _rel(gd, gx))


@pytest.mark.parametrize("nq,n_extra,n", [(20, 0, 3000), (28, 2, 1001), (16, 0, 17),
                                           (40, 2, 5000), (64, 3, 333)])
def test_fused_inference_matches_forward(gpu, nq, n_extra, n):
    """sqdk_mlp_infer (one kernel, activations in registers) against the
    training path's forward pass (encode, three GEMMs, GELUs) and the oracle."""
    from rqutils_amd.startnet import StartNet
    net = StartNet(nq, hidden=256, device=gpu, seed=5, n_extra=n_extra)
    rng = np.random.default_rng(nq + n)
    for i in (1, 3, 5):
        net.params[i].copy_(_f32(0.1 * rng.normal(size=tuple(net.params[i].shape)), gpu))
    net._sync_bf16()
    hi = np.uint64(0xFFFFFFFFFFFFFFFF) if nq == 64 else np.uint64((1 << nq) - 1)
    keys = rng.integers(0, np.iinfo(np.uint64).max, size=n, dtype=np.uint64, endpoint=True) & hi
    extra = rng.uniform(0, 3, size=(n, n_extra)).astype(np.float32)
    kt = torch.from_numpy(keys.view(np.int64)).to(gpu)
    et = _f32(extra, gpu) if n_extra else None
    y_ref = net.forward(kt, et).cpu().numpy()
    y = net.infer(kt, et).cpu().numpy()
    assert np.max(np.abs(y - y_ref)) <= 1e-5 * max(1.0, np.max(np.abs(y_ref)))
    p = [t.cpu().numpy().astype(np.float64) for t in net.params]
    pb = [mo.bf16_round(p[0]), p[1], mo.bf16_round(p[2]), p[3], mo.bf16_round(p[4]), p[5]]
    x = mo.encode(keys, nq, net.kpad, extra=mo.bf16_round(extra) if n_extra else None)
    _, _, y_o, _ = mo.loss_and_grads(pb, x, np.zeros((n, 3)) + [0, 0, 1], lam=0.0,
                                     rounding=mo.bf16_round)
    assert _rel(y.astype(np.float64), y_o) < 1e-4


def test_startnet_start_beats_gaussian(gpu):
    """Train on six disordered 16-qubit Heisenberg chains (ground vectors from
    the FP64 LOBPCG), then start the FP32/FP64 LOBPCG on two held-out chains
    from the network and from a Gaussian vector: both reach the FP64 energy
    within 1e-6 relative; the network start needs at most 3/4 of the
    iterations of the Gaussian start and no more than the one-hot start at the
    smallest diagonal entry (the reference's default)."""
    from rqutils_amd import workloads as W
    from rqutils_amd.sqd import SQDProblem
    from rqutils_amd.startnet import StartNet, gaussian_start, ground_targets

    def problem(seed):
        w = W.heisenberg28(seed=seed, n_states=None, n_random=200, num_qubits=16)
        return SQDProblem(w.hamiltonian, w.keys, 16, device=gpu)

    keys, tgts = [], []
    for seed in range(100, 106):
        prob = problem(seed)
        t, _, info = ground_targets(prob)
        assert info.converged
        keys.append(prob.space.keys)
        tgts.append(t)
    net = StartNet(16, device=gpu, seed=0)
    losses = net.train(keys, tgts, steps=3000, lr=3e-3, log_every=500)
    print("losses", losses, flush=True)
    assert losses[-1] < 0.1 * losses[0]
    for seed in (200, 201):
        prob = problem(seed)
        t, psi, info = ground_targets(prob)
        e_ref = info.eigval
        x_net = net.start_vector(prob.space.keys)
        overlap = abs(torch.vdot(psi, x_net).item()) / psi.norm().item()
        acc = net.sign_accuracy(prob.space.keys, t)
        iters = {}
        for name, x0 in (("net", x_net), ("gauss", gaussian_start(prob.n, gpu, seed)),
                         ("min_diag", prob.vinit_min_diag())):
            _, inf = prob.solve(x0, mode="fused", precision="c64", criterion="relative",
                                tol=1e-5, maxiter=2000)
            assert inf.converged, name
            assert abs(inf.eigval - e_ref) <= 1e-6 * abs(e_ref), (name, inf.eigval, e_ref)
            iters[name] = inf.niter
        print(f"seed {seed}: overlap {overlap:.4f} sign acc {acc:.4f} iters {iters}", flush=True)
        assert overlap > 0.8
        assert iters["net"] <= 0.75 * iters["gauss"]
        assert iters["net"] <= iters["min_diag"]
