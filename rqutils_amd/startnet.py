"""Learned LOBPCG starting vectors: a BF16 MLP trained on ground vectors.

``StartNet`` maps a bitstring (bits as +-1, optionally followed by per-state
features such as the scaled diagonal energy, ``diag_feature``) through
256 -> 256 -> 2 units (GELU) to a log-amplitude and a sign logit. Its weights
live in FP32 with
BF16 copies; every product runs on the matrix cores through
``sqdk_gemm_bf16`` (FP32 accumulation), the activations, loss and Adam in
the library's elementwise kernels. It is trained on ground vectors that the
FP64 LOBPCG itself produced (``targets_from_vector``) and then proposes
``start_vector(keys)`` for new Hamiltonians; a plain Gaussian start remains
the fallback.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from . import _lib

__all__ = ["StartNet", "targets_from_vector", "ground_targets", "gaussian_start",
           "diag_feature", "energy_features", "mix_gaussian", "DEFAULT_MIX"]

# Norm of the Gaussian part a network start carries by default (the network's
# part keeps norm sqrt(1 - DEFAULT_MIX^2)). A network that puts its start on
# the wrong few states could otherwise leave LOBPCG at an excited state; with
# it the ground vector's component is at least of order DEFAULT_MIX / sqrt(n).
DEFAULT_MIX = 0.1


def _p(t):
    return 0 if t is None else t.data_ptr()


def targets_from_vector(psi: torch.Tensor) -> torch.Tensor:
    """Training targets (n x 3: log-amplitude, sign probability, weight) of a
    ground vector: the global phase is fixed so that the largest component is
    real and positive, the sign is that of the real part, the log-amplitude is
    centred by its |psi|^2-weighted mean and the weight is n |psi|^2."""
    psi = psi.to(torch.complex128)
    j = int(torch.argmax(psi.abs()).item())
    psi = psi * (psi[j].conj() / psi[j].abs())
    amp = psi.abs()
    w = amp ** 2
    w = w / w.sum()
    la = torch.log(amp + 1e-300)
    la = la - (w * la).sum()
    n = psi.numel()
    tgt = torch.stack([la, (psi.real > 0).to(torch.float64), w * n], dim=1)
    return tgt.to(torch.float32).contiguous()


def gaussian_start(n: int, device, seed: int = 0) -> torch.Tensor:
    """Seeded complex Gaussian start vector (the method without the network),
    drawn on the device."""
    dev = torch.device(device)
    g = torch.Generator(device=dev).manual_seed(seed)
    return torch.complex(torch.randn(n, generator=g, dtype=torch.float64, device=dev),
                         torch.randn(n, generator=g, dtype=torch.float64, device=dev))


def mix_gaussian(x: torch.Tensor, mix: float, seed: int = 0) -> torch.Tensor:
    """sqrt(1 - mix^2) x / |x| + mix g / |g| with a seeded Gaussian g,
    normalised (complex128)."""
    if not 0.0 <= mix < 1.0:
        raise ValueError("mix must lie in [0, 1)")
    x = x.to(torch.complex128)
    x = x / torch.linalg.vector_norm(x)
    if mix > 0.0:
        g = gaussian_start(x.numel(), x.device, seed)
        x = math.sqrt(1.0 - mix * mix) * x + mix * g / torch.linalg.vector_norm(g)
        x = x / torch.linalg.vector_norm(x)
    return x


def ground_targets(prob, *, tol: float = 1e-10, maxiter: int = 5000, seed: int = 0,
                   precond: str | None = None):
    """Ground vector of an ``SQDProblem`` from the FP64 fused LOBPCG (Gaussian
    start; ``precond="diag"`` for the preconditioned iteration) and its
    training targets. Returns (targets, vector, info)."""
    x0 = gaussian_start(prob.n, prob.space.keys.device, seed)
    vec, info = prob.solve(x0, mode="fused", precision="c128", criterion="relative", tol=tol,
                           maxiter=maxiter, precond=precond)
    return targets_from_vector(vec), vec, info


def diag_feature(prob) -> torch.Tensor:
    """Per-state feature (n x 1, FP32): the diagonal energy <b|H|b> over the
    subspace, shifted by its minimum and divided by its standard deviation,
    so that the lowest-energy configuration of every subspace reads 0."""
    d = prob.op.diagonal_of_group0().real.to(torch.float64)
    sd = float(d.std().item()) if d.numel() > 1 else 1.0
    f = (d - d.min()) / (sd if sd > 0 else 1.0)
    return f.to(torch.float32).reshape(-1, 1).contiguous()


def energy_features(prob) -> torch.Tensor:
    """Per-state features (n x 2, FP32): the scaled diagonal energy f of
    ``diag_feature`` and -log10(f + 1e-6) / 6, which is 1 at the subspace's
    lowest-energy configuration and at most about 0.2 wherever f >= 0.01, so
    that configurations within a small fraction of a standard deviation of
    the minimum are told apart from it."""
    f = diag_feature(prob).to(torch.float64)
    g = -torch.log10(f + 1e-6) / 6.0
    return torch.cat([f, g], dim=1).to(torch.float32).contiguous()


class StartNet:
    """BF16 MLP (num_qubits -> hidden -> hidden -> 2) with MFMA kernels."""

    def __init__(self, num_qubits: int, hidden: int = 256, *, device=None, seed: int = 0,
                 lam: float = 1.0, n_extra: int = 0):
        self.nq = int(num_qubits)
        self.n_extra = int(n_extra)
        self.n_in = self.nq + self.n_extra
        self.kpad = 16 * math.ceil(self.n_in / 16)
        self.h = int(hidden)
        self.lam = float(lam)
        dev = torch.device(device if device is not None else "cuda")
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        g = torch.Generator(device="cpu").manual_seed(seed)

        def init(fan_in, shape):
            bound = 1.0 / math.sqrt(fan_in)
            return ((torch.rand(shape, generator=g, dtype=torch.float64) * 2 - 1) * bound).to(
                torch.float32)

        w1 = init(self.n_in, (self.kpad, self.h))
        w1[self.n_in:] = 0
        self.params = [w1, torch.zeros(self.h), init(self.h, (self.h, self.h)),
                       torch.zeros(self.h), init(self.h, (self.h, 2)), torch.zeros(2)]
        self.params = [p.to(self.device).contiguous() for p in self.params]
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.grads = [torch.zeros_like(p) for p in self.params]
        self.bf16 = [torch.empty(p.shape, dtype=torch.int16, device=self.device)
                     if i % 2 == 0 else None for i, p in enumerate(self.params)]
        self._sync_bf16()
        self.step_count = 0
        self._ws = {}

    # -- helpers ---------------------------------------------------------------

    def _stream(self):
        return _lib.stream_ptr(self.device)

    def _sync_bf16(self):
        for p, q in zip(self.params, self.bf16):
            if q is not None:
                _lib.check(_lib.lib().sqdk_mlp_act(2, p.numel(), p.data_ptr(), None, q.data_ptr(),
                                                   None, self._stream()), "sqdk_mlp_act")

    def _gemm(self, M, N, K, A, lda, ta, B, ldb, tb, Cout, ldc, bias=None, ksplit=1):
        L = _lib.lib()
        work = None
        nbytes = 0
        if ksplit > 1:
            nb = C.c_size_t(0)
            _lib.check(L.sqdk_gemm_bf16_workspace(M, N, ksplit, C.byref(nb)), "gemm workspace")
            key = ("gemm", nb.value)
            if key not in self._ws:
                self._ws[key] = torch.empty(nb.value, dtype=torch.uint8, device=self.device)
            work = self._ws[key]
            nbytes = nb.value
        _lib.check(L.sqdk_gemm_bf16(M, N, K, A.data_ptr(), lda, ta, B.data_ptr(), ldb, tb,
                                    Cout.data_ptr(), ldc, _p(bias), ksplit, _p(work), nbytes,
                                    self._stream()), "sqdk_gemm_bf16")

    def _act(self, op, z, g=None, want16=True, want32=False):
        out16 = torch.empty(z.shape, dtype=torch.int16, device=self.device) if want16 else None
        out32 = torch.empty(z.shape, dtype=torch.float32, device=self.device) if want32 else None
        _lib.check(_lib.lib().sqdk_mlp_act(op, z.numel(), z.data_ptr(), _p(g), _p(out16), _p(out32),
                                           self._stream()), "sqdk_mlp_act")
        return out16, out32

    def _check_extra(self, keys, extra):
        if self.n_extra == 0:
            if extra is not None:
                raise ValueError("this network takes no extra features")
            return None
        if extra is None or tuple(extra.shape) != (keys.numel(), self.n_extra):
            raise ValueError(f"extra features must be ({keys.numel()}, {self.n_extra})")
        if extra.dtype != torch.float32 or not extra.is_contiguous() or extra.device != self.device:
            raise ValueError("extra features must be contiguous float32 on the network's device")
        return extra

    def encode(self, keys: torch.Tensor, extra: torch.Tensor | None = None) -> torch.Tensor:
        n = keys.numel()
        extra = self._check_extra(keys, extra)
        x = torch.empty((n, self.kpad), dtype=torch.int16, device=self.device)
        _lib.check(_lib.lib().sqdk_mlp_encode(keys.data_ptr(), n, self.nq, _p(extra),
                                              self.n_extra, self.kpad, x.data_ptr(),
                                              self._stream()), "sqdk_mlp_encode")
        return x

    # -- passes ----------------------------------------------------------------

    def forward(self, keys: torch.Tensor, extra: torch.Tensor | None = None, keep: bool = False):
        """Outputs y (n x 2, FP32) for uint64 keys (torch int64 tensor) and,
        with ``n_extra`` > 0, their n x n_extra features."""
        n = keys.numel()
        w1, b1, w2, b2, w3, b3 = self.params
        x = self.encode(keys, extra)
        z1 = torch.empty((n, self.h), dtype=torch.float32, device=self.device)
        self._gemm(n, self.h, self.kpad, x, self.kpad, 0, self.bf16[0], self.h, 0, z1, self.h, b1)
        h1, _ = self._act(0, z1)
        z2 = torch.empty_like(z1)
        self._gemm(n, self.h, self.h, h1, self.h, 0, self.bf16[2], self.h, 0, z2, self.h, b2)
        h2, _ = self._act(0, z2)
        y = torch.empty((n, 2), dtype=torch.float32, device=self.device)
        self._gemm(n, 2, self.h, h2, self.h, 0, self.bf16[4], 2, 0, y, 2, b3)
        if keep:
            self._cache = (x, z1, h1, z2, h2)
        return y

    def loss_grads(self, keys: torch.Tensor, target: torch.Tensor, wsum: float | None = None,
                   extra: torch.Tensor | None = None):
        """Loss and gradients (into ``self.grads``) of one batch.

        ``target`` is n x 3 (log-amplitude, sign probability, weight) in FP32;
        the loss is normalised by ``wsum`` (default: the batch's weight sum).
        Returns the loss as a 0-d device tensor (no host synchronisation)."""
        n = keys.numel()
        y = self.forward(keys, extra, keep=True)
        x, z1, h1, z2, h2 = self._cache
        if wsum is None:
            wsum = float(target[:, 2].sum().item())
        nblocks = 64
        dy32 = torch.empty((n, 2), dtype=torch.float32, device=self.device)
        dy16 = torch.empty((n, 2), dtype=torch.int16, device=self.device)
        part = torch.empty(nblocks, dtype=torch.float64, device=self.device)
        L = _lib.lib()
        s = self._stream()
        _lib.check(L.sqdk_mlp_loss(n, y.data_ptr(), 2, target.data_ptr(), wsum, self.lam,
                                   dy32.data_ptr(), dy16.data_ptr(), part.data_ptr(), nblocks, s),
                   "sqdk_mlp_loss")
        gw1, gb1, gw2, gb2, gw3, gb3 = self.grads
        def ks(m, cols, k):
            # K-split of the batch for the weight gradients (m x cols outputs):
            # about 256 workgroups of 64 x 64, at least 128 samples each
            tiles = -(-m // 64) * -(-cols // 64)
            return max(1, min(k // 128, 256 // tiles))
        # output layer: dW3 = H2^T dY, db3 = colsum dY, dH2 = dY W3^T
        self._gemm(self.h, 2, n, h2, self.h, 1, dy16, 2, 0, gw3, 2, ksplit=ks(self.h, 2, n))
        _lib.check(L.sqdk_colsum(n, 2, dy32.data_ptr(), gb3.data_ptr(), s), "sqdk_colsum")
        dh2 = torch.empty((n, self.h), dtype=torch.float32, device=self.device)
        self._gemm(n, self.h, 2, dy16, 2, 0, self.bf16[4], 2, 1, dh2, self.h)
        dz2_16, dz2 = self._act(1, z2, dh2, want16=True, want32=True)
        # hidden layer: dW2 = H1^T dZ2, db2 = colsum dZ2, dH1 = dZ2 W2^T
        self._gemm(self.h, self.h, n, h1, self.h, 1, dz2_16, self.h, 0, gw2, self.h,
                   ksplit=ks(self.h, self.h, n))
        _lib.check(L.sqdk_colsum(n, self.h, dz2.data_ptr(), gb2.data_ptr(), s), "sqdk_colsum")
        dh1 = torch.empty_like(dh2)
        self._gemm(n, self.h, self.h, dz2_16, self.h, 0, self.bf16[2], self.h, 1, dh1, self.h)
        dz1_16, dz1 = self._act(1, z1, dh1, want16=True, want32=True)
        # input layer: dW1 = X^T dZ1, db1 = colsum dZ1
        self._gemm(self.kpad, self.h, n, x, self.kpad, 1, dz1_16, self.h, 0, gw1, self.h,
                   ksplit=ks(self.kpad, self.h, n))
        _lib.check(L.sqdk_colsum(n, self.h, dz1.data_ptr(), gb1.data_ptr(), s), "sqdk_colsum")
        self._cache = None
        return part.sum()

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                  eps: float = 1e-8):
        """One Adam update of every parameter (refreshes the BF16 copies).
        The padding rows of W1 see zero inputs, hence zero gradients, and stay 0."""
        self.step_count += 1
        L = _lib.lib()
        for i, (p, g, m, v) in enumerate(zip(self.params, self.grads, self.m, self.v)):
            _lib.check(L.sqdk_adam(p.numel(), p.data_ptr(), g.data_ptr(), m.data_ptr(),
                                   v.data_ptr(), _p(self.bf16[i]), lr, beta1, beta2, eps,
                                   self.step_count, self._stream()), "sqdk_adam")

    def train(self, keys_list, targets_list, extra_list=None, *, steps: int = 3000,
              batch: int = 4096, lr: float = 3e-3, seed: int = 0, log_every: int = 0,
              floor: float = 0.05):
        """Adam with a cosine learning-rate schedule.

        Ground vectors are often concentrated on few states, so half of each
        batch is drawn with probability proportional to the target weight and
        half uniformly over the concatenated instances; each sample carries the
        importance weight (w + floor) / (w/2 + 1/2) (weights have mean 1), an
        unbiased estimate of the weighted loss plus ``floor`` times a uniform
        one that keeps every state's amplitude supervised. Samples are drawn 64
        steps at a time and the losses stay on the device, so the step loop
        waits on the GPU only where ``log_every`` prints. Returns the losses
        (host floats) of every ``max(1, log_every)``-th step and the last."""
        keys = torch.cat([k.to(self.device, torch.int64) for k in keys_list]).contiguous()
        tgts = torch.cat([t.to(self.device, torch.float32) for t in targets_list]).contiguous()
        extras = None
        if self.n_extra:
            if extra_list is None:
                raise ValueError("this network needs extra features for every instance")
            extras = torch.cat([e.to(self.device, torch.float32) for e in extra_list]).contiguous()
        n_all = keys.numel()
        w_all = tgts[:, 2].to(torch.float64)
        mean_w = float(w_all.mean().item())
        cdf = torch.cumsum(w_all, 0)
        total = float(cdf[-1].item())
        gen = torch.Generator(device=self.device).manual_seed(seed)
        half = batch // 2
        every = max(1, log_every)
        loss_buf = torch.zeros(steps, dtype=torch.float64, device=self.device)
        chunk = 64
        for c0 in range(0, steps, chunk):
            cn = min(chunk, steps - c0)
            # one draw for a chunk of steps: no host round trip inside the loop
            u = torch.rand((cn, half), generator=gen, device=self.device,
                           dtype=torch.float64) * total
            i1 = torch.searchsorted(cdf, u.reshape(-1)).clamp_(max=n_all - 1).view(cn, half)
            i2 = torch.randint(0, n_all, (cn, batch - half), generator=gen, device=self.device)
            idx_all = torch.cat([i1, i2], dim=1)
            wi_all = w_all[idx_all] / mean_w
            imp_all = ((wi_all + floor) / (0.5 * wi_all + 0.5)).to(torch.float32)
            for j in range(cn):
                step = c0 + j
                idx = idx_all[j]
                tb = tgts[idx]
                tb[:, 2] = imp_all[j]
                loss = self.loss_grads(
                    keys[idx], tb, wsum=float(batch),
                    extra=None if extras is None else extras[idx].contiguous())
                self.adam_step(lr=lr * 0.5 * (1.0 + math.cos(math.pi * step / steps)))
                loss_buf[step] = loss
                if log_every and step % every == 0:
                    print(f"[startnet] step {step} loss {float(loss.item()):.4f}", flush=True)
        keep = sorted(set(range(0, steps, every)) | {steps - 1}) if steps else []
        return loss_buf[keep].tolist()

    def infer(self, keys: torch.Tensor, extra: torch.Tensor | None = None) -> torch.Tensor:
        """Outputs y (n x 2, FP32) through the fused inference kernel
        (``sqdk_mlp_infer``; 256 hidden units): one launch, the same values as
        ``forward``."""
        if self.h != 256:
            raise ValueError("the fused inference kernel is built for 256 hidden units")
        keys = keys.contiguous()
        extra = self._check_extra(keys, extra)
        n = keys.numel()
        w1t = self.bf16[0].t().contiguous()
        w2t = self.bf16[2].t().contiguous()
        w3t = torch.zeros((16, self.h), dtype=torch.int16, device=self.device)
        w3t[:2] = self.bf16[4].t()
        y = torch.empty((n, 2), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().sqdk_mlp_infer(
            keys.data_ptr(), n, self.nq, _p(extra), self.n_extra, self.kpad, w1t.data_ptr(),
            self.params[1].data_ptr(), w2t.data_ptr(), self.params[3].data_ptr(),
            w3t.data_ptr(), self.params[5].data_ptr(), y.data_ptr(), self._stream()),
            "sqdk_mlp_infer")
        return y

    def predict(self, keys: torch.Tensor, extra: torch.Tensor | None = None,
                chunk: int = 1 << 18) -> torch.Tensor:
        """Outputs y (n x 2, FP32) over any number of states: the fused
        inference kernel (256 hidden units), else the training path in chunks."""
        if self.h == 256 and keys.numel() > 0:
            return self.infer(keys, extra)
        outs = []
        for s in range(0, keys.numel(), chunk):
            e = None if extra is None else extra[s:s + chunk].contiguous()
            outs.append(self.forward(keys[s:s + chunk].contiguous(), e))
        return torch.cat(outs)

    def sign_accuracy(self, keys: torch.Tensor, target: torch.Tensor,
                      extra: torch.Tensor | None = None) -> float:
        """Weighted fraction of states whose predicted sign matches the target."""
        y = self.predict(keys, extra)
        ok = ((y[:, 1] > 0) == (target[:, 1] > 0.5)).to(torch.float64)
        w = target[:, 2].to(torch.float64)
        return float((ok * w).sum() / w.sum())

    def start_vector(self, keys: torch.Tensor, extra: torch.Tensor | None = None,
                     chunk: int = 1 << 18, mix: float = 0.0, seed: int = 0) -> torch.Tensor:
        """Normalised complex128 start vector exp(y0) sign(y1) over the keys.

        ``mix`` > 0 adds a seeded Gaussian vector of norm ``mix`` (the network's
        part keeps norm sqrt(1 - mix^2)): the start then overlaps every
        eigenvector at least as a Gaussian start does, so a network that
        concentrates its start on the wrong few states cannot leave LOBPCG near
        an excited state."""
        if not 0.0 <= mix < 1.0:
            raise ValueError("mix must lie in [0, 1)")
        y = self.predict(keys, extra, chunk).to(torch.float64)
        a = y[:, 0] - y[:, 0].max()
        x = torch.exp(a) * torch.where(y[:, 1] > 0, 1.0, -1.0)
        return mix_gaussian(x, mix, seed)

    def features(self, prob) -> torch.Tensor | None:
        """The per-state features this network was built for over ``prob``'s
        subspace (anything with ``.op``): none, ``diag_feature`` (1) or
        ``energy_features`` (2)."""
        if self.n_extra == 0:
            return None
        if self.n_extra == 1:
            return diag_feature(prob)
        if self.n_extra == 2:
            return energy_features(prob)
        raise ValueError(f"no feature set of width {self.n_extra}")

    def start_for(self, prob, *, mix: float = DEFAULT_MIX, seed: int = 0) -> torch.Tensor:
        """Start vector over ``prob``'s subspace (an ``SQDProblem`` or anything
        with ``.space.keys`` and ``.op``) with a Gaussian part of norm ``mix``."""
        if prob.space.num_qubits != self.nq:
            raise ValueError(f"network built for {self.nq} qubits, subspace has "
                             f"{prob.space.num_qubits}")
        return self.start_vector(prob.space.keys, self.features(prob), mix=mix, seed=seed)
