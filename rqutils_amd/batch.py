"""Many independent SQD instances solved in one persistent launch per step.

``SQDBatch`` builds each instance's subspace and Pauli tables (the same
``Subspace`` / ``PauliOperator`` as a single solve, with one common rows-per-
thread and bucket cap), then lays them end to end on the device with global
row, group and term indices plus a descriptor per instance. One
``sqdk_batch_matvec`` launch then covers every instance, and
``sqdk_batch_locg_solve`` runs the fused FP32/FP64 LOBPCG on all of them at
once with HIP-graph-captured iterations; converged instances drop out of the
work while the others continue.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .ground_locg import LocgInfo, _CRITERIA, _PRECISIONS
from .operator import DEFAULT_BUCKET_CAP, PauliOperator, Subspace

__all__ = ["SQDBatch", "autotune_batch"]


def _i64(a):
    return torch.as_tensor(np.ascontiguousarray(a)).to(torch.int64)


@dataclass
class BatchSolve:
    vectors: torch.Tensor          # concatenated eigenvectors (solver precision)
    infos: list                    # LocgInfo per instance
    seconds: float = 0.0
    judged: list = field(default_factory=list)   # instances the Gaussian judge solve won


class SQDBatch:
    """A batch of (hamiltonian, keys, num_qubits) instances on one GPU."""

    def __init__(self, instances, *, device=None, bucket_cap: int = DEFAULT_BUCKET_CAP,
                 rows_per_thread: int = 4, group_split: int = 1, compact: int = 0,
                 dense: bool | None = False):
        """``dense``: False searches every instance's targets; None gives each
        instance whose subspace suits it (``operator.dense_fits``) the dense
        bitmap lookup, and one matvec is then two launches (dense, search)."""
        dev = torch.device(device if device is not None else "cuda")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        if not instances:
            raise ValueError("empty batch")
        if len(instances) > 65535:
            raise ValueError("at most 65535 instances per batch")
        self.spaces, self.ops = [], []
        for ham, keys, nq in instances:
            sp = Subspace(keys, nq, device=dev, bucket_cap=bucket_cap,
                          rows_per_thread=rows_per_thread, dense=dense)
            self.spaces.append(sp)
            self.ops.append(PauliOperator(ham, sp))
        self.rows_per_thread = rows_per_thread
        self._concat()
        self.tune = _lib.SqdkTune(rows_per_thread, int(group_split), 1, 0, int(bool(compact)))
        self._work = {}
        self.use_cache = False   # products (and solves) through the batch's hit table
        self._caches = {}
        self.cache_info = {}

    def _concat(self):
        dev = self.device
        keys, bstart, fdir, items, item_inst, fbits = [], [], [], [], [], []
        xmask, grp_off, class_hi, class_off = [], [], [], []
        gz, gc, gc32, dz, dc, dc32 = [], [], [], [], [], []
        descs = (_lib.SqdkInst * len(self.spaces))()
        row_base = [0]
        item_base = [0]
        # Term layout: every instance's group terms first (so that group ranges
        # stay contiguous across instances), then every instance's X = 0 terms.
        total_group_terms = sum(op.n_terms - op.diag_terms for op in self.ops)
        b_off = f_off = g_off = gt_off = dt_off = c_off = fb_off = 0
        max_bucket = max_fine = max_bits = dense_fine = 0
        for i, (sp, op) in enumerate(zip(self.spaces, self.ops)):
            r0 = row_base[-1]
            keys.append(sp.keys)
            bstart.append(sp.bstart.to(torch.int64) + r0)
            fdir.append(sp.fdir)
            items.append(sp.items[: sp.n_items].to(torch.int64) + r0)
            item_inst.append(torch.full((sp.n_items,), i, dtype=torch.int32, device=dev))
            nd = op.diag_terms
            xmask.append(op.xmask[: op.n_groups])
            grp_off.append(op.grp_off[:-1].to(torch.int64) - nd + gt_off)
            gz.append(op.zmask[nd:op.n_terms])
            gc.append(op.coef[2 * nd:2 * op.n_terms])
            gc32.append(op.coef32[2 * nd:2 * op.n_terms])
            dz.append(op.zmask[:nd])
            dc.append(op.coef[:2 * nd])
            dc32.append(op.coef32[:2 * nd])
            class_hi.append(op.class_hi[: op.n_classes])
            class_off.append(op.class_off[:-1].to(torch.int64) + g_off)
            d = descs[i]
            d.bstart_base = b_off
            d.fdir_base = f_off
            d.class_base = c_off
            d.n_classes = op.n_classes
            d.diag_t0 = total_group_terms + dt_off
            d.diag_t1 = total_group_terms + dt_off + nd
            d.bucket_bits = sp.bucket_bits
            d.fine_bits = sp.fine_bits
            if sp.dense:
                d.fbits_base = fb_off
                fbits.append(sp.fbits)
                fb_off += sp.fbits.numel()
                dense_fine = max(dense_fine, sp.fine_bits)
            else:
                d.fbits_base = -1           # search lookup
                max_bucket = max(max_bucket, sp.max_bucket)
                max_fine = max(max_fine, sp.fine_bits)
            b_off += sp.bstart.numel()
            f_off += sp.fdir.numel()
            g_off += op.n_groups
            gt_off += op.n_terms - nd
            dt_off += nd
            c_off += op.n_classes
            row_base.append(r0 + sp.n)
            item_base.append(item_base[-1] + sp.n_items)
            max_bits = max(max_bits, sp.bucket_bits)
        zmask = gz + dz
        coef = gc + dc
        coef32 = gc32 + dc32
        t_off = gt_off + dt_off
        # closing entries of the ragged arrays
        grp_off.append(torch.tensor([gt_off], dtype=torch.int64, device=dev))
        class_off.append(torch.tensor([g_off], dtype=torch.int64, device=dev))
        if row_base[-1] >= 2 ** 31:
            raise ValueError("batch holds 2^31 rows or more")
        self.n = int(row_base[-1])
        self.row_base_host = np.asarray(row_base, np.int64)
        self.keys = torch.cat(keys)
        self.bstart = torch.cat(bstart).to(torch.int32)
        self.fdir = torch.cat(fdir)
        self.items = torch.cat(items).to(torch.int32)
        self.item_inst = torch.cat(item_inst).to(torch.int16)
        self.item_base = torch.tensor(item_base, dtype=torch.int32, device=dev)
        self.row_base = torch.tensor(row_base, dtype=torch.int64, device=dev)
        self.xmask = torch.cat(xmask) if g_off else torch.zeros(1, dtype=torch.int64, device=dev)
        self.grp_off = torch.cat(grp_off).to(torch.int32)
        self.zmask = torch.cat(zmask) if t_off else torch.zeros(1, dtype=torch.int64, device=dev)
        self.coef = torch.cat(coef) if t_off else torch.zeros(2, dtype=torch.float64, device=dev)
        self.coef32 = torch.cat(coef32) if t_off else torch.zeros(2, dtype=torch.float32,
                                                                   device=dev)
        self.class_hi = torch.cat(class_hi) if c_off else torch.zeros(1, dtype=torch.int32,
                                                                       device=dev)
        self.class_off = torch.cat(class_off).to(torch.int32)
        raw = bytes(descs)
        self.inst = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(dev)
        self.n_items = int(item_base[-1])
        self.n_inst = len(self.spaces)
        s = _lib.SqdkBatch()
        s.keys = self.keys.data_ptr()
        s.bstart = self.bstart.data_ptr()
        s.fdir = self.fdir.data_ptr()
        s.items = self.items.data_ptr()
        s.item_inst = self.item_inst.data_ptr()
        s.item_base = self.item_base.data_ptr()
        s.row_base = self.row_base.data_ptr()
        s.xmask = self.xmask.data_ptr()
        s.grp_off = self.grp_off.data_ptr()
        s.zmask = self.zmask.data_ptr()
        s.coef = self.coef.data_ptr()
        s.coef32 = self.coef32.data_ptr()
        s.class_hi = self.class_hi.data_ptr()
        s.class_off = self.class_off.data_ptr()
        s.inst = self.inst.data_ptr()
        s.n = self.n
        s.n_inst = self.n_inst
        s.n_items = self.n_items
        s.chunk_rows = 256 * self.rows_per_thread
        s.max_bucket = max(1, max_bucket)
        s.max_fine_bits = max_fine
        s.max_bucket_bits = max_bits
        self.fbits = torch.cat(fbits) if fbits else None
        s.fbits = self.fbits.data_ptr() if fbits else None
        s.dense_fine_bits = dense_fine
        self.n_dense = len(fbits)
        self.cstruct = s

    # -- views -------------------------------------------------------------------

    def segment(self, vec: torch.Tensor, i: int) -> torch.Tensor:
        return vec[self.row_base_host[i]:self.row_base_host[i + 1]]

    def set_tune(self, group_split: int = 1, compact: int = 0):
        self.tune.group_split = int(group_split)
        self.tune.compact = int(bool(compact))
        self._work.clear()

    def _workspace(self, prec):
        key = (prec, self.tune.group_split)
        if key not in self._work:
            nb = C.c_size_t(0)
            _lib.check(_lib.lib().sqdk_batch_matvec_workspace(
                C.byref(self.cstruct), C.byref(self.tune), prec, C.byref(nb)),
                "sqdk_batch_matvec_workspace")
            self._work[key] = torch.empty(nb.value, dtype=torch.uint8, device=self.device) \
                if nb.value else None
        return self._work[key]

    def build_cache(self, precision: int) -> dict:
        """The hit table of the whole batch in ``precision`` (one table over
        the concatenated rows; see ``PauliOperator.build_cache``); products and
        solves read it while ``use_cache`` is set. Returns its size."""
        if precision in self._caches:
            return self.cache_info[precision]
        L = _lib.lib()
        dev = self.device
        self.cstruct.cache = C.POINTER(_lib.SqdkCache)()
        ns = C.c_int64(0)
        _lib.check(L.sqdk_batch_cache_slices(C.byref(self.cstruct), C.byref(ns)),
                   "sqdk_batch_cache_slices")
        ns = ns.value
        slot_row = torch.empty(64 * ns, dtype=torch.int32, device=dev)
        sbase = torch.empty(ns + 1, dtype=torch.int64, device=dev)
        rlen = torch.empty(self.n, dtype=torch.int32, device=dev)
        # the count's record for the fill (class masks, rows' first hits)
        max_classes = max((op.n_classes for op in self.ops), default=0)
        nw = C.c_size_t(0)
        _lib.check(L.sqdk_batch_cache_workspace(C.byref(self.cstruct), max_classes,
                                                C.byref(nw)), "sqdk_batch_cache_workspace")
        cmask = torch.empty(max(nw.value, 1), dtype=torch.uint8, device=dev)
        stream = _lib.stream_ptr(dev)
        _lib.check(L.sqdk_batch_cache_count(C.byref(self.cstruct), slot_row.data_ptr(),
                                            sbase.data_ptr(), rlen.data_ptr(), max_classes,
                                            cmask.data_ptr(), cmask.numel(), stream),
                   "sqdk_batch_cache_count")
        entries = int(sbase[-1].item())
        cdt = torch.complex128 if precision == _lib.C128 else torch.complex64
        col = torch.empty(max(entries, 1), dtype=torch.int32, device=dev)
        coef = torch.empty(max(entries, 1), dtype=cdt, device=dev)
        diag = torch.empty(self.n, dtype=cdt, device=dev)
        _lib.check(L.sqdk_batch_cache_fill(C.byref(self.cstruct), precision, sbase.data_ptr(),
                                           slot_row.data_ptr(), col.data_ptr(), coef.data_ptr(),
                                           diag.data_ptr(), rlen.data_ptr(), max_classes,
                                           cmask.data_ptr(), cmask.numel(), stream),
                   "sqdk_batch_cache_fill")
        spw = 4 if entries <= 8 * self.n else 1
        cs = _lib.SqdkCache(self.n, ns, slot_row.data_ptr(), sbase.data_ptr(), col.data_ptr(),
                            coef.data_ptr(), diag.data_ptr(), rlen.data_ptr(), precision, spw)
        self._caches[precision] = (cs, slot_row, sbase, col, coef, diag, rlen, cmask)
        esz = 16 if precision == _lib.C128 else 8
        self.cache_info[precision] = {"slices": ns, "entries": entries, "slices_per_wave": spw,
                                      "bytes": entries * (4 + esz) + self.n * (esz + 4)
                                      + ns * (8 + 64 * 4)}
        return self.cache_info[precision]

    def _cache_struct(self, precision: int):
        if not self.use_cache:
            return None
        self.build_cache(precision)
        return self._caches[precision][0]

    def matvec(self, v: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Every instance's H_i v_i in one launch (concatenated complex vectors)."""
        prec = _lib.C128 if v.dtype == torch.complex128 else _lib.C64
        if v.dtype not in (torch.complex128, torch.complex64) or v.numel() != self.n:
            raise ValueError("vector must be complex of the batch's total length")
        v = v.contiguous()
        out = torch.empty_like(v) if out is None else out
        cs = self._cache_struct(prec)
        if cs is not None:
            _lib.check(_lib.lib().sqdk_cache_matvec(C.byref(cs), v.data_ptr(), out.data_ptr(),
                                                    _lib.stream_ptr(self.device)),
                       "sqdk_cache_matvec")
            return out
        work = self._workspace(prec)
        _lib.check(_lib.lib().sqdk_batch_matvec(
            C.byref(self.cstruct), prec, v.data_ptr(), out.data_ptr(), C.byref(self.tune),
            0 if work is None else work.data_ptr(), 0 if work is None else work.numel(),
            _lib.stream_ptr(self.device)), "sqdk_batch_matvec")
        return out

    def start_vector(self, start="gaussian", *, mix: float | None = None,
                     seed: int = 0) -> torch.Tensor:
        """Concatenated start vectors: each instance's ``sqd.start_vector``
        (``"gaussian"``, ``"min_diag"`` or a ``StartNet``) with seed ``seed +
        i``. A network is applied to the instances of its qubit count; the
        others start from the Gaussian, the method without the network."""
        from .sqd import start_vector
        from .startnet import StartNet
        parts = []
        for i, op in enumerate(self.ops):
            st = start
            if isinstance(start, StartNet) and start.nq != op.space.num_qubits:
                st = "gaussian"
            parts.append(start_vector(op, st, mix=mix, seed=seed + i))
        return torch.cat(parts)

    def solve(self, x0: torch.Tensor | None = None, *, precision: str = "c64",
              criterion: str = "relative", tol: float = 1e-6, maxiter: int = 1000,
              refresh_every: int = 16, use_graph: bool = True, start=None,
              mix: float | None = None, seed: int = 0, judge: bool = False) -> BatchSolve:
        """Fused LOBPCG on every instance from the concatenated start vectors
        ``x0``, or from ``start_vector(start, mix=mix, seed=seed)``.

        ``judge``: solve the batch once more from Gaussian starts and keep, per
        instance, the lower of the two energies (both are Rayleigh quotients,
        upper bounds of the ground energy): the plain method judges the chosen
        start, and the result is at least as good as the plain method's.
        ``BatchSolve.judged`` lists the instances the Gaussian solve won."""
        if x0 is None:
            x0 = self.start_vector("gaussian" if start is None else start, mix=mix, seed=seed)
        elif start is not None:
            raise ValueError("give either x0 or start")
        res = self._solve(x0, precision, criterion, tol, maxiter, refresh_every, use_graph)
        if judge:
            t_first = res.seconds
            ref = self._solve(self.start_vector("gaussian", seed=seed), precision, criterion,
                              tol, maxiter, refresh_every, use_graph)
            for i, (a, b) in enumerate(zip(res.infos, ref.infos)):
                if b.converged and (not a.converged or b.eigval < a.eigval):
                    res.infos[i] = b
                    self.segment(res.vectors, i).copy_(self.segment(ref.vectors, i))
                    res.judged.append(i)
            res.seconds = t_first + ref.seconds
        return res

    def _solve(self, x0, precision, criterion, tol, maxiter, refresh_every,
               use_graph) -> BatchSolve:
        L = _lib.lib()
        prec = _PRECISIONS[precision]
        dtype = torch.complex128 if prec == _lib.C128 else torch.complex64
        if x0.numel() != self.n:
            raise ValueError("start vector must have the batch's total length")
        x = x0.reshape(-1).to(self.device, dtype).clone()
        opts = _lib.SqdkLocgOpts()
        opts.mode = _lib.LOCG_FUSED
        opts.precision = prec
        opts.criterion = _CRITERIA[criterion]
        opts.maxiter = int(maxiter)
        opts.tol = float(tol)
        opts.n_for_tol = 0
        opts.check_every = int(refresh_every)
        opts.refresh_every = int(refresh_every)
        opts.use_graph = 1 if use_graph else 0
        cs = self._cache_struct(prec)
        self.cstruct.cache = C.pointer(cs) if cs is not None else C.POINTER(_lib.SqdkCache)()
        nb = C.c_size_t(0)
        _lib.check(L.sqdk_batch_locg_workspace(C.byref(self.cstruct), C.byref(self.tune),
                                               C.byref(opts), C.byref(nb)),
                   "sqdk_batch_locg_workspace")
        work = torch.empty(nb.value, dtype=torch.uint8, device=self.device)
        res = (_lib.SqdkLocgResult * self.n_inst)()
        import time
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        try:
            _lib.check(L.sqdk_batch_locg_solve(C.byref(self.cstruct), C.byref(self.tune),
                                               x.data_ptr(), C.byref(opts), work.data_ptr(),
                                               nb.value, res, _lib.stream_ptr(self.device)),
                       "sqdk_batch_locg_solve")
        finally:
            self.cstruct.cache = C.POINTER(_lib.SqdkCache)()
        torch.cuda.synchronize(self.device)
        dt = time.perf_counter() - t0
        infos = [LocgInfo(r.eigval, r.niter, bool(r.converged), r.matvecs, r.resid_norm,
                          r.ax_norm, r.status) for r in res]
        return BatchSolve(x, infos, dt)


def autotune_batch(instances, *, device=None, caps=(2048, 8192), rows=(4,), splits=(1, 2, 4),
                   compact=(0, 1), dense=(False, None), reps: int = 3):
    """Fastest (bucket cap, rows per thread, group split, compacted hits, dense
    lookup) of the batched matvec; returns ((cap, rows, split, ms, compact,
    dense), table). ``dense`` None lets every suitable instance use the dense
    lookup."""
    best = None
    table = []
    shapes = [(cap, r, dn) for cap in caps for r in rows for dn in dense]
    for cap, r, dn in shapes:
        b = SQDBatch(instances, device=device, bucket_cap=cap, rows_per_thread=r, dense=dn)
        g = torch.Generator(device="cpu").manual_seed(0)
        v = torch.complex(torch.randn(b.n, generator=g), torch.randn(b.n, generator=g)).to(
            torch.complex64).to(b.device)
        for cm in compact:
            for s in splits:
                try:
                    b.set_tune(group_split=s, compact=cm)
                    out = b.matvec(v)
                except (ValueError, RuntimeError):
                    continue        # LDS budget exceeded for this shape
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    b.matvec(v, out)
                e1.record()
                e1.synchronize()
                ms = e0.elapsed_time(e1) / reps
                table.append((cap, r, s, round(ms, 4), cm, dn))
                if best is None or ms < best[3]:
                    best = (cap, r, s, ms, cm, dn)
        del b
    return best, table
