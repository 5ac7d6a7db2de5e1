"""Autotuner of the bucket-join matvec launch.

Three knobs decide the matvec's speed on a given problem:

* the bucket cap (largest bucket staged in LDS), which fixes the bucket bits m,
  the number of classes and the LDS footprint (hence occupancy);
* the rows per thread (a row chunk is 256 x this), which trades staging reuse
  against idle lanes on small buckets;
* the group split, which deals a chunk's groups to several workgroups so that
  small subspaces still fill the 256 CUs (partial sums are reduced after);
* the lookup: a search in the staged keys (at a few sub-bucket sizes of its
  directory: a coarser one stages fewer bytes per class, a finer one needs
  fewer search steps), or, where its directory fits (``dense_fits``), the
  dense bitmap lookup at a few bucket widths;
* how hits are applied: by every lane for all its rows (register mode), or
  through a per-wave LDS queue, one lane per hit (compacted mode), which pays
  off once a wave has hits in most groups.

``autotune`` times every candidate with HIP events on the real problem (a few
matvecs each) and returns the fastest; results are memoised per problem
signature in this process and, if ``RQSQD_TUNE_CACHE`` names a JSON file, on
disk.
"""
from __future__ import annotations

import hashlib
import json
import os
from dataclasses import asdict, dataclass, field

import numpy as np
import torch

from . import _lib
from .operator import PauliOperator, Subspace, dense_fits
from .paulis.symplectic import PauliSumXZ

__all__ = ["TuneResult", "autotune", "signature", "DEFAULT_CAPS", "DEFAULT_ROWS"]

DEFAULT_CAPS = (512, 1024, 2048, 4096, 8192)
DEFAULT_ROWS = (4, 8)
DEFAULT_DENSE_BITS = (15, 16, 17)
DEFAULT_FINE_SHIFTS = (-1, -2, -3)
CU_COUNT = 256

_MEMO: dict = {}


@dataclass
class TuneResult:
    bucket_cap: int
    rows_per_thread: int
    group_split: int
    ms: float
    precision: str
    table: list = field(default_factory=list)   # (cap, rows, split, ms, dense bits, threads, compact, fine shift)
    dense_bits: int = 0                         # bucket bits of the dense lookup, 0: search
    block_threads: int = 256
    compact: int = 0                            # hits through the per-wave LDS queue
    fine_shift: int = -2                        # search directory (Subspace.fine_shift)

    def space_kw(self) -> dict:
        kw = {"rows_per_thread": self.rows_per_thread, "block_threads": self.block_threads}
        if self.dense_bits:
            return {**kw, "dense": True, "bucket_bits": self.dense_bits}
        return {**kw, "bucket_cap": self.bucket_cap, "dense": False,
                "fine_shift": self.fine_shift}


def signature(hamiltonian: PauliSumXZ, keys: torch.Tensor, num_qubits: int, precision: str) -> str:
    """Hash of what the timing depends on: qubits, keys, groups and precision."""
    h = hashlib.sha1()
    h.update(f"{num_qubits}:{keys.numel()}:{precision}:".encode())
    k = keys if keys.numel() <= 4096 else keys[:: max(1, keys.numel() // 4096)]
    h.update(k.cpu().numpy().tobytes())
    h.update(np.ascontiguousarray(hamiltonian.x).tobytes())
    h.update(str(hamiltonian.c.shape).encode())
    return h.hexdigest()


def _load_disk():
    path = os.environ.get("RQSQD_TUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save_disk(sig, res: TuneResult):
    path = os.environ.get("RQSQD_TUNE_CACHE")
    if not path:
        return
    data = _load_disk()
    data[sig] = asdict(res)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh)
    os.replace(tmp, path)


def _time_matvec(op: PauliOperator, v: torch.Tensor, reps: int) -> float:
    out = torch.empty_like(v)
    op.matvec(v, out)                      # warm-up (and workspace allocation)
    stream = torch.cuda.current_stream(v.device)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        op.matvec(v, out)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _split_candidates(n_items: int):
    # Enough workgroups to cover the CUs a few times over; never more than 64 shares.
    out = [1]
    s = 2
    while s <= 64 and n_items * (s // 2) < 4 * CU_COUNT:
        out.append(s)
        s *= 2
    return out


def autotune(hamiltonian, keys, num_qubits: int, *, device=None, precision: str = "c64",
             caps=DEFAULT_CAPS, rows=DEFAULT_ROWS, dense_bits=DEFAULT_DENSE_BITS,
             threads=(256,), compact=(0, 1), fine_shifts=DEFAULT_FINE_SHIFTS, reps: int = 3,
             use_cache: bool = True, verbose: bool = False) -> TuneResult:
    """Fastest (lookup, bucket cap or width, rows per thread, group split) for
    this problem. ``threads=(256, 512)`` also times 512-thread workgroups
    (off by default: measured faster only for complex64 on the sparse 40-qubit
    subspace, and slower for complex128 there)."""
    if not isinstance(hamiltonian, PauliSumXZ):
        hamiltonian = PauliSumXZ.from_paulisum(hamiltonian, force_real=True, add_padding=True)
    dev = torch.device(device if device is not None else "cuda")
    if isinstance(keys, torch.Tensor):
        kt = keys.to(dev)
    else:
        kt = torch.from_numpy(np.ascontiguousarray(np.asarray(keys, np.uint64)).view(np.int64)).to(dev)
    sig = signature(hamiltonian, kt, num_qubits, precision)
    if use_cache:
        if sig in _MEMO:
            return _MEMO[sig]
        disk = _load_disk()
        if sig in disk:
            res = TuneResult(**disk[sig])
            _MEMO[sig] = res
            return res
    n = kt.numel()
    dtype = torch.complex64 if precision == "c64" else torch.complex128
    g = torch.Generator(device="cpu").manual_seed(0)
    v = torch.complex(torch.randn(n, generator=g), torch.randn(n, generator=g)).to(dtype).to(dev)
    table = []
    best = None
    seen_m = {}
    # (cap, rows, dense bits, threads); 512-thread workgroups only at 4 rows
    shapes = [(r, 256) for r in rows] + ([(4, 512)] if 4 in rows and 512 in threads else [])
    candidates = [(int(c), r, 0, t, fs) for c in sorted(set(int(c) for c in caps))
                  for r, t in shapes for fs in fine_shifts]
    if dense_fits(num_qubits, n):
        lo_m, hi_m = max(5, num_qubits - 26), min(num_qubits, 19)
        candidates += [(0, r, int(m), t, 0) for m in dense_bits if lo_m <= m <= hi_m
                       for r, t in shapes]
    for cap, r, dm, nt, fs in candidates:
        try:
            if dm:
                sp = Subspace(kt, num_qubits, device=dev, dense=True, bucket_bits=dm,
                              rows_per_thread=r, block_threads=nt)
            else:
                sp = Subspace(kt, num_qubits, device=dev, bucket_cap=cap, rows_per_thread=r,
                              block_threads=nt, fine_shift=fs)
        except ValueError:
            continue            # keys too concentrated for this cap
        key = (sp.bucket_bits, sp.fine_bits, r, sp.dense, nt)
        if seen_m.get(key):
            continue            # same bucket level as a smaller cap: same kernel
        seen_m[key] = True
        op = PauliOperator(hamiltonian, sp)
        for cm in compact:
            if cm and r not in (4, 8):
                continue
            for split in _split_candidates(sp.n_items):
                op.set_tune(group_split=split, compact=cm)
                try:
                    ms = _time_matvec(op, v, reps)
                except (ValueError, RuntimeError):
                    continue    # the compacted areas exceed the LDS for this shape
                table.append((cap, r, split, round(ms, 4), dm, nt, cm, fs))
                if verbose:
                    print(f"[autotune] {'dense' if dm else 'cap=%d' % cap} m={sp.bucket_bits} "
                          f"fine={sp.fine_bits} rows={r} threads={nt} split={split} "
                          f"compact={cm}: {ms:.3f} ms", flush=True)
                if best is None or ms < best[3]:
                    best = (cap, r, split, ms, dm, nt, cm, fs)
        del op, sp
    if best is None:
        raise ValueError("no matvec configuration fits this subspace")
    res = TuneResult(best[0], best[1], best[2], best[3], precision, table, best[4], best[5],
                     best[6], best[7])
    _MEMO[sig] = res
    _save_disk(sig, res)
    return res


def check_library():
    """The tuner times the HIP kernels; fail early if the library is missing."""
    _lib.lib()
