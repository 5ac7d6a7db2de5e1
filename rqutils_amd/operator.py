"""Device-resident subspace and Pauli-sum operator for the SQD matvec.

``Subspace`` holds the sorted uint64 keys of the sampled bitstrings in HBM with
the directories the bucket-join matvec needs: the bucket directory (first key
of every aligned bucket ``key >> m``), the fine directory inside each bucket
and the row chunks one workgroup processes. ``PauliOperator`` holds a Pauli
sum grouped by X mask, the groups ordered by class (``x >> m``) so that one
staged target bucket serves every group of a class. Both are plain torch
tensors handed to the C-ABI library by pointer.

These replace the reference's packed state array and cached source indices
(rqutils/sqd.py:316-351): nothing of size J x N is ever stored.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .paulis.symplectic import PauliSumXZ, _bits_to_u64

__all__ = ["Subspace", "PauliOperator", "dense_fits", "states_to_keys", "keys_to_states",
           "full_space_matmul", "DEFAULT_BUCKET_CAP", "DEFAULT_ROWS_PER_THREAD"]

DEFAULT_BUCKET_CAP = 8192       # keys per bucket staged in LDS (32 KiB as uint32)
DEFAULT_ROWS_PER_THREAD = 4     # rows per thread; a chunk is 256 x this
DEFAULT_FINE_SHIFT = -2         # sub-buckets of about 4 keys (log2 of mean bucket - 2 fine bits)
MAX_DIR_BITS = 26               # at most 2^26 buckets
DENSE_BUCKET_BITS = 16          # default bucket width of the dense lookup (16 KiB of LDS)


def dense_fits(num_qubits: int, n: int) -> bool:
    """Whether the dense lookup is worth its 6 * 2^(num_qubits - 5) bytes: a
    subspace of at least 2^16 keys, and at most 64 MiB or 48 bytes per key."""
    if num_qubits < 10 or num_qubits - 5 > 26 or n < (1 << 16):
        return False
    nbytes = 6 * (1 << (num_qubits - 5))
    return nbytes <= (64 << 20) or nbytes <= 48 * n


def states_to_keys(states: np.ndarray) -> np.ndarray:
    """uint64 key ``sum_q s_q 2^q`` of each binary row (column 0 = qubit n-1)."""
    states = np.asarray(states)
    if states.ndim != 2:
        raise ValueError("states must be a 2-d binary array")
    return _bits_to_u64(states.astype(np.uint8))


def keys_to_states(keys: np.ndarray, num_qubits: int) -> np.ndarray:
    """Binary rows (column 0 = qubit n-1) of uint64 keys."""
    keys = np.asarray(keys, dtype=np.uint64)
    shifts = np.arange(num_qubits - 1, -1, -1, dtype=np.uint64)
    return ((keys[:, None] >> shifts[None, :]) & np.uint64(1)).astype(np.uint8)


def _as_i64(t: torch.Tensor) -> torch.Tensor:
    return t if t.dtype == torch.int64 else t.view(torch.int64)


def _device(device) -> torch.device:
    dev = torch.device(device if device is not None else "cuda")
    if dev.type != "cuda":
        raise ValueError(f"the SQD kernels run on a GPU device, not {dev}")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


class Subspace:
    """Sorted, distinct uint64 keys in HBM plus the bucket-join directories.

    ``keys``: numpy uint64 array or a torch int64/uint64 tensor, strictly
    ascending, each below ``2**num_qubits``. ``bucket_bits`` defaults to the
    largest m whose biggest bucket holds at most ``bucket_cap`` keys.

    ``dense=True`` builds the dense lookup instead: 32-value sub-buckets
    (``fine_bits = bucket_bits - 5``) with an occupancy bitmap each, so the
    matvec finds a target with one LDS read and a popcount. It costs
    6 * 2^(num_qubits - 5) bytes of HBM whatever the subspace size
    (``dense_fits``); ``bucket_bits`` then defaults to 16. ``dense=None``
    picks it when it fits.

    ``block_threads=512`` (with 4 rows per thread) makes 2,048-row chunks, so
    one staged target table serves twice the rows of a 256-thread chunk.
    ``fine_shift`` sets the search's directory when ``fine_bits`` is not
    given: ``round(log2(mean non-empty bucket)) + fine_shift`` fine bits
    (default -2: about 4 keys per sub-bucket).
    """

    def __init__(self, keys, num_qubits: int, *, device=None, bucket_bits: int | None = None,
                 bucket_cap: int = DEFAULT_BUCKET_CAP,
                 rows_per_thread: int = DEFAULT_ROWS_PER_THREAD, fine_bits: int | None = None,
                 dense: bool | None = False, block_threads: int = 256,
                 fine_shift: int | None = None):
        L = _lib.lib()
        self.device = _device(device)
        if not 1 <= num_qubits <= 64:
            raise ValueError(f"num_qubits={num_qubits} outside 1..64")
        self.num_qubits = int(num_qubits)
        if isinstance(keys, torch.Tensor):
            kt = _as_i64(keys.to(self.device)).contiguous()
        else:
            arr = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64))
            kt = torch.from_numpy(arr.view(np.int64)).to(self.device)
        n = kt.numel()
        if n < 1 or n >= 2 ** 31:
            raise ValueError(f"subspace size {n} outside 1..2^31-1")
        self.keys = kt
        self.n = int(n)
        self.fine_shift = DEFAULT_FINE_SHIFT if fine_shift is None else int(fine_shift)
        stream = _lib.stream_ptr(self.device)
        with torch.cuda.device(self.device):
            self.pc_mask = torch.zeros(2, dtype=torch.int64, device=self.device)
            err = torch.zeros(1, dtype=torch.int32, device=self.device)
            _lib.check(L.sqdk_space_scan(kt.data_ptr(), n, self.num_qubits,
                                         self.pc_mask.data_ptr(), err.data_ptr(), stream),
                       "sqdk_space_scan")
            n_err = int(err.item())
            if n_err != 0:
                raise ValueError(f"keys must be strictly ascending and below 2^{num_qubits} "
                                 f"({n_err} violations)")
            self.rows_per_thread = int(rows_per_thread)
            if block_threads not in (256, 512) or (block_threads == 512 and rows_per_thread != 4):
                raise ValueError("block_threads must be 256, or 512 with 4 rows per thread")
            self.block_threads = int(block_threads)
            self.chunk_rows = self.block_threads * self.rows_per_thread
            if dense is None:
                dense = dense_fits(self.num_qubits, self.n)
            self.dense = bool(dense)
            self.fbits = None
            if self.dense:
                self._plan_dense(bucket_bits)
            else:
                self._plan(bucket_bits, bucket_cap, fine_bits)

    # -- planning --------------------------------------------------------------

    def _dir(self, m: int) -> torch.Tensor:
        L = _lib.lib()
        nb = 1 << (self.num_qubits - m)
        d = torch.empty(nb + 1, dtype=torch.int32, device=self.device)
        _lib.check(L.sqdk_space_dir(self.keys.data_ptr(), self.n, self.num_qubits, m,
                                    d.data_ptr(), _lib.stream_ptr(self.device)),
                   "sqdk_space_dir")
        return d

    def _widest_bits(self, cap: int) -> int:
        """Largest m for which no bucket ``key >> m`` holds more than ``cap`` keys.

        The keys are strictly ascending, so a bucket holds more than ``cap`` of
        them exactly when some window ``keys[i] .. keys[i + cap]`` lies in it,
        i.e. when the highest set bit of ``keys[i] ^ keys[i + cap]`` is below m.
        The answer is the smallest such highest bit over all windows (64 when
        there is no window): one pass and one host read, where a search over m
        would build a directory per probe.
        """
        if self.n <= cap:
            return 64
        x = torch.bitwise_xor(self.keys[cap:], self.keys[:-cap])
        # (x >> 1) as an unsigned value, so the minimum is the unsigned one;
        # x > 0, and its highest bit is the bit length of x >> 1.
        v = int(((x >> 1) & 0x7FFFFFFFFFFFFFFF).min().item())
        return v.bit_length()

    def _plan_dense(self, bucket_bits):
        nq = self.num_qubits
        if nq < 5:
            raise ValueError("the dense lookup needs at least 5 qubits")
        lo_m = max(5, nq - MAX_DIR_BITS)
        m = min(nq, DENSE_BUCKET_BITS) if bucket_bits is None else int(bucket_bits)
        m = max(m, lo_m)
        if not lo_m <= m <= min(nq, 19):
            raise ValueError(f"bucket_bits={m} outside {lo_m}..{min(nq, 19)} for the dense lookup")
        # Buckets must hold fewer than 2^16 keys (16-bit offsets).
        m = max(lo_m, min(m, self._widest_bits(65535)))
        self._plan(m, None, m - 5)
        L = _lib.lib()
        self.fbits = torch.empty(self.n_buckets << self.fine_bits, dtype=torch.int32,
                                 device=self.device)
        _lib.check(L.sqdk_space_bits(self.keys.data_ptr(), self.n, nq, self.bucket_bits,
                                     self.fine_bits, self.bstart.data_ptr(), self.fdir.data_ptr(),
                                     self.fbits.data_ptr(), _lib.stream_ptr(self.device)),
                   "sqdk_space_bits")

    def _plan(self, bucket_bits, cap, fine_bits):
        L = _lib.lib()
        nq = self.num_qubits
        lo_m = max(0, nq - MAX_DIR_BITS)
        hi_m = min(nq, 63)
        if bucket_bits is None:
            # Largest m with max bucket <= cap (the max bucket shrinks with m).
            widest = self._widest_bits(cap)
            if widest < lo_m:
                raise ValueError(f"keys too concentrated: a bucket at the finest level 2^{lo_m} "
                                 f"holds more than {cap} keys")
            bucket_bits = min(widest, hi_m)
        if not lo_m <= bucket_bits <= hi_m:
            raise ValueError(f"bucket_bits={bucket_bits} outside {lo_m}..{hi_m}")
        self.bucket_bits = int(bucket_bits)
        self.bstart = self._dir(self.bucket_bits)
        self.n_buckets = 1 << (nq - self.bucket_bits)
        sizes = self.bstart[1:] - self.bstart[:-1]
        max_bucket, nonempty = torch.stack([sizes.max(), torch.count_nonzero(sizes)]).tolist()
        self.max_bucket = int(max_bucket)
        if self.max_bucket >= 65536:
            raise ValueError(f"largest bucket holds {self.max_bucket} keys (limit 65535)")
        if fine_bits is None:
            # Sized by the mean non-empty bucket: 2^-fine_shift of its keys per
            # sub-bucket on average (4 by default). A coarser directory stages
            # fewer bytes per class; the search then takes a few more LDS steps.
            mean = float(self.n) / max(1, int(nonempty))
            fine_bits = min(self.bucket_bits, 13,
                            max(0, int(round(np.log2(max(mean, 1.0)))) + self.fine_shift))
            # and at most max(2^22, 8 n) directory entries in all (few keys in
            # many buckets would otherwise make it huge)
            budget = max(1 << 22, 8 * self.n)
            while fine_bits > 0 and (self.n_buckets << fine_bits) > budget:
                fine_bits -= 1
        self.fine_bits = int(fine_bits)
        self.fdir = torch.empty(self.n_buckets << self.fine_bits, dtype=torch.int16,
                                device=self.device)
        stream = _lib.stream_ptr(self.device)
        _lib.check(L.sqdk_space_fine(self.keys.data_ptr(), self.n, nq, self.bucket_bits,
                                     self.fine_bits, self.bstart.data_ptr(),
                                     self.fdir.data_ptr(), stream), "sqdk_space_fine")
        max_items = self.n_buckets + -(-self.n // self.chunk_rows)
        self.items = torch.empty(max_items, dtype=torch.int32, device=self.device)
        n_items = torch.zeros(1, dtype=torch.int32, device=self.device)
        _lib.check(L.sqdk_space_items(self.bstart.data_ptr(), self.n_buckets, self.chunk_rows,
                                      self.items.data_ptr(), max_items, n_items.data_ptr(),
                                      stream), "sqdk_space_items")
        n_items_h, pc0, pc1 = torch.cat([n_items.to(torch.int64), self.pc_mask]).tolist()
        self.n_items = int(n_items_h)
        popcounts = int(pc0) & ((1 << 64) - 1)
        popcounts |= (int(pc1) & 1) << 64
        self.popcounts = popcounts

    # -- C view ----------------------------------------------------------------

    def c_struct(self) -> _lib.SqdkSpace:
        s = _lib.SqdkSpace()
        s.keys = self.keys.data_ptr()
        s.bstart = self.bstart.data_ptr()
        s.fdir = self.fdir.data_ptr()
        s.items = self.items.data_ptr()
        s.pc_mask = self.pc_mask.data_ptr()
        s.n = self.n
        s.num_qubits = self.num_qubits
        s.bucket_bits = self.bucket_bits
        s.fine_bits = self.fine_bits
        s.chunk_rows = self.chunk_rows
        s.n_items = self.n_items
        s.max_bucket = self.max_bucket
        s.fbits = self.fbits.data_ptr() if self.fbits is not None else None
        return s

    def keys_numpy(self) -> np.ndarray:
        return self.keys.cpu().numpy().view(np.uint64)

    def single_parity(self) -> bool:
        """True if every key has the same popcount parity."""
        even = sum((self.popcounts >> w) & 1 for w in range(0, 65, 2))
        odd = sum((self.popcounts >> w) & 1 for w in range(1, 65, 2))
        return even == 0 or odd == 0

    def xsource(self, xmask: int) -> torch.Tensor:
        """Source index of ``key ^ xmask`` for every key, or -1 (int32, on device)."""
        out = torch.empty(self.n, dtype=torch.int32, device=self.device)
        sp = self.c_struct()
        _lib.check(_lib.lib().sqdk_xsource(C.byref(sp), int(xmask) & ((1 << 64) - 1),
                                           out.data_ptr(), _lib.stream_ptr(self.device)),
                   "sqdk_xsource")
        return out


@dataclass
class _HostTables:
    xmask: np.ndarray
    grp_off: np.ndarray
    zmask: np.ndarray
    coef: np.ndarray
    class_hi: np.ndarray
    class_off: np.ndarray
    diag_terms: int
    pruned_groups: int


class PauliOperator:
    """A Pauli sum bound to a subspace: ``H`` restricted to ``S``.

    ``ham``: a ``PauliSumXZ`` (padded or not) or ``(labels, coeffs)``. Groups
    that can never hit a key of ``S`` (odd X weight when every key has the
    same popcount parity) are dropped: they contribute nothing.
    """

    def __init__(self, ham, space: Subspace, *, prune: bool = True):
        if not isinstance(ham, PauliSumXZ):
            ham = PauliSumXZ.from_paulisum(ham, force_real=True, add_padding=True)
        if ham.num_qubits != space.num_qubits:
            raise ValueError(f"Hamiltonian has {ham.num_qubits} qubits, subspace "
                             f"{space.num_qubits}")
        self.ham = ham
        self.space = space
        self.num_qubits = ham.num_qubits
        self.is_complex = bool(np.iscomplexobj(ham.c))
        tables = self._tables(ham, space, prune)
        self.host = tables
        dev = space.device
        as_t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev).view(dt)  # noqa: E731
        self.xmask = as_t(tables.xmask.view(np.int64), torch.int64)
        self.grp_off = as_t(tables.grp_off.astype(np.int32), torch.int32)
        self.zmask = as_t(tables.zmask.view(np.int64), torch.int64)
        c = np.ascontiguousarray(tables.coef.astype(np.complex128))
        self.coef = torch.from_numpy(c.view(np.float64)).to(dev)
        self.coef32 = torch.from_numpy(c.astype(np.complex64).view(np.float32)).to(dev)
        self.class_hi = as_t(tables.class_hi.astype(np.uint32).view(np.int32), torch.int32)
        self.class_off = as_t(tables.class_off.astype(np.int32), torch.int32)
        self.n_groups = int(tables.xmask.shape[0])
        self.n_terms = int(tables.zmask.shape[0])
        self.n_classes = int(tables.class_hi.shape[0])
        self.diag_terms = tables.diag_terms
        self.tune = _lib.SqdkTune(space.rows_per_thread, 1, 1, 0)
        self._cstruct = None
        self._work = {}
        self.use_cache = False     # products through the cached hit table (build_cache)
        self.cache_budget = None   # largest table (bytes) built on first use; None: any
        self._caches = {}          # precision -> (SqdkCache, tensors...)
        self._cache_too_big = set()
        self.cache_info = {}

    def set_tune(self, group_split: int = 1, stage_v: int = 0, xcd_remap: int = 1,
                 compact: int = 0):
        """Launch options of the matvec (rows per thread are fixed by the subspace).
        ``compact``: apply hits through the per-wave LDS queue (4 or 8 rows per
        thread); same bits as the default register mode."""
        if not 1 <= int(group_split) <= 1024:
            raise ValueError(f"group_split={group_split} outside 1..1024")
        if compact and self.space.rows_per_thread not in (4, 8):
            raise ValueError("compact hits need 4 or 8 rows per thread")
        self.tune.group_split = int(group_split)
        self.tune.stage_v = int(stage_v)
        self.tune.xcd_remap = int(xcd_remap)
        self.tune.compact = int(bool(compact))
        self._work.clear()

    def work_bytes(self, precision: int) -> int:
        nbytes = C.c_size_t(0)
        s = self.space.c_struct()
        _lib.check(_lib.lib().sqdk_matvec_workspace(C.byref(s), C.byref(self.tune), precision,
                                                    C.byref(nbytes)), "sqdk_matvec_workspace")
        return nbytes.value

    def _workspace(self, precision: int):
        key = (precision, self.tune.group_split)
        if key not in self._work:
            nb = self.work_bytes(precision)
            self._work[key] = torch.empty(nb, dtype=torch.uint8, device=self.space.device) \
                if nb else None
        return self._work[key]

    @staticmethod
    def _tables(ham: PauliSumXZ, space: Subspace, prune: bool) -> _HostTables:
        xm, toff, tz, tc = ham.ragged()
        m = space.bucket_bits
        is_diag = xm == 0
        diag_idx = np.nonzero(is_diag)[0]
        off_idx = np.nonzero(~is_diag)[0]
        pruned = 0
        if prune and space.single_parity() and off_idx.size:
            weights = np.bitwise_count(xm[off_idx])
            keep = (weights & 1) == 0
            pruned = int(np.count_nonzero(~keep))
            off_idx = off_idx[keep]
        hi = (xm[off_idx] >> np.uint64(m)) if m < 64 else np.zeros(off_idx.size, np.uint64)
        order = np.lexsort((xm[off_idx], hi))
        off_idx = off_idx[order]
        hi = hi[order]
        # the X = 0 terms first, then the off-diagonal groups in class order;
        # within a group its nonzero terms in slot order
        groups = np.concatenate([diag_idx, off_idx])
        per_group = toff[groups + 1] - toff[groups]
        # the groups' term ranges laid end to end
        ends = np.cumsum(per_group)
        take = np.arange(int(ends[-1]) if ends.size else 0, dtype=np.int64)
        take += np.repeat(toff[groups] - (ends - per_group), per_group)
        zmask = tz[take]
        coefs = tc[take]
        diag_terms = int(per_group[:diag_idx.size].sum())
        offs = diag_terms + np.concatenate([[0], np.cumsum(per_group[diag_idx.size:])])
        if hi.size:
            starts = np.concatenate([[0], np.nonzero(hi[1:] != hi[:-1])[0] + 1])
            class_hi = hi[starts]
            class_off = np.concatenate([starts, [hi.size]])
        else:
            class_hi = np.zeros(0, np.uint64)
            class_off = np.zeros(1, np.int64)
        return _HostTables(xm[off_idx].astype(np.uint64), np.asarray(offs, np.int64),
                           zmask.astype(np.uint64), coefs, class_hi, class_off, diag_terms,
                           pruned)

    def c_struct(self) -> _lib.SqdkHam:
        h = _lib.SqdkHam()
        h.xmask = self.xmask.data_ptr()
        h.grp_off = self.grp_off.data_ptr()
        h.zmask = self.zmask.data_ptr() if self.n_terms else 0
        h.coef = self.coef.data_ptr() if self.n_terms else 0
        h.coef32 = self.coef32.data_ptr() if self.n_terms else 0
        h.class_hi = self.class_hi.data_ptr() if self.n_classes else 0
        h.class_off = self.class_off.data_ptr()
        h.n_groups = self.n_groups
        h.n_terms = self.n_terms
        h.n_classes = self.n_classes
        h.diag_t0 = 0
        h.diag_t1 = self.diag_terms
        h.num_qubits = self.num_qubits
        h.bucket_bits = self.space.bucket_bits
        return h

    def structs(self):
        """Cached (ham, space, tune) ctypes structs (keep the operator alive while used)."""
        if self._cstruct is None:
            self._cstruct = (self.c_struct(), self.space.c_struct(), self.tune)
        return self._cstruct

    def build_cache(self, precision: int, max_bytes: int | None = None) -> bool:
        """Build the hit table of this operator in ``precision`` (``_lib.C128``
        or ``_lib.C64``): every row's (source row, group coefficient) pairs,
        found once by the bucket join, in 64-row column-major slices, plus the
        diagonal and each row's hit count. Later products read the table instead of searching
        (``use_cache``). Returns False, building nothing, when the table would
        take more than ``max_bytes``. The sparse counterpart of the reference's
        source-index cache (rqutils/sqd.py:353-371)."""
        if precision in self._caches:
            return True
        L = _lib.lib()
        dev = self.space.device
        h, s, _ = self.structs()
        ns = C.c_int64(0)
        _lib.check(L.sqdk_cache_slices(C.byref(s), C.byref(ns)), "sqdk_cache_slices")
        ns = ns.value
        stream = _lib.stream_ptr(dev)
        if self._caches:
            # the layout (slot map, slice bases, row lengths, class mask) does
            # not depend on the precision: the other precision's count is reused
            other = next(iter(self._caches.values()))
            slot_row, sbase, rlen, cmask = other[1], other[2], other[6], other[7]
            entries = next(iter(self.cache_info.values()))["entries"]
        else:
            slot_row = torch.empty(64 * ns, dtype=torch.int32, device=dev)
            sbase = torch.empty(ns + 1, dtype=torch.int64, device=dev)
            rlen = torch.empty(self.space.n, dtype=torch.int32, device=dev)
            # the count's record for the fill (class masks, rows' first hits)
            nw = C.c_size_t(0)
            _lib.check(L.sqdk_cache_workspace(C.byref(h), C.byref(s), C.byref(nw)),
                       "sqdk_cache_workspace")
            cmask = torch.empty(max(nw.value, 1), dtype=torch.uint8, device=dev)
            _lib.check(L.sqdk_cache_count(C.byref(h), C.byref(s), slot_row.data_ptr(),
                                          sbase.data_ptr(), rlen.data_ptr(), cmask.data_ptr(),
                                          cmask.numel(), stream), "sqdk_cache_count")
            entries = int(sbase[-1].item())
        cdt = torch.complex128 if precision == _lib.C128 else torch.complex64
        esz = 16 if precision == _lib.C128 else 8
        nbytes = entries * (4 + esz) + self.space.n * (esz + 4) + ns * (8 + 64 * 4)
        if max_bytes is not None and nbytes > max_bytes:
            return False
        col = torch.empty(max(entries, 1), dtype=torch.int32, device=dev)
        coef = torch.empty(max(entries, 1), dtype=cdt, device=dev)
        diag = torch.empty(self.space.n, dtype=cdt, device=dev)
        _lib.check(L.sqdk_cache_fill(C.byref(h), C.byref(s), precision, sbase.data_ptr(),
                                     slot_row.data_ptr(), col.data_ptr(), coef.data_ptr(),
                                     diag.data_ptr(), rlen.data_ptr(), cmask.data_ptr(),
                                     cmask.numel(), stream), "sqdk_cache_fill")
        # few entries per row: the product is latency-bound, a wave takes 4 slices
        spw = 4 if entries <= 8 * self.space.n else 1
        cs = _lib.SqdkCache(self.space.n, ns, slot_row.data_ptr(), sbase.data_ptr(),
                            col.data_ptr(), coef.data_ptr(), diag.data_ptr(), rlen.data_ptr(),
                            precision, spw)
        self._caches[precision] = (cs, slot_row, sbase, col, coef, diag, rlen, cmask)
        self.cache_info[precision] = {"slices": ns, "entries": entries, "bytes": nbytes,
                                      "slices_per_wave": spw}
        return True

    def cache_struct(self, precision: int):
        """The SqdkCache of ``precision`` when products go through the table
        (building it on first use, within ``cache_budget``), else None (the
        bucket join)."""
        if not self.use_cache or precision in self._cache_too_big:
            return None
        if precision not in self._caches and not self.build_cache(precision, self.cache_budget):
            self._cache_too_big.add(precision)
            return None
        return self._caches[precision][0]

    def drop_cache(self):
        self._caches = {}
        self._cache_too_big = set()
        self.cache_info = {}

    def matvec(self, v: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """``H v`` for a complex128 or complex64 device vector of length ``n``."""
        if v.dtype == torch.complex128:
            prec = _lib.C128
        elif v.dtype == torch.complex64:
            prec = _lib.C64
        else:
            raise ValueError(f"vector dtype {v.dtype} not complex128/complex64")
        if v.numel() != self.space.n or v.device != self.space.device:
            raise ValueError("vector does not match the subspace size/device")
        v = v.contiguous()
        if out is None:
            out = torch.empty_like(v)
        cs = self.cache_struct(prec)
        if cs is not None:
            _lib.check(_lib.lib().sqdk_cache_matvec(C.byref(cs), v.data_ptr(), out.data_ptr(),
                                                    _lib.stream_ptr(self.space.device)),
                       "sqdk_cache_matvec")
            return out
        h, s, t = self.structs()
        work = self._workspace(prec)
        _lib.check(_lib.lib().sqdk_matvec(C.byref(h), C.byref(s), prec, v.data_ptr(),
                                          out.data_ptr(), C.byref(t),
                                          0 if work is None else work.data_ptr(),
                                          0 if work is None else work.numel(),
                                          _lib.stream_ptr(self.space.device)), "sqdk_matvec")
        return out

    def diagonal_of_group0(self) -> torch.Tensor:
        """Coefficient vector of the X = 0 group (complex128; zeros if absent)."""
        out = torch.zeros(self.space.n, dtype=torch.complex128, device=self.space.device)
        if self.diag_terms:
            _lib.check(_lib.lib().sqdk_diagonal(
                self.zmask.data_ptr(), self.coef.data_ptr(), 0, self.diag_terms,
                self.space.keys.data_ptr(), self.space.n, out.data_ptr(),
                _lib.stream_ptr(self.space.device)), "sqdk_diagonal")
        return out


def full_space_matmul(ham: PauliSumXZ, rhs):
    """``H @ rhs`` over the whole ``2**n`` space (rqutils/paulis/symplectic.py:131-150)."""
    n = ham.num_qubits
    is_numpy = not isinstance(rhs, torch.Tensor)
    vec = torch.as_tensor(np.asarray(rhs)) if is_numpy else rhs
    if vec.shape[0] != 2 ** n:
        raise ValueError(f"RHS axis 0 size {vec.shape[0]} is incompatible with num_qubits={n}")
    if n > 31:
        raise ValueError("full-space products are limited to 31 qubits")
    dev = _device(None if is_numpy or vec.device.type != "cuda" else vec.device)
    # The whole-space subspace and the operator are built once per
    # Hamiltonian and device and kept on the Hamiltonian object.
    ops = ham.__dict__.setdefault("_fullspace_ops", {})
    op = ops.get(str(dev))
    if op is None:
        space = Subspace(torch.arange(2 ** n, dtype=torch.int64, device=dev), n, device=dev,
                         dense=None)
        op = PauliOperator(ham, space, prune=False)
        ops.clear()
        ops[str(dev)] = op
    cols = vec.reshape(vec.shape[0], -1)
    out = torch.empty(cols.shape, dtype=torch.complex128, device=dev)
    for j in range(cols.shape[1]):
        col = cols[:, j].to(dev, torch.complex128).contiguous()
        out[:, j] = op.matvec(col)
    out = out.reshape(vec.shape)
    if not (np.iscomplexobj(ham.c) or torch.is_complex(vec)):
        out = out.real
    if is_numpy:
        return out.cpu().numpy()
    return out.to(vec.device)
