"""Cached hit table (sqdk_cache_*): the hits of every group found once and
applied by a sliced sparse product. Its entries must be exactly the oracle's
off-diagonal matrix elements (sort-join construction), its diagonal the X = 0
group's, and its products must equal the bucket join's to the precision's
rounding, for the search and the dense lookup, every chunk size, both
precisions, Hamiltonians without off-diagonal hits or without a diagonal, and
inside the LOBPCG."""
import numpy as np
import pytest
import torch

from rqutils_amd import _lib
from tests.helpers import OracleProblem, small_problem

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _problem(p, **kw):
    from rqutils_amd.sqd import SQDProblem
    return SQDProblem((p.labels, p.coeffs), p.keys, p.n, **kw)


def _table_to_csr(op, prec, n):
    """Host CSR of the hit table (padding dropped), its diagonal and its hit
    count; checks the layout: every row in exactly one slot, the slots of a
    chunk sorted by hit count (most first, equal counts in row order), padding
    only after a row's hits, slots without a row all padding, slice widths
    equal to their largest count, row lengths."""
    import scipy.sparse as sps
    cs, slot_row, sbase, col, coef, diag, rlen, _ = op._caches[prec]
    slot_row = slot_row.cpu().numpy().view(np.uint32).reshape(-1, 64)
    sbase = sbase.cpu().numpy()
    col = col.cpu().numpy().view(np.uint32)
    coef = coef.cpu().numpy()
    rlen = rlen.cpu().numpy()
    seen = np.zeros(n, bool)
    rows, cols, vals = [], [], []
    per_chunk = op.space.chunk_rows // 64
    win = 256 // 64                                          # slices of a sort window
    for c0 in range(0, slot_row.shape[0], per_chunk):
        chunk = slot_row[c0:c0 + per_chunk].reshape(-1)
        real = chunk[chunk != 0xFFFFFFFF].astype(np.int64)
        assert np.all(chunk[real.size:] == 0xFFFFFFFF)      # empty slots last
        rows0 = real.min() if real.size else 0
        for w0 in range(0, per_chunk, win):
            part = slot_row[c0 + w0:c0 + w0 + win].reshape(-1)
            part = part[part != 0xFFFFFFFF].astype(np.int64)
            # the window holds its own 256 rows, sorted: count desc, row asc
            assert np.all((part - rows0) // 256 == w0 // win)
            order = np.lexsort((part, -rlen[part].astype(np.int64)))
            assert np.array_equal(order, np.arange(part.size))
    for sid in range(slot_row.shape[0]):
        b0, b1 = int(sbase[sid]), int(sbase[sid + 1])
        assert (b1 - b0) % 64 == 0
        width = (b1 - b0) // 64
        lanes = slot_row[sid]
        lens = [int(rlen[r]) for r in lanes if r != 0xFFFFFFFF]
        assert width == (max(lens) if lens else 0)
        for r in lanes:
            if r != 0xFFFFFFFF:
                assert not seen[r]
                seen[r] = True
        if width == 0:
            continue
        blk_c = col[b0:b1].reshape(width, 64)
        blk_v = coef[b0:b1].reshape(width, 64)
        for lane in range(64):
            real = blk_c[:, lane] != 0xFFFFFFFF
            r = int(lanes[lane])
            if r == 0xFFFFFFFF:
                assert not real.any()
                continue
            k = int(real.sum())
            assert rlen[r] == k
            assert real[:k].all() and not real[k:].any()
            assert np.all(blk_v[k:, lane] == 0)
            rows += [r] * k
            cols += blk_c[:k, lane].tolist()
            vals += blk_v[:k, lane].tolist()
    assert seen.all()
    m = sps.csr_matrix((np.asarray(vals, np.complex128), (rows, cols)), shape=(n, n))
    return m, diag.cpu().numpy().astype(np.complex128), len(rows)


@pytest.mark.parametrize("seed,dense_lookup,rows", [(0, False, 4), (1, True, 4), (2, False, 8),
                                                    (3, False, 16), (4, True, 8)])
def test_table_entries_match_oracle(gpu, seed, dense_lookup, rows):
    p = small_problem(seed, n=12, terms=90, nstates=1500)
    kw = dict(rows_per_thread=rows)
    if dense_lookup:
        kw.update(dense=True, bucket_bits=8)
    prob = _problem(p, cache=True, **kw)
    assert prob.op.build_cache(_lib.C128)
    import scipy.sparse as sps
    m, diag, real_entries = _table_to_csr(prob.op, _lib.C128, p.keys.size)
    ref = p.csr().astype(np.complex128).tocsr()
    off = ref - sps.diags(ref.diagonal())
    assert abs(m - off).max() < 1e-14 * max(1.0, abs(off).max())
    assert np.allclose(diag, ref.diagonal(), rtol=0, atol=1e-13)
    # exactly one entry per (row, off-diagonal group) hit of the oracle's sort join
    offg = [g for g in range(p.x.shape[0]) if np.any(p.x[g] != 0)]
    hits = int(sum((p.xsources[g] >= 0).sum() for g in offg))
    assert real_entries == hits
    assert prob.op.cache_info[_lib.C128]["entries"] % 64 == 0


@pytest.mark.parametrize("dense_lookup", [False, True])
@pytest.mark.parametrize("prec,tol", [("c128", 1e-13), ("c64", 3e-7)])
def test_cached_matvec_matches_join(gpu, dense_lookup, prec, tol):
    p = small_problem(11, n=12, terms=150, nstates=2500)
    kw = dict(dense=True, bucket_bits=8) if dense_lookup else {}
    join = _problem(p, **kw)
    cached = _problem(p, cache=True, **kw)
    rng = np.random.default_rng(3)
    v = rng.normal(size=p.keys.size) + 1j * rng.normal(size=p.keys.size)
    dt = torch.complex128 if prec == "c128" else torch.complex64
    vt = torch.from_numpy(v).to("cuda", dt)
    a = join.matvec(vt).cpu().numpy().astype(np.complex128)
    b = cached.matvec(vt).cpu().numpy().astype(np.complex128)
    ref = p.matvec(vt.cpu().numpy().astype(np.complex128))
    assert _rel(b, ref) < tol
    assert _rel(b, a) < 2 * tol


def test_cache_diagonal_only_and_offdiagonal_only(gpu):
    rng = np.random.default_rng(5)
    keys = np.sort(rng.choice(2 ** 10, size=600, replace=False)).astype(np.uint64)
    v = rng.normal(size=600) + 1j * rng.normal(size=600)
    for labels in (["ZZIIIIIIII", "IZIIZIIIII", "IIIIIIIIIZ"], ["XXIIIIIIII", "IYYIIIIIII"]):
        p = OracleProblem(labels, rng.normal(size=len(labels)), keys, 10)
        prob = _problem(p, cache=True)
        got = prob.matvec(torch.from_numpy(v).cuda()).cpu().numpy()
        assert _rel(got, p.matvec(v)) < 1e-13


def test_cached_solve_same_energy(gpu):
    from rqutils_amd import workloads as W
    w = W.heisenberg28(seed=2, n_states=1 << 14, n_random=400, num_qubits=20)
    from rqutils_amd.sqd import SQDProblem
    g = torch.Generator(device="cpu").manual_seed(1)
    x0 = torch.complex(torch.randn(w.n, generator=g, dtype=torch.float64),
                       torch.randn(w.n, generator=g, dtype=torch.float64)).cuda()
    res = {}
    for cache in (False, True):
        prob = SQDProblem(w.hamiltonian, w.keys, w.num_qubits, cache=cache)
        for mode, prec, tol in (("fused", "c64", 1e-6), ("reference", "c128", 1e-10)):
            _, info = prob.solve(x0, mode=mode, precision=prec, criterion="relative", tol=tol,
                                 maxiter=3000)
            assert info.converged
            res[(cache, mode)] = info
    for mode in ("fused", "reference"):
        a, b = res[(False, mode)], res[(True, mode)]
        assert abs(a.eigval - b.eigval) / abs(a.eigval) < (1e-7 if mode == "fused" else 1e-12)
        assert abs(a.niter - b.niter) <= 3


def test_sqd_cache_levels_agree(gpu):
    from rqutils_amd.sqd import sqd
    rng = np.random.default_rng(8)
    n = 10
    labels = ["".join("IXYZ"[i] for i in row) for row in rng.integers(0, 4, (60, n))]
    coeffs = rng.uniform(-1, 1, 60)
    states = rng.integers(0, 2, (900, n)).astype(np.uint8)
    e0 = sqd((labels, coeffs), states, cache_level=(0, 0), return_eigvec=False)
    e1 = sqd((labels, coeffs), states, cache_level=(1, 0), return_eigvec=False)
    assert abs(e0 - e1) < 1e-10 * abs(e0)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_slices_per_wave_same_product(gpu, prec):
    """Both launch shapes of the product give the same bits, in both
    precisions, on rows with enough hits for the unrolled steps (4 entries
    per step for complex128, 8 for complex64) and for the remainder loop."""
    from rqutils_amd import _lib as L
    import ctypes as C
    p = small_problem(9, n=12, terms=400, nstates=3000)
    prob = _problem(p, cache=True)
    pc = L.C128 if prec == "c128" else L.C64
    assert prob.op.build_cache(pc)
    cs = prob.op._caches[pc][0]
    rlen = prob.op._caches[pc][6].cpu().numpy()
    assert rlen.max() >= 17 and np.any(rlen % 8 != 0)
    dt = torch.complex128 if prec == "c128" else torch.complex64
    v = torch.randn(p.keys.size, dtype=dt, device="cuda")
    outs = []
    for spw in (1, 4):
        cs.slices_per_wave = spw
        o = torch.empty_like(v)
        L.check(L.lib().sqdk_cache_matvec(C.byref(cs), v.data_ptr(), o.data_ptr(),
                                          L.stream_ptr(v.device)), "cache matvec")
        outs.append(o.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    ref = p.matvec(v.cpu().numpy().astype(np.complex128))
    assert _rel(outs[0].astype(np.complex128), ref) < (1e-13 if prec == "c128" else 3e-7)


def test_cache_budget_falls_back(gpu):
    p = small_problem(6, n=11, terms=80, nstates=900)
    prob = _problem(p, cache=True)
    prob.op.cache_budget = 16         # far too small: products keep the bucket join
    v = torch.randn(p.keys.size, dtype=torch.complex128, device="cuda")
    got = prob.matvec(v).cpu().numpy()
    assert not prob.op._caches
    assert _rel(got, p.matvec(v.cpu().numpy())) < 1e-12


def _batch_instances(k=6, seed=0):
    probs = []
    for i in range(k):
        n = 10 + (i % 4)
        probs.append(small_problem(seed + i, n=n, terms=40 + 15 * i,
                                   nstates=min(2 ** n - 1, 300 + 200 * i)))
    return probs


@pytest.mark.parametrize("dense,rows", [(False, 4), (None, 4), (False, 8)])
def test_batch_cache_matches_join(gpu, dense, rows):
    """One table over a batch (search and, with dense=None, dense-lookup
    instances mixed): every instance's product equals its join product."""
    from rqutils_amd.batch import SQDBatch
    probs = _batch_instances()
    inst = [((p.labels, p.coeffs), p.keys, p.n) for p in probs]
    b = SQDBatch(inst, bucket_cap=512, rows_per_thread=rows, dense=dense)
    rng = np.random.default_rng(4)
    v = torch.from_numpy(rng.normal(size=b.n) + 1j * rng.normal(size=b.n)).cuda()
    join = b.matvec(v).cpu().numpy()
    b.use_cache = True
    info = b.build_cache(_lib.C128)
    assert info["entries"] > 0
    cached = b.matvec(v).cpu().numpy()
    for i, p in enumerate(probs):
        lo, hi = b.row_base_host[i], b.row_base_host[i + 1]
        want = p.matvec(v.cpu().numpy()[lo:hi])
        assert _rel(cached[lo:hi], want) < 1e-13, i
        assert _rel(cached[lo:hi], join[lo:hi]) < 1e-13, i
    # the build workspace changes no entry: a fill without it writes the same bits
    import ctypes as C
    cs, slot_row, sbase, col, coef, diag, rlen, _ = b._caches[_lib.C128]
    ref = (torch.empty_like(col), torch.empty_like(coef), torch.empty_like(diag),
           torch.empty_like(rlen))
    _lib.check(_lib.lib().sqdk_batch_cache_fill(
        C.byref(b.cstruct), _lib.C128, sbase.data_ptr(), slot_row.data_ptr(), ref[0].data_ptr(),
        ref[1].data_ptr(), ref[2].data_ptr(), ref[3].data_ptr(), 0, None, 0,
        _lib.stream_ptr(col.device)), "sqdk_batch_cache_fill")
    for got, want in zip((col, coef, diag, rlen), ref):
        assert torch.equal(got.view(torch.uint8), want.view(torch.uint8))


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_cache_solve(gpu, use_graph):
    from rqutils_amd.batch import SQDBatch
    probs = _batch_instances(k=5, seed=70)
    b = SQDBatch([((p.labels, p.coeffs), p.keys, p.n) for p in probs], bucket_cap=512)
    rng = np.random.default_rng(5)
    x0 = torch.from_numpy(rng.normal(size=b.n) + 1j * rng.normal(size=b.n)).cuda()
    plain = b.solve(x0, precision="c128", tol=1e-9, maxiter=2000, use_graph=use_graph)
    b.use_cache = True
    cached = b.solve(x0, precision="c128", tol=1e-9, maxiter=2000, use_graph=use_graph)
    for i, p in enumerate(probs):
        a, c = plain.infos[i], cached.infos[i]
        assert a.converged and c.converged
        ev = p.ground()
        assert abs(c.eigval - ev) / abs(ev) < 1e-10
        assert abs(a.eigval - c.eigval) / abs(ev) < 1e-11


def _fill_without_work(op, prec):
    """The fill pass of op's table again, into fresh buffers, without the build
    workspace (every class joined for every row); returns (col, coef, diag,
    row_len)."""
    import ctypes as C
    cs, slot_row, sbase, col, coef, diag, rlen, _ = op._caches[prec]
    h, s, _ = op.structs()
    out = (torch.empty_like(col), torch.empty_like(coef), torch.empty_like(diag),
           torch.empty_like(rlen))
    _lib.check(_lib.lib().sqdk_cache_fill(C.byref(h), C.byref(s), prec, sbase.data_ptr(),
                                          slot_row.data_ptr(), out[0].data_ptr(),
                                          out[1].data_ptr(), out[2].data_ptr(),
                                          out[3].data_ptr(), None, 0,
                                          _lib.stream_ptr(col.device)), "sqdk_cache_fill")
    return out


@pytest.mark.parametrize("prec", [_lib.C128, _lib.C64])
@pytest.mark.parametrize("kind", ["sparse32", "mixed28", "dense", "small"])
def test_build_workspace_gives_the_same_table(gpu, prec, kind):
    """The counting pass's record (class masks, wave flags, each row's first two
    source rows) lets the fill take short rows' entries from the record and
    skip (chunk, class) pairs and waves without hits: the table must be bit for
    bit the one a fill without the workspace writes. The record itself is
    checked against the table: a wave's class bit is set iff one of its rows
    has an entry whose source lies in the class's target bucket, its flag iff
    one of its rows has more than two entries, and a row's recorded sources are
    its first entries."""
    from rqutils_amd import workloads as W
    from rqutils_amd.sqd import SQDProblem
    if kind in ("sparse32", "mixed28"):
        # 2^15 states of the 16-ones (14-ones) sector: 0.01 (0.13) hits per row,
        # no row (a dozen rows) with more than two
        nq = 32 if kind == "sparse32" else 28
        w = W.random40(seed=3, n_states=1 << 15, n_terms=3000, pool=800, num_qubits=nq)
        prob = SQDProblem(w.hamiltonian, w.keys, w.num_qubits, cache=True, bucket_cap=512)
        n = w.n
    elif kind == "dense":
        p = small_problem(21, n=12, terms=120, nstates=2000)
        prob = _problem(p, cache=True, dense=True, bucket_bits=8)
        n = p.keys.size
    else:
        p = small_problem(22, n=12, terms=90, nstates=1500)
        prob = _problem(p, cache=True, rows_per_thread=8)
        n = p.keys.size
    op = prob.op
    assert op.build_cache(prec)
    cs, slot_row, sbase, col, coef, diag, rlen, work = op._caches[prec]
    ref = _fill_without_work(op, prec)
    for got, want in zip((col, coef, diag, rlen), ref):
        assert torch.equal(got.view(torch.uint8), want.view(torch.uint8))
    # the record against the table: wave k of chunk i covers the rows
    # items[i] + 64 R k .. +64 R (R rows per thread); a class's hits land in
    # its target bucket, bucket(row) ^ class_hi
    sp = op.space
    nc = op.n_classes
    stride = (nc + 7) // 8
    items = sp.items[: sp.n_items].cpu().numpy().astype(np.int64)
    keys = sp.keys.cpu().numpy().view(np.uint64)
    m = sp.bucket_bits
    chi = op.class_hi[:nc].cpu().numpy().astype(np.uint64)
    raw = work.cpu().numpy()
    nmask = sp.n_items * (stride + 1)
    mask_bytes = -(-4 * nmask // 256) * 256
    words = raw[: 4 * nmask].view(np.uint32).reshape(sp.n_items, stride + 1)
    rowhits = raw[mask_bytes: mask_bytes + 8 * n].view(np.uint32).reshape(n, 2)
    tab, _, _ = _table_to_csr(op, prec, n)
    tab = tab.tocsr()
    hcount = rlen.cpu().numpy().astype(np.int64)
    R = sp.chunk_rows // 256
    bucket_of = keys >> np.uint64(m)
    skipped = longwaves = shortwaves = 0
    for i, r0 in enumerate(items):
        r1 = int(items[i + 1]) if i + 1 < sp.n_items else n
        r1 = min(r1, r0 + sp.chunk_rows)
        b = bucket_of[r0]
        flags = 0
        for k in range(4):
            lo, hi = r0 + 64 * R * k, min(r0 + 64 * R * (k + 1), r1)
            if lo < hi and np.any(hcount[lo:hi] > 2):
                flags |= 1 << k
            longwaves += lo < hi and bool(flags >> k & 1)
            shortwaves += lo < hi and not flags >> k & 1
        assert int(words[i, stride]) == flags, (i, int(words[i, stride]), flags)
        for c in range(nc):
            tgt = b ^ chi[c]
            nib = (int(words[i, c // 8]) >> (4 * (c % 8))) & 0xF
            want = 0
            for k in range(4):
                lo, hi = r0 + 64 * R * k, min(r0 + 64 * R * (k + 1), r1)
                if lo >= hi:
                    continue
                srcs = tab[lo:hi].indices
                if srcs.size and np.any(bucket_of[srcs] == tgt):
                    want |= 1 << k
            assert nib == want, (i, c, nib, want)
            skipped += nib == 0
    # the recorded sources of every row with at most two hits, in table order
    cols = col.cpu().numpy().view(np.uint32)
    sr = slot_row.cpu().numpy().view(np.uint32)
    sb = sbase.cpu().numpy()
    for slot in np.nonzero(sr != 0xFFFFFFFF)[0]:
        row = int(sr[slot])
        e0 = int(sb[slot // 64]) + slot % 64
        for j in range(min(2, int(hcount[row]))):
            assert rowhits[row, j] == cols[e0 + 64 * j], (row, j)
    if kind == "sparse32":
        assert skipped > 0.2 * sp.n_items * nc and longwaves == 0
    if kind == "mixed28":
        assert longwaves > 0 and shortwaves > longwaves
