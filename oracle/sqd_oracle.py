"""NumPy FP64 restatement of the reference's SQD operator path (test oracle).

Test infrastructure only: the product never imports this module. Every
function restates one reference function in the reference's own data layout
(bit-packed ``uint8`` rows with a leading pad bit) and cites the lines it
follows. Where the reference has a known defect, the default here is the
mathematically intended behaviour and the reference's behaviour is available
behind a flag (documented per function).
"""
from __future__ import annotations

import warnings

import numpy as np

__all__ = [
    "from_paulisum", "pack_states", "uniquify_states", "get_xsource",
    "xsource_by_lookup", "get_diagonal", "apply_xgrp", "apply_h_x_cached",
    "apply_h", "apply_h_rows", "vinit_from_min_diag", "hproj_coo", "packed_to_keys",
    "get_diag_signs", "compute_diagonal",
    "PHASES",
]

# (-i)^k for k = popcount(x & z) mod 4 -- rqutils/paulis/symplectic.py:110-111
PHASES = np.array([1.0, -1.0j, -1.0, 1.0j])


def from_paulisum(paulis, coeffs, force_real=False, add_padding=False):
    """Group Pauli strings by X signature; tuple-input path.

    Follows rqutils/paulis/symplectic.py:58-76 (parse, drop zero inputs, merge
    duplicates), :93-96 (force_real), :98-115 (group by unique X, phase-fold
    c * (-i)^{popcount(x&z)}, downcast to real when every phased coefficient
    is real), :117-124 (pad bit, packbits).

    Returns ``(x u8[J, Bp], z u8[J, Kmax, Bp], c [J, Kmax], num_qubits)``.
    """
    if len(paulis) != len(coeffs):
        raise ValueError("Lengths of Pauli and coeff lists do not match")
    if len(set(len(p) for p in paulis)) != 1:
        raise ValueError("Pauli strings have non-uniform lengths")
    coeffs = np.array(coeffs)
    chars = np.array([list(p.upper()) for p in paulis])
    nonzero = np.nonzero(coeffs)
    coeffs = coeffs[nonzero]
    chars = chars[nonzero]
    chars, inverse = np.unique(chars, axis=0, return_inverse=True)
    inverse = inverse.reshape(-1)
    merged = np.zeros(chars.shape[0], dtype=np.result_type(coeffs.dtype, np.float64))
    np.add.at(merged, inverse, coeffs)
    coeffs = merged
    xbits = np.logical_or(chars == "X", chars == "Y")
    zbits = np.logical_or(chars == "Y", chars == "Z")
    num_qubits = chars.shape[1]

    if force_real:
        if np.any(np.imag(coeffs) != 0.0):
            warnings.warn("Found nonzero imaginary part when force_real=True")
        coeffs = np.real(coeffs)

    xuniq, indices, counts = np.unique(xbits, axis=0, return_inverse=True,
                                       return_counts=True)
    indices = indices.reshape(-1)
    xsig = xuniq.astype(np.uint8)
    shape = (xsig.shape[0], int(np.max(counts)))
    zsig = np.zeros(shape + (num_qubits,), dtype=np.uint8)
    phc = np.zeros(shape, dtype=np.complex128)
    for g in range(xsig.shape[0]):
        members = np.nonzero(indices == g)[0]
        zs = zbits[members].astype(np.uint8)
        zsig[g, : counts[g]] = zs
        ph = np.sum(xsig[g] & zs, axis=1) & 3
        phc[g, : counts[g]] = coeffs[members] * PHASES[ph]
    if np.all(phc.imag == 0.0):
        phc = phc.real
    if add_padding:
        xsig = np.pad(xsig, ((0, 0), (1, 0)))
        zsig = np.pad(zsig, ((0, 0), (0, 0), (1, 0)))
    return np.packbits(xsig, axis=-1), np.packbits(zsig, axis=-1), phc, num_qubits


def pack_states(states):
    """Bit-pack binary rows behind a zero pad bit -- rqutils/sqd.py:237."""
    states = np.asarray(states).astype(np.uint8)
    return np.packbits(np.pad(states, ((0, 0), (1, 0))), axis=1)


def uniquify_states(states_p, states_size):
    """Sorted unique rows, filled with 0xFF to ``states_size`` rows.

    Restates rqutils/sqd.py:414-445 (multi-key sort, adjacent-unique,
    compaction, fill value 255). Returns ``(states_u, n_unique)``.
    """
    uniq = np.unique(states_p, axis=0)
    out = np.full((states_size, states_p.shape[1]), 255, dtype=np.uint8)
    out[: uniq.shape[0]] = uniq
    return out, uniq.shape[0]


def get_xsource(xsig, states):
    """Index of ``S[i] ^ x`` in ``S`` or -1, by the reference's sort-join.

    Restates rqutils/sqd.py:476-500 step by step: stack ``[S; S^x]``, stable
    multi-key sort carrying the stack index, equal-neighbour test, then the
    ``cumsum(bincount(cumsum))`` compaction back to row order. This keeps the
    reference's behaviour on repeated fill rows as well.
    """
    size = states.shape[0]
    mapped = np.bitwise_xor(states, xsig[None, :])
    joined = np.concatenate([states, mapped], axis=0)
    # np.lexsort is stable; the last key is primary, so reverse the columns.
    order = np.lexsort(joined.T[::-1]).astype(np.int64)
    jsorted = joined[order]
    equal = np.all(jsorted[:-1] == jsorted[1:], axis=1)
    source_idx = np.where(equal, order[1:] - size, -1)
    is_state = (order < size).astype(np.int64)
    tposition = np.cumsum(np.bincount(np.cumsum(is_state), minlength=size)[:size])
    xsource = np.full(size, -1, dtype=np.int64)
    ok = tposition < source_idx.shape[0]
    xsource[ok] = source_idx[tposition[ok]]
    # Negative values other than -1 are out-of-range gathers in the reference
    # (mode='fill', wrap_negative_indices=False): they read as "no source".
    xsource[xsource < 0] = -1
    return xsource.astype(np.int32)


def xsource_by_lookup(xsig, states):
    """Brute-force dictionary lookup of ``S[i] ^ x``; pins ``get_xsource``."""
    table = {row.tobytes(): i for i, row in enumerate(states)}
    mapped = np.bitwise_xor(states, xsig[None, :])
    return np.array([table.get(row.tobytes(), -1) for row in mapped], dtype=np.int32)


_POPCOUNT8 = np.array([bin(i).count("1") for i in range(256)], dtype=np.uint8)


def get_diag_signs(zsigs, states):
    """Packed sign bits of one group's terms -- rqutils/sqd.py:506-528.

    Term k's parity of ``states & z_k`` goes to bit ``7 - k % 8`` of byte
    ``k // 8`` (bits and bytes counted from the left). Returns
    ``uint8[N, ceil(K/8)]``.
    """
    k = zsigs.shape[0]
    out = np.zeros((states.shape[0], -(-k // 8)), dtype=np.uint8)
    for t in range(k):
        par = (np.bitwise_count(states & zsigs[t]).sum(axis=1) & 1).astype(np.uint8)
        out[:, t // 8] += par << np.uint8(7 - t % 8)
    return out


def compute_diagonal(diag_signs, coeffs, reference_bit_index=False):
    """Diagonal from sign bits and coefficients -- rqutils/sqd.py:531-551.

    Stops at the first zero coefficient as the reference does. The reference
    reads term k's bit at ``7 - (k & 255)`` (sqd.py:544), which for k >= 8
    shifts by a negative amount; XLA's shift then yields 0, i.e. a +1 sign
    (quirk Q1). ``reference_bit_index=True`` reproduces that; the default
    reads ``7 - k % 8``, the bit ``get_diag_signs`` wrote.
    """
    out = np.zeros(diag_signs.shape[0], dtype=np.result_type(coeffs.dtype, np.float64))
    for t in range(coeffs.shape[0]):
        if coeffs[t] == 0:
            break
        if reference_bit_index and t >= 8:
            bit = np.zeros(diag_signs.shape[0], dtype=np.int64)
        else:
            bit = (diag_signs[:, t // 8] >> (7 - t % 8)) & 1
        out = out + coeffs[t] * (1.0 - 2.0 * bit)
    return out


def get_diagonal(zsigs, coeffs, states, stop_at_zero=False):
    """Composed diagonal of one X group -- rqutils/sqd.py:554-575.

    ``C[i] = sum_k c_k (-1)^{popcount(S[i] & z_k)}``. The reference stops at
    the first zero coefficient (sqd.py:562-564); with ``stop_at_zero=False``
    (default) zero terms are skipped instead, which is the intended result.
    """
    diag = np.zeros(states.shape[0], dtype=np.result_type(coeffs.dtype, np.float64))
    for k in range(coeffs.shape[0]):
        if coeffs[k] == 0:
            if stop_at_zero:
                break
            continue
        parity = np.sum(_POPCOUNT8[states & zsigs[k][None, :]], axis=1, dtype=np.int64) & 1
        diag = diag + coeffs[k] * (1.0 - 2.0 * parity)
    return diag


def apply_xgrp(xsource, diagonal, vec):
    """Gather ``v[xsource]`` with fill 0, times the diagonal -- sqd.py:578-587."""
    gathered = np.where(xsource >= 0, vec[np.maximum(xsource, 0)], 0)
    return gathered * diagonal


def apply_h_x_cached(vec, xsources, zsigs, coeffs, states):
    """``Hv`` from cached source indices -- rqutils/sqd.py:641-654.

    Accumulates the groups in their stored order (ascending X signature).
    """
    out = np.zeros_like(vec, dtype=np.result_type(vec.dtype, coeffs.dtype))
    for g in range(xsources.shape[0]):
        out = out + apply_xgrp(xsources[g], get_diagonal(zsigs[g], coeffs[g], states), vec)
    return out


def apply_h(vec, xsigs, zsigs, coeffs, states):
    """``Hv`` recomputing the source indices per call -- rqutils/sqd.py:590-605."""
    xs = np.stack([get_xsource(xsigs[g], states) for g in range(xsigs.shape[0])])
    return apply_h_x_cached(vec, xs, zsigs, coeffs, states)


def apply_h_rows(rows, vec, xsigs, zsigs, coeffs, states):
    """Rows ``rows`` of ``apply_h_x_cached`` (rqutils/sqd.py:641-654), restated
    row by row for subspaces too large for the J x N source cache.

    For each group in stored order, the source of row i is the position of
    ``S[i] ^ x`` among the sorted packed rows ``S`` -- the same index the
    sort-join of ``get_xsource`` (sqd.py:476-500) leaves, found here by a
    binary search over the rows compared as byte strings (lexicographic, as
    the reference's multi-key sort orders them) -- or -1; its coefficient is
    ``get_diagonal`` (sqd.py:562-575) on that row.
    """
    states = np.ascontiguousarray(states, dtype=np.uint8)
    nb = states.shape[1]
    key = states.view(f"V{nb}").ravel()
    rows = np.asarray(rows, dtype=np.int64)
    srows = states[rows]
    n = states.shape[0]
    out = np.zeros(rows.size, dtype=np.result_type(vec.dtype, coeffs.dtype))
    for g in range(xsigs.shape[0]):
        t = np.ascontiguousarray(srows ^ xsigs[g])
        tv = t.view(f"V{nb}").ravel()
        j = np.searchsorted(key, tv)
        jc = np.minimum(j, n - 1)
        src = np.where((j < n) & (key[jc] == tv), j, -1).astype(np.int32)
        out = out + apply_xgrp(src, get_diagonal(zsigs[g], coeffs[g], srows), vec)
    return out


def vinit_from_min_diag(xsigs, zsigs, coeffs, states_u, n_unique):
    """Default start vector of ``run_sqd`` -- rqutils/sqd.py:373-398.

    One-hot at the argmin of the real group-0 diagonal over non-fill rows when
    group 0 is the identity X signature; else one-hot at index 0.
    """
    size = states_u.shape[0]
    dtype = np.complex128 if np.iscomplexobj(coeffs) else np.float64
    v = np.zeros(size, dtype=dtype)
    if np.all(xsigs[0] == 0):
        diag = np.real(get_diagonal(zsigs[0], coeffs[0], states_u))
        diag = np.where(states_u[:, 0] == 255, np.max(diag), diag)
        v[int(np.argmin(diag))] = 1
    else:
        v[0] = 1
    return v


def hproj_coo(xsigs, zsigs, coeffs, states):
    """Projected matrix as COO triplets, intended semantics of sqd.py:251-296.

    The reference packs the states without the pad bit while using a padded
    Hamiltonian (sqd.py:277 vs :280); here both use the pad bit, so the result
    is the projector-sandwiched matrix ``P H P``.
    """
    rows, cols, vals = [], [], []
    for g in range(xsigs.shape[0]):
        xs = get_xsource(xsigs[g], states)
        d = get_diagonal(zsigs[g], coeffs[g], states)
        ok = xs >= 0
        rows.append(np.nonzero(ok)[0])
        cols.append(xs[ok])
        vals.append(d[ok])
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def fullspace_matmul(xsigs, zsigs, coeffs, num_qubits, rhs, groups=None):
    """``H @ rhs`` over the whole 2^n space (rqutils/paulis/symplectic.py:131-150):
    for every group, out[i] += rhs[i ^ x] * sum_k c_k (-1)^{popcount(i & z_k)}
    (the sign from the output index i, zero-padded terms contributing 0).
    ``xsigs`` / ``zsigs`` are the padded packed signatures of ``from_paulisum``;
    ``groups`` restricts the sum to those group indices (for timing samples)."""
    n = num_qubits
    rhs = np.asarray(rhs)
    idx = np.arange(2 ** n, dtype=np.uint64)
    xm = packed_to_keys(np.asarray(xsigs), n)
    out = np.zeros(rhs.shape, dtype=np.result_type(rhs.dtype, np.asarray(coeffs).dtype))
    for g in (range(xm.size) if groups is None else groups):
        zm = packed_to_keys(np.asarray(zsigs[g]), n)
        diag = np.zeros(idx.size, dtype=out.dtype)
        for k in range(zm.size):
            if coeffs[g][k] == 0:
                continue
            par = _popcount64(idx & zm[k]) & np.uint64(1)
            diag += coeffs[g][k] * (1.0 - 2.0 * par.astype(np.float64))
        out += rhs[(idx ^ xm[g]).astype(np.int64)] * (diag if rhs.ndim == 1 else diag[:, None])
    return out


def _popcount64(a):
    return np.bitwise_count(np.asarray(a, dtype=np.uint64)).astype(np.uint64)


def packed_to_keys(states_p, num_qubits):
    """Integer key ``sum_q s_q 2^q`` of each padded packed row (pad bit must be 0).

    Used by tests to relate the reference layout to the uint64-key layout.
    """
    bits = np.unpackbits(states_p, axis=1)[:, 1 : 1 + num_qubits].astype(np.uint64)
    weights = np.uint64(1) << np.arange(num_qubits - 1, -1, -1, dtype=np.uint64)
    return (bits * weights[None, :]).sum(axis=1, dtype=np.uint64)
