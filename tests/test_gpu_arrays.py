"""The reference-signature operator functions on the reference's own packed
arrays (rqutils/sqd.py:414-684), through the C ABI, against the oracle's
restatement on the same arrays: integer results bit for bit, complex128
products within 1e-12, over packed rows of 1 to 8 bytes with fill rows."""
import numpy as np
import pytest
import torch

from oracle import sqd_oracle as so
from tests.helpers import OracleProblem, random_labels

pytestmark = pytest.mark.gpu


def _case(n, seed, terms=80, nstates=500, full=False):
    rng = np.random.default_rng(seed)
    labels = random_labels(rng, n, terms, 1, min(4, n), full=full)
    coeffs = rng.normal(size=terms)
    space = 2 ** n
    if n <= 20:
        keys = np.sort(rng.choice(space, size=min(nstates, space - 1), replace=False))
    else:
        base = rng.integers(0, 2 ** min(n, 62), size=nstates, dtype=np.uint64)
        keys = np.unique(np.concatenate([base, base ^ np.uint64(0b11), base ^ np.uint64(0b101)]))
        keys = keys[keys < np.uint64(2) ** np.uint64(n)]
    p = OracleProblem(labels, coeffs, keys.astype(np.uint64), n)
    rows = p.packed[rng.permutation(np.concatenate([np.arange(p.packed.shape[0])] * 2))]
    size = p.packed.shape[0] + 13
    su, nu = so.uniquify_states(rows, size)
    return p, rows, su, nu, rng


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n", [7, 10, 15, 23, 40, 63])
def test_uniquify_states_bit_exact(gpu, n):
    from rqutils_amd.sqd import uniquify_states
    p, rows, su, nu, _ = _case(n, n)
    got = uniquify_states(rows, su.shape[0]).cpu().numpy()
    assert np.array_equal(got, su), n
    # states_size equal to the distinct count: no fill rows; numpy input works too
    got = uniquify_states(rows, nu).cpu().numpy()
    assert np.array_equal(got, su[:nu])
    # a one-row input with fill
    got = uniquify_states(rows[:1], 3).cpu().numpy()
    assert np.array_equal(got[0], rows[0]) and np.all(got[1:] == 255)


@pytest.mark.parametrize("n,full", [(10, False), (12, True), (40, False)])
def test_products_on_reference_arrays(gpu, n, full):
    from rqutils_amd import sqd as S
    p, rows, su, nu, rng = _case(n, 100 + n, full=full)
    nrow = su.shape[0]
    v = rng.normal(size=nrow) + 1j * rng.normal(size=nrow)
    v[nu:] = 0.0
    xs = np.stack([so.get_xsource(p.x[g], su) for g in range(p.x.shape[0])])
    diags = np.stack([so.get_diagonal(p.z[g], p.c[g], su) for g in range(p.x.shape[0])])
    signs = np.stack([so.get_diag_signs(p.z[g], su) for g in range(p.x.shape[0])])
    want = so.apply_h_x_cached(v, xs, p.z, p.c, su)[:nu]
    vt = torch.from_numpy(v).cuda()
    outs = {
        "apply_h": S.apply_h(vt, p.x, p.z, p.c, su),
        "apply_h_x_cached": S.apply_h_x_cached(vt, xs, p.z, p.c, su),
        "apply_h_z_cached": S.apply_h_z_cached(vt, p.x, su, diags),
        "apply_h_s_cached": S.apply_h_s_cached(vt, p.x, su, signs, p.c),
        "apply_h_xs_cached": S.apply_h_xs_cached(vt, xs, signs, p.c),
        "apply_h_xz_cached": S.apply_h_xz_cached(vt, xs, diags),
    }
    for name, out in outs.items():
        got = out.cpu().numpy()
        assert _rel(got[:nu], want) < 1e-12, name
        if name not in ("apply_h_xs_cached", "apply_h_xz_cached"):   # these get no states
            assert np.all(got[nu:] == 0), name          # fill rows are outside the subspace
    # the source indices and sign bits of the device match the oracle's
    for g in range(p.x.shape[0]):
        got = S.get_xsource(p.x[g], su).cpu().numpy()
        assert np.array_equal(got[:nu], xs[g][:nu]), g
        assert np.array_equal(S.get_diag_signs(p.z[g], su).cpu().numpy(), signs[g]), g
        d = S.compute_diagonal(signs[g], p.c[g]).cpu().numpy()
        assert _rel(d, diags[g]) < 1e-13, g
    # apply_xgrp of one group
    g = p.x.shape[0] // 2
    got = S.apply_xgrp(xs[g], diags[g], vt).cpu().numpy()
    assert _rel(got, so.apply_xgrp(xs[g], diags[g], v)) < 1e-15
    # complex64 vectors: FP64 sums, rounded once
    o64 = S.apply_h(vt.to(torch.complex64), p.x, p.z, p.c, su).cpu().numpy()
    assert _rel(o64[:nu], want) < 1e-6


def test_real_vectors_stay_real(gpu):
    """Real coefficients (no Y) and a real vector give a real product, as the
    reference's do."""
    from rqutils_amd import sqd as S
    rng = np.random.default_rng(5)
    n = 9
    labels = ["".join("IXZ"[i] for i in rng.integers(0, 3, n)) for _ in range(40)]
    coeffs = rng.normal(size=40)
    keys = np.sort(rng.choice(2 ** n, size=300, replace=False)).astype(np.uint64)
    p = OracleProblem(labels, coeffs, keys, n)
    assert not np.iscomplexobj(p.c)
    su, nu = so.uniquify_states(p.packed, p.packed.shape[0] + 5)
    v = rng.normal(size=su.shape[0])
    v[nu:] = 0
    out = S.apply_h(torch.from_numpy(v).cuda(), p.x, p.z, p.c, su)
    assert not out.is_complex()
    want = so.apply_h(v, p.x, p.z, p.c, su)[:nu]
    assert _rel(out.cpu().numpy()[:nu], want) < 1e-12


def test_shape_errors(gpu):
    from rqutils_amd import sqd as S
    p, rows, su, nu, rng = _case(8, 3, terms=30)
    v = torch.zeros(su.shape[0] + 1, dtype=torch.complex128).cuda()
    with pytest.raises(ValueError):
        S.apply_h(v, p.x, p.z, p.c, su)
    with pytest.raises(ValueError):
        S.uniquify_states(np.zeros((4, 9), np.uint8), 4)     # 9-byte rows
