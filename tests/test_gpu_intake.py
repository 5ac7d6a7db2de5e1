"""GPU intake (pack, radix sort + unique, unpack) against numpy, bit for bit.

The oracle for packing is the host ``states_to_keys`` (itself pinned against
the reference's packed-row order in test_oracle.py); for sort + unique it is
``numpy.unique`` on the same uint64 keys.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from rqutils_amd import _lib
from rqutils_amd.intake import pack_states, unique_keys, uniquify_states, unpack_keys
from rqutils_amd.operator import keys_to_states, states_to_keys

pytestmark = pytest.mark.gpu


def _u64(t):
    return t.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("nq", [1, 5, 28, 40, 64])
def test_pack_matches_host(gpu, nq):
    rng = np.random.default_rng(nq)
    states = rng.integers(0, 2, size=(3001, nq)).astype(np.uint8)
    states[::7] *= rng.integers(1, 256, size=states[::7].shape).astype(np.uint8)  # nonzero = 1
    got = _u64(pack_states(states, gpu))
    assert np.array_equal(got, states_to_keys(states != 0))
    assert np.array_equal(_u64(pack_states(states != 0, gpu)), got)     # bool input


def test_pack_row_stride(gpu):
    rng = np.random.default_rng(3)
    wide = rng.integers(0, 2, size=(777, 37)).astype(np.uint8)
    t = torch.from_numpy(wide).to(gpu)
    keys = torch.empty(777, dtype=torch.int64, device=gpu)
    _lib.check(_lib.lib().sqdk_intake_pack(t.data_ptr(), 777, 30, 37, keys.data_ptr(),
                                           _lib.stream_ptr(gpu)), "pack")
    assert np.array_equal(_u64(keys), states_to_keys(wide[:, :30]))


@pytest.mark.parametrize("n,bits,pool", [(1, 1, 1), (2, 3, 2), (4095, 12, 3000), (4096, 28, 100),
                                         (4097, 28, 4097), (100003, 40, 20000),
                                         (65537, 64, 65537), (70000, 7, 128),
                                         ((1 << 20) + 3, 33, 1 << 19)])
def test_unique_matches_numpy(gpu, n, bits, pool):
    rng = np.random.default_rng(n + bits)
    hi = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    values = rng.integers(0, np.iinfo(np.uint64).max, size=pool, dtype=np.uint64,
                          endpoint=True) & hi
    keys = values[rng.integers(0, pool, size=n)]
    kt = torch.from_numpy(keys.view(np.int64)).to(gpu)
    got = _u64(unique_keys(kt, bits))
    assert np.array_equal(got, np.unique(keys))


@pytest.mark.parametrize("order", ["sorted", "reversed", "constant", "top_bit"])
def test_unique_special_orders(gpu, order):
    n = 50000
    if order == "sorted":
        keys = np.sort(np.random.default_rng(1).integers(0, 1 << 20, n).astype(np.uint64))
    elif order == "reversed":
        keys = np.arange(n, dtype=np.uint64)[::-1].copy()
    elif order == "constant":
        keys = np.full(n, 12345, dtype=np.uint64)
    else:
        keys = (np.arange(n, dtype=np.uint64) % np.uint64(97)) | np.uint64(1 << 63)
    kt = torch.from_numpy(keys.view(np.int64)).to(gpu)
    got = _u64(unique_keys(kt, 64))
    assert np.array_equal(got, np.unique(keys))


def test_unpack_and_uniquify_roundtrip(gpu):
    rng = np.random.default_rng(9)
    nq = 23
    base = rng.integers(0, 2, size=(5000, nq)).astype(np.uint8)
    states = base[rng.integers(0, 5000, size=20000)]
    keys = uniquify_states(states, gpu)
    ref = np.unique(states_to_keys(states))
    assert np.array_equal(_u64(keys), ref)
    rows = unpack_keys(keys, nq).cpu().numpy()
    assert np.array_equal(rows, keys_to_states(ref, nq))
    # the distinct rows in the reference's lexicographic order
    assert np.array_equal(rows, np.unique(states, axis=0))


def test_unique_rejects_small_workspace(gpu):
    t = torch.zeros(10, dtype=torch.int64, device=gpu)
    out = torch.empty(10, dtype=torch.int64, device=gpu)
    cnt = torch.empty(1, dtype=torch.int64, device=gpu)
    work = torch.empty(16, dtype=torch.uint8, device=gpu)
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().sqdk_intake_unique(t.data_ptr(), 10, 8, out.data_ptr(),
                                                 cnt.data_ptr(), work.data_ptr(), 16,
                                                 _lib.stream_ptr(gpu)), "unique")


def test_unique_full_size_property(gpu):
    """2^24 keys of 40 bits with repeats: sorted, distinct, and the same set."""
    rng = np.random.default_rng(24)
    pool = rng.integers(0, 1 << 40, size=1 << 23, dtype=np.uint64)
    keys = pool[rng.integers(0, pool.size, size=1 << 24)]
    got = _u64(unique_keys(torch.from_numpy(keys.view(np.int64)).to(gpu), 40))
    assert np.array_equal(got, np.unique(keys))
