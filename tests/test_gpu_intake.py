"""GPU intake (pack, bitmap or radix-sort unique, unpack) against numpy, bit
for bit.

The oracle for packing is the host ``states_to_keys`` (itself pinned against
the reference's packed-row order in test_oracle.py); for sort + unique it is
``numpy.unique`` on the same uint64 keys. Key widths up to 28 bits whose
bitmap is small next to the keys take the bitmap way (one slice of 2^20 keys
up to 20 bits, several above, each built by one or more workgroups), the rest
the radix sort; the cases below cover each of these, and the same keys
through both ways.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from rqutils_amd import _lib
from rqutils_amd.intake import pack_states, unique_keys, unique_state_keys, unpack_keys
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
                                         ((1 << 20) + 3, 33, 1 << 19), (5000, 22, 3000),
                                         ((1 << 19) + 5, 24, 1 << 18), (300000, 21, 1 << 20)])
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
    keys = unique_state_keys(states, gpu)
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


@pytest.mark.parametrize("order", ["sorted", "reversed", "constant", "dense"])
def test_unique_bitmap_and_radix_agree(gpu, order):
    """The same keys below 2^20 through the bitmap with one slice (key_bits
    20) and with 16 slices (key_bits 24), and the radix sort (key_bits 40,
    far too wide for a bitmap)."""
    n = 70001
    rng = np.random.default_rng(2)
    if order == "sorted":
        keys = np.sort(rng.integers(0, 1 << 20, n).astype(np.uint64))
    elif order == "reversed":
        keys = np.arange(n, dtype=np.uint64)[::-1].copy()
    elif order == "constant":
        keys = np.full(n, (1 << 20) - 1, dtype=np.uint64)
    else:                                   # every key of the space, shuffled, plus repeats
        keys = np.concatenate([np.arange(1 << 16, dtype=np.uint64),
                               rng.integers(0, 1 << 16, n - (1 << 16)).astype(np.uint64)])
        rng.shuffle(keys)
    kt = torch.from_numpy(keys.view(np.int64)).to(gpu)
    ref = np.unique(keys)
    assert np.array_equal(_u64(unique_keys(kt, 20)), ref)
    assert np.array_equal(_u64(unique_keys(kt, 24)), ref)
    assert np.array_equal(_u64(unique_keys(kt, 40)), ref)


@pytest.mark.parametrize("bits", [20, 40])
def test_unique_rejects_out_of_range_keys(gpu, bits):
    keys = np.arange(5000, dtype=np.uint64)
    keys[4321] = np.uint64(1 << bits)
    kt = torch.from_numpy(keys.view(np.int64)).to(gpu)
    with pytest.raises(ValueError, match="not below"):
        unique_keys(kt, bits)
    assert np.array_equal(_u64(unique_keys(kt, bits + 1)), np.unique(keys))


def test_workspace_depends_on_key_width():
    """The bitmap way needs 2^key_bits / 8 bytes, the radix way two key
    buffers; the query reflects the way the call will take."""
    L = _lib.lib()
    nb = C.c_size_t(0)
    _lib.check(L.sqdk_intake_unique_workspace(1000, 20, C.byref(nb)), "ws")
    assert nb.value >= (1 << 20) // 8
    _lib.check(L.sqdk_intake_unique_workspace(1 << 22, 28, C.byref(nb)), "ws")
    assert nb.value >= (1 << 28) // 8 + 8 * (1 << 22)       # bitmap + keys grouped by slice
    _lib.check(L.sqdk_intake_unique_workspace(1000, 40, C.byref(nb)), "ws")
    assert 16 * 1000 <= nb.value < 1 << 20


def test_pack_unpack_unaligned(gpu):
    """Rows that do not start on a 4-byte boundary take the direct-load path;
    an unaligned output takes byte stores."""
    rng = np.random.default_rng(4)
    n, nq = 1001, 13
    flat = torch.from_numpy(rng.integers(0, 2, size=n * nq + 1).astype(np.uint8)).to(gpu)
    states = flat[1:].view(n, nq)
    assert states.data_ptr() % 4 != 0
    keys = pack_states(states, gpu)
    ref = states_to_keys(states.cpu().numpy())
    assert np.array_equal(_u64(keys), ref)
    out = torch.zeros(n * nq + 3, dtype=torch.uint8, device=gpu)
    _lib.check(_lib.lib().sqdk_intake_unpack(keys.data_ptr(), n, nq, out.data_ptr() + 3,
                                             _lib.stream_ptr(gpu)), "unpack")
    got = out.cpu().numpy()
    assert got[:3].tolist() == [0, 0, 0]
    assert np.array_equal(got[3:].reshape(n, nq), keys_to_states(ref, nq))


def test_unique_bitmap_full_size(gpu):
    """The bench shape: 2^24 sampled keys of 28 bits (2^22 distinct), through
    the bitmap."""
    rng = np.random.default_rng(28)
    pool = np.unique(rng.integers(0, 1 << 28, size=1 << 22, dtype=np.uint64))
    keys = pool[rng.integers(0, pool.size, size=1 << 24)]
    got = _u64(unique_keys(torch.from_numpy(keys.view(np.int64)).to(gpu), 28))
    assert np.array_equal(got, np.unique(keys))
