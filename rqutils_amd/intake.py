"""GPU intake of sampled bitstrings (rqutils/sqd.py:237, 246, 414-445).

The reference packs sampled state rows behind a pad bit, sorts them with a
multi-key ``lax.sort`` and removes duplicates (``uniquify_states``), filling
the tail with 0xFF rows; ``sqd`` unpacks the result into ``basis_states``.
Here rows become uint64 keys on the device (``pack_states``), an occupancy
bitmap (keys of at most 28 bits whose bitmap is small enough) or a stable
radix sort plus compaction gives the sorted distinct keys (``unique_keys``;
``sqd`` does not materialise the fill rows: ``states_size`` only enters the
stopping rule), and ``unpack_keys`` turns keys back into uint8 rows. The
reference-signature ``uniquify_states(states_p, states_size)`` on packed rows,
fill rows included, is in :mod:`rqutils_amd.sqd_arrays`. All three run in the HIP
library (``sqdk_intake_*``) and are bit-exact.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

__all__ = ["pack_states", "unique_keys", "unique_state_keys", "unpack_keys"]


def _dev(device):
    return torch.device(device if device is not None else "cuda")


def pack_states(states, device=None) -> torch.Tensor:
    """Keys (int64 tensor holding the uint64 bits) of binary state rows
    ``[N, n]`` (numpy or torch, bool or uint8; column 0 = qubit n-1; any
    nonzero byte is a 1)."""
    dev = _dev(device)
    if isinstance(states, torch.Tensor):
        st = states.to(dev)
        if st.dtype == torch.bool:
            st = st.to(torch.uint8)
    else:
        arr = np.asarray(states)
        if arr.dtype == np.bool_:
            arr = arr.view(np.uint8)
        st = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.uint8)).to(dev)
    if st.dtype != torch.uint8 or st.dim() != 2:
        raise ValueError("states must be a 2-d binary uint8 or bool array")
    st = st.contiguous()
    n, nq = st.shape
    if not 1 <= nq <= 64:
        raise ValueError(f"{nq} columns: keys hold 1 to 64 qubits")
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    if n:
        _lib.check(_lib.lib().sqdk_intake_pack(st.data_ptr(), n, nq, nq, keys.data_ptr(),
                                               _lib.stream_ptr(dev)), "sqdk_intake_pack")
    return keys


def unique_keys(keys: torch.Tensor, key_bits: int = 64) -> torch.Tensor:
    """Sorted distinct keys (every key below ``2**key_bits``, else
    ``ValueError``). Narrow key spaces go through an occupancy bitmap, wide
    ones through a radix sort; both give the same keys."""
    dev = keys.device
    n = keys.numel()
    if n == 0:
        return keys.clone()
    keys = keys.contiguous()
    L = _lib.lib()
    nb = C.c_size_t(0)
    _lib.check(L.sqdk_intake_unique_workspace(n, int(key_bits), C.byref(nb)),
               "sqdk_intake_unique_workspace")
    work = torch.empty(nb.value, dtype=torch.uint8, device=dev)
    out = torch.empty(n, dtype=torch.int64, device=dev)
    count = torch.empty(1, dtype=torch.int64, device=dev)
    _lib.check(L.sqdk_intake_unique(keys.data_ptr(), n, int(key_bits), out.data_ptr(),
                                    count.data_ptr(), work.data_ptr(), nb.value,
                                    _lib.stream_ptr(dev)), "sqdk_intake_unique")
    m = int(count.item())
    if m < 0:
        raise ValueError(f"a key is not below 2**{key_bits}")
    return out[:m]


def unique_state_keys(states, device=None) -> torch.Tensor:
    """Sorted distinct keys of sampled binary state rows ``[N, n]`` (the
    distinct rows of ``uniquify_states``, rqutils/sqd.py:414-445)."""
    nq = states.shape[1]
    return unique_keys(pack_states(states, device), nq)


def unpack_keys(keys: torch.Tensor, num_qubits: int) -> torch.Tensor:
    """uint8 rows ``[n, num_qubits]`` of keys (column 0 = qubit n-1)."""
    n = keys.numel()
    out = torch.empty((n, num_qubits), dtype=torch.uint8, device=keys.device)
    if n:
        _lib.check(_lib.lib().sqdk_intake_unpack(keys.contiguous().data_ptr(), n, num_qubits,
                                                 out.data_ptr(), _lib.stream_ptr(keys.device)),
                   "sqdk_intake_unpack")
    return out
