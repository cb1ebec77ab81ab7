"""Morton encoding (src/kernels/morton.py).

``compute_codes[blocks, threads](pos_x, pos_y, pos_z, codes, indices,
n_bodies)`` as in the reference (morton.py:65-91), running the gfx950 kernel
``cosmo_compute_codes``.  codes: int64/uint64 tensor (bit pattern of the
uint64 code); indices: int32/uint32 tensor or None.  ``mask_set`` selects
MASK_FULL21 (21 bits per axis, the 63-bit code the reference documents) or
MASK_REF10 (the reference's masks as written, morton.py:33-36, which keep
each axis' low 10 bits).
"""
from __future__ import annotations

from ._hip import (MASK_FULL21, MASK_REF10, U32, U64, check, kernel, launch_extent, load, ptr,
                   stream_handle)

__all__ = ["compute_codes", "MASK_FULL21", "MASK_REF10", "key_bits"]


def key_bits(mask_set: int) -> int:
    """Number of low code bits that can vary: 63 (FULL21) or 30 (REF10)."""
    return 30 if mask_set == MASK_REF10 else 63


@kernel
def compute_codes(pos_x, pos_y, pos_z, codes, indices, n_bodies, mask_set=MASK_FULL21,
                  _launch=None):
    n = launch_extent(_launch, int(n_bodies))
    if n <= 0:
        return
    import torch
    args = [ptr(t, torch.float32, nm, numel=n) for t, nm in
            ((pos_x, "pos_x"), (pos_y, "pos_y"), (pos_z, "pos_z"))]
    c = ptr(codes, U64, "codes", numel=n)
    ix = None if indices is None else ptr(indices, U32, "indices", numel=n)
    check(load().cosmo_compute_codes(*args, c, ix, n, int(mask_set), stream_handle()),
          "compute_codes")
