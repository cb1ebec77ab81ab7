"""Particle state and the Morton sort (src/core/state.py).

``SimulationState(n_bodies, pos_init, vel_init)`` keeps the reference's
attributes (pos and vel as 3-tuples of float32 arrays, force (n, 3), mass
(n,) = 1/n, codes, indices) as torch tensors on the HIP device.
``sort_buffer_by_morton(config)`` encodes Morton codes, sorts them on the key
(code, index) and gathers positions, velocities and masses into the
permuted order (state.py:20-40).  Instead of allocating eight new arrays per
step it permutes into a second set of buffers and swaps; ``sort_idx`` keeps
the last permutation.
"""
from __future__ import annotations

import numpy as np
import torch

from src.kernels import _hip
from src.kernels.morton import MASK_FULL21, MASK_REF10, compute_codes, key_bits


def _to_device(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device)


def mask_set_of(config) -> int:
    name = str(getattr(config, "MORTON_MASKS", "full21")).lower()
    return MASK_REF10 if name == "ref10" else MASK_FULL21


class SimulationState:
    def __init__(self, n_bodies, pos_init, vel_init, device="cuda"):
        self.n = int(n_bodies)
        dev = torch.device(device)
        self.pos = tuple(_to_device(p, dev) for p in pos_init)
        self.vel = tuple(_to_device(v, dev) for v in vel_init)
        for t in self.pos + self.vel:
            if t.shape != (self.n,):
                raise ValueError(f"expected arrays of shape ({self.n},), got {tuple(t.shape)}")
        self.force = torch.zeros((self.n, 3), dtype=torch.float32, device=dev)
        self.mass = torch.full((self.n,), 1.0 / self.n, dtype=torch.float32, device=dev)
        self.codes = torch.zeros(self.n, dtype=torch.int64, device=dev)
        self.indices = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self.sort_idx = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self._alt = [torch.empty_like(t) for t in self.pos + self.vel + (self.mass,)]
        self._ctx = None

    @property
    def ctx(self):
        if self._ctx is None:
            self._ctx = _hip.Context(self.n)
        return self._ctx

    def sort_buffer_by_morton(self, config):
        ms = mask_set_of(config)
        compute_codes(*self.pos, self.codes, self.indices, self.n, mask_set=ms)
        self.sort_permutation(key_bits(ms))
        self.apply_permutation()

    def sort_permutation(self, bits=63):
        """Sorts self.codes in place; self.sort_idx receives the permutation."""
        L = _hip.load()
        _hip.check(L.cosmo_sort_by_code(self.ctx.handle, _hip.ptr(self.codes, _hip.U64, "codes"),
                                        _hip.ptr(self.codes, _hip.U64, "codes"),
                                        _hip.ptr(self.sort_idx, _hip.U32, "sort_idx"), self.n,
                                        int(bits), _hip.stream_handle()), "sort_by_code")

    def apply_permutation(self):
        """pos, vel, mass <- x[sort_idx] (state.py:37-39), ping-pong buffers."""
        import ctypes as C
        src = list(self.pos + self.vel + (self.mass,))
        dst = self._alt
        SrcArr = C.c_void_p * 7
        s = SrcArr(*[t.data_ptr() for t in src])
        d = SrcArr(*[t.data_ptr() for t in dst])
        _hip.check(_hip.load().cosmo_permute(_hip.ptr(self.sort_idx, _hip.U32, "sort_idx"), 7, s,
                                             d, self.n, _hip.stream_handle()), "permute")
        self.pos = tuple(dst[0:3])
        self.vel = tuple(dst[3:6])
        self.mass = dst[6]
        self._alt = src
