"""The reference's inference-time hash-grid encoder on the jax-tcnn layout.

`hashgrid_encode` mirrors jaxtcnn.hashgrid_encode (deps/jax-tcnn/src/jaxtcnn/
hashgrid_tcnn/__init__.py): coordinates [3, N] in the unit cube, a table with
unpadded level sizes, encodings [L*F, N]; it is differentiable (table and
coordinates) through the `tcnn_hashgrid_encode_backward` target.
`TCNNHashGridEncoder` mirrors models/encoders.py:259-305: positions in
[-bound, bound]^3 are mapped to the unit cube, the level offsets are the unpadded
prefix sums, and it returns (encodings [N, L*F], 0).

Its table layout differs from HashGridEncoder's (encoders.py:76-103 rounds every dense
level up to a multiple of 8 rows; this one does not), so a table trained with one is
not laid out for the other: from level 2 on (res 23^3 = 12167 rows vs 12168) the
level offsets disagree.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib


def unpadded_offsets(L: int, T: int, N_min: int, b: float, dim: int = 3) -> np.ndarray:
    """Level offsets of encoders.py:268-283: res^dim rows for dense levels, T otherwise."""
    offsets = [0]
    for i in range(L):
        scale = N_min * (b ** i) - 1
        res = math.ceil(scale) + 1
        n = res ** dim
        offsets.append(offsets[-1] + (n if n <= T else T))
    return np.asarray(offsets, np.uint32)


class _TcnnEncode(torch.autograd.Function):
    @staticmethod
    def forward(ctx, offsets, coords, params, L, F, N_min, pls):
        n = coords.shape[1]
        enc = torch.empty(L * F, n, dtype=torch.float32, device=coords.device)
        dy = torch.empty(L * F, n, 3, dtype=torch.float32, device=coords.device)
        desc = _lib.tcnn_hashgrid_descriptor(n, L, F, N_min, pls)
        if n:
            _lib.call("tcnn_hashgrid_encode", [offsets, coords, params, enc, dy], desc)
        ctx.save_for_backward(offsets, coords, dy)
        ctx.shape = (params.shape, desc)
        return enc

    @staticmethod
    def backward(ctx, g):
        offsets, coords, dy = ctx.saved_tensors
        pshape, desc = ctx.shape
        dparams = torch.empty(pshape, dtype=torch.float32, device=coords.device)
        dcoords = torch.empty_like(coords)
        _lib.call("tcnn_hashgrid_encode_backward", [offsets, coords, g.contiguous(), dy, dparams, dcoords], desc)
        return None, dcoords, dparams, None, None, None, None


def hashgrid_encode(L: int, F: int, N_min: int, per_level_scale: float, offset_table_data: torch.Tensor,
                    coords_rm: torch.Tensor, params: torch.Tensor) -> torch.Tensor:
    """jaxtcnn.hashgrid_encode: offsets u32[L+1], coords f32[3, N], params f32[offsets[L], F]
    -> encodings f32[L*F, N] (the shape and dtype rules of hashgrid_tcnn/abstract.py)."""
    if coords_rm.dim() != 2 or coords_rm.shape[0] != 3:
        raise NotImplementedError(f"expected coordinates of shape (3, n_coords), got {tuple(coords_rm.shape)}")
    if tuple(offset_table_data.shape) != (L + 1,) or offset_table_data.dtype != torch.uint32:
        raise RuntimeError("offset_table_data must be uint32 of shape (L + 1,)")
    if coords_rm.dtype != torch.float32 or params.dtype != torch.float32:
        raise NotImplementedError("coordinates and parameters must be float32")
    if params.dim() != 2 or params.shape[1] != F:
        raise ValueError(f"params must have shape (n_params, {F}), got {tuple(params.shape)}")
    return _TcnnEncode.apply(offset_table_data.contiguous(), coords_rm.contiguous(), params, L, F, N_min,
                             float(per_level_scale))


class TCNNHashGridEncoder(torch.nn.Module):
    """encoders.py:259-305 on the jax-tcnn kernels."""

    def __init__(self, L: int = 16, T: int = 2 ** 19, F: int = 2, N_min: int = 16, N_max: int = 2048,
                 device=None, generator: torch.Generator | None = None):
        super().__init__()
        self.L, self.T, self.F, self.N_min, self.N_max = L, T, F, N_min, N_max
        self.b = math.exp((math.log(N_max) - math.log(N_min)) / (L - 1))
        self.offsets_host = unpadded_offsets(L, T, N_min, self.b)
        self.register_buffer("offsets", torch.from_numpy(self.offsets_host.astype(np.int32)).view(torch.uint32))
        n = int(self.offsets_host[-1])
        self.table = torch.nn.Parameter(torch.empty(n, F).uniform_(-1e-4, 1e-4, generator=generator))
        if device is not None:
            self.to(device)

    def forward(self, pos: torch.Tensor, bound: float):
        p = (pos + bound) / (2 * bound)
        enc = hashgrid_encode(self.L, self.F, self.N_min, self.b, self.offsets, p.t().contiguous(), self.table)
        return enc.t(), 0
