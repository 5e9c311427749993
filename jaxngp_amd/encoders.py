"""HashGridEncoder: drop-in mirror of the reference's flax module
(models/encoders.py:57-256) on the gfx950 hash-grid kernels.

Same fields (L, T, F, N_min, N_max, tv_scale, param_dtype), same parameter name
(``"latent codes stored on grid vertices"``, shape ``[offsets[-1], F]``), same
call ``encoder(pos, bound) -> (encodings [N, L*F], tv)``.  The level table
(growth factor b, scales, resolutions, 8-aligned dense sizes, row offsets) is
computed on the host in float64 exactly as encoders.py:76-103 does and handed to
the kernels in the descriptor; scales are rounded to f32 there (encoders.py:216).

Differentiation goes through ``torch.autograd.Function`` with the hand-written
table-gradient kernel.  Positions get no gradient: in the reference's training
step they come from ``march_rays``, which is not differentiable.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import torch

from . import _lib

PARAM_NAME = "latent codes stored on grid vertices"


def next_multiple(value: int, step: int) -> int:
    """utils/common.py:336-337."""
    return (value + step - 1) // step * step


@dataclass(frozen=True)
class LevelTable:
    b: float
    scales: tuple  # float64 host values
    resolutions: tuple
    offsets: tuple  # L + 1 row offsets
    first_hash_level: int

    @property
    def n_rows(self) -> int:
        return self.offsets[-1]


def level_table(L: int, T: int, N_min: int, N_max: int, dim: int = 3) -> LevelTable:
    """encoders.py:76-103: b = exp((ln N_max - ln N_min)/(L-1)), scale_l = N_min b^l - 1,
    res_l = ceil(scale_l) + 1, dense size next_multiple(res_l^dim, 8) while it fits in T."""
    if L < 2:
        raise ValueError("HashGridEncoder needs L >= 2 (b divides by L - 1)")
    b = math.exp((math.log(N_max) - math.log(N_min)) / (L - 1))
    scales, res, offsets, first = [], [], [0], 0
    for i in range(L):
        scale = N_min * (b ** i) - 1
        r = math.ceil(scale) + 1
        n = next_multiple(r ** dim, 8)
        if n <= T:
            first += 1
        else:
            n = T
        scales.append(scale)
        res.append(r)
        offsets.append(offsets[-1] + n)
    return LevelTable(b, tuple(scales), tuple(res), tuple(offsets), first)


class _HashGridFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, table, desc_fn, n_out):
        N = pos.shape[0]
        enc = torch.empty(N, n_out, dtype=torch.float32, device=pos.device)
        if N:
            _lib.call("hashgrid_encode", [pos, table.contiguous(), enc], desc_fn(N))
        ctx.save_for_backward(pos)
        ctx.desc_fn = desc_fn
        ctx.table_shape = table.shape
        ctx.table_dtype = table.dtype
        return enc

    @staticmethod
    def backward(ctx, g):
        (pos,) = ctx.saved_tensors
        dtable = torch.empty(ctx.table_shape, dtype=torch.float32, device=pos.device)
        desc = ctx.desc_fn(pos.shape[0])
        ws = torch.empty(_lib.hashgrid_backward_workspace_bytes(desc), dtype=torch.uint8, device=pos.device)
        _lib.call("hashgrid_encode_backward", [pos, g.float().contiguous(), dtable, ws], desc)
        return None, dtable.to(ctx.table_dtype), None, None


class HashGridEncoder(torch.nn.Module):
    """Multiresolution hash encoding (models/encoders.py:57-256).

    ``forward(pos, bound)`` maps positions in [-bound, bound]^3 (f32 [N, 3]) to
    encodings f32 [N, L*F], level-major per point, and returns ``(encodings, tv)``.
    """

    def __init__(self, L: int, T: int, F: int, N_min: int, N_max: int, tv_scale: float = 0.0,
                 param_dtype: torch.dtype = torch.float32, device=None, generator: torch.Generator | None = None):
        super().__init__()
        if param_dtype not in (torch.float32, torch.float16):
            raise NotImplementedError(f"param_dtype must be float32 or float16, got {param_dtype}")
        if F not in (1, 2, 4, 8):
            raise NotImplementedError(f"F must be 1, 2, 4 or 8, got {F}")
        if L > _lib.HASHGRID_MAX_LEVELS:
            raise NotImplementedError(f"at most {_lib.HASHGRID_MAX_LEVELS} levels, got {L}")
        self.L, self.T, self.F, self.N_min, self.N_max = L, T, F, N_min, N_max
        self.tv_scale = float(tv_scale)
        self.param_dtype = param_dtype
        self.levels = level_table(L, T, N_min, N_max)
        # paper init U(-1e-4, 1e-4) (encoders.py:105-114)
        init = torch.empty(self.levels.n_rows, F, dtype=torch.float32, device=device)
        init.uniform_(-1e-4, 1e-4, generator=generator)
        self.register_parameter(PARAM_NAME, torch.nn.Parameter(init.to(param_dtype)))

    @property
    def b(self) -> float:
        return self.levels.b

    @property
    def table(self) -> torch.nn.Parameter:
        return getattr(self, PARAM_NAME)

    def descriptor(self, n_points: int, bound: float, enc_layout: int = _lib.ENC_POINT_MAJOR) -> bytes:
        """HashGridDescriptor bytes; enc_layout ENC_LEVEL_MAJOR selects the [L, N, F]
        encoding layout the fused training step uses."""
        lv = self.levels
        dt = _lib.DTYPE_F16 if self.table.dtype == torch.float16 else _lib.DTYPE_F32
        return _lib.hashgrid_descriptor(n_points, self.L, self.F, self.T, lv.first_hash_level, dt, lv.n_rows,
                                        float(bound), [float(s) for s in lv.scales], lv.resolutions, lv.offsets,
                                        enc_layout)

    def forward(self, pos: torch.Tensor, bound: float):
        if pos.dim() != 2 or pos.shape[1] != 3:
            raise ValueError(f"pos must be [N, 3], got {tuple(pos.shape)}")
        if pos.dtype != torch.float32:
            raise NotImplementedError(f"pos must be float32, got {pos.dtype}")
        if self.tv_scale > 0:
            raise NotImplementedError("total-variation regularisation (tv_scale > 0) is not implemented")
        desc = lambda n, _b=float(bound): self.descriptor(n, _b)
        enc = _HashGridFn.apply(pos.contiguous(), self.table, desc, self.L * self.F)
        return enc, 0.0
