"""The learned proposal network (no counterpart in the reference).

A hash grid (L=5, F=2, T=2^17, N_min=16, N_max=512) and a 10 -> 16 -> 1 MLP give a
density at 64 equal bins of each ray's segment [t_start, t_end).  It is trained
online against the NeRF's own sample weights from the occupancy-grid marcher
(histogram-bound loss, ``proposal_loss``), and at render time places 48 samples per
ray at the quantiles of its bin weights (``proposal_sample``); rays whose predicted
mass is below the renderer's fallback threshold (``render_image_proposal``'s
``mass_threshold``, 0.98 by default) are rendered by the occupancy-grid marcher.
With ``TrainConfig(proposal_guided=True)`` the field itself trains on samples the
proposal places (``proposal_train_samples``), and the proposal on the field's
weights at them; such a proposal renders with 32 samples spread over the support of
its predicted mass and falls back below 0.5 (``Trainer.proposal_render_settings``).
"""
from __future__ import annotations

import math

import torch

from . import _lib
from .encoders import HashGridEncoder

N_PARAMS = _lib.PROPOSAL_N_PARAMS
BINS = _lib.PROPOSAL_BINS


def glorot_uniform_proposal_(flat: torch.Tensor, generator=None) -> torch.Tensor:
    """glorot-uniform init of the two Dense kernels (10x16, 16x1)."""
    o = 0
    for fi, fo in ((10, 16), (16, 1)):
        a = math.sqrt(6.0 / (fi + fo))
        flat[o:o + fi * fo].uniform_(-a, a, generator=generator)
        o += fi * fo
    return flat


class ProposalNet(torch.nn.Module):
    def __init__(self, bound: float = 1.0, device=None, seed: int = 0, table_dtype: torch.dtype = torch.float32):
        super().__init__()
        self.bound = float(bound)
        gen = torch.Generator(device="cpu").manual_seed(seed)
        self.position_encoder = HashGridEncoder(L=5, T=2 ** 17, F=2, N_min=16, N_max=512, param_dtype=table_dtype,
                                                generator=gen)
        self.mlp_params = torch.nn.Parameter(glorot_uniform_proposal_(torch.empty(N_PARAMS), gen))
        self.mlp_half: torch.Tensor | None = None  # f16 copy of the weights (precision path)
        if device is not None:
            self.to(device)

    def weights(self):
        """(weights tensor, NGP dtype) the MLP kernels read."""
        if self.mlp_half is not None:
            return self.mlp_half, _lib.DTYPE_F16
        return self.mlp_params.detach(), _lib.DTYPE_F32

    @torch.no_grad()
    def density(self, xyz: torch.Tensor) -> torch.Tensor:
        """Proposal density at points xyz [N, 3] (f32)."""
        N = xyz.shape[0]
        out = torch.empty(N, dtype=torch.float32, device=xyz.device)
        if N == 0:
            return out
        pe = self.position_encoder
        enc = torch.empty(pe.L, N, pe.F, dtype=torch.float32, device=xyz.device)
        _lib.call("hashgrid_encode", [xyz.contiguous(), pe.table.detach(), enc],
                  pe.descriptor(N, self.bound, _lib.ENC_LEVEL_MAJOR))
        w, pd = self.weights()
        _lib.call("proposal_mlp_forward", [enc, w, out], _lib.proposal_mlp_descriptor(N, 0, _lib.ENC_LEVEL_MAJOR, pd))
        return out

    @torch.no_grad()
    def bin_density(self, rays_o, rays_d, t_starts, t_ends, chunk: int = 1 << 16) -> torch.Tensor:
        """Density at the 64 bin centres of every ray: [n_rays, 64]."""
        n = rays_o.shape[0]
        out = torch.empty(n, BINS, dtype=torch.float32, device=rays_o.device)
        pts = torch.empty(min(n, chunk), BINS, 3, dtype=torch.float32, device=rays_o.device)
        for a in range(0, n, chunk):
            m = min(chunk, n - a)
            p = pts[:m]
            _lib.call("proposal_points", [rays_o[a:a + m], rays_d[a:a + m], t_starts[a:a + m], t_ends[a:a + m], p],
                      _lib.proposal_rays_descriptor(m))
            out[a:a + m] = self.density(p.reshape(-1, 3)).reshape(m, BINS)
        return out
