"""The reference's background model for unbounded scenes on the sky_sphere_bg targets.

`SkySphereBg` mirrors models/nerfs.py:133-213 and `make_skysphere_background_model_ngp`
(nerfs.py:410-419: radius 4 bound, SH degree 2 of the hit point, SH degree 4 of the
direction, a 32->32->3 MLP without biases, sigmoid).  Called as
bg(rays_o [n, 3], rays_d [n, 3], appearance) it returns colours [n, 3] and is
differentiable in its parameters and in per-ray appearance embeddings, the
gradients jax.grad takes through `bg_fn` in app/nerf/_utils.py:128-133.  As in the
reference, the hit points are divided by the norm of the whole batch, so a colour
depends on the other rays of the call, and one ray that misses the sphere makes
every colour NaN.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

DENSE_SHAPES = lambda E: (("Dense_0", 20 + E, 32), ("Dense_1", 32, 32), ("Dense_2", 32, 3))  # noqa: E731


def _appearance(app: torch.Tensor | None, n: int, E: int, dev) -> tuple[torch.Tensor, bool]:
    if E == 0:
        return torch.empty(0, device=dev), True
    if app is None:
        raise ValueError(f"the model takes {E} appearance dims per ray")
    app = app.to(device=dev, dtype=torch.float32)
    if app.dim() == 1:
        if app.shape[0] != E:
            raise ValueError(f"appearance must have {E} values, got {tuple(app.shape)}")
        return app.contiguous(), False
    if tuple(app.shape) != (n, E):
        raise ValueError(f"appearance must be [{n}, {E}] or [{E}], got {tuple(app.shape)}")
    return app.contiguous(), True


class _SkySphere(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rays_o, rays_d, app, params, E, per_ray, radius):
        n = rays_o.shape[0]
        dev = rays_o.device
        colors = torch.empty(n, 3, dtype=torch.float32, device=dev)
        desc = _lib.sky_sphere_bg_descriptor(n, E, per_ray, radius)
        if n:
            ws = torch.empty(_lib.sky_sphere_bg_workspace_bytes(desc), dtype=torch.uint8, device=dev)
            _lib.call("sky_sphere_bg", [rays_o, rays_d, app, params, colors, ws], desc)
        ctx.save_for_backward(rays_o, rays_d, app, params)
        ctx.meta = (E, per_ray, radius)
        return colors

    @staticmethod
    def backward(ctx, g):
        rays_o, rays_d, app, params = ctx.saved_tensors
        E, per_ray, radius = ctx.meta
        if E and not per_ray:
            raise NotImplementedError("gradients need per-ray appearance embeddings ([n, E])")
        n = rays_o.shape[0]
        dev = rays_o.device
        desc = _lib.sky_sphere_bg_descriptor(n, E, 1, radius)
        dparams = torch.empty_like(params)
        dapp = torch.empty(n, E, dtype=torch.float32, device=dev) if E else torch.empty(0, device=dev)
        ws = torch.empty(_lib.sky_sphere_bg_workspace_bytes(desc), dtype=torch.uint8, device=dev)
        _lib.call("sky_sphere_bg_backward", [rays_o, rays_d, app, params, g.contiguous(), dparams, dapp, ws], desc)
        return None, None, (dapp if E else None), dparams, None, None, None


class SkySphereBg(torch.nn.Module):
    """nerfs.py:133-213 with make_skysphere_background_model_ngp's shapes.  The
    parameters are one flat f32 vector of the flax Dense kernels ([in, out],
    row-major): rgb_mlp/Dense_0 [20+E, 32], Dense_1 [32, 32], Dense_2 [32, 3]."""

    def __init__(self, radius: float, n_extra_dims: int = 0, device=None, seed: int = 0):
        super().__init__()
        if not 0 <= n_extra_dims <= _lib.SKY_MAX_EXTRA_DIMS:
            raise ValueError(f"n_extra_dims must be in [0, {_lib.SKY_MAX_EXTRA_DIMS}]")
        self.radius = float(radius)
        self.n_extra_dims = int(n_extra_dims)
        # flax's glorot_uniform per kernel: U(-l, l), l = sqrt(6 / (fan_in + fan_out))
        g = torch.Generator().manual_seed(seed)
        parts = []
        for _, fi, fo in DENSE_SHAPES(self.n_extra_dims):
            lim = (6.0 / (fi + fo)) ** 0.5
            parts.append((torch.rand(fi * fo, generator=g, dtype=torch.float64) * 2 - 1) * lim)
        self.params = torch.nn.Parameter(torch.cat(parts).float().to(device))
        assert self.params.numel() == _lib.sky_sphere_bg_n_params(self.n_extra_dims)

    def forward(self, rays_o: torch.Tensor, rays_d: torch.Tensor,
                appearance: torch.Tensor | None = None) -> torch.Tensor:
        if rays_o.shape != rays_d.shape or rays_o.dim() != 2 or rays_o.shape[1] != 3:
            raise ValueError(f"rays must be [n, 3], got {tuple(rays_o.shape)} and {tuple(rays_d.shape)}")
        dev = self.params.device
        o = rays_o.to(device=dev, dtype=torch.float32).contiguous()
        d = rays_d.to(device=dev, dtype=torch.float32).contiguous()
        app, per_ray = _appearance(appearance, o.shape[0], self.n_extra_dims, dev)
        return _SkySphere.apply(o, d, app, self.params, self.n_extra_dims, per_ray, self.radius)

    # flax parameter tree under params["bg"] (the checkpoint layout)
    def param_tree(self) -> dict:
        flat = self.params.detach().cpu().numpy()
        tree, o = {}, 0
        for name, fi, fo in DENSE_SHAPES(self.n_extra_dims):
            tree[name] = {"kernel": flat[o:o + fi * fo].reshape(fi, fo).copy()}
            o += fi * fo
        return {"rgb_mlp": tree}

    def load_param_tree(self, tree: dict) -> None:
        parts = []
        for name, fi, fo in DENSE_SHAPES(self.n_extra_dims):
            k = np.asarray(tree["rgb_mlp"][name]["kernel"], np.float32)
            if k.shape != (fi, fo):
                raise ValueError(f"rgb_mlp/{name}/kernel has shape {k.shape}, expected {(fi, fo)}")
            parts.append(k.reshape(-1))
        with torch.no_grad():
            self.params.copy_(torch.from_numpy(np.concatenate(parts)).to(self.params.device))


def make_skysphere_background_model_ngp(bound: float, n_extra_dims: int = 0, device=None,
                                        seed: int = 0) -> SkySphereBg:
    """nerfs.py:410-419: radius 4 bound."""
    return SkySphereBg(radius=bound * 4, n_extra_dims=n_extra_dims, device=device, seed=seed)
