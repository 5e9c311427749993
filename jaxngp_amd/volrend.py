"""Drop-in mirror of the reference's ``volrendjax`` package on torch device tensors.

Same function names, argument order, argument meaning and return tuples as
deps/volume-rendering-jax/src/volrendjax/{marching,integrating,packbits,morton3d}/__init__.py;
the shape/dtype rules of the reference's abstract evaluation
(``*/abstract.py``) are checked here and raise the same exception kinds
(ValueError for shapes, NotImplementedError for dtypes).  Each function launches
the gfx950 kernels of libngp_hip.so on the current stream; there is no CPU path.

``integrate_rays`` is differentiable through ``torch.autograd.Function`` (the
analogue of the reference's ``jax.custom_vjp``, integrating/impl.py:317-413): it
returns gradients for ``bgs``, ``z_vals`` and ``drgbs`` only, like the reference.
"""
from __future__ import annotations

from typing import Tuple

import torch

from . import _lib

u32 = torch.uint32


def _u32_to_i64(t: torch.Tensor) -> torch.Tensor:
    """Exact widening of a uint32 tensor (through a same-size int32 view)."""
    return t.view(torch.int32).to(torch.int64) & 0xFFFFFFFF


def _as_u32(t: torch.Tensor) -> torch.Tensor:
    """uint32 view / conversion of an integer tensor (values must fit in 32 bits)."""
    if t.dtype == u32:
        return t.contiguous()
    if t.dtype == torch.int32:
        return t.contiguous().view(u32)
    return t.to(torch.int64).to(torch.int32).contiguous().view(u32)


def _need(cond: bool, msg: str) -> None:
    if not cond:
        raise ValueError(msg)


def _need_dtype(t: torch.Tensor, dtype: torch.dtype, what: str) -> None:
    if t.dtype != dtype:
        raise NotImplementedError(f"{what} is only implemented for {dtype}, got {t.dtype}")


def _f32(x, n: int, device, what: str, width: int | None = None) -> torch.Tensor:
    """Broadcasts a scalar / tensor to [n] (or [n, width]) f32 on `device`, contiguous."""
    shape = (n,) if width is None else (n, width)
    t = torch.as_tensor(x, dtype=torch.float32, device=device)
    try:
        t = torch.broadcast_to(t, shape)
    except RuntimeError as e:
        raise ValueError(f"{what}: cannot broadcast shape {tuple(t.shape)} to {shape}") from e
    return t.contiguous()


# ---------------------------------------------------------------------------
# marching (volrendjax/marching/__init__.py)
# ---------------------------------------------------------------------------

def march_rays(
    total_samples: int,
    diagonal_n_steps: int,
    K: int,
    G: int,
    bound: float,
    stepsize_portion: float,
    rays_o: torch.Tensor,
    rays_d: torch.Tensor,
    t_starts: torch.Tensor,
    t_ends: torch.Tensor,
    noises,
    occupancy_bitfield: torch.Tensor,
):
    """Generates samples along each ray through occupied cells (marching/__init__.py:8-93).

    Returns ``(measured_batch_size_before_compaction, ray_is_valid, rays_n_samples,
    rays_sample_startidx, idcs, xyzs, dirs, dss, z_vals)`` with the reference's
    dtypes and shapes.  Sample ranges are assigned in ray order (exclusive prefix
    sum), so the output is deterministic; ``measured_batch_size_before_compaction``
    is a 0-d uint32 device tensor.
    """
    _need(rays_o.dim() == 2 and rays_o.shape[1] == 3, f"rays_o must be [n_rays, 3], got {tuple(rays_o.shape)}")
    n_rays = rays_o.shape[0]
    _need(tuple(rays_d.shape) == (n_rays, 3), f"rays_d must be [{n_rays}, 3], got {tuple(rays_d.shape)}")
    _need(tuple(t_starts.shape) == (n_rays,) and tuple(t_ends.shape) == (n_rays,), "t_starts/t_ends must be [n_rays]")
    _need(tuple(occupancy_bitfield.shape) == (K * G * G * G // 8,),
          f"occupancy_bitfield must be [{K * G * G * G // 8}], got {tuple(occupancy_bitfield.shape)}")
    _need_dtype(occupancy_bitfield, torch.uint8, "occupancy_bitfield")
    _need(total_samples > 0 and diagonal_n_steps > 0 and K > 0 and G > 0 and bound > 0 and stepsize_portion >= 0,
          "total_samples, diagonal_n_steps, K, G, bound must be positive and stepsize_portion non-negative")
    _need_dtype(rays_o, torch.float32, "march_rays (input coordinates)")
    dev = rays_o.device
    noises = _f32(noises, n_rays, dev, "noises")
    S = int(total_samples)
    next_loc = torch.empty(1, dtype=u32, device=dev)
    exceeded = torch.empty(1, dtype=u32, device=dev)
    valid = torch.empty(n_rays, dtype=torch.bool, device=dev)
    n_samples = torch.empty(n_rays, dtype=u32, device=dev)
    startidx = torch.empty(n_rays, dtype=u32, device=dev)
    idcs = torch.empty(S, dtype=u32, device=dev)
    xyzs = torch.empty(S, 3, dtype=torch.float32, device=dev)
    dirs = torch.empty(S, 3, dtype=torch.float32, device=dev)
    dss = torch.empty(S, dtype=torch.float32, device=dev)
    z_vals = torch.empty(S, dtype=torch.float32, device=dev)
    _lib.call(
        "march_rays",
        [rays_o.contiguous(), rays_d.float().contiguous(), t_starts.float().contiguous(), t_ends.float().contiguous(),
         noises, occupancy_bitfield.contiguous(), next_loc, exceeded, valid, n_samples, startidx, idcs, xyzs, dirs,
         dss, z_vals],
        _lib.marching_descriptor(n_rays, S, diagonal_n_steps, K, G, float(bound), float(stepsize_portion)),
    )
    measured = (_u32_to_i64(next_loc) - _u32_to_i64(exceeded))[0].to(torch.int32).view(u32)
    return measured, valid, n_samples, startidx, idcs, xyzs, dirs, dss, z_vals


def march_rays_inference(
    diagonal_n_steps: int,
    K: int,
    G: int,
    march_steps_cap: int,
    bound: float,
    stepsize_portion: float,
    rays_o: torch.Tensor,
    rays_d: torch.Tensor,
    t_starts: torch.Tensor,
    t_ends: torch.Tensor,
    occupancy_bitfield: torch.Tensor,
    next_ray_index_in: torch.Tensor,
    terminated: torch.Tensor,
    indices: torch.Tensor,
):
    """One marching round of the inference ray pool (marching/__init__.py:96-157).

    Returns ``(next_ray_index, indices, n_samples, t_starts, xyzs, dss, z_vals)``;
    ``t_starts`` is the input array with the marched rays' values replaced
    (``t_starts.at[indices].set(...)``, out-of-range indices dropped).
    """
    _need(rays_o.dim() == 2 and rays_o.shape[1] == 3, "rays_o must be [n_total_rays, 3]")
    n_total = rays_o.shape[0]
    _need(terminated.dim() == 1, "terminated must be [n_rays]")
    n_rays = terminated.shape[0]
    _need(tuple(rays_d.shape) == (n_total, 3), "rays_d must be [n_total_rays, 3]")
    _need(tuple(t_starts.shape) == (n_total,) and tuple(t_ends.shape) == (n_total,), "t_starts/t_ends must be [n_total_rays]")
    _need(tuple(occupancy_bitfield.shape) == (K * G * G * G // 8,), "occupancy_bitfield has the wrong size")
    _need_dtype(occupancy_bitfield, torch.uint8, "occupancy_bitfield")
    _need(tuple(next_ray_index_in.shape) == (1,), "next_ray_index_in must be [1]")
    _need(tuple(indices.shape) == (n_rays,), "indices must be [n_rays]")
    _need(march_steps_cap > 0, "march_steps_cap must be positive")
    dev = rays_o.device
    cap = int(march_steps_cap)
    next_out = torch.empty(1, dtype=u32, device=dev)
    idx_out = torch.empty(n_rays, dtype=u32, device=dev)
    n_samples = torch.empty(n_rays, dtype=u32, device=dev)
    t_out = torch.empty(n_rays, dtype=torch.float32, device=dev)
    xyzs = torch.empty(n_rays, cap, 3, dtype=torch.float32, device=dev)
    dss = torch.empty(n_rays, cap, dtype=torch.float32, device=dev)
    z_vals = torch.empty(n_rays, cap, dtype=torch.float32, device=dev)
    _lib.call(
        "march_rays_inference",
        [rays_o.float().contiguous(), rays_d.float().contiguous(), t_starts.float().contiguous(),
         t_ends.float().contiguous(), occupancy_bitfield.contiguous(), _as_u32(next_ray_index_in),
         terminated.to(torch.bool).contiguous(), _as_u32(indices), next_out, idx_out, n_samples, t_out,
         xyzs, dss, z_vals],
        _lib.marching_inference_descriptor(n_total, n_rays, diagonal_n_steps, K, G, cap, float(bound),
                                           float(stepsize_portion)),
    )
    idx64 = _u32_to_i64(idx_out)
    keep = idx64 < n_total
    t_new = t_starts.clone()
    t_new[idx64[keep]] = t_out[keep]
    return next_out, idx_out, n_samples, t_new, xyzs, dss, z_vals


# ---------------------------------------------------------------------------
# integrating (volrendjax/integrating/__init__.py, impl.py)
# ---------------------------------------------------------------------------

def _check_integrate_inputs(rays_sample_startidx, rays_n_samples, bgs, dss, z_vals, drgbs):
    _need(rays_sample_startidx.dim() == 1, "rays_sample_startidx must be [n_rays]")
    n_rays = rays_sample_startidx.shape[0]
    _need(dss.dim() == 1, "dss must be [total_samples]")
    S = dss.shape[0]
    _need(tuple(rays_n_samples.shape) == (n_rays,), "rays_n_samples must be [n_rays]")
    _need(tuple(bgs.shape) == (n_rays, 3), f"bgs must be [{n_rays}, 3], got {tuple(bgs.shape)}")
    _need(tuple(z_vals.shape) == (S,), "z_vals must be [total_samples]")
    _need(tuple(drgbs.shape) == (S, 4), f"drgbs must be [{S}, 4], got {tuple(drgbs.shape)}")
    _need_dtype(drgbs, torch.float32, "integrate_rays (input prediction)")
    return n_rays, S


class _IntegrateRays(torch.autograd.Function):
    """integrate_rays with the reference's hand-written VJP (integrating/impl.py:340-413)."""

    @staticmethod
    def forward(ctx, near_distance, startidx, n_samples, bgs, dss, z_vals, drgbs):
        n_rays, S = _check_integrate_inputs(startidx, n_samples, bgs, dss, z_vals, drgbs)
        dev = drgbs.device
        measured = torch.empty(1, dtype=u32, device=dev)
        rgbds = torch.empty(n_rays, 4, dtype=torch.float32, device=dev)
        opac = torch.empty(n_rays, dtype=torch.float32, device=dev)
        ins = [_as_u32(startidx), _as_u32(n_samples), bgs.float().contiguous(),
               dss.float().contiguous(), z_vals.float().contiguous(), drgbs.contiguous()]
        _lib.call("integrate_rays", ins + [measured, rgbds, opac], _lib.integrating_descriptor(n_rays, S))
        ctx.save_for_backward(*ins, rgbds, opac)
        ctx.near_distance = float(near_distance)
        ctx.mark_non_differentiable(measured)
        return measured, rgbds, opac

    @staticmethod
    def backward(ctx, _g_measured, g_rgbds, _g_opac):
        startidx, n_samples, bgs, dss, z_vals, drgbs, rgbds, opac = ctx.saved_tensors
        n_rays, S = startidx.shape[0], dss.shape[0]
        dev = drgbs.device
        if g_rgbds is None:
            g_rgbds = torch.zeros(n_rays, 4, dtype=torch.float32, device=dev)
        dbgs = torch.empty(n_rays, 3, dtype=torch.float32, device=dev)
        dz = torch.empty(S, dtype=torch.float32, device=dev)
        ddrgbs = torch.empty(S, 4, dtype=torch.float32, device=dev)
        _lib.call(
            "integrate_rays_backward",
            [startidx, n_samples, bgs, dss, z_vals, drgbs, rgbds, opac, g_rgbds.float().contiguous(), dbgs, dz,
             ddrgbs],
            _lib.integrating_backward_descriptor(n_rays, S, ctx.near_distance),
        )
        # (near_distance, startidx, n_samples) get no gradient, nor does dss (impl.py:393-408)
        return None, None, None, dbgs, None, dz, ddrgbs


def integrate_rays(
    near_distance: float,
    rays_sample_startidx: torch.Tensor,
    rays_n_samples: torch.Tensor,
    bgs,
    dss: torch.Tensor,
    z_vals: torch.Tensor,
    drgbs: torch.Tensor,
) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Front-to-back compositing with early stop (integrating/__init__.py:11-59).

    Returns ``(measured_batch_size, final_rgbds, final_opacities)``; gradients flow
    to ``bgs``, ``z_vals`` and ``drgbs`` through the reference's custom VJP.
    ``bgs`` may be anything broadcastable to [n_rays, 3].
    """
    n_rays = rays_sample_startidx.shape[0]
    bgs_t = torch.as_tensor(bgs, dtype=torch.float32, device=drgbs.device)
    bgs_b = torch.broadcast_to(bgs_t, (n_rays, 3))
    measured, rgbds, opac = _IntegrateRays.apply(near_distance, rays_sample_startidx, rays_n_samples, bgs_b, dss,
                                                 z_vals, drgbs)
    return measured[0], rgbds, opac


def integrate_rays_inference(
    rays_bg: torch.Tensor,
    rays_rgbd: torch.Tensor,
    rays_T: torch.Tensor,
    n_samples: torch.Tensor,
    indices: torch.Tensor,
    dss: torch.Tensor,
    z_vals: torch.Tensor,
    drgbs: torch.Tensor,
):
    """Continues compositing of the rays in the pool (integrating/__init__.py:62-110).

    Returns ``(terminate_cnt, terminated, rays_rgbd, rays_T)`` where ``rays_rgbd`` /
    ``rays_T`` are the inputs with the pooled rays' entries replaced.
    """
    _need(rays_rgbd.dim() == 2 and rays_rgbd.shape[1] == 4, "rays_rgbd must be [n_total_rays, 4]")
    n_total = rays_rgbd.shape[0]
    _need(dss.dim() == 2, "dss must be [n_rays, march_steps_cap]")
    n_rays, cap = dss.shape
    _need(tuple(rays_bg.shape) == (n_total, 3), "rays_bg must be [n_total_rays, 3]")
    _need(tuple(rays_T.shape) == (n_total,), "rays_T must be [n_total_rays]")
    _need(tuple(n_samples.shape) == (n_rays,) and tuple(indices.shape) == (n_rays,), "n_samples/indices must be [n_rays]")
    _need(tuple(z_vals.shape) == (n_rays, cap), "z_vals must be [n_rays, march_steps_cap]")
    _need(tuple(drgbs.shape) == (n_rays, cap, 4), "drgbs must be [n_rays, march_steps_cap, 4]")
    dev = rays_rgbd.device
    cnt = torch.empty(1, dtype=u32, device=dev)
    term = torch.empty(n_rays, dtype=torch.bool, device=dev)
    rgbd_out = torch.empty(n_rays, 4, dtype=torch.float32, device=dev)
    T_out = torch.empty(n_rays, dtype=torch.float32, device=dev)
    idx = _as_u32(indices)
    _lib.call(
        "integrate_rays_inference",
        [rays_bg.float().contiguous(), rays_rgbd.float().contiguous(), rays_T.float().contiguous(),
         _as_u32(n_samples), idx, dss.float().contiguous(), z_vals.float().contiguous(),
         drgbs.float().contiguous(), cnt, term, rgbd_out, T_out],
        _lib.integrating_inference_descriptor(n_total, n_rays, cap),
    )
    idx64 = _u32_to_i64(idx)
    keep = idx64 < n_total
    rgbd_new = rays_rgbd.clone()
    T_new = rays_T.clone()
    rgbd_new[idx64[keep]] = rgbd_out[keep]
    T_new[idx64[keep]] = T_out[keep]
    return cnt[0], term, rgbd_new, T_new


# ---------------------------------------------------------------------------
# packbits, morton3d
# ---------------------------------------------------------------------------

def packbits(density_threshold, density_grid: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """Occupancy mask and LSB-first bitfield of ``density_grid > density_threshold``
    (packbits/__init__.py:8-28)."""
    _need(density_grid.dim() == 1, "density_grid must be rank 1")
    n_bits = density_grid.shape[0]
    if n_bits % 8 != 0:
        raise ValueError(f"pack_density_into_bits expects size of density grid to be divisible by 8, got {n_bits}")
    _need_dtype(density_grid, torch.float32, "pack_density_into_bits (densities)")
    n_bytes = n_bits // 8
    _need(n_bytes > 0, "expected n_bytes to be a positive integer, got 0")
    dev = density_grid.device
    thr = _f32(density_threshold, n_bits, dev, "density_threshold")
    mask = torch.empty(n_bits, dtype=torch.bool, device=dev)
    bits = torch.empty(n_bytes, dtype=torch.uint8, device=dev)
    _lib.call("pack_density_into_bits", [thr, density_grid.contiguous(), mask, bits], _lib.packbits_descriptor(n_bytes))
    return mask, bits


def morton3d(xyzs: torch.Tensor) -> torch.Tensor:
    """Morton code of uint32 [length, 3] coordinates (morton3d/__init__.py:6-7)."""
    _need(xyzs.dim() == 2 and xyzs.shape[1] == 3, "xyzs must be [length, 3]")
    _need_dtype(xyzs, u32, "morton3d (input coordinates)")
    n = xyzs.shape[0]
    out = torch.empty(n, dtype=u32, device=xyzs.device)
    if n:
        _lib.call("morton3d", [xyzs.contiguous(), out], _lib.morton3d_descriptor(n))
    return out


def morton3d_invert(idcs: torch.Tensor) -> torch.Tensor:
    """Inverse Morton code of uint32 [length] indices (morton3d/__init__.py:9-10)."""
    _need(idcs.dim() == 1, "idcs must be [length]")
    _need_dtype(idcs, u32, "morton3d_invert (input indices)")
    n = idcs.shape[0]
    out = torch.empty(n, 3, dtype=u32, device=idcs.device)
    if n:
        _lib.call("morton3d_invert", [idcs.contiguous(), out], _lib.morton3d_descriptor(n))
    return out
