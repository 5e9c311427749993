"""Ray generation and renderers (models/renderers/cuda.py) on the gfx950 kernels.

``render_rays_train`` is the differentiable training renderer (cuda.py:100-162):
march_rays -> NeRF -> integrate_rays, with gradients through the hand-written
integrate_rays VJP, the fused MLP backward and the hash-grid table gradient.

``render_image_inference`` renders a full frame with the reference's ray-pool
scheme (cuda.py:176-373): a pool of up to 8192 ray slots, each round marches at
most ``march_steps_cap`` occupied steps per ray (march_rays_inference), evaluates
the field on those samples and composites them (integrate_rays_inference);
terminated slots take the next unrendered rays.  Unlike the reference it keeps the
pool on the device between rounds and only synchronises once per outer round to
read the termination count.
"""
from __future__ import annotations

import math

import torch

from . import _lib, volrend
from .scene import Camera, f32_to_u8


def make_rays_worldspace(camera: Camera, pose: torch.Tensor):
    """Pixel-centre rays of every pixel (cuda.py:21-53, types.py:400-439).
    pose: camera-to-world [3, 4] (rotation | translation)."""
    dev = pose.device
    idx = torch.arange(camera.n_pixels, device=dev)
    x = (idx % camera.width).float()
    y = torch.div(idx, camera.width, rounding_mode="floor").float()
    dx = ((x + 0.5) - camera.cx) / camera.fx
    dy = ((y + 0.5) - camera.cy) / camera.fy
    d = torch.stack([dx, -dy, -torch.ones_like(dx)], -1)
    d = d / torch.linalg.norm(d, dim=-1, keepdim=True)
    R = pose[:, :3].float()
    o = pose[:, 3].float().expand(camera.n_pixels, 3).contiguous()
    return o, (d @ R.T).contiguous()


def make_near_far_from_bound(bound: float, o: torch.Tensor, d: torch.Tensor):
    """AABB entry/exit distances (cuda.py:56-97)."""
    eps = 1e-15
    d = torch.where(torch.signbit(d), torch.clamp(d, max=-eps), torch.clamp(d, min=eps))
    t0 = (-bound - o) / d
    t1 = (bound - o) / d
    ts = torch.minimum(t0, t1).amax(-1)
    te = torch.maximum(t0, t1).amin(-1)
    return torch.clamp(ts, min=0.0).contiguous(), te.contiguous()


def render_rays_train(o_world, d_world, bg, total_samples: int, model, occupancy_bitfield, *, K=1, G=128,
                      diagonal_n_steps=1024, stepsize_portion=0.0, near=0.3, perturb=True, generator=None):
    """Differentiable training render (cuda.py:100-162).  Returns (batch_metrics, final_rgbds, tv)."""
    bound = model.bound
    t_starts, t_ends = make_near_far_from_bound(bound, o_world, d_world)
    noises = (torch.rand(t_starts.shape, device=o_world.device, generator=generator) if perturb else 0.0)
    measured0, valid, ns, start, idcs, xyzs, dirs, dss, z_vals = volrend.march_rays(
        total_samples, diagonal_n_steps, K, G, bound, stepsize_portion, o_world, d_world, t_starts, t_ends, noises,
        occupancy_bitfield)
    drgbs, tv = model(xyzs, dirs, None)
    effective, final_rgbds, final_op = volrend.integrate_rays(near, start, ns, bg, dss, z_vals, drgbs)
    metrics = {
        "n_valid_rays": valid.sum(),
        "ray_is_valid": valid,
        "measured_batch_size_before_compaction": measured0,
        "measured_batch_size": effective,
    }
    return metrics, final_rgbds, tv


@torch.no_grad()
def march_and_integrate_inference(next_ray_index, terminated, indices, rays_rgbd, rays_T, t_starts, *, model,
                                  rays_o, rays_d, t_ends, rays_bg, occupancy_bitfield, march_steps_cap, K, G,
                                  diagonal_n_steps, stepsize_portion):
    """One pool round (cuda.py:176-241)."""
    next_ray_index, indices, n_samples, t_starts, xyzs, dss, z_vals = volrend.march_rays_inference(
        diagonal_n_steps, K, G, march_steps_cap, model.bound, stepsize_portion, rays_o, rays_d, t_starts, t_ends,
        occupancy_bitfield, next_ray_index, terminated, indices)
    # The reference evaluates the field on every [slot, step] entry, zero padding
    # included (cuda.py:213-223); here only the marched samples go through the
    # encoder and MLP, and the padding entries get drgbs = 0 (integrate_rays_inference
    # reads only the first n_samples of each slot).
    n_slots, cap = dss.shape
    live = (torch.arange(cap, device=dss.device)[None, :] < n_samples.view(torch.int32)[:, None]).reshape(-1)
    pos = live.nonzero().squeeze(1)  # one host sync per round
    drgbs = torch.zeros(n_slots * cap, 4, dtype=torch.float32, device=dss.device)
    if pos.numel():
        slot = torch.div(pos, cap, rounding_mode="floor")
        ray = indices.view(torch.int32)[slot].long()
        out, _ = model(xyzs.reshape(-1, 3)[pos], rays_d[ray], None)
        drgbs[pos] = out
    terminate_cnt, terminated, rays_rgbd, rays_T = volrend.integrate_rays_inference(
        rays_bg, rays_rgbd, rays_T, n_samples, indices, dss, z_vals, drgbs.reshape(n_slots, cap, 4))
    return terminate_cnt, next_ray_index, terminated, indices, rays_rgbd, rays_T, t_starts


@torch.no_grad()
def render_image_inference(model, camera: Camera, pose: torch.Tensor, occupancy_bitfield, bg=(1.0, 1.0, 1.0), *,
                           K=1, G=128, diagonal_n_steps=1024, stepsize_portion=0.0, march_steps_cap=8,
                           pool: int = 1 << 18, bg_model=None, appearance=None):
    """Renders one frame (cuda.py:244-373).  Returns (bg [H,W,3] f32, rgb u8 [H,W,3],
    depth u8 [H,W]).

    Same ray-pool rounds as the reference: march_rays_inference, the field on the
    marched samples, integrate_rays_inference, scatter of the per-slot results,
    and one host sync per outer round to count finished rays.  The image does not
    depend on the pool size (every ray is marched in the same chunks of
    march_steps_cap steps whichever slot holds it); the reference uses 8192
    slots, the default here is 2^18, which fills the GPU in fewer rounds.  The
    buffers are allocated once per frame and every launch is pre-bound; the field
    runs fused (nerf_field_chunks) on the marched samples only.  With `bg_model`
    (a SkySphereBg) the per-ray background is the model's colour for the frame's
    rays and `appearance` (cuda.py:283-288), instead of the constant `bg`.
    """
    o, d = make_rays_worldspace(camera, pose)
    if bg_model is not None:
        with torch.no_grad():
            bg = bg_model(o, d, appearance)
    rays_rgbd, rays_bg = render_rays_inference(model, o, d, occupancy_bitfield, bg, K=K, G=G,
                                               diagonal_n_steps=diagonal_n_steps, stepsize_portion=stepsize_portion,
                                               march_steps_cap=march_steps_cap, pool=pool)
    return _image(camera, rays_rgbd, rays_bg)


def _image(camera: Camera, rays_rgbd: torch.Tensor, rays_bg: torch.Tensor):
    rgb = f32_to_u8(rays_rgbd[:, :3]).reshape(camera.height, camera.width, 3)
    depth = rays_rgbd[:, 3]
    depth_u8 = f32_to_u8((depth - depth.min()) / (depth.max() - depth.min() + 1e-15)).reshape(camera.height,
                                                                                              camera.width)
    return rays_bg.reshape(camera.height, camera.width, 3), rgb, depth_u8


@torch.no_grad()
def render_rays_inference(model, o: torch.Tensor, d: torch.Tensor, occupancy_bitfield, bg=(1.0, 1.0, 1.0), *, K=1,
                          G=128, diagonal_n_steps=1024, stepsize_portion=0.0, march_steps_cap=8, pool: int = 1 << 18):
    """The ray-pool loop of render_image_inference on explicit rays o, d [n, 3].
    Returns (rays_rgbd [n, 4], rays_bg [n, 3])."""
    dev = o.device
    t_starts, t_ends = make_near_far_from_bound(model.bound, o, d)
    n_px = o.shape[0]
    rays_rgbd = torch.zeros(n_px, 4, device=dev)
    rays_T = torch.ones(n_px, device=dev)
    rays_bg = torch.as_tensor(bg, dtype=torch.float32, device=dev).expand(n_px, 3).contiguous()
    n = min(pool, n_px)
    cap = int(march_steps_cap)
    u32 = torch.uint32
    nxt = torch.zeros(2, dtype=torch.int32, device=dev).view(u32)
    terminated = torch.ones(n, dtype=torch.bool, device=dev)
    idx = torch.zeros(2, n, dtype=torch.int32, device=dev).view(u32)
    n_samples = torch.empty(n, dtype=u32, device=dev)
    t_out = torch.empty(n, device=dev)
    xyzs = torch.empty(n, cap, 3, device=dev)
    dss = torch.empty(n, cap, device=dev)
    zs = torch.empty(n, cap, device=dev)
    drgbs = torch.empty(n, cap, 4, device=dev)
    cnt = torch.empty(1, dtype=u32, device=dev)
    rgbd_out = torch.empty(n, 4, device=dev)
    T_out = torch.empty(n, device=dev)
    total = torch.zeros((), dtype=torch.int64, device=dev)
    pe = model.position_encoder
    half = getattr(model, "mlp_half", None) is not None
    params = model.mlp_half if half else model.mlp_params.detach()
    fdesc = _lib.field_chunks_descriptor(pe.descriptor(0, model.bound), n, cap, n_px,
                                         _lib.DTYPE_F16 if half else _lib.DTYPE_F32)
    mdesc = _lib.marching_inference_descriptor(n_px, n, diagonal_n_steps, K, G, cap, model.bound, stepsize_portion)
    P = _lib.Prepared
    rounds = []
    for a, b in ((0, 1), (1, 0)):  # next index / slot indices ping-pong between rounds
        ib = idx[b]
        rounds.append((
            P("march_rays_inference", [o, d, t_starts, t_ends, occupancy_bitfield, nxt[a:a + 1], terminated, idx[a],
                                       nxt[b:b + 1], ib, n_samples, t_out, xyzs, dss, zs], mdesc),
            P("pool_scatter", [ib, t_out, t_starts], _lib.pool_scatter_descriptor(n, n_px, 1)),
            P("nerf_field_chunks", [xyzs, n_samples, ib, d, pe.table.detach(), params, drgbs], fdesc),
            P("integrate_rays_inference", [rays_bg, rays_rgbd, rays_T, n_samples, ib, dss, zs, drgbs, cnt,
                                           terminated, rgbd_out, T_out],
              _lib.integrating_inference_descriptor(n_px, n, cap)),
            P("pool_scatter", [ib, rgbd_out, rays_rgbd], _lib.pool_scatter_descriptor(n, n_px, 4)),
            P("pool_scatter", [ib, T_out, rays_T], _lib.pool_scatter_descriptor(n, n_px, 1)),
        ))
    cnt32 = cnt.view(torch.int32)
    n_rendered, r = 0, 0
    while n_rendered < n_px:
        iters = max(1, (n_px - n_rendered) // n)
        iters = 2 ** int(math.log2(iters) + 1)
        for _ in range(iters):
            for call in rounds[r & 1]:
                call()
            total += cnt32[0]
            r += 1
        _lib.check()
        n_rendered = int(total.item())
    return rays_rgbd, rays_bg


_SIDE_STREAMS: dict = {}


def _side_stream(dev) -> torch.cuda.Stream:
    """A second stream per device (default priority) for work that overlaps the ray pool."""
    key = torch.device(dev).index
    if key not in _SIDE_STREAMS:
        _SIDE_STREAMS[key] = torch.cuda.Stream(device=dev)
    return _SIDE_STREAMS[key]


@torch.no_grad()
def render_image_proposal(model, proposal, camera: Camera, pose: torch.Tensor, occupancy_bitfield,
                          bg=(1.0, 1.0, 1.0), *, n_per_ray: int = 48, mass_threshold: float = 0.98,
                          placement: int = _lib.PROPOSAL_QUANTILES, K=1, G=128,
                          diagonal_n_steps=1024, stepsize_portion=0.0, march_steps_cap=8, pool: int = 1 << 18,
                          fallback_steps_cap: int = 64, fallback_pool: int = 1 << 16,
                          empty_mass: float = 0.0, stats: dict | None = None):
    """Renders one frame with samples placed by the learned proposal: its density at
    64 bins of every ray gives 48 samples, by default at the quantiles of the ray's
    predicted mass (placement PROPOSAL_QUANTILES; PROPOSAL_SUPPORT spreads them evenly
    over the bins that hold mass instead), minus those in cells the occupancy grid
    marks empty (proposal_render_samples, which evaluates the proposal too); the field
    runs on them (nerf_field_chunks) and integrate_rays composites them.  Rays whose
    predicted mass is below `mass_threshold` are rendered by the occupancy-grid ray
    pool (render_rays_inference) instead.  The proposal is trained to bound the
    field's weights from above, so its mass over-reads; 0.98 is the default because
    rays predicted between 0.9 and 0.98 are where its samples miss the field's mass.
    The fallback marches `fallback_steps_cap` samples per ray and round: a ray's
    samples and colour do not depend on that cap (march_rays_inference resumes from
    the t it stopped at and defers a far-plane sample to the next round), so its
    pixels equal the plain render's, in fewer rounds.  The one exception is the
    reference's own: when a ray's last regular sample ends exactly on t_end (an f32
    tie) and the far-plane sample fills a round's last slot, the ray survives the
    round (integrating.cu:300) and the next round writes the far-plane sample again
    (marching.cu:317, 367-394), so such a ray depends on where the rounds fall.  About
    one ray in 10^5.  The fallback's ray pool has `fallback_pool` slots: its rays
    (about a third of a frame) finish sooner in a pool of 2^16 than in the plain
    render's 2^18, whose later rounds sweep mostly idle slots (measured 14.3 against
    15.8 ms per 800x800 frame, same pixels; the plain render of all rays is fastest
    at 2^18).  `march_steps_cap` and `pool` are accepted for the same keyword set as
    render_image_inference.  Same return as
    render_image_inference; `stats` (if given) receives the fallback count.  Rays whose
    predicted mass is below `empty_mass` (0: none) do not fall back: they get no
    samples and composite to the background, as the plain render does where the field
    holds no mass along the ray."""
    dev = pose.device
    o, d = make_rays_worldspace(camera, pose)
    t0, t1 = make_near_far_from_bound(model.bound, o, d)
    n = o.shape[0]
    u32 = torch.uint32
    start = torch.empty(n, dtype=u32, device=dev)
    cnt = torch.empty(n, dtype=u32, device=dev)
    fb = torch.empty(n, dtype=torch.bool, device=dev)
    xyzs = torch.empty(n, n_per_ray, 3, device=dev)
    dss = torch.empty(n, n_per_ray, device=dev)
    zs = torch.empty(n, n_per_ray, device=dev)
    ppe = proposal.position_encoder
    pw, ppd = proposal.weights()
    _lib.call("proposal_render_samples", [o, d, t0, t1, ppe.table.detach(), pw, occupancy_bitfield, start, cnt, fb,
                                          xyzs, dss, zs],
              _lib.proposal_render_descriptor(ppe.descriptor(0, proposal.bound),
                                              _lib.proposal_sample_descriptor(n, n_per_ray, mass_threshold, placement,
                                                                              1e-3, diagonal_n_steps, K, G, model.bound,
                                                                              stepsize_portion, empty_mass), ppd))
    drgbs = torch.empty(n, n_per_ray, 4, device=dev)
    pe = model.position_encoder
    half = getattr(model, "mlp_half", None) is not None
    params = model.mlp_half if half else model.mlp_params.detach()
    ident = torch.arange(n, dtype=torch.int32, device=dev).view(u32)
    bgs = torch.as_tensor(bg, dtype=torch.float32, device=dev).expand(n, 3).contiguous()
    rays_rgbd = torch.empty(n, 4, device=dev)
    op = torch.empty(n, device=dev)
    measured = torch.empty(1, dtype=u32, device=dev)
    idx = fb.nonzero().squeeze(1)
    # The proposal rays' field and compositing (side stream) and the fallback rays'
    # ray pool (this stream) touch disjoint rays: they run concurrently, and the
    # pool's host checks no longer leave the GPU idle.  Every buffer is allocated on
    # this stream before the fork and outlives the join.
    main = torch.cuda.current_stream(dev)
    side = _side_stream(dev)
    side.wait_stream(main)
    with torch.cuda.stream(side):
        _lib.call("nerf_field_chunks", [xyzs, cnt, ident, d, pe.table.detach(), params, drgbs],
                  _lib.field_chunks_descriptor(pe.descriptor(0, model.bound), n, n_per_ray, n,
                                               _lib.DTYPE_F16 if half else _lib.DTYPE_F32))
        _lib.call("integrate_rays", [start, cnt, bgs, dss.reshape(-1), zs.reshape(-1), drgbs.reshape(-1, 4),
                                     measured, rays_rgbd, op], _lib.integrating_descriptor(n, n * n_per_ray))
    rgbd_fb = None
    if idx.numel():
        rgbd_fb, _ = render_rays_inference(model, o[idx].contiguous(), d[idx].contiguous(), occupancy_bitfield, bg,
                                           K=K, G=G, diagonal_n_steps=diagonal_n_steps,
                                           stepsize_portion=stepsize_portion, march_steps_cap=fallback_steps_cap,
                                           pool=fallback_pool)
    main.wait_stream(side)
    if rgbd_fb is not None:
        rays_rgbd[idx] = rgbd_fb
    if stats is not None:
        stats["fallback_rays"] = int(idx.numel())
        stats["proposal_rays"] = int((cnt.view(torch.int32) > 0).sum().item())
    return _image(camera, rays_rgbd, bgs)
