"""Synthetic procedural scene for training and evaluation.

Geometry: Gaussian density blobs in [-1, 1]^3,
    sigma(x) = sum_k s_k exp(-|x - c_k|^2 / (2 r_k^2)),
peak s_k ~ U(20, 200), radius r_k ~ U(0.1, 0.4), centre c_k ~ U(-0.6, 0.6)^3,
albedo ~ U[0,1]^3; the colour at x is the density-weighted albedo.  Cameras sit on
a sphere of radius 4 looking at the origin (pinhole, focal 1111 px, 800x800).

Ground truth is rendered in float64 inside the AABB [-bound, bound]^3 (the region
the field models): along a unit ray, each blob's density is a 1-D Gaussian in t,
so the optical depth has a closed form in erf and the transmittance T(t) is exact
at every step boundary; each step contributes (T_i - T_{i+1}) times the
density-weighted albedo at its midpoint.  Views are stored as RGBA u8 with
straight (non-premultiplied) colour and alpha = 1 - T(far), like NeRF-synthetic
PNGs, so training blends them with the background as utils/data.py:443-464 does.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch


@dataclass
class Camera:
    width: int = 800
    height: int = 800
    fx: float = 1111.0
    fy: float = 1111.0
    cx: float = 400.0
    cy: float = 400.0
    near: float = 0.3

    @property
    def n_pixels(self) -> int:
        return self.width * self.height


@dataclass
class BlobScene:
    centers: np.ndarray  # [K, 3]
    radii: np.ndarray  # [K]
    peaks: np.ndarray  # [K]
    albedo: np.ndarray  # [K, 3]
    bound: float = 1.0

    @classmethod
    def random(cls, seed: int = 3, n_blobs: int = 8, bound: float = 1.0) -> "BlobScene":
        rng = np.random.default_rng(seed)
        return cls(centers=rng.uniform(-0.6, 0.6, (n_blobs, 3)), radii=rng.uniform(0.1, 0.4, n_blobs),
                   peaks=rng.uniform(20.0, 200.0, n_blobs), albedo=rng.uniform(0.0, 1.0, (n_blobs, 3)), bound=bound)

    def density(self, x: np.ndarray) -> np.ndarray:
        d2 = ((x[:, None, :] - self.centers[None]) ** 2).sum(-1)
        return (self.peaks[None] * np.exp(-d2 / (2 * self.radii[None] ** 2))).sum(-1)


def look_at_poses(n: int, radius: float = 4.0, seed: int = 0, golden: bool = True) -> np.ndarray:
    """Camera-to-world [n, 3, 4] (rotation columns = camera x, y, z axes in world;
    the camera looks along its -z axis, CG convention) on a sphere of `radius`."""
    rng = np.random.default_rng(seed)
    if golden:
        k = np.arange(n) + 0.5
        zc = 1 - 2 * k / n
        phi = math.pi * (1 + 5 ** 0.5) * k + rng.uniform(0, 2 * math.pi)
        dirs = np.stack([np.sqrt(1 - zc ** 2) * np.cos(phi), np.sqrt(1 - zc ** 2) * np.sin(phi), zc], 1)
    else:
        dirs = rng.normal(size=(n, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    poses = np.zeros((n, 3, 4))
    for i, dvec in enumerate(dirs):
        eye = dvec * radius
        zaxis = dvec / np.linalg.norm(dvec)  # camera looks along -z, i.e. towards the origin
        up = np.array([0.0, 0.0, 1.0]) if abs(zaxis[2]) < 0.99 else np.array([0.0, 1.0, 0.0])
        xaxis = np.cross(up, zaxis)
        xaxis /= np.linalg.norm(xaxis)
        yaxis = np.cross(zaxis, xaxis)
        poses[i, :, 0], poses[i, :, 1], poses[i, :, 2], poses[i, :, 3] = xaxis, yaxis, zaxis, eye
    return poses


def pinhole_rays(cam: Camera, pose: np.ndarray, dtype=torch.float64, device="cpu"):
    """All pixel rays of one view (utils/types.py:400-439 + models/renderers/cuda.py:21-53)."""
    idx = torch.arange(cam.n_pixels, device=device)
    x = (idx % cam.width).to(dtype)
    y = torch.div(idx, cam.width, rounding_mode="floor").to(dtype)
    dx = ((x + 0.5) - cam.cx) / cam.fx
    dy = ((y + 0.5) - cam.cy) / cam.fy
    d = torch.stack([dx, -dy, -torch.ones_like(dx)], -1)
    d = d / torch.linalg.norm(d, dim=-1, keepdim=True)
    R = torch.as_tensor(pose[:, :3], dtype=dtype, device=device)
    t = torch.as_tensor(pose[:, 3], dtype=dtype, device=device)
    return t.expand(cam.n_pixels, 3), d @ R.T


def near_far(o: torch.Tensor, d: torch.Tensor, bound: float):
    """AABB entry/exit times (models/renderers/cuda.py:56-97)."""
    eps = 1e-15
    d = torch.where(torch.signbit(d), torch.clamp(d, max=-eps), torch.clamp(d, min=eps))
    t0 = (-bound - o) / d
    t1 = (bound - o) / d
    ts = torch.minimum(t0, t1).amax(-1)
    te = torch.maximum(t0, t1).amin(-1)
    return torch.clamp(ts, min=0.0), te


def render_gt(scene: BlobScene, o: torch.Tensor, d: torch.Tensor, n_steps: int = 512, chunk: int = 1 << 14):
    """float64 (premultiplied rgb [N,3], opacity [N]) of rays o + t d inside the AABB."""
    dev = o.device
    o = o.to(torch.float64)
    d = d.to(torch.float64)
    C = torch.as_tensor(scene.centers, dtype=torch.float64, device=dev)
    Rr = torch.as_tensor(scene.radii, dtype=torch.float64, device=dev)
    A = torch.as_tensor(scene.peaks, dtype=torch.float64, device=dev)
    alb = torch.as_tensor(scene.albedo, dtype=torch.float64, device=dev)
    rgb_out = torch.zeros(o.shape[0], 3, dtype=torch.float64, device=dev)
    op_out = torch.zeros(o.shape[0], dtype=torch.float64, device=dev)
    s2 = math.sqrt(2.0)
    for a in range(0, o.shape[0], chunk):
        oo, dd = o[a:a + chunk], d[a:a + chunk]
        t0, t1 = near_far(oo, dd, scene.bound)
        hit = t1 > t0
        t1 = torch.where(hit, t1, t0)
        oc = C[None] - oo[:, None]  # [n, K, 3]
        tc = (oc * dd[:, None]).sum(-1)  # closest-approach t per blob
        h2 = (oc * oc).sum(-1) - tc * tc
        amp = A[None] * torch.exp(-torch.clamp(h2, min=0.0) / (2 * Rr[None] ** 2))  # [n, K]
        u = torch.linspace(0.0, 1.0, n_steps + 1, dtype=torch.float64, device=dev)
        tb = t0[:, None] + (t1 - t0)[:, None] * u[None]  # [n, S+1] step boundaries
        # optical depth of blob k from t0 to t: amp r sqrt(pi/2) [erf((t-tc)/(r sqrt2)) - erf((t0-tc)/(r sqrt2))]
        z = (tb[:, :, None] - tc[:, None, :]) / (Rr[None, None] * s2)
        er = torch.special.erf(z)
        tau = ((er - er[:, :1]) * (amp * Rr[None] * math.sqrt(math.pi / 2))[:, None, :]).sum(-1)  # [n, S+1]
        T = torch.exp(-tau)
        w = T[:, :-1] - T[:, 1:]  # [n, S]
        tm = 0.5 * (tb[:, :-1] + tb[:, 1:])
        sig_k = amp[:, None, :] * torch.exp(-((tm[:, :, None] - tc[:, None, :]) ** 2) / (2 * Rr[None, None] ** 2))
        col = (sig_k @ alb) / torch.clamp(sig_k.sum(-1, keepdim=True), min=1e-300)
        rgb_out[a:a + chunk] = (w[:, :, None] * col).sum(1)
        op_out[a:a + chunk] = torch.where(hit, 1.0 - T[:, -1], torch.zeros_like(T[:, -1]))
    return rgb_out, op_out


def f32_to_u8(x: torch.Tensor) -> torch.Tensor:
    """utils/data.py:41-43."""
    return torch.clamp(torch.round(x * 255), 0, 255).to(torch.uint8)


def psnr_u8(a: torch.Tensor, b: torch.Tensor) -> float:
    """utils/data.py:406-410: clip(20 log10(255 / sqrt(mse + 1e-15)), 0, 100) on u8 images."""
    mse = ((a.double() - b.double()) ** 2).mean().item()
    return float(min(max(20 * math.log10(255 / math.sqrt(mse + 1e-15)), 0.0), 100.0))


def blend_u8(rgba: torch.Tensor, bg) -> torch.Tensor:
    """utils/data.py:443-464 for u8 input: blend in float, re-quantise."""
    rgb = rgba[..., :3].double() / 255
    a = rgba[..., 3:].double() / 255
    bg = torch.as_tensor(bg, dtype=torch.float64, device=rgba.device)
    return f32_to_u8(rgb * a + bg * (1 - a))


@dataclass
class SceneData:
    """Training / evaluation views of a BlobScene on one device."""
    scene: BlobScene
    camera: Camera
    poses: np.ndarray  # [V, 3, 4] camera-to-world
    rgba: torch.Tensor  # u8 [V, H, W, 4]
    transforms: torch.Tensor  # f32 [V, 12] (R row-major, t)

    @property
    def n_views(self) -> int:
        return self.poses.shape[0]


def make_views(scene: BlobScene, cam: Camera, poses: np.ndarray, device, n_steps: int = 512,
               progress=None) -> SceneData:
    V = poses.shape[0]
    rgba = torch.empty(V, cam.height, cam.width, 4, dtype=torch.uint8, device=device)
    for v in range(V):
        if progress is not None and v % 10 == 0:
            progress(v)
        o, d = pinhole_rays(cam, poses[v], device=device)
        rgb, op = render_gt(scene, o, d, n_steps=n_steps)
        straight = rgb / torch.clamp(op, min=1e-12)[:, None]
        img = torch.cat([torch.clamp(straight, 0, 1), op[:, None]], -1)
        rgba[v] = f32_to_u8(img).reshape(cam.height, cam.width, 4)
    tr = np.concatenate([poses[:, :, :3].reshape(V, 9), poses[:, :, 3]], 1)
    return SceneData(scene, cam, poses, rgba, torch.as_tensor(tr, dtype=torch.float32, device=device).contiguous())
