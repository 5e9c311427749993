"""CPU checks of the synthetic scene: the float64 ground-truth renderer against
closed forms, camera poses, PSNR and blending helpers (utils/data.py:401-464)."""
import math

import numpy as np
import torch

from jaxngp_amd import scene as sc


def test_single_blob_through_centre_closed_form():
    """A ray through the centre of a lone blob: total optical depth A r sqrt(2 pi),
    colour = albedo (1 - T)."""
    blob = sc.BlobScene(centers=np.zeros((1, 3)), radii=np.array([0.15]), peaks=np.array([7.0]),
                        albedo=np.array([[0.2, 0.6, 0.9]]), bound=1.0)
    o = torch.tensor([[0.0, 0.0, 4.0], [0.0, 3.0, 0.0]], dtype=torch.float64)
    d = torch.tensor([[0.0, 0.0, -1.0], [0.0, -1.0, 0.0]], dtype=torch.float64)
    rgb, op = sc.render_gt(blob, o, d, n_steps=256)
    T = math.exp(-7.0 * 0.15 * math.sqrt(2 * math.pi))
    np.testing.assert_allclose(op.numpy(), 1 - T, rtol=1e-9)
    np.testing.assert_allclose(rgb.numpy(), np.outer([1, 1], [0.2, 0.6, 0.9]) * (1 - T), rtol=1e-9)


def test_offset_ray_and_mixture_against_quadrature():
    rng = np.random.default_rng(0)
    blobs = sc.BlobScene.random(seed=5, n_blobs=4)
    o = torch.tensor([[0.3, -4.0, 0.2]], dtype=torch.float64)
    dvec = np.array([0.05, 1.0, -0.03])
    dvec /= np.linalg.norm(dvec)
    d = torch.tensor([dvec], dtype=torch.float64)
    rgb, op = sc.render_gt(blobs, o, d, n_steps=512)
    # brute-force midpoint quadrature of the volume-rendering integral
    t0, t1 = sc.near_far(o, d, 1.0)
    t = np.linspace(float(t0), float(t1), 400001)
    tm = 0.5 * (t[1:] + t[:-1])
    x = o.numpy()[0] + tm[:, None] * dvec
    d2 = ((x[:, None] - blobs.centers[None]) ** 2).sum(-1)
    sk = blobs.peaks[None] * np.exp(-d2 / (2 * blobs.radii[None] ** 2))
    sig = sk.sum(1)
    dt = np.diff(t)
    tau = np.concatenate([[0], np.cumsum(sig * dt)])
    w = np.exp(-tau[:-1]) - np.exp(-tau[1:])
    col = (sk @ blobs.albedo) / np.maximum(sig, 1e-300)[:, None]
    np.testing.assert_allclose(op.numpy()[0], 1 - np.exp(-tau[-1]), rtol=1e-6)
    np.testing.assert_allclose(rgb.numpy()[0], (w[:, None] * col).sum(0), rtol=1e-4, atol=1e-6)


def test_poses_look_at_origin_and_are_rotations():
    P = sc.look_at_poses(20, radius=4.0, seed=1)
    for p in P:
        R, t = p[:, :3], p[:, 3]
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(R) - 1) < 1e-12
        np.testing.assert_allclose(np.linalg.norm(t), 4.0)
        # the viewing direction (-z) points at the origin
        np.testing.assert_allclose(-R[:, 2], -t / 4.0, atol=1e-12)


def test_centre_pixel_ray_hits_origin():
    cam = sc.Camera(width=800, height=800)
    P = sc.look_at_poses(3, seed=2)
    o, d = sc.pinhole_rays(cam, P[0])
    # pixel (400, 400) has its centre at +0.5: direction within half a pixel of the axis
    k = 400 * 800 + 400
    axis = -P[0][:, 2]
    assert float(torch.dot(d[k], torch.as_tensor(axis))) > math.cos(1.0 / 1111)


def test_psnr_and_blend():
    a = torch.zeros(10, 10, 3, dtype=torch.uint8)
    b = torch.full((10, 10, 3), 2, dtype=torch.uint8)
    assert abs(sc.psnr_u8(a, b) - 20 * math.log10(255 / 2)) < 1e-9
    assert sc.psnr_u8(a, a) == 100.0
    rgba = torch.tensor([[[255, 0, 0, 128]]], dtype=torch.uint8)
    out = sc.blend_u8(rgba, (0.0, 0.0, 1.0))
    assert out.tolist() == [[[128, 0, 127]]]
