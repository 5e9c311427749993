"""GPU checks of the inference renderer (models/renderers/cuda.py:244-373): the frame
does not depend on the ray-pool size (each ray is marched in the same chunks of
march_steps_cap steps whichever slot holds it), and a trained field renders the
held-out view close to the float64 ground truth."""
import torch
import pytest

from jaxngp_amd import scene as sc
from jaxngp_amd.renderer import render_image_inference
from jaxngp_amd.train import TrainConfig, Trainer

pytestmark = pytest.mark.gpu


def test_render_pool_invariance_and_quality():
    dev = torch.device("cuda:0")
    cam = sc.Camera(width=96, height=96, fx=133.0, fy=133.0, cx=48.0, cy=48.0)
    blobs = sc.BlobScene.random(seed=3)
    data = sc.make_views(blobs, cam, sc.look_at_poses(16, seed=1), dev, n_steps=256)
    tr = Trainer(data, TrainConfig(total_samples=1 << 16))
    for _ in range(400):
        tr.step()
    poses = sc.look_at_poses(2, seed=5, golden=False)
    gt = sc.make_views(blobs, cam, poses, dev, n_steps=256)
    kw = dict(K=tr.K, G=128, diagonal_n_steps=1024, stepsize_portion=tr.portion, march_steps_cap=8)
    for i in range(2):
        pose = torch.as_tensor(poses[i], dtype=torch.float32, device=dev)
        _, a, da = render_image_inference(tr.model, cam, pose, tr.grid.bitfield, pool=300, **kw)
        _, b, db = render_image_inference(tr.model, cam, pose, tr.grid.bitfield, pool=1 << 14, **kw)
        assert torch.equal(a, b) and torch.equal(da, db)
        psnr = sc.psnr_u8(a, sc.blend_u8(gt.rgba[i], (1.0, 1.0, 1.0)))
        print(f"view {i}: PSNR {psnr:.2f} dB")
        assert psnr > 22.0
