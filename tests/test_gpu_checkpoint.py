"""A trained field saved in the reference's checkpoint format and restored into a
fresh Trainer renders the same image bit for bit, and training goes on from the
restored step (FP32 and the f16 path)."""
import pytest
import torch

from jaxngp_amd import scene as sc
from jaxngp_amd.checkpoint import load_checkpoint, restore_trainer, save_checkpoint
from jaxngp_amd.renderer import render_image_inference
from jaxngp_amd.train import TrainConfig, Trainer

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_checkpoint_round_trip_renders_identically(tmp_path, precision):
    dev = torch.device("cuda:0")
    cam = sc.Camera(width=64, height=64, fx=90.0, fy=90.0, cx=32.0, cy=32.0)
    blobs = sc.BlobScene.random(seed=3)
    data = sc.make_views(blobs, cam, sc.look_at_poses(8, seed=1), dev, n_steps=128)
    cfg = TrainConfig(total_samples=1 << 15, precision=precision)
    tr = Trainer(data, cfg)
    for _ in range(60):
        tr.step()
    path = save_checkpoint(str(tmp_path), tr)
    assert path.endswith("checkpoint_60")
    assert load_checkpoint(str(tmp_path))["step"] == 60
    tr2 = Trainer(data, cfg)
    assert restore_trainer(tr2, str(tmp_path)) == 60
    pose = torch.as_tensor(sc.look_at_poses(1, seed=7, golden=False)[0], dtype=torch.float32, device=dev)
    kw = dict(K=tr.K, G=128, diagonal_n_steps=1024, stepsize_portion=tr.portion, march_steps_cap=8)
    _, a, da = render_image_inference(tr.model, cam, pose, tr.grid.bitfield, **kw)
    _, b, db = render_image_inference(tr2.model, cam, pose, tr2.grid.bitfield, **kw)
    assert torch.equal(a, b) and torch.equal(da, db)
    assert torch.equal(tr.grid.bitfield, tr2.grid.bitfield)
    for _ in range(5):
        tr2.step()
    assert tr2.step_count == 65 and torch.isfinite(torch.tensor(tr2.metrics()["loss"]))
