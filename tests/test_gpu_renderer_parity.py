"""Renderer parity against the oracle's chain (models/renderers/cuda.py).

Inference (render_image_inference's ray pool, cuda.py:244-373): a trained field's
48x48 frame rendered by the HIP pool (fused encoder + MLP per round) against the
oracle's pool loop over march_rays_inference / integrate_rays_inference with the
oracle's field (the C encoder restatement and the float64 MLP, rounded to f32).
The marches are the same (they read only the bitfield); the composited colours
differ only by the MLP's f32 rounding, except where a ray's transmittance crosses
1e-4 one sample apart, which moves its colour by at most the remaining 1e-4.

Training (render_rays_train, cuda.py:100-162): march_rays -> NeRF -> integrate_rays
and the backward through all three against the oracle's march, the float64
encoder + MLP, the f32 integrator and its VJP, the float64 MLP backward and the
float64 table scatter.
"""
import numpy as np
import pytest
import torch

import oracle
from jaxngp_amd import scene as sc
from jaxngp_amd.renderer import make_near_far_from_bound, make_rays_worldspace, render_rays_inference, render_rays_train
from jaxngp_amd.train import TrainConfig, Trainer
from tests.test_gpu_hashgrid import _c_forward

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(dev):
    cam = sc.Camera(width=96, height=96, fx=133.0, fy=133.0, cx=48.0, cy=48.0)
    data = sc.make_views(sc.BlobScene.random(seed=3), cam, sc.look_at_poses(16, seed=1), dev, n_steps=256)
    tr = Trainer(data, TrainConfig(total_samples=1 << 16))
    for _ in range(300):
        tr.step()
    torch.cuda.synchronize()
    return tr


def _oracle_field(model, xyz, dirs):
    """drgbs f32 of the oracle's field: the C encoder (bit-exact with the kernels)
    and the float64 MLP."""
    table = model.position_encoder.table.detach().cpu().numpy()
    enc = _c_forward(model.position_encoder, xyz, table, model.bound)
    drgbs, _ = oracle.mlp_forward64(enc, dirs, model.mlp_params.detach().cpu().numpy())
    return drgbs.astype(np.float32)


def _oracle_pool_render(model, o, d, t0, t1, bits, cap, pool, K, portion):
    n_tot = o.shape[0]
    rgbd_all = np.zeros((n_tot, 4), np.float32)
    T_all = np.ones(n_tot, np.float32)
    bg = np.ones((n_tot, 3), np.float32)
    ts = t0.copy()
    terminated = np.ones(pool, bool)
    indices = np.zeros(pool, np.uint32)
    nri, done, samples = np.uint32(0), 0, np.zeros(n_tot, np.int64)
    for _ in range(100000):
        out = oracle.march_rays_inference(1024, K, 128, cap, model.bound, portion, o, d, ts, t1, bits, nri,
                                          terminated, indices)
        nri, indices = out["next_ray_index"], out["indices"]
        live = indices < n_tot
        r = indices[live]
        ts[r] = out["t_starts_out"][live]
        k = out["n_samples"]
        samples[r] += k[live]
        mask = np.arange(cap)[None, :] < k[:, None]
        mask &= live[:, None]
        drgbs = np.zeros((pool, cap, 4), np.float32)
        if mask.any():
            sl, st = np.nonzero(mask)
            drgbs[sl, st] = _oracle_field(model, out["xyzs"][sl, st], d[indices[sl]])
        cnt, terminated, rgbd, T = oracle.integrate_rays_inference(bg, rgbd_all, T_all, k, indices, out["dss"],
                                                                   out["z_vals"], drgbs)
        rgbd_all[r] = rgbd[live]
        T_all[r] = T[live]
        done += int(cnt)
        if done >= n_tot:
            return rgbd_all, T_all, samples
    raise AssertionError("pool loop did not finish")


def test_inference_frame_matches_oracle_pool(dev, trained):
    tr = trained
    model = tr.model
    cam = sc.Camera(width=48, height=48, fx=66.5, fy=66.5, cx=24.0, cy=24.0)
    pose = torch.as_tensor(sc.look_at_poses(1, seed=7, golden=False)[0], dtype=torch.float32, device=dev)
    o, d = make_rays_worldspace(cam, pose)
    t0, t1 = make_near_far_from_bound(model.bound, o, d)
    bits = tr.grid.bitfield
    print(f"inference parity: {cam.n_pixels} rays, pool 512, cap 8", flush=True)
    got, _ = render_rays_inference(model, o, d, bits, K=tr.K, G=128, diagonal_n_steps=1024,
                                   stepsize_portion=tr.portion, march_steps_cap=8, pool=512)
    got = got.cpu().numpy()
    want, T_end, samples = _oracle_pool_render(model, o.cpu().numpy(), d.cpu().numpy(), t0.cpu().numpy(),
                                               t1.cpu().numpy(), bits.cpu().numpy(), 8, 512, tr.K, tr.portion)
    hit = samples > 0
    # a frame with surfaces, some rays stopping early (T <= 1e-4) and many not
    assert hit.mean() > 0.2 and (T_end[hit] <= 1e-4).sum() > 50 and (T_end[hit] > 1e-3).sum() > 50
    err = np.abs(got - want)
    # colours: f32 rounding of the field, or one sample more or less past T = 1e-4
    assert (err[:, :3] <= 3e-4).all(), err[:, :3].max()
    # depth: the same, times z (at most about 5 here)
    assert (err[:, 3] <= 2e-3).all(), err[:, 3].max()
    # rays far from the stop threshold agree to the field's rounding
    clear = T_end > 1e-3
    assert (err[clear, :3] <= 2e-5).all(), err[clear, :3].max()
    diff_u8 = sc.f32_to_u8(torch.from_numpy(got[:, :3])) != sc.f32_to_u8(torch.from_numpy(want[:, :3]))
    assert diff_u8.float().mean() < 1e-2


def test_render_rays_train_matches_oracle_chain(dev, trained):
    tr = trained
    model = tr.model
    rng = np.random.default_rng(12)
    n = 1024
    S = 1 << 16
    pose = torch.as_tensor(sc.look_at_poses(1, seed=9, golden=False)[0], dtype=torch.float32, device=dev)
    cam = sc.Camera(width=32, height=32, fx=44.0, fy=44.0, cx=16.0, cy=16.0)
    o, d = make_rays_worldspace(cam, pose)
    bg = torch.from_numpy(rng.uniform(0, 1, (n, 3)).astype(np.float32)).to(dev)
    bits = tr.grid.bitfield
    model.zero_grad(set_to_none=True)
    print(f"training-renderer parity: {n} rays, {S} sample slots", flush=True)
    metrics, rgbd, tv = render_rays_train(o, d, bg, S, model, bits, K=tr.K, G=128, diagonal_n_steps=1024,
                                          stepsize_portion=tr.portion, near=0.3, perturb=False)
    g = rng.normal(size=(n, 4)).astype(np.float32)
    (rgbd * torch.from_numpy(g).to(dev)).sum().backward()
    dtable = model.position_encoder.table.grad.detach().cpu().numpy().astype(np.float64)
    dparams = model.mlp_params.grad.detach().cpu().numpy().astype(np.float64)
    # the oracle's chain
    O, D = o.cpu().numpy(), d.cpu().numpy()
    t0, t1 = make_near_far_from_bound(model.bound, o, d)
    m = oracle.march_rays(S, 1024, tr.K, 128, model.bound, tr.portion, O, D, t0.cpu().numpy(), t1.cpu().numpy(),
                          np.zeros(n, np.float32), bits.cpu().numpy())
    assert int(metrics["measured_batch_size_before_compaction"].view(torch.int32)) == int(m["measured"])
    np.testing.assert_array_equal(metrics["ray_is_valid"].cpu().numpy(), m["ray_is_valid"])
    M = int(m["measured"])
    assert M > 10000 and int(metrics["n_valid_rays"]) == int(m["ray_is_valid"].sum()) > 100
    xyz, dirs = m["xyzs"][:M], m["dirs"][:M]
    table = model.position_encoder.table.detach().cpu().numpy()
    params = model.mlp_params.detach().cpu().numpy().astype(np.float64)
    lv = oracle.hashgrid_levels(16, 1 << 19, 16, 2048)
    enc64 = oracle.hashgrid_forward64(xyz, table, model.bound, lv, 1 << 19)
    drgbs64, _ = oracle.mlp_forward64(enc64, dirs, params)
    drgbs = np.zeros((S, 4), np.float32)
    drgbs[:M] = drgbs64
    bgs = bg.cpu().numpy()
    cnt, rgbd_o, op_o = oracle.integrate_rays(m["rays_sample_startidx"], m["rays_n_samples"], bgs, m["dss"],
                                              m["z_vals"], drgbs)
    err = np.abs(rgbd.detach().cpu().numpy() - rgbd_o)
    assert (err[:, :3] <= 3e-4).all() and (err[:, 3] <= 2e-3).all(), err.max(0)
    assert int(metrics["measured_batch_size"].view(torch.int32)) == pytest.approx(int(cnt), abs=n // 100)
    _, _, dd = oracle.integrate_rays_backward(0.3, m["rays_sample_startidx"], m["rays_n_samples"], bgs, m["dss"],
                                              m["z_vals"], drgbs, rgbd_o, op_o, g)
    dX, dP = oracle.mlp_backward64(enc64, dirs, params, dd[:M].astype(np.float64))
    dT = oracle.hashgrid_backward64(xyz, dX, model.bound, lv, 1 << 19, 2)
    assert (np.abs(dparams - dP) <= 2e-3 * np.abs(dP).max()).all(), np.abs(dparams - dP).max() / np.abs(dP).max()
    absT = oracle.hashgrid_backward64(xyz, np.abs(dX), model.bound, lv, 1 << 19, 2)
    assert (np.abs(dtable - dT) <= 2e-3 * absT + 1e-3 * np.abs(dT).max()).all()
    assert np.abs(dT).max() > 0
