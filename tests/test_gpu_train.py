"""GPU tests of the training harness: the fused ray generation against a torch
restatement, the loss gradient, Adam against the optax update rule, and that a few
hundred steps on the synthetic scene reduce the loss."""
import math

import numpy as np
import pytest
import torch

from jaxngp_amd import _lib, scene as sc
from jaxngp_amd.train import TrainConfig, Trainer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_data():
    dev = torch.device("cuda:0")
    cam = sc.Camera(width=100, height=100, fx=139.0, fy=139.0, cx=50.0, cy=50.0)
    blobs = sc.BlobScene.random(seed=3)
    poses = sc.look_at_poses(12, seed=1)
    return sc.make_views(blobs, cam, poses, dev, n_steps=256)


def test_make_rays_matches_torch(small_data):
    d = small_data
    dev = d.rgba.device
    n = 4096
    out = {k: torch.empty(n, w, device=dev) for k, w in [("o", 3), ("d", 3), ("bgs", 3), ("gt", 3)]}
    t0, t1, noise = (torch.empty(n, device=dev) for _ in range(3))
    pix = torch.empty(n, dtype=torch.int32, device=dev).view(torch.uint32)
    cam = d.camera
    _lib.call("train_make_rays", [d.transforms, d.rgba, out["o"], out["d"], t0, t1, noise, out["bgs"], out["gt"], pix],
              _lib.train_rays_descriptor(n, d.n_views, cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy, 1.0,
                                         123, 7, _lib.TRAIN_PERTURB | _lib.TRAIN_RANDOM_BG, (1.0, 1.0, 1.0)))
    p = pix.view(torch.int32).long()
    view, px = p // cam.n_pixels, p % cam.n_pixels
    assert (view < d.n_views).all()
    from jaxngp_amd.renderer import make_near_far_from_bound, make_rays_worldspace
    for v in range(d.n_views):
        sel = view == v
        if not sel.any():
            continue
        o_all, d_all = make_rays_worldspace(cam, torch.as_tensor(d.poses[v], dtype=torch.float32, device=dev))
        torch.testing.assert_close(out["o"][sel], o_all[px[sel]], rtol=0, atol=1e-6)
        torch.testing.assert_close(out["d"][sel], d_all[px[sel]], rtol=0, atol=2e-6)
    ts, te = make_near_far_from_bound(1.0, out["o"], out["d"])
    torch.testing.assert_close(t0, ts, rtol=1e-5, atol=1e-5)
    torch.testing.assert_close(t1, te, rtol=1e-5, atol=1e-5)
    assert ((noise >= 0) & (noise < 1)).all() and ((out["bgs"] >= 0) & (out["bgs"] < 1)).all()
    rgba = d.rgba.reshape(-1, 4)[p].float() / 255
    want = rgba[:, :3] * rgba[:, 3:] + out["bgs"] * (1 - rgba[:, 3:])
    torch.testing.assert_close(out["gt"], want, rtol=1e-6, atol=1e-6)
    # uniform over pixels and views
    assert abs(float(noise.mean()) - 0.5) < 0.03 and abs(float(view.float().mean()) - (d.n_views - 1) / 2) < 0.5


def test_loss_grad_and_adam(small_data):
    dev = small_data.rgba.device
    rng = np.random.default_rng(0)
    n = 1000
    rgbd = torch.from_numpy(rng.uniform(0, 1, (n, 4)).astype(np.float32)).to(dev)
    gt = torch.from_numpy(rng.uniform(0, 1, (n, 3)).astype(np.float32)).to(dev)
    valid = torch.from_numpy(rng.uniform(size=n) < 0.7).to(dev)
    nv = torch.empty(1, dtype=torch.int32, device=dev).view(torch.uint32)
    g = torch.empty(n, 4, device=dev)
    loss = torch.empty(1, device=dev)
    _lib.call("train_loss_grad", [rgbd, gt, valid, nv, g, loss], _lib.loss_descriptor(n, 0.1))
    pred = rgbd[:, :3].clone().requires_grad_(True)
    err = pred - gt
    hub = torch.where(err.abs() <= 0.1, 0.5 * err ** 2, 0.1 * (err.abs() - 0.05))
    ref = torch.where(valid, hub.mean(-1), torch.zeros(n, device=dev)).sum() / valid.sum()
    ref.backward()
    assert int(nv.view(torch.int32)[0]) == int(valid.sum())
    torch.testing.assert_close(loss[0], ref.detach(), rtol=1e-5, atol=1e-7)
    torch.testing.assert_close(g[:, :3], pred.grad, rtol=1e-5, atol=1e-9)
    assert (g[:, 3] == 0).all()
    # Adam: optax scale_by_adam + scale(-lr) + add_decayed_weights
    N = 5000
    p = torch.from_numpy(rng.normal(size=N).astype(np.float32)).to(dev)
    m = torch.zeros(N, device=dev)
    v = torch.zeros(N, device=dev)
    p_ref = p.double().clone()
    m_ref = torch.zeros(N, dtype=torch.float64, device=dev)
    v_ref = torch.zeros(N, dtype=torch.float64, device=dev)
    for t in range(1, 6):
        gr = torch.from_numpy(rng.normal(size=N).astype(np.float32)).to(dev)
        gr[::7] = 0
        _lib.call("adam_update", [p, gr, m, v], _lib.adam_descriptor(N, 1e-2, 0.9, 0.99, 1e-15, 1e-15, 0.9 ** t,
                                                                     0.99 ** t, 1e-6))
        m_ref = 0.9 * m_ref + 0.1 * gr.double()
        v_ref = 0.99 * v_ref + 0.01 * gr.double() ** 2
        u = -1e-2 * (m_ref / (1 - 0.9 ** t)) / (torch.sqrt(v_ref / (1 - 0.99 ** t) + 1e-15) + 1e-15) + 1e-6 * p_ref
        p_ref = p_ref + u
        torch.testing.assert_close(p.double(), p_ref, rtol=1e-5, atol=1e-6)


def test_training_reduces_loss(small_data):
    cfg = TrainConfig(total_samples=1 << 16)
    tr = Trainer(small_data, cfg)
    losses = []
    for i in range(300):
        tr.step()
        if i % 50 == 0 or i == 299:
            m = tr.metrics()
            losses.append(m["loss"])
            assert m["n_valid_rays"] > 0 and m["measured_batch_size_before_compaction"] > 0
            assert math.isfinite(m["loss"])
    assert losses[-1] < 0.5 * losses[0], losses
    # the occupancy grid was thresholded: some cells empty, some occupied
    occ = tr.grid.occ_mask.float().mean().item()
    assert 0.0 < occ < 1.0


def _state(dev, scale):
    return _lib.loss_scale_state(dev, scale)


def _scale_of(state):
    return float(state[:4].view(torch.float32)[0].item()), int(state[4:8].view(torch.int32)[0].item()), \
        int(state[8:12].view(torch.int32)[0].item())


def test_loss_scaling_targets(small_data):
    """Scaled loss gradient = scale x unscaled; the check flags a non-finite gradient;
    Adam with the state undoes the scale, or changes nothing after an overflow; the
    update halves after an overflow and doubles after growth_interval finite steps."""
    dev = small_data.rgba.device
    rng = np.random.default_rng(1)
    n = 777
    rgbd = torch.from_numpy(rng.uniform(0, 1, (n, 4)).astype(np.float32)).to(dev)
    gt = torch.from_numpy(rng.uniform(0, 1, (n, 3)).astype(np.float32)).to(dev)
    valid = torch.from_numpy(rng.uniform(size=n) < 0.8).to(dev)
    nv = torch.empty(1, dtype=torch.int32, device=dev).view(torch.uint32)
    loss = torch.empty(1, device=dev)
    g1, g2 = torch.empty(n, 4, device=dev), torch.empty(n, 4, device=dev)
    st = _state(dev, 1024.0)
    _lib.call("train_loss_grad", [rgbd, gt, valid, nv, g1, loss], _lib.loss_descriptor(n, 0.1))
    l1 = loss.clone()
    _lib.call("train_loss_grad", [rgbd, gt, valid, nv, g2, loss, st], _lib.loss_descriptor(n, 0.1, _lib.LOSS_SCALED))
    torch.testing.assert_close(g2, g1 * 1024.0, rtol=0, atol=0)
    torch.testing.assert_close(loss, l1, rtol=1e-6, atol=0)  # summed with atomics in any order
    # Adam: scaled gradient + state == unscaled gradient without it
    N = 3000
    p0 = torch.from_numpy(rng.normal(size=N).astype(np.float32)).to(dev)
    gr = torch.from_numpy(rng.normal(size=N).astype(np.float32)).to(dev)
    none = torch.empty(0, device=dev)
    pa, ma, va = p0.clone(), torch.zeros(N, device=dev), torch.zeros(N, device=dev)
    pb, mb, vb = p0.clone(), torch.zeros(N, device=dev), torch.zeros(N, device=dev)
    desc = lambda f: _lib.adam_descriptor(N, 1e-2, 0.9, 0.99, 1e-15, 1e-15, 0.9, 0.99, 0.0, flags=f)
    _lib.call("adam_update", [pa, gr, ma, va], desc(0))
    _lib.call("adam_update", [pb, gr * 1024.0, mb, vb, none, st], desc(_lib.LOSS_SCALED))
    torch.testing.assert_close(pb, pa, rtol=0, atol=0)
    # check: finite gradients leave the flag clear, one inf sets it
    _lib.call("loss_scale_check", [gr, st], _lib.loss_scale_check_descriptor(N))
    assert _scale_of(st)[1] == 0
    bad = gr.clone()
    bad[N - 5] = float("inf")
    _lib.call("loss_scale_check", [bad, st], _lib.loss_scale_check_descriptor(N))
    assert _scale_of(st)[1] == 1
    before = pb.clone()
    _lib.call("adam_update", [pb, bad, mb, vb, none, st], desc(_lib.LOSS_SCALED))
    torch.testing.assert_close(pb, before, rtol=0, atol=0)  # skipped
    _lib.call("loss_scale_update", [st], _lib.loss_scale_update_descriptor(3))
    assert _scale_of(st) == (512.0, 0, 0)
    for k in range(3):
        _lib.call("loss_scale_update", [st], _lib.loss_scale_update_descriptor(3))
    assert _scale_of(st) == (1024.0, 0, 0)


def test_fp16_training_path(small_data):
    """f16 table + f16 MLP + dynamic loss scaling trains the scene like f32: loss
    falls by half, and the final loss is within 1.5x of the f32 run's."""
    finals = {}
    for prec in ("fp32", "fp16"):
        cfg = TrainConfig(total_samples=1 << 16, precision=prec)
        tr = Trainer(small_data, cfg)
        first = None
        for i in range(300):
            tr.step()
            if i == 0:
                first = tr.metrics()["loss"]
        m = tr.metrics()
        assert math.isfinite(m["loss"]) and m["loss"] < 0.5 * first, (prec, first, m)
        finals[prec] = m["loss"]
        if prec == "fp16":
            assert tr.model.position_encoder.table.dtype == torch.float16
            assert 1.0 <= m["loss_scale"] <= 2.0 ** 24
    assert finals["fp16"] < 1.5 * finals["fp32"], finals


def test_proposal_guided_training(small_data):
    """TrainConfig(proposal_guided=True): the field trains on the proposal's
    stratified samples (proposal_train_samples) and the proposal on the field's
    weights at them.  The loss falls by half within 1.5x of the marcher-trained
    run's and the steps fill their sample slots."""
    finals = {}
    for guided in (False, True):
        cfg = TrainConfig(total_samples=1 << 16, precision="fp16", proposal=True, proposal_guided=guided)
        tr = Trainer(small_data, cfg)
        first = None
        for i in range(300):
            tr.step()
            if i == 0:
                first = tr.metrics()
        m = tr.metrics()
        assert math.isfinite(m["loss"]) and m["loss"] < 0.5 * first["loss"], (guided, first, m)
        finals[guided] = m["loss"]
        if guided:
            assert tr.n_rays == (1 << 16) // 32
            assert m["measured_batch_size_before_compaction"] > 0.25 * (1 << 16), m
            assert m["n_valid_rays"] > 0 and math.isfinite(m["proposal_loss"])
    assert finals[True] < 1.5 * finals[False], finals


def test_background_model_training(tmp_path):
    """TrainConfig(background_model=True): the reference's loss_fn with a sky-sphere
    background (app/nerf/_utils.py:128-150).  Within a step: the per-ray backgrounds
    are the model's colours (float64 oracle), the ground truth is blended over them
    and the background gradient is dL/dbgs - dL/dpred (1 - alpha) (both bit for bit
    against float32 torch), and the parameter gradient is the oracle's VJP of that
    gradient at the pre-step parameters.  Training lowers the loss and moves the
    background parameters."""
    import oracle
    dev = torch.device("cuda:0")
    cam = sc.Camera(width=64, height=64, fx=89.0, fy=89.0, cx=32.0, cy=32.0)
    blobs = sc.BlobScene.random(seed=3)
    # cameras inside the sky sphere (radius 4 bound), as in the unbounded scenes it is for
    data = sc.make_views(blobs, cam, sc.look_at_poses(12, radius=3.0, seed=1), dev, n_steps=256)
    tr = Trainer(data, TrainConfig(total_samples=1 << 15, background_model=True))
    p0 = tr.bg_model.params.detach().clone()
    tr.step()
    torch.cuda.synchronize()
    b, bb = tr.b, tr.bb
    o, d = b["o"].cpu().numpy(), b["d"].cpu().numpy()
    none = np.zeros((o.shape[0], 0))
    want_bg = oracle.sky_sphere_bg64(o, d, none, p0.cpu().numpy(), 4.0, 0)
    assert np.abs(b["bgs"].cpu().numpy() - want_bg).max() <= 2e-6
    # float32 restatements in numpy (IEEE division; torch divides by a scalar through
    # its reciprocal)
    pix = b["pix"].view(torch.int32).long()
    rgba = data.rgba.reshape(-1, 4)[pix].cpu().numpy().astype(np.float32) / np.float32(255)
    alpha = rgba[:, 3:]
    one = np.float32(1)
    bgs = b["bgs"].cpu().numpy()
    np.testing.assert_array_equal(b["gt"].cpu().numpy(), rgba[:, :3] * alpha + bgs * (one - alpha))
    dbg = b["dbgs"].cpu().numpy() - b["g_rgbd"].cpu().numpy()[:, :3] * (one - alpha)
    np.testing.assert_array_equal(bb["dbg"].cpu().numpy(), dbg)
    assert np.abs(dbg).max() > 0
    dp64, _ = oracle.sky_sphere_bg_backward64(o, d, none, p0.cpu().numpy(), 4.0, 0, dbg)
    assert np.abs(bb["dparams"].cpu().numpy() - dp64).max() <= 2e-5 * np.abs(dp64).max()
    first = tr.metrics()["loss"]
    losses = []
    for i in range(300):
        tr.step()
        if i >= 280:
            losses.append(tr.metrics()["loss"])
    assert np.mean(losses) < 0.5 * first
    assert not torch.equal(tr.bg_model.params.detach(), p0)
    assert torch.isfinite(tr.bg_model.params).all()
    from jaxngp_amd.checkpoint import restore_trainer, save_checkpoint
    path = save_checkpoint(str(tmp_path), tr)
    tr2 = Trainer(data, TrainConfig(total_samples=1 << 15, background_model=True, seed=5))
    assert not torch.equal(tr2.bg_model.params, tr.bg_model.params)
    restore_trainer(tr2, path)
    assert torch.equal(tr2.bg_model.params, tr.bg_model.params)


def test_overlapped_update_matches_in_order(small_data):
    """The optimizer update on the side stream (overlapping the next step's ray
    generation and march) waits for the backward and is waited for by the next
    step's encoder.  From the same seeds: the MLP after one step is bit-identical
    to the in-order update's (its gradient and update are fixed-order), and the
    second step's MLP gradient agrees up to the table gradient's rounding (its f64
    sums take their terms in any order).  An update started before the backward
    finished, or a second step that read the table before the update landed (the
    table moves from its 1e-4 init by lr = 1e-2 per step), would change them
    grossly.  (Longer runs are no sharper test: Adam with eps = 1e-15 moves a
    gradient element near zero by a full +-lr whichever rounding it got.)"""
    res = {}
    for ov in (True, False):
        tr = Trainer(small_data, TrainConfig(total_samples=1 << 16, overlap_update=ov, prefetch_march=False))
        tr.step()
        m1 = tr.model.mlp_params.detach().clone()
        tr.step()
        torch.cuda.synchronize()
        res[ov] = (m1, tr.b["dparams"].clone())
    assert torch.equal(res[True][0], res[False][0])
    g1, g0 = res[True][1], res[False][1]
    assert float(g0.abs().max()) > 0
    assert float((g1 - g0).abs().max()) <= 1e-3 * float(g0.abs().max())


@pytest.mark.parametrize("guided", [False, True])
def test_prefetched_samples_equal_a_fresh_run(small_data, guided):
    """prefetch_march: the next step's rays and samples, made on the prefetch stream
    while a step computes (into the second buffer set), equal train_make_rays and the
    sampler run afresh, alone, for the same step index on the same bitfield (and, in
    guided mode, the same proposal).  The step then takes them."""
    cfg = TrainConfig(total_samples=1 << 16, precision="fp16", proposal=guided, proposal_guided=guided)
    tr = Trainer(small_data, cfg)
    for _ in range(60):
        tr.step()
        if tr._pf_pending[tr._cur] is not None:
            break
    nxt = tr._cur
    assert tr._pf_pending[nxt] is not None, "no step was prefetched"
    bits, s = tr._pf_pending[nxt]
    assert bits is tr.grid.bitfield and s == tr.step_count
    torch.cuda.synchronize()
    got = tr._sets[nxt]
    fresh = tr._second_set(got)
    tr._rays_and_samples(tr._prepare(fresh), s, None)
    torch.cuda.synchronize()
    for key in Trainer.RAY_KEYS:
        assert torch.equal(got[key], fresh[key]), key
    for key in ("next_loc", "exceeded"):
        assert torch.equal(got[key], fresh[key]), key
    assert int(got["next_loc"].view(torch.int32)) > 0
    tr.step()
    torch.cuda.synchronize()
    m = tr.metrics()
    assert math.isfinite(m["loss"]) and m["n_valid_rays"] > 0
