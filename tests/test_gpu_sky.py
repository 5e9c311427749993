"""GPU parity of the sky-sphere background targets against the float64 oracle
(models/nerfs.py:133-213).  Tolerances: colours within 2e-6 absolute (f32
geometry, SH and MLP against float64; colours lie in (0, 1)); gradients within
2e-5 of the largest magnitude of the same output (f32 sums over up to 2^18 rays
against float64)."""
import numpy as np
import pytest
import torch

import oracle
from jaxngp_amd import _lib
from jaxngp_amd.background import SkySphereBg, make_skysphere_background_model_ngp

pytestmark = pytest.mark.gpu
R = 4.0


def _inputs(seed, n, E, per_ray=True):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    app = rng.normal(size=(n, E) if per_ray else (E,)).astype(np.float32)
    p = rng.uniform(-0.4, 0.4, (20 + E) * 32 + 1120).astype(np.float32)
    g = rng.normal(size=(n, 3)).astype(np.float32)
    return o, d, app, p, g


def _dev(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _forward(dev, o, d, app, p, E, per_ray):
    n = o.shape[0]
    desc = _lib.sky_sphere_bg_descriptor(n, E, per_ray, R)
    colors = torch.full((n, 3), -1.0, device=dev)
    ws = torch.empty(_lib.sky_sphere_bg_workspace_bytes(desc), dtype=torch.uint8, device=dev)
    a = _dev(app, dev) if E else torch.empty(0, device=dev)
    _lib.call("sky_sphere_bg", [_dev(o, dev), _dev(d, dev), a, _dev(p, dev), colors, ws], desc)
    return colors.cpu().numpy()


def _backward(dev, o, d, app, p, g, E):
    n = o.shape[0]
    desc = _lib.sky_sphere_bg_descriptor(n, E, 1, R)
    dp = torch.full((p.size,), np.nan, device=dev)
    da = torch.full((n, E), np.nan, device=dev) if E else torch.empty(0, device=dev)
    ws = torch.empty(_lib.sky_sphere_bg_workspace_bytes(desc), dtype=torch.uint8, device=dev)
    a = _dev(app, dev) if E else torch.empty(0, device=dev)
    _lib.call("sky_sphere_bg_backward", [_dev(o, dev), _dev(d, dev), a, _dev(p, dev), _dev(g, dev), dp, da, ws],
              desc)
    return dp.cpu().numpy(), (da.cpu().numpy() if E else np.zeros((n, 0)))


@pytest.mark.parametrize("n,E", [(1, 0), (1000, 0), (4099, 3), (777, 16)])
def test_forward_and_backward_match_oracle(dev, n, E):
    o, d, app, p, g = _inputs(n + E, n, E)
    want = oracle.sky_sphere_bg64(o, d, app, p, R, E)
    got = _forward(dev, o, d, app, p, E, True)
    assert np.abs(got - want).max() <= 2e-6
    dp, da = _backward(dev, o, d, app, p, g, E)
    dp64, da64 = oracle.sky_sphere_bg_backward64(o, d, app, p, R, E, g)
    assert np.abs(dp - dp64).max() <= 2e-5 * np.abs(dp64).max()
    if E:
        assert np.abs(da - da64).max() <= 2e-5 * np.abs(da64).max()


def test_broadcast_appearance(dev):
    o, d, app, p, _ = _inputs(5, 3000, 8, per_ray=False)
    got = _forward(dev, o, d, app, p, 8, False)
    assert np.abs(got - oracle.sky_sphere_bg64(o, d, app, p, R, 8)).max() <= 2e-6
    np.testing.assert_array_equal(got, _forward(dev, o, d, np.tile(app, (3000, 1)), p, 8, True))


def test_training_batch_size_and_determinism(dev):
    """2^18 + 77 rays (a ragged last workgroup): the colours match the oracle, the
    batch norm sees every ray, and two backward calls agree bit for bit."""
    n = (1 << 18) + 77
    o, d, app, p, g = _inputs(9, n, 0)
    got = _forward(dev, o, d, app, p, 0, True)
    assert np.abs(got - oracle.sky_sphere_bg64(o, d, app, p, R, 0)).max() <= 2e-6
    dp1, _ = _backward(dev, o, d, app, p, g, 0)
    dp2, _ = _backward(dev, o, d, app, p, g, 0)
    np.testing.assert_array_equal(dp1, dp2)
    dp64, _ = oracle.sky_sphere_bg_backward64(o, d, app, p, R, 0, g)
    assert np.abs(dp1 - dp64).max() <= 2e-5 * np.abs(dp64).max()


def test_a_miss_poisons_the_batch(dev):
    o, d, app, p, g = _inputs(11, 500, 0)
    o[123] = [6.0, 0.0, 0.0]  # the line x = 6, z = 0 never meets the sphere
    d[123] = [0.0, 1.0, 0.0]
    got = _forward(dev, o, d, app, p, 0, True)
    np.testing.assert_array_equal(got[123], 0.0)
    assert np.isnan(np.delete(got, 123, axis=0)).all()
    dp, _ = _backward(dev, o, d, app, p, g, 0)
    # the exact NaN pattern of jax.grad (relu's VJP is a select)
    np.testing.assert_array_equal(dp, oracle.sky_sphere_bg_backward64(o, d, app, p, R, 0, g)[0])
    assert np.isnan(dp[640:]).all() and np.isnan(dp[32:128]).all()


def test_no_rays_gives_zero_gradient(dev):
    dp, _ = _backward(dev, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), np.zeros((0, 0), np.float32),
                      np.ones(1760, np.float32), np.zeros((0, 3), np.float32), 0)
    np.testing.assert_array_equal(dp, 0.0)


def test_module_autograd_and_param_tree(dev):
    bg = make_skysphere_background_model_ngp(1.0, n_extra_dims=4, device=dev, seed=3)
    assert bg.radius == 4.0 and bg.params.numel() == 24 * 32 + 1120
    o, d, app, _, g = _inputs(13, 2048, 4)
    a = _dev(app, dev).requires_grad_(True)
    colors = bg(_dev(o, dev), _dev(d, dev), a)
    (colors * _dev(g, dev)).sum().backward()
    p = bg.params.detach().cpu().numpy()
    assert np.abs(colors.detach().cpu().numpy() - oracle.sky_sphere_bg64(o, d, app, p, R, 4)).max() <= 2e-6
    dp64, da64 = oracle.sky_sphere_bg_backward64(o, d, app, p, R, 4, g)
    assert np.abs(bg.params.grad.cpu().numpy() - dp64).max() <= 2e-5 * np.abs(dp64).max()
    assert np.abs(a.grad.cpu().numpy() - da64).max() <= 2e-5 * np.abs(da64).max()
    other = SkySphereBg(4.0, n_extra_dims=4, device=dev, seed=99)
    other.load_param_tree(bg.param_tree())
    assert torch.equal(other.params, bg.params)


def test_renderer_background_model(dev):
    """render_image_inference with bg_model (cuda.py:283-288): through an empty
    occupancy grid every pixel is the model's colour for its ray; with a partly
    occupied grid it equals the render given those colours as the per-ray bg."""
    from jaxngp_amd import scene as sc
    from jaxngp_amd.nerfs import NeRF
    from jaxngp_amd.renderer import make_rays_worldspace, render_image_inference
    from jaxngp_amd.scene import f32_to_u8
    cam = sc.Camera(width=64, height=48, fx=70.0, fy=70.0, cx=32.0, cy=24.0)
    # the camera inside the sky sphere (radius 4), as in the unbounded scenes the model is for
    pose = torch.as_tensor(sc.look_at_poses(1, radius=3.0, seed=2, golden=False)[0], dtype=torch.float32, device=dev)
    model = NeRF(bound=1.0, device=dev, seed=5)
    bg = make_skysphere_background_model_ngp(1.0, device=dev, seed=7)
    o, d = make_rays_worldspace(cam, pose)
    with torch.no_grad():
        colors = bg(o, d)
    empty = torch.zeros(128 ** 3 // 8, dtype=torch.uint8, device=dev)
    _, rgb, _ = render_image_inference(model, cam, pose, empty, bg_model=bg)
    assert torch.equal(rgb, f32_to_u8(colors).reshape(48, 64, 3))
    assert len(torch.unique(rgb.reshape(-1, 3), dim=0)) > 10  # the background varies over the frame
    bits = torch.randint(0, 256, (128 ** 3 // 8,), dtype=torch.uint8, device=dev) & 0x11
    _, a, _ = render_image_inference(model, cam, pose, bits, bg_model=bg)
    _, b, _ = render_image_inference(model, cam, pose, bits, bg=colors)
    assert torch.equal(a, b)
