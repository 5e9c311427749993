"""Pins the float64 sky-sphere restatement (models/nerfs.py:133-213, 383-419).

The reference has no test or fixture for the background model.  Each part is
checked against what the mathematics fixes: the hit point lies on the sphere and is
the larger root that numpy.roots finds; SH degree 2 is the first four values of the
degree-4 encoding (pinned by orthonormality in test_oracle_pins); the hit points are
divided by the norm of the whole batch (duplicating the batch divides them by
sqrt(2)); one ray that misses poisons every colour with NaN; the MLP equals torch's
float64 layers; the gradients equal central finite differences.
"""
import numpy as np
import pytest
import torch

import oracle

R = 4.0


def _rays(rng, n, inside=True):
    o = rng.uniform(-1.5, 1.5, (n, 3)) if inside else rng.uniform(5.0, 6.0, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d


def _params(rng, E):
    return rng.uniform(-0.4, 0.4, (20 + E) * 32 + 1120)


def test_hit_point_is_the_far_root_on_the_sphere(rng):
    o, d = _rays(rng, 200)
    d *= rng.uniform(0.5, 2.0, (200, 1))  # non-unit directions: a != 1
    t = oracle.sky_sphere_hit64(o, d, R)
    assert np.all(np.isfinite(t)) and np.all(t > 0)  # origin inside: one root on each side
    np.testing.assert_allclose(np.linalg.norm(o + t[:, None] * d, axis=1), R, rtol=1e-12)
    for i in range(0, 200, 17):
        roots = np.roots([(d[i] ** 2).sum(), 2 * (o[i] * d[i]).sum(), (o[i] ** 2).sum() - R * R])
        assert abs(t[i] - roots.real.max()) <= 1e-12 * abs(t[i])


def test_miss_is_nan_and_poisons_the_batch(rng):
    o, d = _rays(rng, 50)
    app = np.zeros((50, 0))
    p = _params(rng, 0)
    ok = oracle.sky_sphere_bg64(o, d, app, p, R, 0)
    assert np.all(np.isfinite(ok)) and np.all((ok > 0) & (ok < 1))
    # ray 7's line x = 6, z = 0 never meets the sphere (a ray that leaves it
    # instead has real roots behind its origin and a finite t)
    o[7] = [6.0, 0.0, 0.0]
    d[7] = [0.0, 1.0, 0.0]
    assert np.isnan(oracle.sky_sphere_hit64(o[7:8], d[7:8], R)[0])
    got = oracle.sky_sphere_bg64(o, d, app, p, R, 0)
    np.testing.assert_array_equal(got[7], 0.0)
    assert np.all(np.isnan(np.delete(got, 7, axis=0)))


def test_positions_are_divided_by_the_batch_norm(rng):
    o, d = _rays(rng, 30)
    x1, _ = oracle.sky_sphere_inputs64(o, d, np.zeros((30, 0)), R, 0)
    x2, _ = oracle.sky_sphere_inputs64(np.concatenate([o, o]), np.concatenate([d, d]), np.zeros((60, 0)), R, 0)
    np.testing.assert_allclose(x2[:30, 1:4], x1[:, 1:4] / np.sqrt(2.0), rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(x2[:30, 4:], x1[:, 4:])  # the direction part does not change
    # one ray: the hit point is normalised to the unit sphere, SH1 = -c (y, -z, x)
    x, _ = oracle.sky_sphere_inputs64(o[:1], d[:1], np.zeros((1, 0)), R, 0)
    t = oracle.sky_sphere_hit64(o[:1], d[:1], R)[0]
    u = (o[0] + t * d[0]) / R
    c = 0.48860251190291987
    np.testing.assert_allclose(x[0, 1:4], [-c * u[1], c * u[2], -c * u[0]], rtol=1e-12)
    np.testing.assert_allclose(x[0, :4], oracle.sh4_64(u[None])[0, :4], rtol=1e-12)


@pytest.mark.parametrize("E", [0, 5])
def test_mlp_matches_torch_float64(rng, E):
    n = 64
    o, d = _rays(rng, n)
    app = rng.normal(size=(n, E))
    p = _params(rng, E)
    x, fin = oracle.sky_sphere_inputs64(o, d, app, R, E)
    assert fin.all()
    np.testing.assert_array_equal(x[:, 20:], app)
    W0, W1, W2 = (torch.from_numpy(w) for w in oracle.sky_sphere_split(p, E))
    X = torch.from_numpy(x)
    y = torch.sigmoid(torch.relu(torch.relu(X @ W0) @ W1) @ W2).numpy()
    np.testing.assert_allclose(oracle.sky_sphere_bg64(o, d, app, p, R, E), y, rtol=1e-13, atol=1e-15)


def test_broadcast_appearance_equals_per_ray_copies(rng):
    o, d = _rays(rng, 20)
    a = rng.normal(size=4)
    p = _params(rng, 4)
    np.testing.assert_array_equal(oracle.sky_sphere_bg64(o, d, a, p, R, 4),
                                  oracle.sky_sphere_bg64(o, d, np.tile(a, (20, 1)), p, R, 4))


@pytest.mark.parametrize("E", [0, 3])
def test_backward_matches_finite_differences(rng, E):
    n = 24
    o, d = _rays(rng, n)
    o[3] = [9.0, 9.0, 9.0]  # a ray that misses would poison the batch; keep this
    d[3] = -o[3] / np.linalg.norm(o[3])  # one outside but aimed through the sphere
    app = rng.normal(size=(n, E))
    p = _params(rng, E)
    g = rng.normal(size=(n, 3))
    dp, dapp = oracle.sky_sphere_bg_backward64(o, d, app, p, R, E, g)
    loss = lambda pp, aa: float((oracle.sky_sphere_bg64(o, d, aa, pp, R, E) * g).sum())
    h = 1e-6
    for k in rng.choice(p.size, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (loss(pp, app) - loss(pm, app)) / (2 * h)
        assert abs(fd - dp[k]) <= 1e-7 + 1e-6 * abs(fd), (k, fd, dp[k])
    for i, e in [(0, 0), (5, E - 1), (n - 1, 1)] if E else []:
        ap, am = app.copy(), app.copy()
        ap[i, e] += h
        am[i, e] -= h
        fd = (loss(p, ap) - loss(p, am)) / (2 * h)
        assert abs(fd - dapp[i, e]) <= 1e-7 + 1e-6 * abs(fd)
    assert dapp.shape == (n, E)


def test_gradient_of_a_poisoned_batch(rng):
    """A ray that misses gets colour 0 and its cotangent is dropped, but the NaN
    norm makes every other ray's hit-point encoding, activations and colour NaN.
    With relu's VJP a select (jax.nn.relu: select(z > 0, g, 0)), the deltas behind
    a NaN pre-activation are 0: dW2 and dW1 are NaN (NaN activations times the
    deltas), dW0 is NaN only on the three hit-point SH rows (SH2's constant row and
    the direction rows meet zero deltas) and the appearance gradient is 0."""
    o, d = _rays(rng, 6)
    o[2] = [0.0, 7.0, 0.0]
    d[2] = [1.0, 0.0, 0.0]
    app = rng.normal(size=(6, 2))
    p = _params(rng, 2)
    g = rng.normal(size=(6, 3))
    dp, dapp = oracle.sky_sphere_bg_backward64(o, d, app, p, R, 2, g)
    W0 = dp[:22 * 32].reshape(22, 32)
    assert np.isnan(dp[22 * 32:]).all()
    assert np.isnan(W0[1:4]).all()
    np.testing.assert_array_equal(W0[0], 0.0)
    np.testing.assert_array_equal(W0[4:], 0.0)
    np.testing.assert_array_equal(dapp, 0.0)


def test_relu_gradient_at_zero_is_zero():
    """jax.nn.relu's derivative at z = 0 is 0 (select(z > 0, ...)): a unit whose
    pre-activation is exactly 0 passes no gradient."""
    o = np.array([[0.0, 0.0, 0.0]])
    d = np.array([[0.0, 0.0, 1.0]])
    p = np.zeros(20 * 32 + 1120)
    p[20 * 32 + 1024:] = 1.0  # W2 = 1: dL/dh2 nonzero, but z2 = 0 everywhere
    dp, _ = oracle.sky_sphere_bg_backward64(o, d, np.zeros((1, 0)), p, R, 0, np.ones((1, 3)))
    np.testing.assert_array_equal(dp, 0.0)
