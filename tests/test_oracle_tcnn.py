"""Pins of the oracle's restatement of the jax-tcnn grid encoding (tiny-cuda-nn v1.6's
linear grid with the coherent-prime hash, not vendored): closed forms, the index
formula against big-integer arithmetic, the coordinate derivative against finite
differences, and the adjoint identity of the table gradient."""
import math

import numpy as np

import oracle
from jaxngp_amd.tcnn import unpadded_offsets

CFG = dict(L=4, F=2, N_min=4, per_level_scale=2.0)  # resolutions 4..32: dense levels and one hashed level


def _setup(T=2 ** 12, seed=0):
    L, F, N_min, b = CFG["L"], CFG["F"], CFG["N_min"], CFG["per_level_scale"]
    offs = unpadded_offsets(L, T, N_min, b)
    rng = np.random.default_rng(seed)
    table = rng.uniform(-1, 1, (int(offs[-1]), F)).astype(np.float32)
    return offs, table, rng


def test_unpadded_offsets_and_the_jax_layout():
    """encoders.py:268-283: dense levels hold res^3 rows (no rounding up to 8), so at
    the encoder workload the table has 12 rows fewer than HashGridEncoder's."""
    b = math.exp((math.log(2048) - math.log(16)) / 15)
    o = unpadded_offsets(16, 2 ** 19, 16, b)
    lv = oracle.hashgrid_levels(16, 2 ** 19, 16, 2048)
    assert int(o[-1]) == 6098108 and lv["n_rows"] == 6098120
    res = [int(r) for r in lv["resolutions"]]
    for l in range(5):
        assert o[l + 1] - o[l] == res[l] ** 3
    assert all(o[l + 1] - o[l] == 2 ** 19 for l in range(5, 16))


def test_index_formula():
    rng = np.random.default_rng(1)
    x, y, z = (rng.integers(0, 3000, 500).astype(np.uint32) for _ in range(3))
    # dense: res^3 <= size
    got = oracle.tcnn_index(23 ** 3, 23, x % 24, y % 24, z % 24)
    want = [((int(a) % 24) + (int(b) % 24) * 23 + (int(c) % 24) * 23 * 23) % 23 ** 3 for a, b, c in zip(x, y, z)]
    np.testing.assert_array_equal(got, want)
    # hashed: res^2 > size stops the dense stride
    T = 2 ** 19
    got = oracle.tcnn_index(T, 2049, x, y, z)
    want = [((int(a) ^ (int(b) * 2654435761 % 2 ** 32) ^ (int(c) * 805459861 % 2 ** 32)) % 2 ** 32) % T
            for a, b, c in zip(x, y, z)]
    np.testing.assert_array_equal(got, want)


def test_constant_table_and_linear_field():
    offs, table, rng = _setup()
    L, F = CFG["L"], CFG["F"]
    pts = rng.uniform(0, 1, (3, 2000)).astype(np.float32)
    # x below (res - 1.5) / scale on every level: the +0.5 offset then keeps the x + 1
    # corner inside the grid (above it, dense indices wrap into the next row)
    pts[0] *= np.float32(0.8)
    enc, _ = oracle.tcnn_grid_forward(pts, np.full_like(table, 0.25), offs, **CFG)
    np.testing.assert_allclose(enc, 0.25, rtol=0, atol=2e-7)
    # a dense level whose rows hold their vertex's x coordinate interpolates
    # pos = fma(scale, x, 0.5) itself
    levels = oracle.tcnn_levels(L, CFG["N_min"], CFG["per_level_scale"])
    lin = np.zeros_like(table)
    for l, (scale, res) in enumerate(levels):
        size = int(offs[l + 1] - offs[l])
        if res ** 3 <= size:
            r = np.arange(size)
            lin[offs[l]:offs[l + 1], 0] = r % res
    enc, _ = oracle.tcnn_grid_forward(pts, lin, offs, **CFG)
    for l, (scale, res) in enumerate(levels):
        if res ** 3 <= int(offs[l + 1] - offs[l]):
            want = np.float64(scale) * pts[0].astype(np.float64) + 0.5
            np.testing.assert_allclose(enc[l * F], want, rtol=2e-6, atol=2e-6)


def test_coordinate_derivative_matches_finite_differences():
    offs, table, rng = _setup(seed=2)
    L, F = CFG["L"], CFG["F"]
    pts = rng.uniform(0.05, 0.95, (3, 300)).astype(np.float32)
    _, dy = oracle.tcnn_grid_forward(pts, table, offs, **CFG)
    h = 1e-6
    scales = [s for s, _ in oracle.tcnn_levels(L, CFG["N_min"], CFG["per_level_scale"])]
    for d in range(3):
        p1, p0 = pts.astype(np.float64).copy(), pts.astype(np.float64).copy()
        p1[d] += h
        p0[d] -= h
        fd = (oracle.tcnn_grid_forward64(p1, table, offs, **CFG) - oracle.tcnn_grid_forward64(p0, table, offs, **CFG)) / (2 * h)
        for l, s in enumerate(scales):
            pos = float(s) * pts[d].astype(np.float64) + 0.5
            away = np.abs(pos - np.round(pos)) > 1e-3  # no cell boundary within the step
            for f in range(F):
                k = l * F + f
                np.testing.assert_allclose(dy[k, away, d], fd[k, away], rtol=1e-4, atol=1e-4 * float(s))


def test_table_gradient_is_the_adjoint():
    offs, table, rng = _setup(seed=3)
    L, F = CFG["L"], CFG["F"]
    pts = rng.uniform(0, 1, (3, 700)).astype(np.float32)
    g = rng.normal(size=(L * F, 700))
    enc64 = oracle.tcnn_grid_forward64(pts, table, offs, **CFG)
    _, dy = oracle.tcnn_grid_forward(pts, table, offs, **CFG)
    dt, dc = oracle.tcnn_grid_backward64(pts, g, dy, offs, **CFG, n_rows=int(offs[-1]))
    assert abs((enc64 * g).sum() - (table.astype(np.float64) * dt).sum()) < 1e-6 * np.abs(enc64 * g).sum()
    np.testing.assert_allclose(dc, np.einsum("kn,knd->dn", g, dy.astype(np.float64)), rtol=1e-12)
