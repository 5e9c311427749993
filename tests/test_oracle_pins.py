"""Pins the CPU oracle independently of the reference.

The reference has no tests, golden vectors or fixtures for this path, and cannot
run here (CUDA-only kernels, JAX-only encoder).  Each oracle function is therefore
checked against what the mathematics fixes: closed forms, brute force on small
inputs, invariants, library routines, and the known-answer values recorded from
the reference's own formulas (tests/golden/).  Each check is chosen so that a
plausible slip (a dropped term, a wrong index or sign, a transposed operand)
fails it.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.helpers import (SQRT3, blob_bitfield, camera_rays, integrate_inputs, march_all_rounds, march_inputs,
                           near_far_np)

GOLDEN = Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------------------
# deterministic exp
# ---------------------------------------------------------------------------

def test_expf_accuracy_and_specials(rng):
    xs = np.concatenate([rng.uniform(-100, 88, 20000), rng.uniform(-1, 1, 5000), [0.0, -0.0, 1.0, -1.0, 88.7]])
    for x in xs.astype(np.float32):
        got = oracle.expf(float(x))
        want = math.exp(float(x))
        if want < 1.1754944e-38:  # subnormal range: absolute error of one subnormal ulp
            assert abs(got - want) <= 1.5e-45 * 2
        else:
            assert abs(got - want) <= 2.5e-7 * want, (x, got, want)
    assert oracle.expf(0.0) == 1.0
    assert oracle.expf(-200.0) == 0.0
    assert math.isinf(oracle.expf(100.0))
    assert math.isnan(oracle.expf(float("nan")))


# ---------------------------------------------------------------------------
# Morton / packbits
# ---------------------------------------------------------------------------

def _interleave(x, y, z):
    out = 0
    for i in range(10):
        out |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
    return out


def test_morton_matches_bit_interleave(rng):
    xyz = rng.integers(0, 1024, (2000, 3)).astype(np.uint32)
    got = oracle.morton3d(xyz)
    want = np.array([_interleave(*map(int, p)) for p in xyz], np.uint32)
    np.testing.assert_array_equal(got, want)


def test_morton_invert_round_trip_all_cells():
    G = 128
    idx = np.arange(G ** 3, dtype=np.uint32)
    xyz = oracle.morton3d_invert(idx)
    assert xyz.max() == G - 1
    np.testing.assert_array_equal(oracle.morton3d(xyz), idx)
    # bijection onto the grid
    flat = xyz[:, 0].astype(np.int64) * G * G + xyz[:, 1] * G + xyz[:, 2]
    assert len(np.unique(flat)) == G ** 3


def test_packbits_matches_numpy(rng):
    dens = rng.normal(size=8 * 1000).astype(np.float32)
    thr = rng.normal(size=8 * 1000).astype(np.float32)
    mask, bits = oracle.packbits(thr, dens)
    np.testing.assert_array_equal(mask, dens > thr)
    np.testing.assert_array_equal(bits, np.packbits(dens > thr, bitorder="little"))
    mask, bits = oracle.packbits(np.float32(0.25), dens)
    np.testing.assert_array_equal(bits, np.packbits(dens > 0.25, bitorder="little"))


# ---------------------------------------------------------------------------
# hash-grid level table and indices
# ---------------------------------------------------------------------------

def test_level_table_known_answers():
    kat = json.loads((GOLDEN / "hashgrid_levels.json").read_text())
    for case in kat["cases"]:
        lv = oracle.hashgrid_levels(case["L"], case["T"], case["N_min"], case["N_max"])
        assert abs(lv["b"] - case["b"]) < 1e-6
        assert lv["first_hash_level"] == case["first_hash_level"]
        assert lv["n_rows"] == case["n_rows"]
        sizes = np.diff(lv["offsets"])[: case["first_hash_level"]].tolist()
        assert sizes == case["dense_sizes"]
        # resolutions enter the indices of dense levels only (hashed levels ignore them;
        # there the float64 host loop may round 16*b^15-1 up to 2047.0000000000002)
        fhl = case["first_hash_level"]
        assert lv["resolutions"].tolist()[:fhl] == case["resolutions"][:fhl]


def _vertex_index_bigint(p, bound, scale32, res, dense, T, offset, n_rows, corner):
    """Integer-exact index of one corner with Python ints (f32 steps done in numpy f32)."""
    b = np.float32(bound)
    ps = (np.float32(p) + b) / (np.float32(2) * b) * np.float32(scale32) + np.float32(0.5)
    fl = np.floor(ps)
    v = [int(fl[a] + np.float32(corner[a])) for a in range(3)]
    if dense:
        h = (v[0] + v[1] * res + v[2] * ((res * res) % 2 ** 32)) % 2 ** 32
    else:
        h = (v[0] ^ (v[1] * 2654435761 % 2 ** 32) ^ (v[2] * 805459861 % 2 ** 32)) % 2 ** 32
    return min(h % T + offset, n_rows - 1)


@pytest.mark.parametrize("cfg", [(16, 2 ** 19, 16, 2048), (5, 2 ** 17, 16, 512), (4, 2 ** 6, 2, 6)])
def test_hash_indices_bigint(cfg, rng):
    L, T, nmin, nmax = cfg
    lv = oracle.hashgrid_levels(L, T, nmin, nmax)
    pos = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    pos[:5] = [[-1, -1, -1], [1, 1, 1], [0, 0, 0], [np.nextafter(np.float32(1), 0)] * 3, [-0.5, 0.25, 0.999]]
    idx, _ = oracle.hashgrid_vertices(pos, 1.0, lv, T)
    corners = [((c >> 2) & 1, (c >> 1) & 1, c & 1) for c in range(8)]
    for n in range(0, 300, 7):
        for l in range(L):
            for c in range(8):
                want = _vertex_index_bigint(pos[n], 1.0, lv["scales_f32"][l], int(lv["resolutions"][l]),
                                            l < lv["first_hash_level"], T, int(lv["offsets"][l]), lv["n_rows"],
                                            corners[c])
                assert idx[n, l, c] == want


def test_hash_indices_c_oracle_matches_numpy(rng):
    import ctypes
    L, T = 16, 2 ** 19
    lv = oracle.hashgrid_levels(L, T, 16, 2048)
    pos = rng.uniform(-1, 1, (200, 3)).astype(np.float32)
    idx_np, frac = oracle.hashgrid_vertices(pos, 1.0, lv, T)
    fn = oracle.lib().oracle_hashgrid_level
    out_i = np.zeros(8, np.uint32)
    out_w = np.zeros(8, np.float32)
    for n in range(pos.shape[0]):
        p = np.ascontiguousarray(pos[n])
        for l in range(L):
            fn(p.ctypes.data_as(ctypes.c_void_p), ctypes.c_float(1.0), ctypes.c_float(lv["scales_f32"][l]),
               ctypes.c_uint32(int(lv["resolutions"][l])), ctypes.c_int(int(l < lv["first_hash_level"])),
               ctypes.c_uint32(T), ctypes.c_uint32(int(lv["offsets"][l])), ctypes.c_uint32(lv["n_rows"]),
               out_i.ctypes.data_as(ctypes.c_void_p), out_w.ctypes.data_as(ctypes.c_void_p))
            np.testing.assert_array_equal(out_i.astype(np.int64), idx_np[n, l])
            np.testing.assert_allclose(out_w, oracle.hashgrid_weights64(frac[n:n + 1, l:l + 1])[0, 0], rtol=0,
                                       atol=2e-7)


def test_encoder_partition_of_unity_and_constant_table(rng):
    L, T = 8, 2 ** 12
    lv = oracle.hashgrid_levels(L, T, 4, 64)
    pos = rng.uniform(-1, 1, (500, 3)).astype(np.float32)
    _, frac = oracle.hashgrid_vertices(pos, 1.0, lv, T)
    w = oracle.hashgrid_weights64(frac)
    np.testing.assert_allclose(w.sum(-1), 1.0, atol=1e-12)
    table = np.full((lv["n_rows"], 2), 0.375)
    enc = oracle.hashgrid_forward64(pos, table, 1.0, lv, T)
    np.testing.assert_allclose(enc, 0.375, atol=1e-12)


def test_encoder_reproduces_linear_field_on_dense_levels(rng):
    """A table whose rows hold a linear function of the vertex coordinates is
    interpolated exactly on dense levels (away from the res-overflow vertices)."""
    L, T = 3, 2 ** 20
    lv = oracle.hashgrid_levels(L, T, 8, 32)
    assert lv["first_hash_level"] == L
    table = np.zeros((lv["n_rows"], 1))
    for l in range(L):
        r = int(lv["resolutions"][l])
        v = np.arange(int(lv["offsets"][l + 1] - lv["offsets"][l]))
        x, y, z = v % r, (v // r) % r, v // (r * r)
        table[lv["offsets"][l]:lv["offsets"][l + 1], 0] = 2.0 * x - 3.0 * y + 0.5 * z
    pos = rng.uniform(-0.9, 0.9, (400, 3)).astype(np.float32)
    enc = oracle.hashgrid_forward64(pos, table, 1.0, lv, T)
    for l in range(L):
        ps = ((pos.astype(np.float32) + 1) / 2) * lv["scales_f32"][l] + np.float32(0.5)
        r = int(lv["resolutions"][l])
        inside = (np.floor(ps) + 1 < r).all(axis=1)  # the +1 vertex stays inside the level
        ps = ps.astype(np.float64)
        want = 2.0 * ps[:, 0] - 3.0 * ps[:, 1] + 0.5 * ps[:, 2]
        np.testing.assert_allclose(enc[inside, l], want[inside], atol=1e-9)


def test_encoder_backward_is_the_adjoint(rng):
    L, T, F = 6, 2 ** 10, 2
    lv = oracle.hashgrid_levels(L, T, 4, 48)
    pos = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    theta = rng.normal(size=(lv["n_rows"], F))
    g = rng.normal(size=(300, L * F))
    lhs = (oracle.hashgrid_forward64(pos, theta, 1.0, lv, T) * g).sum()
    rhs = (theta * oracle.hashgrid_backward64(pos, g, 1.0, lv, T, F)).sum()
    assert abs(lhs - rhs) <= 1e-9 * max(1.0, abs(lhs))


def test_c_encoder_matches_float64(rng):
    import ctypes
    L, T, F = 16, 2 ** 14, 2
    lv = oracle.hashgrid_levels(L, T, 16, 512)
    N = 2000
    pos = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    table = rng.uniform(-1e-2, 1e-2, (lv["n_rows"], F)).astype(np.float32)
    enc = np.zeros((N, L * F), np.float32)
    scales = lv["scales_f32"].astype(np.float32)
    res = lv["resolutions"].astype(np.uint32)
    offs = lv["offsets"].astype(np.uint32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    oracle.lib().oracle_hashgrid_fwd(ctypes.c_uint32(N), ctypes.c_uint32(L), ctypes.c_uint32(F), ctypes.c_uint32(T),
                                     ctypes.c_uint32(lv["first_hash_level"]), ctypes.c_uint32(lv["n_rows"]),
                                     ctypes.c_float(1.0), P(scales), P(res), P(offs), P(pos), P(table), P(enc))
    want = oracle.hashgrid_forward64(pos, table, 1.0, lv, T)
    np.testing.assert_allclose(enc, want, rtol=0, atol=1e-5 * 1e-2)


def _c_table_gradient(lv, T, F, pos, g):
    import ctypes
    N, L = pos.shape[0], len(lv["scales_f32"])
    out = np.zeros((lv["n_rows"], F), np.float64)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    scales = np.ascontiguousarray(lv["scales_f32"], np.float32)
    res = np.ascontiguousarray(lv["resolutions"], np.uint32)
    offs = np.ascontiguousarray(lv["offsets"], np.uint32)
    oracle.lib().oracle_hashgrid_bwd(ctypes.c_uint32(N), ctypes.c_uint32(L), ctypes.c_uint32(F), ctypes.c_uint32(T),
                                     ctypes.c_uint32(lv["first_hash_level"]), ctypes.c_uint32(lv["n_rows"]),
                                     ctypes.c_float(1.0), P(scales), P(res), P(offs),
                                     P(np.ascontiguousarray(pos, np.float32)), P(np.ascontiguousarray(g, np.float32)),
                                     P(out))
    return out


@pytest.mark.parametrize("cfg", [dict(L=16, T=2 ** 19, F=2, N_min=16, N_max=2048),
                                 dict(L=4, T=2 ** 14, F=8, N_min=16, N_max=256)], ids=["l16_t2e19", "f8"])
def test_c_table_gradient_checker_is_the_adjoint(cfg, rng):
    """The C table gradient (the checker of the GPU table-gradient tests) is the
    adjoint of the float64 forward, <enc(theta), g> = <theta, dtheta>, for random
    tables theta, and matches the adjoint-pinned float64 scatter row by row.  A wrong
    corner, weight, feature or level index in the C scatter fails one of these."""
    L, T, F = cfg["L"], cfg["T"], cfg["F"]
    lv = oracle.hashgrid_levels(L, T, cfg["N_min"], cfg["N_max"])
    N = 3000
    pos = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    pos[:3] = [[-1, -1, -1], [1, 1, 1], [0.999999, -0.999999, 0.5]]
    g = rng.normal(size=(N, L * F)).astype(np.float32)
    got = _c_table_gradient(lv, T, F, pos, g)
    absw = _c_table_gradient(lv, T, F, pos, np.abs(g))
    # the C weights are f32 products, the float64 scatter's are float64 products of the
    # same f32 fractions: they agree to a few f32 ulps of each contribution
    want = oracle.hashgrid_backward64(pos, g, 1.0, lv, T, F)
    assert (np.abs(got - want) <= 1e-6 * absw + 1e-300).all()
    assert (got[absw == 0] == 0).all()
    for seed in range(3):
        theta = np.random.default_rng(seed).normal(size=(lv["n_rows"], F))
        lhs = float((oracle.hashgrid_forward64(pos, theta, 1.0, lv, T) * g).sum())
        rhs = float((theta * got).sum())
        scale = float((np.abs(theta) * absw).sum())
        assert abs(lhs - rhs) <= 1e-6 * scale, (lhs, rhs, scale)


# ---------------------------------------------------------------------------
# marcher
# ---------------------------------------------------------------------------

def _dense_f32_sequence(t0, t1, ds, noise, limit):
    t = np.float32(t0)
    t = np.float32(t + np.float32(ds * np.float32(noise)))
    out = []
    while len(out) < limit and t < t1:
        out.append(t)
        t = np.float32(t + ds)
    return np.array(out, np.float32)


def test_march_all_occupied_closed_form(rng):
    n = 300
    o, d, t0, t1, noise, _, _ = march_inputs(rng, n, fill=True)
    bits = np.full(128 ** 3 // 8, 255, np.uint8)
    out = oracle.march_rays(10 ** 6, 1024, 1, 128, 1.0, 0.0, o, d, t0, t1, noise, bits)
    assert out["number_of_exceeded_samples"] == 0
    ds = np.float32(np.float32(2 * SQRT3) * np.float32(1.0) / np.float32(1024))
    for i in range(n):
        if not t1[i] > t0[i]:
            assert not out["ray_is_valid"][i]
            continue
        seq = _dense_f32_sequence(t0[i], t1[i], ds, noise[i], 1024)
        s0, k = out["rays_sample_startidx"][i], out["rays_n_samples"][i]
        assert k == len(seq)
        np.testing.assert_array_equal(out["z_vals"][s0:s0 + k], seq)
        np.testing.assert_array_equal(out["dss"][s0:s0 + k], np.full(k, ds, np.float32))
        np.testing.assert_array_equal(out["xyzs"][s0:s0 + k], (o[i] + seq[:, None] * d[i]).astype(np.float32))
        assert (out["idcs"][s0:s0 + k] == i).all()


def test_march_empty_field_and_invalid_rays(rng):
    n = 100
    o, d, t0, t1, noise, _, _ = march_inputs(rng, n)
    t1[::5] = t0[::5]  # t_end <= t_start -> invalid
    bits = np.zeros(128 ** 3 // 8, np.uint8)
    out = oracle.march_rays(1000, 1024, 1, 128, 1.0, 0.0, o, d, t0, t1, noise, bits)
    assert out["measured"] == 0 and out["next_sample_write_location"] == 0
    np.testing.assert_array_equal(out["ray_is_valid"], t1 > t0)
    assert not out["rays_n_samples"].any() and not out["xyzs"].any()


def test_march_single_cell_brute_force(rng):
    """Only one occupied cell: the samples are exactly the points of the dense f32
    step sequence that fall in that cell (boundary-ulp points flagged, not failed)."""
    G = 128
    ix, iy, iz = 70, 60, 65
    flat = np.zeros(G ** 3, bool)
    flat[oracle.morton3d(np.array([[ix, iy, iz]], np.uint32))[0]] = True
    bits = np.packbits(flat, bitorder="little")
    center = (np.array([ix, iy, iz]) + 0.5) / G * 2 - 1
    o, _ = camera_rays(rng, 200)
    tgt = center + rng.uniform(-0.5, 0.5, (200, 3)) / G * 2
    d = (tgt - o)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    t0, t1 = near_far_np(o, d, 1.0)
    noise = rng.uniform(0, 1, 200).astype(np.float32)
    out = oracle.march_rays(100000, 1024, 1, G, 1.0, 0.0, o, d, t0, t1, noise, bits)
    ds = np.float32(np.float32(2 * SQRT3) / np.float32(1024))
    flagged = 0
    for i in range(200):
        seq = _dense_f32_sequence(t0[i], t1[i], ds, noise[i], 10 ** 6)
        pos = (o[i] + seq[:, None] * d[i]).astype(np.float32)
        g = ((pos + np.float32(1)) * np.float32(0.5)) * np.float32(G)
        cell = np.clip(np.floor(g), 0, G - 1).astype(int)
        inside = (cell == [ix, iy, iz]).all(axis=1)
        near_edge = (np.abs(g - np.round(g)) < 1e-3).any(axis=1)
        s0, k = out["rays_sample_startidx"][i], out["rays_n_samples"][i]
        got = set(out["z_vals"][s0:s0 + k].tolist())
        want = set(seq[inside].tolist())
        diff = got ^ want
        edge = set(seq[near_edge].tolist())
        assert diff <= edge, (i, sorted(diff - edge))
        flagged += len(diff)
        for z in got:
            assert z in set(seq.tolist())  # samples sit on the dense sequence
    assert flagged <= 2


def test_march_overflow_semantics(rng):
    n = 400
    o, d, t0, t1, noise, bits, _ = march_inputs(rng, n)
    full = oracle.march_rays(10 ** 6, 1024, 1, 128, 1.0, 0.0, o, d, t0, t1, noise, bits)
    counts = full["rays_n_samples"].astype(np.int64)
    prefix = np.concatenate([[0], np.cumsum(counts)[:-1]])
    for S in [1, 7, int(counts.sum() // 3), int(counts.sum()), int(counts.sum()) - 1]:
        S = max(S, 1)
        out = oracle.march_rays(S, 1024, 1, 128, 1.0, 0.0, o, d, t0, t1, noise, bits)
        want_valid = (prefix < S) & (t1 > t0) & (prefix + counts <= S)
        np.testing.assert_array_equal(out["ray_is_valid"], want_valid)
        measured = int(counts[want_valid].sum())
        assert out["measured"] == measured
        over = np.nonzero((prefix < S) & (prefix + counts > S))[0]
        exc = int(counts[over[0]]) if len(over) else 0
        assert out["number_of_exceeded_samples"] == exc
        assert out["next_sample_write_location"] == measured + exc
        np.testing.assert_array_equal(out["rays_n_samples"], np.where(want_valid, counts, 0))
        # kept rays' samples equal the unconstrained march's
        for i in np.nonzero(want_valid & (counts > 0))[0][:20]:
            a, b = out["rays_sample_startidx"][i], full["rays_sample_startidx"][i]
            np.testing.assert_array_equal(out["z_vals"][a:a + counts[i]], full["z_vals"][b:b + counts[i]])
        assert not out["z_vals"][measured:].any() and not out["idcs"][measured:].any()


def test_march_inference_chunks_equal_training_march(rng):
    """Marching a ray in chunks of `cap` occupied steps from the saved t (no noise)
    visits the same samples as one training march with noise 0, plus the far-plane
    sample when the ray leaves the box in an occupied cell."""
    n = 64
    o, d, t0, t1, _, bits, _ = march_inputs(rng, n)
    full = oracle.march_rays(10 ** 6, 1024, 1, 128, 1.0, 0.0, o, d, t0, t1, np.zeros(n, np.float32), bits)
    cap = 8
    ts = t0.copy()
    terminated = np.ones(n, bool)
    indices = np.zeros(n, np.uint32)
    nri = np.uint32(0)
    got = [[] for _ in range(n)]
    for _ in range(400):
        out = oracle.march_rays_inference(1024, 1, 128, cap, 1.0, 0.0, o, d, ts, t1, bits, nri, terminated, indices)
        nri, indices = out["next_ray_index"], out["indices"]
        for slot in range(n):
            r = indices[slot]
            if r >= n:
                continue
            k = out["n_samples"][slot]
            got[r].extend(out["z_vals"][slot, :k].tolist())
            ts[r] = out["t_starts_out"][slot]
        terminated = out["n_samples"] < cap
        if (indices >= n).all():
            break
    for r in range(n):
        s0, k = full["rays_sample_startidx"][r], full["rays_n_samples"][r]
        train = full["z_vals"][s0:s0 + k].tolist()
        assert got[r][:k] == train
        extra = got[r][k:]
        assert all(z == t1[r] for z in extra)


# ---------------------------------------------------------------------------
# integrator
# ---------------------------------------------------------------------------

def test_integrate_constant_medium_closed_form():
    n_rays, m = 6, 400
    sigmas = [0.5, 3.0, 20.0, 80.0, 150.0, 400.0]
    ds = 0.01
    start = (np.arange(n_rays) * m).astype(np.uint32)
    ns = np.full(n_rays, m, np.uint32)
    S = n_rays * m
    dss = np.full(S, ds, np.float32)
    z = np.tile(3.0 + np.arange(m) * ds, n_rays).astype(np.float32)
    c = np.array([0.2, 0.5, 0.9], np.float32)
    drgbs = np.zeros((S, 4), np.float32)
    drgbs[:, 1:] = c
    for i, s in enumerate(sigmas):
        drgbs[i * m:(i + 1) * m, 0] = s
    bg = np.array([1.0, 1.0, 1.0], np.float32)
    measured, rgbd, op = oracle.integrate_rays(start, ns, bg, dss, z, drgbs)
    tot = 0
    for i, s in enumerate(sigmas):
        s32 = float(np.float32(s) * np.float32(ds))
        k_stop = next((k for k in range(1, m + 1) if math.exp(-s32 * k) <= 1e-4), m)
        boundary = abs(math.exp(-s32 * k_stop) - 1e-4) < 1e-9
        T = math.exp(-s32 * k_stop)
        if T <= 1e-4:
            want = c
        else:
            want = c * (1 - T) + T * bg
        np.testing.assert_allclose(rgbd[i, :3], want, rtol=2e-5, atol=2e-6)
        np.testing.assert_allclose(op[i], 1 - T, rtol=1e-5, atol=1e-6)
        tot += k_stop
        assert not boundary
    assert measured == tot


def test_integrate_f32_matches_float64(rng):
    start, ns, bgs, dss, z, drgbs = integrate_inputs(rng, 300, 80)
    m, rgbd, op = oracle.integrate_rays(start, ns, bgs, dss, z, drgbs)
    rgbd64, op64, stops = oracle.integrate_rays64(start, ns, bgs, dss, z, drgbs)
    # total composited samples (f32 early stops) equals the float64 count on these inputs
    assert int(m) == int(stops.sum())
    np.testing.assert_allclose(rgbd, rgbd64, rtol=5e-5, atol=5e-5)
    np.testing.assert_allclose(op, op64, rtol=5e-5, atol=5e-6)


def test_integrate_inference_chunks_equal_single_pass(rng):
    cap = 8
    start, ns, bgs, dss, z, drgbs = integrate_inputs(rng, 50, 60, total_pad=0)
    _, rgbd_train, op_train = oracle.integrate_rays(start, ns, bgs, dss, z, drgbs)
    n = len(ns)
    rgbd = np.zeros((n, 4), np.float32)
    T = np.ones(n, np.float32)
    done = np.zeros(n, bool)
    for c0 in range(0, int(ns.max()) + cap, cap):
        k = np.clip(ns.astype(np.int64) - c0, 0, cap).astype(np.uint32)
        idx = np.where(done, n, np.arange(n)).astype(np.uint32)
        cd = np.zeros((n, cap)); cz = np.zeros((n, cap)); cc = np.zeros((n, cap, 4))
        for i in range(n):
            s = start[i] + c0
            cd[i, :k[i]] = dss[s:s + k[i]]; cz[i, :k[i]] = z[s:s + k[i]]; cc[i, :k[i]] = drgbs[s:s + k[i]]
        cnt, term, out, Tout = oracle.integrate_rays_inference(bgs, rgbd, T, k, idx, cd, cz, cc)
        for i in range(n):
            if not done[i]:
                rgbd[i], T[i] = out[i], Tout[i]
        done |= term
        if done.all():
            break
    assert done.all()
    np.testing.assert_array_equal(rgbd, rgbd_train)


def _pool_render(o, d, t0, t1, bits, cap, pool, sigma_scale):
    """The inference ray-pool loop (models/renderers/cuda.py:319-361) over the oracle's
    march_rays_inference / integrate_rays_inference with a closed-form field.
    Returns (rays_rgbd, number of far-plane samples, number of full chunks that
    deferred their far-plane sample)."""
    n_tot = o.shape[0]
    rgbd_all = np.zeros((n_tot, 4), np.float32)
    T_all = np.ones(n_tot, np.float32)
    bg = np.ones((n_tot, 3), np.float32)
    ts = t0.copy()
    terminated = np.ones(pool, bool)
    indices = np.zeros(pool, np.uint32)
    nri, done, far, deferred = np.uint32(0), 0, 0, 0
    for _ in range(20000):
        out = oracle.march_rays_inference(1024, 1, 128, cap, 1.0, 0.0, o, d, ts, t1, bits, nri, terminated, indices)
        nri, indices = out["next_ray_index"], out["indices"]
        live = indices < n_tot
        r = indices[live]
        ts[r] = out["t_starts_out"][live]
        k = out["n_samples"]
        full = k[live] == cap
        deferred += int((full & (out["t_starts_out"][live] == t1[r]) & (out["z_vals"][live, -1] < t1[r])).sum())
        far += int(sum((out["z_vals"][i, :k[i]] == t1[indices[i]]).sum() for i in np.nonzero(live)[0]))
        x = out["xyzs"].astype(np.float32)
        drgbs = np.zeros(x.shape[:2] + (4,), np.float32)
        drgbs[..., 0] = np.float32(sigma_scale) * np.abs(np.sin(np.float32(5.0) * x[..., 0]) *
                                                         np.cos(np.float32(3.0) * x[..., 1]))
        drgbs[..., 1:] = np.float32(0.5) + np.float32(0.5) * np.sin(np.float32(4.0) * x)
        cnt, terminated, rgbd, T = oracle.integrate_rays_inference(bg, rgbd_all, T_all, k, indices, out["dss"],
                                                                   out["z_vals"], drgbs)
        rgbd_all[r] = rgbd[live]
        T_all[r] = T[live]
        done += int(cnt)
        if done >= n_tot:
            return rgbd_all, far, deferred
    raise AssertionError("pool loop did not finish")


def test_inference_render_independent_of_steps_cap(rng):
    """A ray's colour from the ray pool does not depend on march_steps_cap: the march
    resumes from the t it stopped at, a far-plane sample that does not fit is written
    in the next round, and the compositing carries (rgbd, T) between rounds in f32 in
    the same order.  The one exception is the exact tie pinned by
    test_inference_far_plane_tie_is_marched_twice.  The proposal renderer's
    occupancy-grid fallback relies on this to march more samples per round than the
    plain render and give the same pixels.  Two scenes: blobs of occupied cells with a
    dense field (early termination), and every cell occupied with a faint field, where
    rays leave the box in an occupied cell (far-plane samples, deferred when the chunk
    is full)."""
    n = 96
    o, d, t0, t1, _, bits, _ = march_inputs(rng, n)
    full_bits = np.full_like(bits, 0xFF)
    for b, sigma_scale, cases in ((bits, 40.0, ((1, 40), (3, 17), (32, 40), (64, 96))),
                                  (full_bits, 0.3, ((2, 96), (5, 48), (32, 96)))):
        ref, far, _ = _pool_render(o, d, t0, t1, b, 8, 40, sigma_scale)
        assert (ref[:, :3] != 1.0).any()  # some rays composite samples
        deferred = 0
        for cap, pool in cases:
            got, far_c, deferred_c = _pool_render(o, d, t0, t1, b, cap, pool, sigma_scale)
            np.testing.assert_array_equal(got, ref)
            assert far_c == far
            deferred += deferred_c
        if b is full_bits:
            assert far > 0 and deferred > 0


def test_inference_far_plane_tie_is_marched_twice():
    """The reference's ray pool writes the far-plane sample twice when the last
    regular sample ends exactly on t_end (z + ds == t_end in f32) and the far-plane
    sample fills the round's last slot: the round returns march_steps_cap samples, so
    the integrator keeps the ray (integrating.cu:300), the next round starts at
    t_start == t_end, which the marcher does not reject (marching.cu:317 rejects only
    t_end < t_start), and its far-plane block (marching.cu:367-394) writes the sample
    again.  With any other cap, or with t_end one ulp later, the sample is written
    once.  The kernels keep this behaviour (it is the reference's output); it is the
    only way a pixel of the pool render depends on march_steps_cap."""
    o = np.array([-0.9, -0.05, 0.1], np.float32)
    d = np.array([1.0, 0.0, 0.0], np.float32)
    bits = np.full(128 ** 3 // 8, 0xFF, np.uint8)
    t0 = np.float32(0.0)
    # the step chain t_{k+1} = t_k + ds of the oracle's own march (all cells occupied)
    chain, _ = march_all_rounds(o, d, t0, np.float32(1.5), bits, 64)
    assert len(chain) > 20 and np.all(np.diff(chain) > 0)
    t_end = chain[7]  # the 7th sample ends exactly here
    z8, rounds8 = march_all_rounds(o, d, t0, t_end, bits, 8)
    z16, rounds16 = march_all_rounds(o, d, t0, t_end, bits, 16)
    # 7 regular samples + the far-plane sample, once at cap 16, twice at cap 8
    assert rounds16 == [8] and rounds8 == [8, 1]
    np.testing.assert_array_equal(z16, np.append(chain[:7], t_end))
    np.testing.assert_array_equal(z8, np.append(z16, t_end))
    # one ulp further: the 8th regular sample fits, the far-plane sample is deferred
    # to the next round at cap 8, and both caps march the same samples
    t_late = np.nextafter(t_end, np.float32(np.inf))
    z8, rounds8 = march_all_rounds(o, d, t0, t_late, bits, 8)
    z16, rounds16 = march_all_rounds(o, d, t0, t_late, bits, 16)
    assert rounds8 == [8, 1] and rounds16 == [9]
    np.testing.assert_array_equal(z8, z16)
    np.testing.assert_array_equal(z16, np.append(chain[:8], t_late))


def _fd_loss(startidx, ns, bgs, dss, z, drgbs, g):
    rgbd, _, _ = oracle.integrate_rays64(startidx, ns, bgs, dss, z, drgbs)
    return float((rgbd * g).sum())


def test_integrate_backward64_matches_finite_differences(rng):
    """Regime where the reference VJP is close to the true gradient: z >= 1,
    z >= near, final transmittance > 1e-4 (no early stop, no renormalisation).

    One difference remains there: the reference's density gradient subtracts the
    background twice for rays that did not terminate (integrating.cu:201-214 uses
    final_rgbd, which already holds T_N*bg, and also subtracts bg*(1-opacity) =
    T_N*bg).  So dL/dsigma_i(reference) = FD - ds_i * T_N * <g_rgb, bg>, which this
    test checks exactly; the other channels equal the finite differences."""
    n = 12
    start, ns, bgs, dss, z, drgbs = integrate_inputs(rng, n, 20, total_pad=0)
    drgbs = drgbs.astype(np.float64)
    drgbs[:, 0] = rng.uniform(0.0, 15.0, len(dss))
    rgbd, op, stops = oracle.integrate_rays64(start, ns, bgs, dss, z, drgbs)
    assert (1 - op > 2e-4).all()
    g = rng.normal(size=(n, 4))
    dbg, dz, dd = oracle.integrate_rays_backward64(0.3, start, ns, bgs, dss, z, drgbs, rgbd, op, g)
    h = 1e-6
    picks = rng.choice(len(dss), size=min(32, len(dss)), replace=False)
    for s in picks:
        for ch in range(4):
            p, m = drgbs.copy(), drgbs.copy()
            p[s, ch] += h
            m[s, ch] -= h
            fd = (_fd_loss(start, ns, bgs, dss, z, p, g) - _fd_loss(start, ns, bgs, dss, z, m, g)) / (2 * h)
            if ch == 0:
                ray = int(np.searchsorted(start.astype(np.int64), s, side="right") - 1)
                while ns[ray] == 0 or not (start[ray] <= s < start[ray] + ns[ray]):
                    ray -= 1
                fd = fd - float(dss[s]) * (1.0 - op[ray]) * float(np.dot(g[ray, :3], bgs[ray]))
            assert abs(fd - dd[s, ch]) <= 1e-6 + 1e-5 * abs(fd), (s, ch, fd, dd[s, ch])
        zp, zm = z.astype(np.float64).copy(), z.astype(np.float64).copy()
        zp[s] += h
        zm[s] -= h
        fd = (_fd_loss(start, ns, bgs, dss, zp, drgbs, g) - _fd_loss(start, ns, bgs, dss, zm, drgbs, g)) / (2 * h)
        assert abs(fd - dz[s]) <= 1e-6 + 1e-5 * abs(fd)
    for i in range(n):
        for ch in range(3):
            bp, bm = bgs.astype(np.float64).copy(), bgs.astype(np.float64).copy()
            bp[i, ch] += h
            bm[i, ch] -= h
            fd = (_fd_loss(start, ns, bp, dss, z, drgbs, g) - _fd_loss(start, ns, bm, dss, z, drgbs, g)) / (2 * h)
            assert abs(fd - dbg[i, ch]) <= 1e-6 + 1e-5 * abs(fd)


def test_integrate_backward_f32_matches_float64(rng):
    start, ns, bgs, dss, z, drgbs = integrate_inputs(rng, 200, 60)
    _, rgbd, op = oracle.integrate_rays(start, ns, bgs, dss, z, drgbs)
    g = rng.normal(size=(200, 4)).astype(np.float32)
    z_near = z.copy()
    dbg, dz, dd = oracle.integrate_rays_backward(3.0, start, ns, bgs, dss, z_near, drgbs, rgbd, op, g)
    _, _, stops = oracle.integrate_rays64(start, ns, bgs, dss, z, drgbs)
    dbg64, dz64, dd64 = oracle.integrate_rays_backward64(3.0, start, ns, bgs, dss, z_near, drgbs, rgbd, op, g,
                                                         stop_at=stops)
    np.testing.assert_allclose(dbg, dbg64, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(dz, dz64, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(dd, dd64, rtol=1e-3, atol=2e-5)
    # the near-plane regulariser is present exactly where sigma > 4e-5 and z < near
    assert ((dd[:, 0] != 0) | (dd64[:, 0] == 0)).all()


# ---------------------------------------------------------------------------
# MLP float64 transcription
# ---------------------------------------------------------------------------

def test_mlp_backward64_matches_finite_differences(rng):
    S = 5
    params = rng.normal(0, 0.3, 9408)
    enc = rng.normal(0, 1, (S, 32))
    dirs = rng.normal(size=(S, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    g = rng.normal(size=(S, 4))
    dX, dP = oracle.mlp_backward64(enc, dirs, params, g)
    f = lambda e, p: float((oracle.mlp_forward64(e, dirs, p)[0] * g).sum())
    h = 1e-6
    for j in rng.choice(9408, 48, replace=False):
        pp, pm = params.copy(), params.copy()
        pp[j] += h
        pm[j] -= h
        fd = (f(enc, pp) - f(enc, pm)) / (2 * h)
        assert abs(fd - dP[j]) <= 1e-5 * max(1, abs(fd)), j
    for s in range(S):
        for k in range(0, 32, 5):
            ep, em = enc.copy(), enc.copy()
            ep[s, k] += h
            em[s, k] -= h
            fd = (f(ep, params) - f(em, params)) / (2 * h)
            assert abs(fd - dX[s, k]) <= 1e-5 * max(1, abs(fd))


def test_sh4_orthonormality():
    """The 16 real SH basis functions are orthonormal on the sphere (Monte Carlo)."""
    rng = np.random.default_rng(7)
    d = rng.normal(size=(400000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Y = oracle.sh4_64(d)
    gram = 4 * np.pi * (Y.T @ Y) / len(d)
    np.testing.assert_allclose(gram, np.eye(16), atol=0.02)


def _real_sh_scipy(d):
    """Real spherical harmonics of degree < 4 from scipy's complex ones (which carry
    the Condon-Shortley phase): column l^2 + l + m holds sqrt(2) Im Y_l^|m| for m < 0,
    Y_l^0 for m = 0 and sqrt(2) Re Y_l^m for m > 0.  This is the sign convention of
    models/encoders.py:384-405 (column 1 = -sqrt(3/4pi) y, column 3 = -sqrt(3/4pi) x)."""
    from scipy.special import sph_harm_y
    theta = np.arccos(np.clip(d[:, 2], -1.0, 1.0))  # polar angle from +z
    phi = np.arctan2(d[:, 1], d[:, 0])  # azimuth from +x
    out = np.zeros((len(d), 16))
    for l in range(4):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            if m < 0:
                v = math.sqrt(2) * y.imag
            elif m == 0:
                v = y.real
            else:
                v = math.sqrt(2) * y.real
            out[:, l * l + l + m] = v
    return out


def test_sh4_matches_scipy_real_harmonics():
    """All 16 SH(4) terms against scipy's spherical harmonics at random directions and
    at the six axis directions and the eight cube diagonals: a flipped sign, a swapped
    pair of terms or a wrong coefficient fails here (orthonormality alone does not
    see signs or permutations)."""
    rng = np.random.default_rng(11)
    d = rng.normal(size=(2000, 3))
    axes = np.concatenate([np.eye(3), -np.eye(3)])
    diag = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], np.float64)
    d = np.concatenate([d, axes, diag])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    np.testing.assert_allclose(oracle.sh4_64(d), _real_sh_scipy(d), rtol=0, atol=1e-12)
    # the reference's own annotations of the first terms, as closed forms
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    c1 = math.sqrt(3) / (2 * math.sqrt(math.pi))
    Y = oracle.sh4_64(d)
    np.testing.assert_allclose(Y[:, 1:4], np.stack([-c1 * y, c1 * z, -c1 * x], 1), atol=1e-15)


# ---- learned proposal (defined by this build; pinned by calculus and closed forms) ----

def test_proposal_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(31)
    N = 7
    enc = rng.normal(0, 0.7, (N, 10))
    p = rng.normal(0, 0.4, oracle.PROP_N_PARAMS)
    g = rng.normal(size=N)
    dX, dP = oracle.proposal_mlp_backward64(enc, p, g)
    f = lambda e, q: float((oracle.proposal_mlp_forward64(e, q)[0] * g).sum())
    h = 1e-6
    for k in rng.choice(oracle.PROP_N_PARAMS, 24, replace=False):
        d = np.zeros_like(p)
        d[k] = h
        assert abs((f(enc, p + d) - f(enc, p - d)) / (2 * h) - dP[k]) <= 1e-6 * (1 + abs(dP[k]))
    for (i, j) in [(0, 0), (3, 9), (6, 4)]:
        d = np.zeros_like(enc)
        d[i, j] = h
        assert abs((f(enc + d, p) - f(enc - d, p)) / (2 * h) - dX[i, j]) <= 1e-6 * (1 + abs(dX[i, j]))


def _prop_rays(rng, n_rays=5):
    ns = rng.integers(1, 120, n_rays).astype(np.uint32)
    start = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.uint32)
    S = int(ns.sum())
    t0 = rng.uniform(2.0, 3.0, n_rays).astype(np.float32)
    t1 = (t0 + rng.uniform(1.0, 2.5, n_rays)).astype(np.float32)
    z = np.zeros(S, np.float32)
    dss = np.zeros(S, np.float32)
    for r in range(n_rays):
        zz = np.sort(rng.uniform(t0[r], t1[r], ns[r]))
        z[start[r]:start[r] + ns[r]] = zz
        dss[start[r]:start[r] + ns[r]] = 0.0034
    drgbs = np.concatenate([rng.uniform(0, 60, (S, 1)), rng.uniform(0, 1, (S, 3))], 1).astype(np.float32)
    return start, ns, dss, z, drgbs, t0, t1


def test_nerf_sample_weights_sum_to_opacity():
    rng = np.random.default_rng(2)
    start, ns, dss, z, drgbs, t0, t1 = _prop_rays(rng)
    w = oracle.nerf_sample_weights64(start, ns, dss, drgbs)
    _, op, _ = oracle.integrate_rays64(start, ns, np.zeros(3), dss, z, drgbs)
    for r in range(len(start)):
        assert abs(w[start[r]:start[r] + ns[r]].sum() - op[r]) <= 1e-12


def test_proposal_loss_gradient_and_closed_forms():
    rng = np.random.default_rng(3)
    start, ns, dss, z, drgbs, t0, t1 = _prop_rays(rng)
    n = len(start)
    valid = np.array([True, True, False, True, True])
    sig = rng.uniform(0, 30, (n, oracle.PROP_BINS))
    loss, grad = oracle.proposal_loss64(start, ns, dss, z, drgbs, t0, t1, sig, valid, 4)
    assert loss > 0 and (grad[2] == 0).all()
    h = 1e-6
    for r, b in [(0, 0), (1, 17), (3, 40), (4, 63), (0, 31)]:
        d = np.zeros_like(sig)
        d[r, b] = h
        fd = (oracle.proposal_loss64(start, ns, dss, z, drgbs, t0, t1, sig + d, valid, 4)[0] -
              oracle.proposal_loss64(start, ns, dss, z, drgbs, t0, t1, sig - d, valid, 4)[0]) / (2 * h)
        assert abs(fd - grad[r, b]) <= 1e-6 * (1 + abs(grad[r, b])), (r, b, fd, grad[r, b])
    # a proposal that is opaque from the first bin bounds nothing after it; one that puts
    # at least the NeRF's weight in every bin has zero loss
    empty = np.zeros(ns.sum(), np.float32)
    l0, g0 = oracle.proposal_loss64(start, ns, dss, z, np.concatenate([empty[:, None], drgbs[:, 1:]], 1), t0, t1,
                                    sig, valid, 4)
    assert l0 == 0 and (g0 == 0).all()


def test_proposal_sample_quantiles():
    """Closed forms of the quantile placement: a constant density (sigma tiny, so the
    bin weights are nearly equal) puts the samples at evenly spaced quantiles of the
    segment; mass below the threshold marks the fallback; the samples are sorted and
    the first and last lie inside the bins holding the 1/(2n) and 1 - 1/(2n) quantiles."""
    n = 48
    t0, t1 = np.array([2.0, 2.0, 1.0]), np.array([4.0, 2.0, 1.5])
    sig = np.zeros((3, 64))
    sig[0] = 400.0  # opaque: mass ~ 1, concentrated in the first bins
    sig[2] = 0.5  # nearly transparent: mass << 0.9
    counts, fb, t, ds = oracle.proposal_sample64(t0, t1, sig, n)
    assert list(counts) == [n, 0, 0] and list(fb) == [False, False, True]
    assert (np.diff(t[0]) >= 0).all() and t[0, 0] >= 2.0 and t[0, -1] <= 4.0
    # uniform weights: sigma such that every bin holds the same weight is not constant;
    # instead use the exact inverse: w_b = 1/64 needs T_b a_b = 1/64
    w = np.full(64, 1.0 / 64)
    T = 1.0 - np.concatenate([[0.0], np.cumsum(w)[:-1]])
    a = np.minimum(w / T, 1.0)
    D = 2.0 / 64
    sig_u = -np.log1p(-np.clip(a, 0, 1 - 1e-15)) / D
    counts, fb, t, ds = oracle.proposal_sample64(np.array([2.0]), np.array([4.0]), sig_u[None], n)
    want = 2.0 + (np.arange(n) + 0.5) / n * 2.0
    assert np.abs(t[0] - want).max() <= 1e-9
    assert np.abs(ds[0, :-1] - 2.0 / n).max() <= 1e-9


def test_proposal_sample_stratified_jitter():
    """Stratified training samples: with equal bin weights, sample j of a ray with
    jitter u sits at t0 + (j + u) / n (t1 - t0); jitter 1/2 is the render placement;
    a zero-mass ray with threshold 0 gets no samples (its fallback flag set) instead
    of dividing by zero."""
    n = 48
    w = np.full(64, 1.0 / 64)
    T = 1.0 - np.concatenate([[0.0], np.cumsum(w)[:-1]])
    a = np.minimum(w / T, 1.0)
    D = 2.0 / 64
    sig_u = -np.log1p(-np.clip(a, 0, 1 - 1e-15)) / D
    jit = np.array([0.0, 0.25, 0.5, 0.999])
    k = len(jit)
    counts, fb, t, ds = oracle.proposal_sample64(np.full(k, 2.0), np.full(k, 4.0), np.tile(sig_u, (k, 1)), n, 0.0,
                                                 jitter=jit)
    assert list(counts) == [n] * k and not fb.any()
    want = 2.0 + (np.arange(n)[None, :] + jit[:, None]) / n * 2.0
    assert np.abs(t - want).max() <= 1e-9
    assert np.abs(ds[:, :-1] - 2.0 / n).max() <= 1e-9
    np.testing.assert_array_equal(t[2], oracle.proposal_sample64(np.array([2.0]), np.array([4.0]), sig_u[None], n)[2][0])
    c0, fb0, _, _ = oracle.proposal_sample64(np.array([2.0]), np.array([4.0]), np.zeros((1, 64)), n, 0.0)
    assert c0[0] == 0 and fb0[0]


def test_proposal_sample_support():
    """Support placement: equal bin weights cover the whole segment, so the samples are
    evenly spaced over [t0, t1); an opaque first bin confines them to that bin."""
    n = 48
    w = np.full(64, 1.0 / 64)
    T = 1.0 - np.concatenate([[0.0], np.cumsum(w)[:-1]])
    a = np.minimum(w / T, 1.0)
    D = 2.0 / 64
    sig_u = -np.log1p(-np.clip(a, 0, 1 - 1e-15)) / D
    sig_o = np.zeros(64)
    sig_o[0] = 1e4
    counts, fb, t, ds = oracle.proposal_sample_support64(np.array([2.0, 2.0]), np.array([4.0, 4.0]),
                                                         np.stack([sig_u, sig_o]), n)
    assert list(counts) == [n, n] and not fb.any()
    assert np.abs(t[0] - (2.0 + (np.arange(n) + 0.5) / n * 2.0)).max() <= 1e-12
    assert np.abs(ds[0] - 2.0 / n).max() <= 1e-12
    assert t[1, 0] > 2.0 and t[1, -1] < 2.0 + D and np.abs(ds[1] - D / n).max() <= 1e-12
