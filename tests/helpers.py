"""Seeded inputs shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

SQRT3 = np.float32(1.7320508075688772)


def near_far_np(o, d, bound):
    """f32 AABB entry/exit times, restating models/renderers/cuda.py:56-97."""
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    eps = np.float32(1e-15)
    d = np.where(np.signbit(d), np.minimum(d, -eps), np.maximum(d, eps)).astype(np.float32)
    b = np.float32(bound)
    t0 = (-b - o) / d
    t1 = (b - o) / d
    ts = np.minimum(t0, t1).max(axis=1)
    te = np.maximum(t0, t1).min(axis=1)
    return np.maximum(np.float32(0), ts).astype(np.float32), te.astype(np.float32)


def camera_rays(rng, n, radius=4.0, spread=0.35):
    """Rays from random points on a sphere of `radius` aimed near the origin."""
    v = rng.normal(size=(n, 3))
    o = (v / np.linalg.norm(v, axis=1, keepdims=True) * radius).astype(np.float32)
    target = rng.uniform(-spread, spread, size=(n, 3))
    d = target - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o, d


def blob_bitfield(rng, G=128, K=1, n_blobs=6, fill=None):
    """Occupancy bitfield (Morton order) of a few random balls; returns (bits, dense bool [K,G,G,G])."""
    import oracle

    occ = np.zeros((K, G, G, G), bool)
    ax = (np.arange(G) + 0.5) / G * 2 - 1
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    for _ in range(n_blobs):
        c = rng.uniform(-0.6, 0.6, 3)
        r = rng.uniform(0.1, 0.35)
        occ[0] |= (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 < r * r
    if fill is not None:
        occ[:] = fill
    for k in range(1, K):
        occ[k] = occ[0]
    # dense [x, y, z] -> Morton order
    xs, ys, zs = np.meshgrid(np.arange(G), np.arange(G), np.arange(G), indexing="ij")
    coords = np.stack([xs.ravel(), ys.ravel(), zs.ravel()], 1).astype(np.uint32)
    m = oracle.morton3d(coords)
    flat = np.zeros(K * G ** 3, bool)
    for k in range(K):
        flat[k * G ** 3 + m] = occ[k].ravel()
    bits = np.packbits(flat.reshape(-1, 8), axis=-1, bitorder="little").reshape(-1)
    return bits, occ


def march_inputs(rng, n_rays, bound=1.0, G=128, K=1, **kw):
    o, d = camera_rays(rng, n_rays)
    t0, t1 = near_far_np(o, d, bound)
    noise = rng.uniform(0, 1, n_rays).astype(np.float32)
    bits, occ = blob_bitfield(rng, G=G, K=K, **kw)
    return o, d, t0, t1, noise, bits, occ


def integrate_inputs(rng, n_rays, max_samples=64, contiguous=True, total_pad=7):
    """Random rays with contiguous sample ranges and random densities/colours."""
    ns = rng.integers(0, max_samples + 1, n_rays).astype(np.uint32)
    start = np.zeros(n_rays, np.uint32)
    start[1:] = np.cumsum(ns)[:-1]
    S = int(ns.sum()) + total_pad
    dss = rng.uniform(0.002, 0.02, S).astype(np.float32)
    z = np.zeros(S, np.float32)
    for i in range(n_rays):
        s0 = start[i]
        z[s0:s0 + ns[i]] = (2.5 + np.cumsum(dss[s0:s0 + ns[i]])).astype(np.float32)
    sig = np.log1p(np.exp(rng.normal(0, 2, S))).astype(np.float32) * 20
    rgb = rng.uniform(0, 1, (S, 3)).astype(np.float32)
    drgbs = np.concatenate([sig[:, None], rgb], 1).astype(np.float32)
    bgs = rng.uniform(0, 1, (n_rays, 3)).astype(np.float32)
    return start, ns, bgs, dss, z, drgbs


def march_all_rounds(o, d, t0, t1, bits, cap, K=1, G=128, diag=1024, bound=1.0, portion=0.0):
    """Every sample the ray pool marches for one ray (the oracle's march), round after round, until the
    integrator's rule for a transparent ray ends it (integrating.cu:300: terminated
    once a round returns fewer than march_steps_cap samples).  Returns the z values
    and the samples per round."""
    import oracle

    ts = np.array([t0], np.float32)
    te = np.array([t1], np.float32)
    term, idx, nri = np.ones(1, bool), np.zeros(1, np.uint32), np.uint32(0)
    zs, per_round = [], []
    for _ in range(100000):
        out = oracle.march_rays_inference(diag, K, G, cap, bound, portion, o[None], d[None], ts, te, bits, nri, term,
                                          idx)
        nri, idx = out["next_ray_index"], out["indices"]
        if idx[0] >= 1:
            return np.array(zs, np.float32), per_round
        k = int(out["n_samples"][0])
        zs += list(out["z_vals"][0, :k])
        per_round.append(k)
        ts[0] = out["t_starts_out"][0]
        term = np.array([k < cap])
    raise AssertionError("ray did not finish")
