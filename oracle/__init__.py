"""CPU oracle of the NeRF hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline; the product
package (jaxngp_amd) never imports it.

Two layers:
  * ``liboracle.so`` (oracle.c): an f32 restatement of the reference kernels under
    a fixed rounding convention (no contraction, IEEE division, a deterministic
    exp), processing rays in index order.  The HIP kernels follow the same
    convention, so integer outputs (sample counts, offsets, cells, early-stop
    indices, hash indices) and most f32 outputs match bit for bit.
  * numpy float64 transcriptions (this file) of the floating-point parts, used to
    bound the error of the f32 results: integration forward/backward, the hash
    encoder (vertices from the f32 path, weights and sums in f64), the NeRF MLP.

Each function cites the reference lines it restates.  How each is pinned
independently of the reference (which cannot run here: CUDA-only kernels and a
JAX-only encoder) is listed in tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes
import math
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.ngp_expf.argtypes = [ctypes.c_float]
        _lib.ngp_expf.restype = ctypes.c_float
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


u32 = ctypes.c_uint32
f32 = ctypes.c_float


# ---------------------------------------------------------------------------
# f32 restatement (liboracle.so)
# ---------------------------------------------------------------------------

def expf(x: float) -> float:
    return lib().ngp_expf(f32(x))


def morton3d(xyzs) -> np.ndarray:
    xyzs = _c(xyzs, np.uint32)
    out = np.empty(xyzs.shape[0], np.uint32)
    lib().oracle_morton3d(u32(xyzs.shape[0]), _p(xyzs), _p(out))
    return out


def morton3d_invert(idcs) -> np.ndarray:
    idcs = _c(idcs, np.uint32)
    out = np.empty((idcs.shape[0], 3), np.uint32)
    lib().oracle_morton3d_invert(u32(idcs.shape[0]), _p(idcs), _p(out))
    return out


def packbits(thr, density):
    density = _c(density, np.float32)
    thr = _c(np.broadcast_to(np.asarray(thr, np.float32), density.shape), np.float32)
    n_bytes = density.shape[0] // 8
    mask = np.empty(density.shape[0], np.uint8)
    bits = np.empty(n_bytes, np.uint8)
    lib().oracle_packbits(u32(n_bytes), _p(thr), _p(density), _p(mask), _p(bits))
    return mask.astype(bool), bits


def march_rays(total_samples, diagonal_n_steps, K, G, bound, stepsize_portion, rays_o, rays_d, t_starts, t_ends,
               noises, occupancy_bitfield):
    """Sequential (ray-order) restatement of marching.cu:101-268 + marching/__init__.py:70-93."""
    rays_o = _c(rays_o, np.float32)
    rays_d = _c(rays_d, np.float32)
    n = rays_o.shape[0]
    t_starts = _c(t_starts, np.float32)
    t_ends = _c(t_ends, np.float32)
    noises = _c(np.broadcast_to(np.asarray(noises, np.float32), (n,)), np.float32)
    bits = _c(occupancy_bitfield, np.uint8)
    S = int(total_samples)
    nxt = np.zeros(1, np.uint32)
    exc = np.zeros(1, np.uint32)
    valid = np.zeros(n, np.uint8)
    nsamp = np.zeros(n, np.uint32)
    start = np.zeros(n, np.uint32)
    idcs = np.zeros(S, np.uint32)
    xyzs = np.zeros((S, 3), np.float32)
    dirs = np.zeros((S, 3), np.float32)
    dss = np.zeros(S, np.float32)
    z = np.zeros(S, np.float32)
    lib().oracle_march_rays(u32(n), u32(S), u32(diagonal_n_steps), u32(K), u32(G), f32(bound), f32(stepsize_portion),
                            _p(rays_o), _p(rays_d), _p(t_starts), _p(t_ends), _p(noises), _p(bits), _p(nxt), _p(exc),
                            _p(valid), _p(nsamp), _p(start), _p(idcs), _p(xyzs), _p(dirs), _p(dss), _p(z))
    measured = np.uint32(int(nxt[0]) - int(exc[0]))
    return dict(measured=measured, next_sample_write_location=nxt[0], number_of_exceeded_samples=exc[0],
                ray_is_valid=valid.astype(bool), rays_n_samples=nsamp, rays_sample_startidx=start, idcs=idcs,
                xyzs=xyzs, dirs=dirs, dss=dss, z_vals=z)


def march_rays_inference(diagonal_n_steps, K, G, march_steps_cap, bound, stepsize_portion, rays_o, rays_d, t_starts,
                         t_ends, occupancy_bitfield, next_ray_index_in, terminated, indices):
    """Slot-order restatement of marching.cu:271-397; t_starts returned raw (per slot)."""
    rays_o = _c(rays_o, np.float32)
    rays_d = _c(rays_d, np.float32)
    nt = rays_o.shape[0]
    terminated = _c(terminated, np.uint8)
    n = terminated.shape[0]
    cap = int(march_steps_cap)
    nri = _c(np.asarray(next_ray_index_in).reshape(1), np.uint32)
    idx_in = _c(indices, np.uint32)
    nxt = np.zeros(1, np.uint32)
    idx = np.zeros(n, np.uint32)
    nsamp = np.zeros(n, np.uint32)
    tout = np.zeros(n, np.float32)
    xyzs = np.zeros((n, cap, 3), np.float32)
    dss = np.zeros((n, cap), np.float32)
    z = np.zeros((n, cap), np.float32)
    lib().oracle_march_rays_inference(u32(nt), u32(n), u32(diagonal_n_steps), u32(K), u32(G), u32(cap), f32(bound),
                                      f32(stepsize_portion), _p(rays_o), _p(rays_d), _p(_c(t_starts, np.float32)),
                                      _p(_c(t_ends, np.float32)), _p(_c(occupancy_bitfield, np.uint8)), _p(nri),
                                      _p(terminated), _p(idx_in), _p(nxt), _p(idx), _p(nsamp), _p(tout), _p(xyzs),
                                      _p(dss), _p(z))
    return dict(next_ray_index=nxt[0], indices=idx, n_samples=nsamp, t_starts_out=tout, xyzs=xyzs, dss=dss, z_vals=z)


def integrate_rays(startidx, nsamp, bgs, dss, z_vals, drgbs):
    """f32 restatement of integrating.cu:24-104."""
    startidx = _c(startidx, np.uint32)
    nsamp = _c(nsamp, np.uint32)
    n = startidx.shape[0]
    bgs = _c(np.broadcast_to(np.asarray(bgs, np.float32), (n, 3)), np.float32)
    dss = _c(dss, np.float32)
    S = dss.shape[0]
    measured = np.zeros(1, np.uint32)
    rgbd = np.zeros((n, 4), np.float32)
    op = np.zeros(n, np.float32)
    lib().oracle_integrate_rays(u32(n), u32(S), _p(startidx), _p(nsamp), _p(bgs), _p(dss),
                                _p(_c(z_vals, np.float32)), _p(_c(drgbs, np.float32)), _p(measured), _p(rgbd), _p(op))
    return measured[0], rgbd, op


def integrate_rays_backward(near_distance, startidx, nsamp, bgs, dss, z_vals, drgbs, final_rgbds, final_opacities,
                            dL_dfinal_rgbds):
    """f32 restatement of integrating.cu:106-240."""
    startidx = _c(startidx, np.uint32)
    nsamp = _c(nsamp, np.uint32)
    n = startidx.shape[0]
    bgs = _c(np.broadcast_to(np.asarray(bgs, np.float32), (n, 3)), np.float32)
    dss = _c(dss, np.float32)
    S = dss.shape[0]
    dbg = np.zeros((n, 3), np.float32)
    dz = np.zeros(S, np.float32)
    dd = np.zeros((S, 4), np.float32)
    lib().oracle_integrate_rays_backward(u32(n), u32(S), f32(near_distance), _p(startidx), _p(nsamp), _p(bgs),
                                         _p(dss), _p(_c(z_vals, np.float32)), _p(_c(drgbs, np.float32)),
                                         _p(_c(final_rgbds, np.float32)), _p(_c(final_opacities, np.float32)),
                                         _p(_c(dL_dfinal_rgbds, np.float32)), _p(dbg), _p(dz), _p(dd))
    return dbg, dz, dd


def integrate_rays_inference(rays_bg, rays_rgbd, rays_T, n_samples, indices, dss, z_vals, drgbs):
    """f32 restatement of integrating.cu:242-322 (per-slot outputs, no scatter)."""
    rays_rgbd = _c(rays_rgbd, np.float32)
    ntot = rays_rgbd.shape[0]
    dss = _c(dss, np.float32)
    n, cap = dss.shape
    cnt = np.zeros(1, np.uint32)
    term = np.zeros(n, np.uint8)
    rgbd = np.zeros((n, 4), np.float32)
    T = np.zeros(n, np.float32)
    lib().oracle_integrate_rays_inference(u32(ntot), u32(n), u32(cap), _p(_c(rays_bg, np.float32)), _p(rays_rgbd),
                                          _p(_c(rays_T, np.float32)), _p(_c(n_samples, np.uint32)),
                                          _p(_c(indices, np.uint32)), _p(dss), _p(_c(z_vals, np.float32)),
                                          _p(_c(drgbs, np.float32)), _p(cnt), _p(term), _p(rgbd), _p(T))
    return cnt[0], term.astype(bool), rgbd, T


# ---------------------------------------------------------------------------
# hash grid
# ---------------------------------------------------------------------------

def hashgrid_levels(L: int, T: int, N_min: int, N_max: int, dim: int = 3):
    """Level table of HashGridEncoder (models/encoders.py:76-103, utils/common.py:336-337):
    float64 b and scales, integer resolutions, 8-aligned dense sizes, row offsets."""
    b = math.exp((math.log(N_max) - math.log(N_min)) / (L - 1))
    scales, res, offsets, first_hash = [], [], [0], 0
    for i in range(L):
        s = N_min * (b ** i) - 1
        r = math.ceil(s) + 1
        n = (r ** dim + 7) // 8 * 8
        if n <= T:
            first_hash += 1
        else:
            n = T
        scales.append(s)
        res.append(r)
        offsets.append(offsets[-1] + n)
    return dict(b=b, scales=np.asarray(scales, np.float64), scales_f32=np.asarray(scales, np.float32),
                resolutions=np.asarray(res, np.uint32), offsets=np.asarray(offsets, np.int64),
                first_hash_level=first_hash, n_rows=offsets[-1])


def hashgrid_vertices(pos, bound, lv, T):
    """f32 vertex indices [N, L, 8] (int64) and f32 fractional parts [N, L, 3] of the
    encoder (encoders.py:87, 116-123, 134-189, 216-218).  Corner c = (cx, cy, cz) with
    cz the fastest bit."""
    pos = np.asarray(pos, np.float32)
    b32 = np.float32(bound)
    p = (pos + b32) / (np.float32(2) * b32)  # [N, 3]
    L = len(lv["scales_f32"])
    ps = p[:, None, :] * lv["scales_f32"][None, :, None] + np.float32(0.5)  # [N, L, 3] f32
    fl = np.floor(ps)
    frac = ps - fl
    corners = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], np.float32)
    vert = (fl[:, :, None, :] + corners[None, None]).astype(np.uint64)  # [N, L, 8, 3]
    res = lv["resolutions"].astype(np.uint64)[None, :, None]
    M = np.uint64(0xFFFFFFFF)
    dense = (vert[..., 0] + vert[..., 1] * res + vert[..., 2] * ((res * res) & M)) & M
    hashed = (vert[..., 0] ^ ((vert[..., 1] * np.uint64(2654435761)) & M) ^ ((vert[..., 2] * np.uint64(805459861)) & M)) & M
    lvl = np.arange(L)[None, :, None]
    idx = np.where(lvl < lv["first_hash_level"], dense, hashed) % np.uint64(T)
    idx = idx.astype(np.int64) + lv["offsets"][:-1][None, :, None]
    idx = np.minimum(idx, lv["n_rows"] - 1)
    return idx, frac


def hashgrid_weights64(frac):
    """float64 trilinear weights [N, L, 8] from f32 fractional parts (encoders.py:191-213)."""
    f = frac.astype(np.float64)
    corners = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], np.float64)
    w = np.where(corners[None, None] == 1, f[:, :, None, :], 1.0 - f[:, :, None, :])
    return w.prod(axis=-1)


def hashgrid_forward64(pos, table, bound, lv, T):
    """Encodings [N, L*F] in float64 (encoders.py:226-233)."""
    idx, frac = hashgrid_vertices(pos, bound, lv, T)
    w = hashgrid_weights64(frac)
    tab = np.asarray(table, np.float64)
    enc = (tab[idx] * w[..., None]).sum(axis=2)  # [N, L, F]
    return enc.reshape(enc.shape[0], -1)


def hashgrid_backward64(pos, g, bound, lv, T, F):
    """Table gradient [n_rows, F] in float64: scatter-add of w * g."""
    idx, frac = hashgrid_vertices(pos, bound, lv, T)
    w = hashgrid_weights64(frac)
    N, L = idx.shape[:2]
    g = np.asarray(g, np.float64).reshape(N, L, F)
    contrib = w[..., None] * g[:, :, None, :]  # [N, L, 8, F]
    out = np.zeros((lv["n_rows"], F), np.float64)
    for f in range(F):
        np.add.at(out[:, f], idx.reshape(-1), contrib[..., f].reshape(-1))
    return out


# ---------------------------------------------------------------------------
# integration in float64
# ---------------------------------------------------------------------------

T_THRESHOLD = 1e-4


def integrate_rays64(startidx, nsamp, bgs, dss, z_vals, drgbs, stop_at=None):
    """float64 transcription of integrating.cu:24-104.  `stop_at` (per ray) forces the
    early-stop index (e.g. the f32 path's), so values can be compared at the same
    composited prefix; None stops where the float64 transmittance crosses 1e-4."""
    n = len(startidx)
    bgs = np.broadcast_to(np.asarray(bgs, np.float64), (n, 3))
    rgbd = np.zeros((n, 4))
    op = np.zeros(n)
    stops = np.zeros(n, np.int64)
    for i in range(n):
        s0, m = int(startidx[i]), int(nsamp[i])
        T, acc, k = 1.0, np.zeros(4), 0
        while k < m and (T > T_THRESHOLD if stop_at is None else k < stop_at[i]):
            s = s0 + k
            a = 1.0 - math.exp(-float(drgbs[s, 0]) * float(dss[s]))
            w = T * a
            acc += w * np.array([drgbs[s, 1], drgbs[s, 2], drgbs[s, 3], z_vals[s]], np.float64)
            T *= 1.0 - a
            k += 1
        stops[i] = k
        op[i] = 1.0 - T
        if T <= T_THRESHOLD:
            rgbd[i] = acc / (1.0 - T)
        else:
            rgbd[i, :3] = acc[:3] + T * bgs[i]
            rgbd[i, 3] = acc[3]
    return rgbd, op, stops


def integrate_rays_backward64(near_distance, startidx, nsamp, bgs, dss, z_vals, drgbs, final_rgbds, final_opacities,
                              g, stop_at=None):
    """float64 transcription of the reference VJP (integrating.cu:158-239)."""
    n = len(startidx)
    bgs = np.broadcast_to(np.asarray(bgs, np.float64), (n, 3))
    S = len(dss)
    dbg = np.zeros((n, 3))
    dz = np.zeros(S)
    dd = np.zeros((S, 4))
    for i in range(n):
        s0, m = int(startidx[i]), int(nsamp[i])
        R = np.asarray(final_rgbds[i], np.float64)
        gr = np.asarray(g[i], np.float64)
        op = float(final_opacities[i])
        term = op >= 1.0 - T_THRESHOLD
        bgk = 0.0 if term else 1.0 - op
        T, cur, k = 1.0, np.zeros(4), 0
        while k < m and (T > T_THRESHOLD if stop_at is None else k < stop_at[i]):
            s = s0 + k
            sigma, ds, z = float(drgbs[s, 0]), float(dss[s]), float(z_vals[s])
            a = 1.0 - math.exp(-sigma * ds)
            w = T * a
            c = np.array([drgbs[s, 1], drgbs[s, 2], drgbs[s, 3]], np.float64)
            cur[:3] += w * c
            cur[3] += w * z
            T *= 1.0 - a
            dz[s] = w * gr[3]
            dsig = ds * (np.sum(gr[:3] * (T * c - (R[:3] - cur[:3]) - bgs[i] * bgk)) + gr[3] * (T * z - (R[3] - cur[3])))
            reg = 1e-4 if (sigma > 4e-5 and z < near_distance) else 0.0
            dd[s, 0] = min(z * z, 1.0) * dsig + reg
            dd[s, 1:] = w * gr[:3]
            k += 1
        if T > T_THRESHOLD:
            dbg[i] = T * gr[:3]
    return dbg, dz, dd


# ---------------------------------------------------------------------------
# NeRF MLP in float64 (models/nerfs.py:39-128, 216-238; encoders.py:365-406)
# ---------------------------------------------------------------------------

MLP_SHAPES = [("density_mlp/Dense_0", (32, 64)), ("density_mlp/Dense_1", (64, 16)),
              ("rgb_mlp/Dense_0", (32, 64)), ("rgb_mlp/Dense_1", (64, 64)), ("rgb_mlp/Dense_2", (64, 3))]


def unpack_mlp(params):
    params = np.asarray(params, np.float64)
    out, o = [], 0
    for _, (a, b) in MLP_SHAPES:
        out.append(params[o:o + a * b].reshape(a, b))
        o += a * b
    return out


def sh4_64(d):
    """Real spherical harmonics up to degree 4 (16 values), encoders.py:384-405."""
    d = np.asarray(d, np.float64)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xy, xz, yz, x2, y2, z2 = x * y, x * z, y * z, x * x, y * y, z * z
    return np.stack([
        np.full_like(x, 0.28209479177387814), -0.48860251190291987 * y, 0.48860251190291987 * z,
        -0.48860251190291987 * x, 1.0925484305920792 * xy, -1.0925484305920792 * yz,
        0.94617469575755997 * z2 - 0.31539156525251999, -1.0925484305920792 * xz,
        0.54627421529603959 * x2 - 0.54627421529603959 * y2, 0.59004358992664352 * y * (-3.0 * x2 + y2),
        2.8906114426405538 * xy * z, 0.45704579946446572 * y * (1.0 - 5.0 * z2),
        0.3731763325901154 * z * (5.0 * z2 - 3.0), 0.45704579946446572 * x * (1.0 - 5.0 * z2),
        1.4453057213202769 * z * (x2 - y2), 0.59004358992664352 * x * (-x2 + 3.0 * y2)], axis=-1)


def mlp_forward64(enc, dirs, params):
    """drgbs [S, 4] (density, rgb) and the intermediates, float64."""
    W1, W2, W3, W4, W5 = unpack_mlp(params)
    X = np.asarray(enc, np.float64)
    h1 = np.maximum(X @ W1, 0.0)
    o = h1 @ W2
    sigma = np.exp(o[:, 0])
    sh = sh4_64(dirs)
    cin = np.concatenate([o, sh], axis=1)
    h2 = np.maximum(cin @ W3, 0.0)
    h3 = np.maximum(h2 @ W4, 0.0)
    r = h3 @ W5
    rgb = 1.0 / (1.0 + np.exp(-r))
    return np.concatenate([sigma[:, None], rgb], axis=1), dict(X=X, h1=h1, o=o, cin=cin, h2=h2, h3=h3, rgb=rgb)


def mlp_relu_masks_f16(enc, dirs, params):
    """ReLU masks of a forward that rounds every layer's input to f16 (encodings,
    h1, [o, SH], h2, h3), as the f16-weight kernels do, products in float64.  Where
    a pre-activation lies within f16 rounding of 0 these differ from the float64
    forward's masks, and the whole unit's gradient with them."""
    W1, W2, W3, W4, W5 = unpack_mlp(params)
    r16 = lambda a: np.asarray(a, np.float64).astype(np.float16).astype(np.float64)
    h1 = np.maximum(r16(enc) @ W1, 0.0)
    o = r16(h1) @ W2
    h2 = np.maximum(r16(np.concatenate([o, sh4_64(dirs)], axis=1)) @ W3, 0.0)
    h3 = np.maximum(r16(h2) @ W4, 0.0)
    return dict(h1=h1 > 0, h2=h2 > 0, h3=h3 > 0)


def mlp_backward64(enc, dirs, params, g, masks=None):
    """(dL/denc [S,32], dL/dparams [9408]) in float64, with trunc_exp's VJP
    exp(clip(x, -15, 15)) (nerfs.py:230-233).  `masks` replaces the ReLU masks of
    the float64 forward (see mlp_relu_masks_f16)."""
    W1, W2, W3, W4, W5 = unpack_mlp(params)
    _, a = mlp_forward64(enc, dirs, params)
    m = masks if masks is not None else dict(h1=a["h1"] > 0, h2=a["h2"] > 0, h3=a["h3"] > 0)
    g = np.asarray(g, np.float64)
    dr = g[:, 1:] * a["rgb"] * (1.0 - a["rgb"])
    dh3 = (dr @ W5.T) * m["h3"]
    dh2 = (dh3 @ W4.T) * m["h2"]
    dcin = dh2 @ W3.T
    do = dcin[:, :16].copy()
    do[:, 0] += g[:, 0] * np.exp(np.clip(a["o"][:, 0], -15.0, 15.0))
    dh1 = (do @ W2.T) * m["h1"]
    dX = dh1 @ W1.T
    grads = [a["X"].T @ dh1, a["h1"].T @ do, a["cin"].T @ dh2, a["h2"].T @ dh3, a["h3"].T @ dr]
    return dX, np.concatenate([gw.reshape(-1) for gw in grads])


# ---------------------------------------------------------------------------
# float32 CPU training step (the timed CPU baseline; same algorithm, host cores)
# ---------------------------------------------------------------------------

def mlp_forward32(enc, dirs, params):
    W = [w.astype(np.float32) for w in unpack_mlp(params)]
    X = np.asarray(enc, np.float32)
    h1 = np.maximum(X @ W[0], 0)
    o = h1 @ W[1]
    cin = np.concatenate([o, sh4_64(dirs).astype(np.float32)], 1)
    h2 = np.maximum(cin @ W[2], 0)
    h3 = np.maximum(h2 @ W[3], 0)
    rgb = 1 / (1 + np.exp(-(h3 @ W[4])))
    return np.concatenate([np.exp(o[:, :1]), rgb], 1), (X, h1, o, cin, h2, h3, rgb, W)


def mlp_backward32(cache, g):
    X, h1, o, cin, h2, h3, rgb, W = cache
    dr = g[:, 1:] * rgb * (1 - rgb)
    dh3 = (dr @ W[4].T) * (h3 > 0)
    dh2 = (dh3 @ W[3].T) * (h2 > 0)
    dcin = dh2 @ W[2].T
    do = dcin[:, :16].copy()
    do[:, 0] += g[:, 0] * np.exp(np.clip(o[:, 0], -15, 15))
    dh1 = (do @ W[1].T) * (h1 > 0)
    dX = dh1 @ W[0].T
    dP = np.concatenate([(X.T @ dh1).ravel(), (h1.T @ do).ravel(), (cin.T @ dh2).ravel(), (h2.T @ dh3).ravel(),
                         (h3.T @ dr).ravel()])
    return dX, dP


def cpu_train_step(rays_o, rays_d, t_starts, t_ends, noises, bits, bgs, gt, table, params, lv, T, total_samples):
    """One plain training step on the CPU with the oracle's ops (marcher, hash
    encoder, MLP, integrator, Huber gradient, backward passes).  Returns the number
    of marched samples (the samples/s numerator)."""
    L, F = len(lv["scales_f32"]), table.shape[1]
    m = march_rays(total_samples, 1024, 1, 128, 1.0, 0.0, rays_o, rays_d, t_starts, t_ends, noises, bits)
    S = total_samples
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    scales = np.ascontiguousarray(lv["scales_f32"], np.float32)
    res = np.ascontiguousarray(lv["resolutions"], np.uint32)
    offs = np.ascontiguousarray(lv["offsets"], np.uint32)
    enc = np.zeros((S, L * F), np.float32)
    lib().oracle_hashgrid_fwd(u32(S), u32(L), u32(F), u32(T), u32(lv["first_hash_level"]), u32(lv["n_rows"]),
                              f32(1.0), P(scales), P(res), P(offs), P(m["xyzs"]), P(table), P(enc))
    drgbs, cache = mlp_forward32(enc, m["dirs"], params)
    drgbs = np.ascontiguousarray(drgbs, np.float32)
    _, rgbd, op = integrate_rays(m["rays_sample_startidx"], m["rays_n_samples"], bgs, m["dss"], m["z_vals"], drgbs)
    valid = m["ray_is_valid"]
    nv = max(int(valid.sum()), 1)
    e = rgbd[:, :3] - gt
    g = np.zeros((len(valid), 4), np.float32)
    g[:, :3] = np.where(valid[:, None], np.clip(e, -0.1, 0.1) / (3 * nv), 0)
    _, _, dd = integrate_rays_backward(0.3, m["rays_sample_startidx"], m["rays_n_samples"], bgs, m["dss"],
                                       m["z_vals"], drgbs, rgbd, op, g)
    dX, _ = mlp_backward32(cache, dd)
    dt = np.zeros((lv["n_rows"], F), np.float64)
    dX = np.ascontiguousarray(dX, np.float32)
    lib().oracle_hashgrid_bwd(u32(S), u32(L), u32(F), u32(T), u32(lv["first_hash_level"]), u32(lv["n_rows"]),
                              f32(1.0), P(scales), P(res), P(offs), P(m["xyzs"]), P(dX), P(dt))
    return int(m["measured"])


# ---------------------------------------------------------------------------
# occupancy-grid update (utils/types.py:1149-1225).  The reference draws its cells
# and points with JAX's PRNG; the build draws them from a counter-based splitmix64
# stream (the same draws the kernel makes), restated here, so the chosen cells and
# points are compared exactly and the densities against float64.
# ---------------------------------------------------------------------------

OGRID_DECAY = np.float32(0.95)  # types.py:1163


def mix64(z):
    """splitmix64 finaliser (Steele, Lea, Flood 2014) on uint64 arrays."""
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def draw(seed, step, i, stream):
    """The kernels' random draw of (seed, step, item i, stream)."""
    i = np.asarray(i, np.uint64)
    with np.errstate(over="ignore"):
        a = mix64(i * np.uint64(4) + np.uint64(stream))
        b = mix64((np.uint64(step) * np.uint64(0x9E3779B97F4A7C15)) ^ a)
    return mix64(np.uint64(int(seed) & 0xFFFFFFFFFFFFFFFF) ^ b)


def below(r, n):
    """Uniform index in [0, n) from a draw: (r >> 32) * n >> 32."""
    return ((np.asarray(r, np.uint64) >> np.uint64(32)) * np.uint64(n)) >> np.uint64(32)


def ogrid_choice(alive, occupied, G, cascade, update_all, seed, step, bound):
    """Cells (Morton index within the cascade, one per update slot) and their random
    points, f32 [n, 3], of one occupancy-grid update (types.py:1165-1207).
    `alive`: the cascade's alive cells, ascending; `occupied`: bool [G^3] of the
    cascade (bits of the last threshold).  update_all: every alive cell once; else
    max(1, M // 2) cells uniform over the alive ones and as many uniform over the
    occupied ones (types.py:1172-1191), M = max(1, n_alive // 2), and uniform over
    the alive ones when none is occupied."""
    alive = np.asarray(alive, np.int64)
    n_alive = alive.shape[0]
    key0 = np.uint64(cascade) << np.uint64(32)
    if update_all:
        cells = alive.copy()
    else:
        M = max(1, n_alive // 2)
        h = max(1, M // 2)
        j = np.arange(2 * h, dtype=np.uint64)
        cells = alive[below(draw(seed, step, key0 | j, 0), n_alive).astype(np.int64)]
        occ = np.flatnonzero(np.asarray(occupied, bool))
        if occ.size:
            cells[h:] = occ[below(draw(seed, step, key0 | j[h:], 1), occ.size).astype(np.int64)]
    n = cells.shape[0]
    j = np.arange(n, dtype=np.uint64)
    r = draw(seed, step, key0 | j, 2)
    u = np.stack([((r >> np.uint64(s)) & np.uint64(0x1FFFFF)).astype(np.float32) for s in (43, 22, 1)], 1)
    u = u * np.float32(1.0 / 2097152.0)
    v = morton3d_invert(cells.astype(np.uint32)).astype(np.float32)
    mip = np.float32(min(float(bound), 2.0 ** cascade))
    hcw = np.float32(mip / np.float32(G))
    c = (v / np.float32(G - 1)) * np.float32(2) - np.float32(1)
    pts = c * (mip - hcw) + (u * np.float32(2) - np.float32(1)) * hcw
    return cells, pts.astype(np.float32)


def ogrid_density64(points, table, params, bound, lv, T):
    """The field's density at `points` in float64: trunc_exp of the density MLP's
    first output on the position encoding (models/nerfs.py:66-77)."""
    enc = hashgrid_forward64(points, table, bound, lv, T)
    W1, W2 = unpack_mlp(params)[:2]
    return np.exp(np.maximum(enc @ W1, 0.0) @ W2[:, 0])


# ---------------------------------------------------------------------------
# Learned proposal (no counterpart in the reference; defined by this build and
# pinned by finite differences and closed forms in tests/test_oracle_pins.py)
# ---------------------------------------------------------------------------

PROP_BINS = 64
PROP_N_PARAMS = 10 * 16 + 16  # Dense 10->16 (ReLU), Dense 16->1, no biases


def proposal_mlp_forward64(enc, params):
    """sigma = exp(relu(enc W1) W2) for enc [N, 10]; returns (sigma [N], hidden, logit)."""
    p = np.asarray(params, np.float64)
    W1, W2 = p[:160].reshape(10, 16), p[160:176].reshape(16, 1)
    h = np.maximum(np.asarray(enc, np.float64) @ W1, 0.0)
    o = (h @ W2)[:, 0]
    return np.exp(o), h, o


def proposal_mlp_backward64(enc, params, g_sigma):
    """(dL/denc [N, 10], dL/dparams [176]) with trunc_exp's VJP exp(clip(o, -15, 15))."""
    p = np.asarray(params, np.float64)
    W1, W2 = p[:160].reshape(10, 16), p[160:176].reshape(16, 1)
    X = np.asarray(enc, np.float64)
    _, h, o = proposal_mlp_forward64(X, p)
    do = np.asarray(g_sigma, np.float64) * np.exp(np.clip(o, -15.0, 15.0))
    dW2 = h.T @ do[:, None]
    dh = (do[:, None] @ W2.T) * (h > 0)
    return dh @ W1.T, np.concatenate([(X.T @ dh).reshape(-1), dW2.reshape(-1)])


def proposal_bin_index(z, t0, t1, n_bins=PROP_BINS):
    """Bin of sample distance z on the ray segment [t0, t1), in the kernels' f32
    arithmetic: floor((z - t0) * (n_bins / (t1 - t0))) clamped to [0, n_bins)."""
    f = np.float32
    inv = f(n_bins) / (f(t1) - f(t0))
    x = (f(z) - f(t0)) * inv
    return int(min(max(math.floor(float(x)), 0), n_bins - 1))


def nerf_sample_weights64(startidx, nsamp, dss, drgbs):
    """w_i = T_i alpha_i of every sample (integrating.cu:58-66), 0 after the ray's
    transmittance falls to 1e-4."""
    w = np.zeros(len(dss))
    for r in range(len(startidx)):
        s0, m = int(startidx[r]), int(nsamp[r])
        T = 1.0
        for k in range(m):
            if T <= T_THRESHOLD:
                break
            a = 1.0 - math.exp(-float(drgbs[s0 + k, 0]) * float(dss[s0 + k]))
            w[s0 + k] = T * a
            T *= 1.0 - a
    return w


def proposal_loss64(startidx, nsamp, dss, z_vals, drgbs, t0, t1, sigma_p, valid, n_valid, eps=1e-7):
    """Histogram-bound loss of the proposal.  Per valid ray, the NeRF's sample
    weights are summed into 64 equal bins of [t0, t1) (W_b); the proposal's density
    sigma_p[r, b] at the bin centres gives bin weights w_b = T_b (1 - exp(-sigma_b D)),
    D = (t1 - t0) / 64; loss = sum_r sum_b max(0, W_b - w_b)^2 / (W_b + eps) / n_valid.
    Returns (loss, dL/dsigma_p [n_rays, 64])."""
    w = nerf_sample_weights64(startidx, nsamp, dss, drgbs)
    n_rays = len(startidx)
    sig = np.asarray(sigma_p, np.float64).reshape(n_rays, PROP_BINS)
    grad = np.zeros_like(sig)
    loss = 0.0
    nv = max(int(n_valid), 1)
    for r in range(n_rays):
        if not valid[r]:
            continue
        W = np.zeros(PROP_BINS)
        s0, m = int(startidx[r]), int(nsamp[r])
        for k in range(m):
            W[proposal_bin_index(z_vals[s0 + k], t0[r], t1[r])] += w[s0 + k]
        D = (float(t1[r]) - float(t0[r])) / PROP_BINS
        a = 1.0 - np.exp(-sig[r] * D)
        T = np.concatenate([[1.0], np.cumprod(1.0 - a)[:-1]])
        wp = T * a
        gap = np.maximum(W - wp, 0.0)
        loss += float((gap ** 2 / (W + eps)).sum()) / nv
        g = -2.0 * gap / (W + eps) / nv  # dL/dwp
        Tn = T * (1.0 - a)  # T_{b+1}
        suffix = np.concatenate([np.cumsum((g * wp)[::-1])[::-1][1:], [0.0]])  # sum_{b' > b} g wp
        grad[r] = D * (g * Tn - suffix)
    return loss, grad


def proposal_sample_support64(t0, t1, sigma_p, n_samples=48, mass_threshold=0.9, eps=1e-3):
    """Support placement: the bins from the first whose cumulative share C_b + p_b
    exceeds eps to the first where it reaches 1 - eps hold the ray's mass; n samples
    sit evenly in that range [t_lo, t_hi), t_j = t_lo + (j + 1/2) (t_hi - t_lo) / n,
    ds_j = (t_hi - t_lo) / n.  Fallback and empty segments as proposal_sample64.
    Returns (counts, fallback, t, ds)."""
    n_rays = len(t0)
    sig = np.asarray(sigma_p, np.float64).reshape(n_rays, PROP_BINS)
    counts = np.zeros(n_rays, np.int64)
    fallback = np.zeros(n_rays, bool)
    t = np.zeros((n_rays, n_samples))
    ds = np.zeros((n_rays, n_samples))
    for r in range(n_rays):
        a0, a1 = float(t0[r]), float(t1[r])
        if not a1 > a0:
            continue
        D = (a1 - a0) / PROP_BINS
        a = 1.0 - np.exp(-sig[r] * D)
        T = np.concatenate([[1.0], np.cumprod(1.0 - a)[:-1]])
        w = T * a
        M = w.sum()
        if not M >= mass_threshold:
            fallback[r] = True
            continue
        inc = np.cumsum(w / M)
        lo = int(np.argmax(inc > eps))
        hi = int(np.argmax(inc >= 1.0 - eps)) if (inc >= 1.0 - eps).any() else PROP_BINS - 1
        hi = max(hi, lo)
        t_lo, t_hi = a0 + lo * D, a0 + (hi + 1) * D
        step = (t_hi - t_lo) / n_samples
        t[r] = t_lo + (np.arange(n_samples) + 0.5) * step
        ds[r] = step
        counts[r] = n_samples
    return counts, fallback, t, ds


def proposal_sample64(t0, t1, sigma_p, n_samples=48, mass_threshold=0.9, jitter=0.5):
    """Samples placed by the proposal, per ray: bin weights w_b = T_b (1 - exp(-sigma_b D))
    over 64 equal bins of [t0, t1); mass M = sum w_b; rays with t1 <= t0 get no samples,
    rays with M < mass_threshold (or M = 0) get none either and are marked for the
    occupancy-grid fallback.  Otherwise sample j sits at the (j + jitter) / n quantile
    of the piecewise constant density w_b / (M D) (jitter 1/2 when rendering, the ray's
    noise in [0, 1) for stratified training samples): t_j = t0 + (b + (u - C_b) / p_b) D
    with p_b = w_b / M, C_b = sum_{b' < b} p_b', b the last bin with C_b <= u;
    ds_j = t_{j+1} - t_j and, for the last sample, min(D / (n p_b), t1 - t_j).
    `jitter` is a scalar or one value per ray.
    Returns (counts [n_rays], fallback [n_rays], t [n_rays, n], ds [n_rays, n])."""
    n_rays = len(t0)
    jit = np.broadcast_to(np.asarray(jitter, np.float64), (n_rays,))
    sig = np.asarray(sigma_p, np.float64).reshape(n_rays, PROP_BINS)
    counts = np.zeros(n_rays, np.int64)
    fallback = np.zeros(n_rays, bool)
    t = np.zeros((n_rays, n_samples))
    ds = np.zeros((n_rays, n_samples))
    for r in range(n_rays):
        a0, a1 = float(t0[r]), float(t1[r])
        if not a1 > a0:
            continue
        D = (a1 - a0) / PROP_BINS
        a = 1.0 - np.exp(-sig[r] * D)
        T = np.concatenate([[1.0], np.cumprod(1.0 - a)[:-1]])
        w = T * a
        M = w.sum()
        if not (M >= mass_threshold and M > 0):
            fallback[r] = True
            continue
        p = w / M
        C = np.concatenate([[0.0], np.cumsum(p)[:-1]])
        for j in range(n_samples):
            u = (j + jit[r]) / n_samples
            b = int(np.nonzero(C <= u)[0].max())
            frac = min(max((u - C[b]) / p[b], 0.0), 1.0) if p[b] > 0 else 0.5
            t[r, j] = a0 + (b + frac) * D
        for j in range(n_samples - 1):
            ds[r, j] = t[r, j + 1] - t[r, j]
        bl = int(np.nonzero(C <= (n_samples - 1 + jit[r]) / n_samples)[0].max())
        ds[r, -1] = min(D / (n_samples * p[bl]) if p[bl] > 0 else D, a1 - t[r, -1])
        counts[r] = n_samples
    return counts, fallback, t, ds


# ---------------------------------------------------------------------------
# jax-tcnn hash grid (tiny-cuda-nn v1.6 kernel_grid as deps/jax-tcnn/lib/impl/
# hashgrid.cu:20-96 launches it; tiny-cuda-nn is not vendored, so this restates its
# published linear / coherent-prime grid encoding).  Test infrastructure only.
# ---------------------------------------------------------------------------

def _log2f(x: float) -> np.float32:
    """The host's libm log2f (what the library's entry computes on the host)."""
    import ctypes.util
    m = ctypes.CDLL(ctypes.util.find_library("m"))
    m.log2f.restype, m.log2f.argtypes = ctypes.c_float, [ctypes.c_float]
    return np.float32(m.log2f(float(x)))


def _fmaf(a, b, c):
    """f32 fused multiply-add (the product of two f32 is exact in f64)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def tcnn_levels(L, N_min, per_level_scale):
    """f32 scales and resolutions: exp2(l log2(b)) N_min - 1, ceil + 1."""
    l2 = _log2f(per_level_scale)
    out = []
    for l in range(L):
        t = np.float32(l) * l2
        e = np.float32(np.exp2(np.float64(t)))
        scale = np.float32(e * np.float32(N_min)) - np.float32(1.0)
        out.append((np.float32(scale), int(np.ceil(scale)) + 1))
    return out


def tcnn_index(size, res, x, y, z):
    """Dense index while the stride stays within the level, else the coherent prime
    hash; both modulo the level's size (u32 arithmetic)."""
    x, y, z = (np.asarray(v, np.uint32) for v in (x, y, z))
    size, res = np.uint32(size), np.uint32(res)
    with np.errstate(over="ignore"):
        stride = np.uint32(1)
        index = np.zeros_like(x)
        for p in (x, y, z):
            if stride > size:
                break
            index = index + p * stride
            stride = stride * res
        if size < stride:
            index = x ^ (y * np.uint32(2654435761)) ^ (z * np.uint32(805459861))
    return index % size


def tcnn_grid_forward(coords, table, offsets, L, F, N_min, per_level_scale):
    """(encodings f32[L*F, N], dy_dcoords f32[L*F, N, 3]) in the kernels' f32 order."""
    coords = np.asarray(coords, np.float32)
    n = coords.shape[1]
    enc = np.zeros((L * F, n), np.float32)
    dy = np.zeros((L * F, n, 3), np.float32)
    for l, (scale, res) in enumerate(tcnn_levels(L, N_min, per_level_scale)):
        lo, hi = int(offsets[l]), int(offsets[l + 1])
        size = hi - lo
        grid = table[lo:hi]
        pos = _fmaf(scale, coords, np.float32(0.5))
        g = np.floor(pos).astype(np.float32)
        frac = (pos - g).astype(np.float32)
        gi = g.astype(np.int64).astype(np.uint32)
        acc = np.zeros((n, F), np.float32)
        for c in range(8):
            w = np.ones(n, np.float32)
            q = []
            for d in range(3):
                bit = (c >> d) & 1
                w = (w * (frac[d] if bit else np.float32(1) - frac[d])).astype(np.float32)
                q.append(gi[d] + np.uint32(bit))
            rows = grid[tcnn_index(size, res, *q)]
            acc = _fmaf(w[:, None], rows, acc)
        enc[l * F:(l + 1) * F] = acc.T
        for gd in range(3):
            grad = np.zeros((n, F), np.float32)
            others = [d for d in range(3) if d != gd]
            for c in range(4):
                w = np.full(n, scale, np.float32)
                q = [None, None, None]
                for k, d in enumerate(others):
                    bit = (c >> k) & 1
                    w = (w * (frac[d] if bit else np.float32(1) - frac[d])).astype(np.float32)
                    q[d] = gi[d] + np.uint32(bit)
                q[gd] = gi[gd]
                left = grid[tcnn_index(size, res, *q)]
                q[gd] = gi[gd] + np.uint32(1)
                right = grid[tcnn_index(size, res, *q)]
                grad = (grad + (w[:, None] * (right - left).astype(np.float32)).astype(np.float32)).astype(np.float32)
            dy[l * F:(l + 1) * F, :, gd] = grad.T
    return enc, dy


def tcnn_grid_forward64(coords, table, offsets, L, F, N_min, per_level_scale):
    """float64 encodings of the same grid (the f32 scales and resolutions), for
    finite differences of dy_dcoords."""
    coords = np.asarray(coords, np.float64)
    n = coords.shape[1]
    enc = np.zeros((L * F, n))
    for l, (scale, res) in enumerate(tcnn_levels(L, N_min, per_level_scale)):
        lo, hi = int(offsets[l]), int(offsets[l + 1])
        grid = np.asarray(table[lo:hi], np.float64)
        pos = float(scale) * coords + 0.5
        g = np.floor(pos)
        frac = pos - g
        gi = g.astype(np.int64).astype(np.uint32)
        acc = np.zeros((n, F))
        for c in range(8):
            w = np.ones(n)
            q = []
            for d in range(3):
                bit = (c >> d) & 1
                w = w * (frac[d] if bit else 1 - frac[d])
                q.append(gi[d] + np.uint32(bit))
            acc += w[:, None] * grid[tcnn_index(hi - lo, res, *q)]
        enc[l * F:(l + 1) * F] = acc.T
    return enc


def tcnn_grid_backward64(coords, dL_dy, dy_dcoords, offsets, L, F, N_min, per_level_scale, n_rows):
    """float64 (dL_dparams [n_rows, F], dL_dcoords [3, N])."""
    coords = np.asarray(coords, np.float32)
    n = coords.shape[1]
    dt = np.zeros((n_rows, F))
    for l, (scale, res) in enumerate(tcnn_levels(L, N_min, per_level_scale)):
        lo, hi = int(offsets[l]), int(offsets[l + 1])
        pos = _fmaf(scale, coords, np.float32(0.5))
        g = np.floor(pos).astype(np.float32)
        frac = (pos - g).astype(np.float32)
        gi = g.astype(np.int64).astype(np.uint32)
        gl = np.asarray(dL_dy[l * F:(l + 1) * F], np.float64).T
        for c in range(8):
            w = np.ones(n, np.float32)
            q = []
            for d in range(3):
                bit = (c >> d) & 1
                w = (w * (frac[d] if bit else np.float32(1) - frac[d])).astype(np.float32)
                q.append(gi[d] + np.uint32(bit))
            idx = lo + tcnn_index(hi - lo, res, *q).astype(np.int64)
            np.add.at(dt, idx, w.astype(np.float64)[:, None] * gl)
    dcoords = np.einsum("kn,knd->dn", np.asarray(dL_dy, np.float64), np.asarray(dy_dcoords, np.float64))
    return dt, dcoords


# ---------------------------------------------------------------------------
# sky-sphere background model (models/nerfs.py:133-213, make_skysphere_background_
# model_ngp 383-419), float64.  Pinned by tests/test_oracle_sky.py.
# ---------------------------------------------------------------------------

SKY_IN0, SKY_HID, SKY_OUT = 20, 32, 3


def sky_sphere_split(params, n_extra_dims):
    """Flat parameters -> rgb_mlp Dense kernels W0 [20+E, 32], W1 [32, 32], W2 [32, 3]."""
    p = np.asarray(params, np.float64)
    K = SKY_IN0 + int(n_extra_dims)
    W0 = p[:K * 32].reshape(K, 32)
    W1 = p[K * 32:K * 32 + 1024].reshape(32, 32)
    W2 = p[K * 32 + 1024:K * 32 + 1120].reshape(32, 3)
    assert p.size == K * 32 + 1120
    return W0, W1, W2


def sky_sphere_hit64(rays_o, rays_d, radius):
    """The larger root t of |o + t d|^2 = r^2 (nerfs.py:172-184); NaN where the ray
    misses (np.maximum propagates NaN like jnp.maximum)."""
    o = np.asarray(rays_o, np.float64)
    d = np.asarray(rays_d, np.float64)
    a = (d * d).sum(-1)
    b = 2.0 * (o * d).sum(-1)
    c = (o * o).sum(-1) - float(radius) ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.sqrt(b * b - 4.0 * a * c)
        return np.maximum((-b + s) / (2.0 * a), (-b - s) / (2.0 * a))


def sky_sphere_inputs64(rays_o, rays_d, appearance, radius, n_extra_dims):
    """MLP inputs [SH2(pos / ||pos||_F or 0), SH4(d), appearance] and finite(t)
    (nerfs.py:186-207).  ||pos||_F is the norm of the whole [n, 3] array."""
    o = np.asarray(rays_o, np.float64)
    d = np.asarray(rays_d, np.float64)
    n = o.shape[0]
    t = sky_sphere_hit64(o, d, radius)
    fin = np.isfinite(t)
    with np.errstate(invalid="ignore", over="ignore"):
        pos = o + t[:, None] * d
        denom = np.sqrt((pos * pos).sum()) + 1e-15
        pd = np.where(fin[:, None], pos / denom, 0.0)
    E = int(n_extra_dims)
    app = np.broadcast_to(np.asarray(appearance, np.float64).reshape(-1, E) if E else np.zeros((1, 0)), (n, E))
    return np.concatenate([sh4_64(pd)[:, :4], sh4_64(d), app], axis=1), fin


def sky_sphere_bg64(rays_o, rays_d, appearance, params, radius, n_extra_dims):
    """colors [n, 3]: sigmoid(relu(relu(x W0) W1) W2) where t is finite, else 0."""
    x, fin = sky_sphere_inputs64(rays_o, rays_d, appearance, radius, n_extra_dims)
    W0, W1, W2 = sky_sphere_split(params, n_extra_dims)
    with np.errstate(invalid="ignore", over="ignore"):
        h1 = np.maximum(x @ W0, 0.0)
        h2 = np.maximum(h1 @ W1, 0.0)
        y = 1.0 / (1.0 + np.exp(-(h2 @ W2)))
    return np.where(fin[:, None], y, 0.0)


def sky_sphere_bg_backward64(rays_o, rays_d, appearance, params, radius, n_extra_dims, dL_dcolors):
    """(dL/dparams flat, dL/dappearance [n, E]): the VJP jax.grad takes through
    SkySphereBg.__call__.  The cotangent is masked by finite(t) and relu's VJP is
    jax.nn.relu's select(z > 0, g, 0), a select rather than a product, so a NaN
    pre-activation passes 0, not NaN."""
    x, fin = sky_sphere_inputs64(rays_o, rays_d, appearance, radius, n_extra_dims)
    W0, W1, W2 = sky_sphere_split(params, n_extra_dims)
    z1 = x @ W0
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W1
    h2 = np.maximum(z2, 0.0)
    y = 1.0 / (1.0 + np.exp(-(h2 @ W2)))
    g = np.where(fin[:, None], np.asarray(dL_dcolors, np.float64), 0.0)
    do = g * y * (1.0 - y)
    with np.errstate(invalid="ignore"):
        dz2 = np.where(z2 > 0, do @ W2.T, 0.0)
        dz1 = np.where(z1 > 0, dz2 @ W1.T, 0.0)
    dx = dz1 @ W0.T
    dparams = np.concatenate([(x.T @ dz1).ravel(), (h1.T @ dz2).ravel(), (h2.T @ do).ravel()])
    return dparams, dx[:, SKY_IN0:]
