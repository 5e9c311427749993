"""GPU parity of the hash-grid encoder (forward and table gradient).

Forward: bit-exact against the oracle's f32 restatement (same vertex indices,
weights and summation order), and within |a-b| <= 1e-5 * max(|b|, 1e-4) of the
float64 transcription (the tolerance of the encoder workload: hash indices exact,
encodings within 1e-5 relative of FP64).
Backward: within 1e-5 of the float64 scatter-add, relative to the sum of the
absolute contributions of each row (the kernels merge runs in f32 and sum the
rest in any order); rows no point touches are exactly 0 (a touched row may
cancel to 0 within the tolerance).
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from jaxngp_amd.encoders import HashGridEncoder

pytestmark = pytest.mark.gpu


def _c_forward(enc_mod, pos, table, bound=1.0):
    lv = enc_mod.levels
    N, L, F = pos.shape[0], enc_mod.L, enc_mod.F
    out = np.zeros((N, L * F), np.float32)
    scales = np.asarray(lv.scales, np.float32)
    res = np.asarray(lv.resolutions, np.uint32)
    offs = np.asarray(lv.offsets, np.uint32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    pos = np.ascontiguousarray(pos, np.float32)
    table = np.ascontiguousarray(table, np.float32)
    oracle.lib().oracle_hashgrid_fwd(ctypes.c_uint32(N), ctypes.c_uint32(L), ctypes.c_uint32(F),
                                     ctypes.c_uint32(enc_mod.T), ctypes.c_uint32(lv.first_hash_level),
                                     ctypes.c_uint32(lv.n_rows), ctypes.c_float(bound), P(scales), P(res), P(offs),
                                     P(pos), P(table), P(out))
    return out


def _c_backward(enc_mod, pos, g, bound=1.0):
    lv = enc_mod.levels
    N, L, F = pos.shape[0], enc_mod.L, enc_mod.F
    out = np.zeros((lv.n_rows, F), np.float64)
    scales = np.asarray(lv.scales, np.float32)
    res = np.asarray(lv.resolutions, np.uint32)
    offs = np.asarray(lv.offsets, np.uint32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    oracle.lib().oracle_hashgrid_bwd(ctypes.c_uint32(N), ctypes.c_uint32(L), ctypes.c_uint32(F),
                                     ctypes.c_uint32(enc_mod.T), ctypes.c_uint32(lv.first_hash_level),
                                     ctypes.c_uint32(lv.n_rows), ctypes.c_float(bound), P(scales), P(res), P(offs),
                                     P(np.ascontiguousarray(pos, np.float32)), P(np.ascontiguousarray(g, np.float32)),
                                     P(out))
    return out


CONFIGS = {
    "l16_t2e19": dict(L=16, T=2 ** 19, F=2, N_min=16, N_max=2048),
    "proposal": dict(L=5, T=2 ** 17, F=2, N_min=16, N_max=512),
    "small_f4": dict(L=8, T=2 ** 12, F=4, N_min=4, N_max=64),
    "f1_dense_only": dict(L=3, T=2 ** 20, F=1, N_min=8, N_max=32),
    "f8": dict(L=4, T=2 ** 14, F=8, N_min=16, N_max=256),
    # dense levels of up to 4.1 M rows: a vertex on a level's upper face addresses up
    # to res^2 + res rows (~26 k) past the level, several 8192-row slices into the next
    "dense_overflow": dict(L=4, T=2 ** 22, F=2, N_min=64, N_max=160),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_hashgrid_forward_backward(dev, name):
    cfg = CONFIGS[name]
    rng = np.random.default_rng(0)
    g_gen = torch.Generator(device="cpu").manual_seed(1)
    enc = HashGridEncoder(**cfg, generator=g_gen).to(dev)
    N = 40000
    pos = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    pos[:4] = [[-1, -1, -1], [1, 1, 1], [0, 0, 0], [0.999999, -0.999999, 0.5]]
    table = enc.table.detach().cpu().numpy()
    pos_t = torch.from_numpy(pos).to(dev)
    out, tv = enc(pos_t, 1.0)
    assert tv == 0.0
    got = out.detach().cpu().numpy()
    np.testing.assert_array_equal(got, _c_forward(enc, pos, table))
    lv = oracle.hashgrid_levels(cfg["L"], cfg["T"], cfg["N_min"], cfg["N_max"])
    want64 = oracle.hashgrid_forward64(pos[:8000], table, 1.0, lv, cfg["T"])
    assert (np.abs(got[:8000] - want64) <= 1e-5 * np.maximum(np.abs(want64), 1e-4)).all()
    g = rng.normal(size=got.shape).astype(np.float32)
    out.backward(torch.from_numpy(g).to(dev))
    dt = enc.table.grad.detach().cpu().numpy().astype(np.float64)
    want = _c_backward(enc, pos, g)
    absw = _c_backward(enc, pos, np.abs(g))
    assert (dt[absw == 0] == 0).all()  # rows no point touches stay exactly 0
    assert (np.abs(dt - want) <= 1e-5 * absw + 1e-30).all()


def test_hashgrid_encoder_workload_full_size(dev):
    """Encoder workload: N = 2^18 points, L=16, F=2, T=2^19, table U(-1e-4, 1e-4)."""
    rng = np.random.default_rng(0)
    enc = HashGridEncoder(**CONFIGS["l16_t2e19"], generator=torch.Generator().manual_seed(1)).to(dev)
    assert enc.levels.n_rows == 6098120 and enc.levels.first_hash_level == 5
    N = 2 ** 18
    xyz = rng.uniform(0, 1, (N, 3)).astype(np.float32)
    pos = (xyz * np.float32(2) - np.float32(1)).astype(np.float32)
    out, _ = enc(torch.from_numpy(pos).to(dev), 1.0)
    got = out.detach().cpu().numpy()
    np.testing.assert_array_equal(got, _c_forward(enc, pos, enc.table.detach().cpu().numpy()))
    g = np.random.default_rng(2).normal(size=(N, 32)).astype(np.float32)
    out.backward(torch.from_numpy(g).to(dev))
    dt = enc.table.grad.detach().cpu().numpy().astype(np.float64)
    want = _c_backward(enc, pos, g)
    absw = _c_backward(enc, pos, np.abs(g))
    assert (dt[absw == 0] == 0).all()  # rows no point touches stay exactly 0
    assert (np.abs(dt - want) <= 1e-5 * absw + 1e-30).all()


def test_hashgrid_fp16_table(dev):
    rng = np.random.default_rng(3)
    enc = HashGridEncoder(**CONFIGS["proposal"], param_dtype=torch.float16,
                          generator=torch.Generator().manual_seed(4)).to(dev)
    with torch.no_grad():
        enc.table.copy_(torch.from_numpy(rng.uniform(-1, 1, tuple(enc.table.shape)).astype(np.float16)))
    pos = rng.uniform(-1, 1, (20000, 3)).astype(np.float32)
    out, _ = enc(torch.from_numpy(pos).to(dev), 1.0)
    table32 = enc.table.detach().float().cpu().numpy()
    np.testing.assert_array_equal(out.detach().cpu().numpy(), _c_forward(enc, pos, table32))


def test_level_major_layout_matches_point_major(dev):
    """The level-major [L, N, F] path (XCD-affine kernels) gives the point-major
    encodings transposed, and the same table gradient from the transposed cotangent."""
    from jaxngp_amd import _lib
    rng = np.random.default_rng(9)
    enc = HashGridEncoder(**CONFIGS["l16_t2e19"], generator=torch.Generator().manual_seed(1)).to(dev)
    with torch.no_grad():
        enc.table.uniform_(-1, 1)
    N = 30001
    pos = torch.from_numpy(rng.uniform(-1, 1, (N, 3)).astype(np.float32)).to(dev)
    L, F = 16, 2
    pm = torch.empty(N, L * F, device=dev)
    lm = torch.empty(L, N, F, device=dev)
    t = enc.table.detach()
    _lib.call("hashgrid_encode", [pos, t, pm], enc.descriptor(N, 1.0))
    _lib.call("hashgrid_encode", [pos, t, lm], enc.descriptor(N, 1.0, _lib.ENC_LEVEL_MAJOR))
    torch.testing.assert_close(lm.permute(1, 0, 2).reshape(N, L * F), pm, rtol=0, atol=0)
    g = torch.from_numpy(rng.normal(size=(N, L * F)).astype(np.float32)).to(dev)
    g_lm = g.reshape(N, L, F).permute(1, 0, 2).contiguous()
    d1 = torch.empty_like(t)
    d2 = torch.empty_like(t)
    for d, gg, layout in ((d1, g, _lib.ENC_POINT_MAJOR), (d2, g_lm, _lib.ENC_LEVEL_MAJOR)):
        desc = enc.descriptor(N, 1.0, layout)
        ws = torch.empty(_lib.hashgrid_backward_workspace_bytes(desc), dtype=torch.uint8, device=dev)
        _lib.call("hashgrid_encode_backward", [pos, gg, d, ws], desc)
    want = _c_backward(enc, pos.cpu().numpy(), g.cpu().numpy())
    absw = _c_backward(enc, pos.cpu().numpy(), np.abs(g.cpu().numpy()))
    for d in (d1, d2):
        dt = d.cpu().numpy().astype(np.float64)
        assert (dt[absw == 0] == 0).all()  # rows no point touches stay exactly 0
        assert (np.abs(dt - want) <= 1e-5 * absw + 1e-30).all()


def _backward_raw(enc, pos_t, g_t):
    from jaxngp_amd import _lib
    N = pos_t.shape[0]
    desc = enc.descriptor(N, 1.0)
    ws = torch.empty(_lib.hashgrid_backward_workspace_bytes(desc), dtype=torch.uint8, device=pos_t.device)
    d = torch.empty(tuple(enc.table.shape), dtype=torch.float32, device=pos_t.device)  # the gradient is f32
    _lib.call("hashgrid_encode_backward", [pos_t, g_t, d, ws], desc)
    return d


def test_backward_scale_free(dev):
    """The table gradient's sums keep relative precision at any magnitude:
    cotangents scaled by 2^-60 or 2^+60 give the gradient scaled by the same power
    of two, within the f32 tolerance (contributions stay in the f32 normal range)."""
    rng = np.random.default_rng(21)
    enc = HashGridEncoder(**CONFIGS["l16_t2e19"], generator=torch.Generator().manual_seed(1)).to(dev)
    N = 50000
    pos = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    g = rng.normal(size=(N, 32)).astype(np.float32)
    pos_t = torch.from_numpy(pos).to(dev)
    want = _c_backward(enc, pos, g)
    absw = _c_backward(enc, pos, np.abs(g))
    for k in (-60, 60):
        gs = np.ldexp(g, k).astype(np.float32)
        d = _backward_raw(enc, pos_t, torch.from_numpy(gs).to(dev)).cpu().numpy().astype(np.float64)
        d = np.ldexp(d, -k)
        assert (d[absw == 0] == 0).all()
        assert (np.abs(d - want) <= 1e-5 * absw + 1e-30).all()


def test_backward_non_finite_reaches_its_rows(dev):
    """An inf or NaN cotangent (a loss-scaling overflow) makes exactly the rows its
    point touches non-finite, as a float sum does; the other rows stay exact."""
    rng = np.random.default_rng(22)
    cfg = CONFIGS["proposal"]
    enc = HashGridEncoder(**cfg, generator=torch.Generator().manual_seed(2)).to(dev)
    N = 20000
    pos = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    g = rng.normal(size=(N, cfg["L"] * cfg["F"])).astype(np.float32)
    g[7, 3] = np.inf
    g[11, 0] = np.nan
    d = _backward_raw(enc, torch.from_numpy(pos).to(dev), torch.from_numpy(g).to(dev)).cpu().numpy()
    want = _c_backward(enc, pos, g)
    bad = ~np.isfinite(want)
    assert bad.any()
    np.testing.assert_array_equal(~np.isfinite(d), bad)
    gf = np.where(np.isfinite(g), g, 0).astype(np.float32)
    absw = _c_backward(enc, pos, np.abs(gf))
    ok = ~bad
    assert (np.abs(d[ok].astype(np.float64) - want[ok]) <= 1e-5 * absw[ok] + 1e-30).all()


def test_fp16_table_backward(dev):
    """f16 tables (the precision path) sort f16-rounded contributions: the gradient
    is within 1e-3 of the float64 scatter relative to each row's absolute
    contributions (f16 keeps 11 bits) plus 1e-6 absolute (contributions below f16's
    smallest subnormal, 6e-8, round to zero; the loss scale keeps training's far
    above it), untouched rows stay 0, and a contribution past the f16 range reaches
    its rows as inf (what the loss-scale check looks for)."""
    from jaxngp_amd import _lib
    rng = np.random.default_rng(31)
    enc = HashGridEncoder(**CONFIGS["l16_t2e19"], param_dtype=torch.float16, generator=torch.Generator().manual_seed(4)).to(dev)
    N = 30000
    pos = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    g = rng.normal(size=(N, 32)).astype(np.float32)
    pos_t = torch.from_numpy(pos).to(dev)
    desc = enc.descriptor(N, 1.0)
    ws = torch.empty(_lib.hashgrid_backward_workspace_bytes(desc), dtype=torch.uint8, device=dev)
    d = torch.full(tuple(enc.table.shape), 5.0, dtype=torch.float32, device=dev)
    _lib.call("hashgrid_encode_backward", [pos_t, torch.from_numpy(g).to(dev), d, ws], desc)
    dt = d.cpu().numpy().astype(np.float64)
    want = _c_backward(enc, pos, g)
    absw = _c_backward(enc, pos, np.abs(g))
    assert (dt[absw == 0] == 0).all()
    assert (np.abs(dt - want) <= 1e-3 * absw + 1e-6).all()
    assert np.abs(dt - want).max() > 0  # the f16 rounding is visible: this is not the f32 path
    g[5, 7] = 1e6  # past 65504 once weighted by anything above 0.066
    _lib.call("hashgrid_encode_backward", [pos_t, torch.from_numpy(g).to(dev), d, ws], desc)
    assert torch.isinf(d).any() and torch.isfinite(d).float().mean() > 0.999


def test_fp16_table_2pow22(dev):
    """The precision workload's larger table: T = 2^22 entries per level (39,474,024
    rows, levels 0-7 dense), f16.  Forward bit-exact against the oracle's f32
    restatement on the f16 values; the table gradient within the f16 bound of the
    float64 scatter (as test_fp16_table_backward), untouched rows 0."""
    from jaxngp_amd import _lib
    rng = np.random.default_rng(41)
    cfg = dict(L=16, T=2 ** 22, F=2, N_min=16, N_max=2048)
    enc = HashGridEncoder(**cfg, param_dtype=torch.float16, generator=torch.Generator().manual_seed(5)).to(dev)
    assert enc.levels.n_rows == 39474024 and enc.levels.first_hash_level == 8
    with torch.no_grad():
        enc.table.copy_(torch.from_numpy(rng.uniform(-1e-2, 1e-2, tuple(enc.table.shape)).astype(np.float16)))
    N = 2 ** 17
    pos = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    pos_t = torch.from_numpy(pos).to(dev)
    print(f"hashgrid f16 T=2^22: N={N}", flush=True)
    out, _ = enc(pos_t, 1.0)
    table32 = enc.table.detach().float().cpu().numpy()
    np.testing.assert_array_equal(out.detach().cpu().numpy(), _c_forward(enc, pos, table32))
    del table32
    g = rng.normal(size=(N, 32)).astype(np.float32)
    desc = enc.descriptor(N, 1.0)
    ws = torch.empty(_lib.hashgrid_backward_workspace_bytes(desc), dtype=torch.uint8, device=dev)
    d = torch.full(tuple(enc.table.shape), 5.0, dtype=torch.float32, device=dev)
    _lib.call("hashgrid_encode_backward", [pos_t, torch.from_numpy(g).to(dev), d, ws], desc)
    dt = d.cpu().numpy().astype(np.float64)
    want = _c_backward(enc, pos, g)
    absw = _c_backward(enc, pos, np.abs(g))
    assert (dt[absw == 0] == 0).all()
    assert (np.abs(dt - want) <= 1e-3 * absw + 1e-6).all()


def test_fused_table_update_matches_gradient_then_adam(dev):
    """hashgrid_encode_backward_adam (the optimizer step taken in the table gradient's
    last pass) against hashgrid_encode_backward followed by adam_update on the same
    state, over three steps: 2^16 points (coarse slices summed in several parts, so
    the combine pass updates them), 20,001 points, and no points (every row takes a
    zero-gradient step).  Both use one element update; the gradient's f64 sums are
    taken in LDS-atomic order, so a row's f32 gradient may differ by its last bit
    between two runs, hence the tolerance; with no points the results are equal."""
    from jaxngp_amd import _lib
    rng = np.random.default_rng(31)
    enc = HashGridEncoder(**CONFIGS["l16_t2e19"], generator=torch.Generator().manual_seed(5)).to(dev)
    L, F = 16, 2
    with torch.no_grad():
        enc.table.uniform_(-1e-2, 1e-2)
    p1 = enc.table.detach().clone()
    m1 = torch.from_numpy(rng.normal(0, 1e-3, tuple(p1.shape)).astype(np.float32)).to(dev)
    v1 = torch.from_numpy(rng.uniform(0, 1e-5, tuple(p1.shape)).astype(np.float32)).to(dev)
    p2, m2, v2 = p1.clone(), m1.clone(), v1.clone()
    n = p1.numel()
    for t, N in ((1, 1 << 16), (2, 20001), (3, 0)):
        print(f"fused table update: step {t}, {N} points", flush=True)
        pos = torch.from_numpy(rng.uniform(-1, 1, (max(N, 1), 3)).astype(np.float32)).to(dev)[:N]
        g = torch.from_numpy(rng.normal(0, 1e-3, (L, max(N, 1), F)).astype(np.float32)).to(dev)[:, :N].contiguous()
        desc = enc.descriptor(N, 1.0, _lib.ENC_LEVEL_MAJOR)
        assert _lib.hashgrid_backward_adam_supported(desc)
        ws = torch.empty(_lib.hashgrid_backward_workspace_bytes(desc) if N else 256, dtype=torch.uint8, device=dev)
        adam = _lib.adam_descriptor(n, 1e-2, 0.9, 0.99, 1e-15, 1e-15, 0.9 ** t, 0.99 ** t, 0.0)
        d = torch.empty_like(p1)
        _lib.call("hashgrid_encode_backward", [pos, g, d, ws], desc)
        _lib.call("adam_update", [p1, d, m1, v1, torch.empty(0, device=dev)], adam)
        _lib.call("hashgrid_encode_backward_adam", [pos, g, p2, m2, v2, ws], _lib.hashgrid_adam_descriptor(desc, adam))
        torch.cuda.synchronize()
        for a, b, name in ((p2, p1, "params"), (m2, m1, "m"), (v2, v1, "v")):
            if N == 0:
                assert torch.equal(a, b), name
            else:
                torch.testing.assert_close(a, b, rtol=1e-5, atol=1e-9, msg=name)
        same = (p2 == p1).float().mean().item()
        print(f"  params equal bit for bit: {same:.6f}", flush=True)
        assert same > 0.999
