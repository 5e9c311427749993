"""GPU parity of the fused NeRF MLP (MFMA, f32 accumulate) against the float64
transcription of models/nerfs.py:39-128 (oracle.mlp_forward64 / mlp_backward64,
themselves pinned by finite differences in test_oracle_pins.py).

Tolerances (f32 MFMA products and sums over K <= 64):
  drgbs: |a-b| <= 2e-5 * (1 + |b|)
  dL/denc, dL/dparams: |a-b| <= 2e-4 * max|b| + 1e-5 * |b| per output (sums over
  up to 2^16 samples for the parameter gradient).
"""
import numpy as np
import pytest
import torch

import oracle
from jaxngp_amd.nerfs import N_PARAMS, glorot_uniform_, nerf_mlp

pytestmark = pytest.mark.gpu


def _inputs(S, seed):
    rng = np.random.default_rng(seed)
    enc = rng.normal(0, 0.5, (S, 32)).astype(np.float32)
    d = rng.normal(size=(S, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    params = glorot_uniform_(torch.empty(N_PARAMS), torch.Generator().manual_seed(seed)).numpy()
    g = rng.normal(size=(S, 4)).astype(np.float32)
    return enc, d, params, g


@pytest.mark.parametrize("S", [1, 15, 16, 17, 1000, 65536 + 5])
def test_mlp_forward_backward(dev, S):
    enc, d, params, g = _inputs(S, S)
    enc_t = torch.from_numpy(enc).to(dev).requires_grad_(True)
    par_t = torch.from_numpy(params).to(dev).requires_grad_(True)
    out = nerf_mlp(enc_t, torch.from_numpy(d).to(dev), par_t)
    want, _ = oracle.mlp_forward64(enc, d, params)
    got = out.detach().cpu().numpy()
    assert (np.abs(got - want) <= 2e-5 * (1 + np.abs(want))).all(), np.abs(got - want).max()
    out.backward(torch.from_numpy(g).to(dev))
    dX, dP = oracle.mlp_backward64(enc, d, params, g)
    gX = enc_t.grad.cpu().numpy()
    gP = par_t.grad.cpu().numpy()
    assert (np.abs(gX - dX) <= 2e-4 * np.abs(dX).max() + 1e-5 * np.abs(dX)).all(), np.abs(gX - dX).max()
    assert (np.abs(gP - dP) <= 2e-4 * np.abs(dP).max() + 1e-5 * np.abs(dP)).all(), np.abs(gP - dP).max()


def test_mlp_backward_deterministic(dev):
    enc, d, params, g = _inputs(20000, 3)
    res = []
    for _ in range(2):
        par_t = torch.from_numpy(params).to(dev).requires_grad_(True)
        out = nerf_mlp(torch.from_numpy(enc).to(dev), torch.from_numpy(d).to(dev), par_t)
        out.backward(torch.from_numpy(g).to(dev))
        res.append(par_t.grad.cpu().numpy())
    np.testing.assert_array_equal(res[0], res[1])


def test_mlp_zero_cotangent_and_padding(dev):
    """Padding samples (x = 0, dir = 0, zero cotangent) contribute nothing."""
    enc, d, params, g = _inputs(100, 4)
    enc[60:] = 0
    d[60:] = 0
    g[60:] = 0
    par_t = torch.from_numpy(params).to(dev).requires_grad_(True)
    out = nerf_mlp(torch.from_numpy(enc).to(dev), torch.from_numpy(d).to(dev), par_t)
    out.backward(torch.from_numpy(g).to(dev))
    _, dP = oracle.mlp_backward64(enc[:60], d[:60], params, g[:60])
    gP = par_t.grad.cpu().numpy()
    assert (np.abs(gP - dP) <= 2e-4 * np.abs(dP).max() + 1e-5 * np.abs(dP)).all()


def test_mlp_level_major_encodings(dev):
    from jaxngp_amd import _lib
    from jaxngp_amd.nerfs import mlp_workspace_blocks
    S = 5000
    enc, d, params, g = _inputs(S, 11)
    enc_t = torch.from_numpy(enc).to(dev)
    lm = enc_t.reshape(S, 16, 2).permute(1, 0, 2).contiguous()
    d_t = torch.from_numpy(d).to(dev)
    p_t = torch.from_numpy(params).to(dev)
    g_t = torch.from_numpy(g).to(dev)
    outs = []
    for layout, e in ((_lib.ENC_POINT_MAJOR, enc_t), (_lib.ENC_LEVEL_MAJOR, lm)):
        out = torch.empty(S, 4, device=dev)
        _lib.call("nerf_mlp_forward", [e, d_t, p_t, out], _lib.nerf_mlp_descriptor(S, 0, layout))
        denc = torch.empty_like(e)
        dp = torch.empty(N_PARAMS, device=dev)
        blocks = mlp_workspace_blocks(S)
        ws = torch.empty(blocks, N_PARAMS, device=dev)
        _lib.call("nerf_mlp_backward", [e, d_t, p_t, g_t, denc, dp, ws], _lib.nerf_mlp_descriptor(S, blocks, layout))
        if layout == _lib.ENC_LEVEL_MAJOR:
            denc = denc.permute(1, 0, 2).reshape(S, 32)
        outs.append((out, denc, dp))
    for a, b in zip(outs[0], outs[1]):
        torch.testing.assert_close(a, b, rtol=0, atol=0)


@pytest.mark.parametrize("half", [False, True])
def test_mlp_no_samples(dev, half):
    """n_samples = 0: the forward does nothing and the backward still writes an all-zero
    weight gradient (its workspace rows are written, not left stale)."""
    from jaxngp_amd import _lib
    p = glorot_uniform_(torch.empty(N_PARAMS), torch.Generator().manual_seed(3)).to(dev)
    p_t = p.half() if half else p
    dtype = _lib.DTYPE_F16 if half else _lib.DTYPE_F32
    e = torch.zeros(1, 32, device=dev)
    d_t = torch.zeros(1, 3, device=dev)
    g_t = torch.zeros(1, 4, device=dev)
    denc = torch.zeros(1, 32, device=dev)
    dp = torch.full((N_PARAMS,), 7.0, device=dev)
    ws = torch.full((4, N_PARAMS), 5.0, device=dev)
    _lib.call("nerf_mlp_backward", [e, d_t, p_t, g_t, denc, dp, ws], _lib.nerf_mlp_descriptor(0, 4, 0, dtype))
    torch.cuda.synchronize()
    assert (dp == 0).all() and (ws == 0).all()
