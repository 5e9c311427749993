"""GPU parity of the learned-proposal kernels against the float64 definitions in
the oracle (proposal_mlp_forward64 / backward64 / proposal_loss64, pinned by finite
differences in test_oracle_pins.py).

Tolerances (f32 arithmetic against float64):
  points:            bit-exact against the same f32 formula in numpy
  MLP sigma:         |a - b| <= 1e-5 * |b|
  MLP gradients:     |a - b| <= 1e-5 * max|b| + 1e-6 * |b|  (sums over up to 2^17 points)
  loss:              relative 1e-4;  dL/dsigma_p: |a - b| <= 1e-4 * max|b|
"""
import numpy as np
import pytest
import torch

import oracle
from jaxngp_amd import _lib
from tests.test_oracle_pins import _prop_rays

pytestmark = pytest.mark.gpu


def test_proposal_points(dev):
    rng = np.random.default_rng(0)
    n = 1000
    o = rng.normal(size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    t0 = rng.uniform(0, 2, n).astype(np.float32)
    t1 = (t0 + rng.uniform(0.1, 3, n)).astype(np.float32)
    pts = torch.empty(n, 64, 3, device=dev)
    T = lambda a: torch.from_numpy(a).to(dev)
    _lib.call("proposal_points", [T(o), T(d), T(t0), T(t1), pts], _lib.proposal_rays_descriptor(n))
    f = np.float32
    D = (t1 - t0) / f(64)
    tc = t0[:, None] + (np.arange(64, dtype=f)[None, :] + f(0.5)) * D[:, None]
    want = o[:, None, :] + tc[:, :, None] * d[:, None, :]
    np.testing.assert_array_equal(pts.cpu().numpy(), want.astype(f))


@pytest.mark.parametrize("N,dtype,layout", [(1, np.float32, 0), (777, np.float32, 1), (1 << 17, np.float32, 1),
                                            (50000, np.float16, 0)])
def test_proposal_mlp(dev, N, dtype, layout):
    rng = np.random.default_rng(N)
    enc = rng.normal(0, 0.6, (N, 10)).astype(np.float32)
    p = rng.normal(0, 0.4, 176).astype(dtype)
    g = rng.normal(size=N).astype(np.float32)
    e = torch.from_numpy(enc).to(dev)
    if layout == _lib.ENC_LEVEL_MAJOR:
        e = e.reshape(N, 5, 2).permute(1, 0, 2).contiguous()
    pt = torch.from_numpy(p).to(dev)
    pd = _lib.DTYPE_F16 if dtype == np.float16 else _lib.DTYPE_F32
    sig = torch.empty(N, device=dev)
    _lib.call("proposal_mlp_forward", [e, pt, sig], _lib.proposal_mlp_descriptor(N, 0, layout, pd))
    p64 = p.astype(np.float64)
    want, _, _ = oracle.proposal_mlp_forward64(enc, p64)
    got = sig.cpu().numpy()
    assert (np.abs(got - want) <= 1e-5 * np.abs(want)).all(), np.abs(got / want - 1).max()
    blocks = max(1, min(512, (N + 255) // 256))
    denc = torch.empty_like(e)
    dp = torch.empty(176, device=dev)
    ws = torch.empty(blocks, 176, device=dev)
    _lib.call("proposal_mlp_backward", [e, pt, torch.from_numpy(g).to(dev), denc, dp, ws],
              _lib.proposal_mlp_descriptor(N, blocks, layout, pd))
    if layout == _lib.ENC_LEVEL_MAJOR:
        denc = denc.permute(1, 0, 2).reshape(N, 10)
    dX, dP = oracle.proposal_mlp_backward64(enc, p64, g)
    gx, gp = denc.cpu().numpy(), dp.cpu().numpy()
    assert (np.abs(gx - dX) <= 1e-5 * np.abs(dX).max() + 1e-6 * np.abs(dX)).all()
    assert (np.abs(gp - dP) <= 1e-5 * np.abs(dP).max() + 1e-6 * np.abs(dP)).all()


@pytest.mark.parametrize("seed,n_rays", [(3, 5), (4, 300)])
def test_proposal_loss(dev, seed, n_rays):
    rng = np.random.default_rng(seed)
    start, ns, dss, z, drgbs, t0, t1 = _prop_rays(rng, n_rays)
    valid = rng.uniform(size=n_rays) < 0.8
    n_valid = int(valid.sum())
    sig = rng.uniform(0, 30, (n_rays, 64)).astype(np.float32)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out = torch.empty(n_rays, 64, device=dev)
    loss = torch.empty(1, device=dev)
    nv = torch.tensor([n_valid], dtype=torch.int32, device=dev).view(torch.uint32)
    _lib.call("proposal_loss", [T(start).view(torch.uint32), T(ns).view(torch.uint32), T(dss), T(z), T(drgbs), T(t0),
                                T(t1), T(sig), T(valid), nv, out, loss], _lib.proposal_loss_descriptor(n_rays))
    want_l, want_g = oracle.proposal_loss64(start, ns, dss, z, drgbs, t0, t1, sig, valid, n_valid)
    got_l, got_g = float(loss.item()), out.cpu().numpy()
    assert abs(got_l - want_l) <= 1e-4 * want_l, (got_l, want_l)
    assert (np.abs(got_g - want_g) <= 1e-4 * np.abs(want_g).max()).all(), np.abs(got_g - want_g).max()
    assert (got_g[~valid] == 0).all()
