"""GPU parity of the jax-tcnn hash-grid targets against the oracle's restatement of
tiny-cuda-nn v1.6's grid encoding (parity with tiny-cuda-nn itself is unpinned: it
is not vendored).  Forward and dy_dcoords: the same f32 operations in the same
order, so they agree to the last bit except where the oracle's emulated fma rounds
twice (rare; bounded by 2 ulp); table gradient within 1e-5 of the float64 scatter
relative to each row's absolute contributions (atomics add in any order)."""
import math

import numpy as np
import pytest
import torch

import oracle
from jaxngp_amd import _lib
from jaxngp_amd.tcnn import TCNNHashGridEncoder, hashgrid_encode, unpadded_offsets

pytestmark = pytest.mark.gpu

CASES = {
    "encoder_workload": dict(L=16, T=2 ** 19, F=2, N_min=16, N_max=2048),
    "f4_small": dict(L=6, T=2 ** 12, F=4, N_min=4, N_max=96),
}


def _run(name, n, seed):
    cfg = CASES[name]
    L, F, N_min = cfg["L"], cfg["F"], cfg["N_min"]
    b = math.exp((math.log(cfg["N_max"]) - math.log(N_min)) / (L - 1))
    offs = unpadded_offsets(L, cfg["T"], N_min, b)
    rng = np.random.default_rng(seed)
    table = rng.uniform(-1, 1, (int(offs[-1]), F)).astype(np.float32)
    pts = rng.uniform(0, 1, (3, n)).astype(np.float32)
    pts[:, :3] = [[0.0, 1.0, 0.5], [0.0, 1.0, 0.999999], [0.0, 1.0, 1e-7]]
    dev = torch.device("cuda:0")
    o_t = torch.from_numpy(offs.astype(np.int32)).to(dev).view(torch.uint32)
    p_t, t_t = torch.from_numpy(pts).to(dev), torch.from_numpy(table).to(dev)
    enc = torch.empty(L * F, n, device=dev)
    dy = torch.empty(L * F, n, 3, device=dev)
    desc = _lib.tcnn_hashgrid_descriptor(n, L, F, N_min, b)
    _lib.call("tcnn_hashgrid_encode", [o_t, p_t, t_t, enc, dy], desc)
    return cfg, b, offs, table, pts, (o_t, p_t, t_t, enc, dy, desc)


@pytest.mark.parametrize("name", list(CASES))
def test_tcnn_forward_and_coordinate_derivative(name):
    cfg, b, offs, table, pts, (o_t, p_t, t_t, enc, dy, desc) = _run(name, 6000, 0)
    want, want_dy = oracle.tcnn_grid_forward(pts, table, offs, cfg["L"], cfg["F"], cfg["N_min"], b)
    got, got_dy = enc.cpu().numpy(), dy.cpu().numpy()
    assert (got == want).mean() > 0.999 and (got_dy == want_dy).mean() > 0.999
    ulp = np.spacing(np.maximum(np.abs(want), 1e-30).astype(np.float32))
    assert (np.abs(got - want) <= 2 * ulp + 1e-30).all()
    np.testing.assert_allclose(got_dy, want_dy, rtol=1e-5, atol=1e-5 * float(np.abs(want_dy).max()))


@pytest.mark.parametrize("name", list(CASES))
def test_tcnn_backward(name):
    cfg, b, offs, table, pts, (o_t, p_t, t_t, enc, dy, desc) = _run(name, 5000, 1)
    L, F = cfg["L"], cfg["F"]
    g = np.random.default_rng(2).normal(size=(L * F, pts.shape[1])).astype(np.float32)
    dev = o_t.device
    dparams = torch.full((int(offs[-1]), F), 9.0, device=dev)  # must be overwritten
    dcoords = torch.empty(3, pts.shape[1], device=dev)
    _lib.call("tcnn_hashgrid_encode_backward", [o_t, p_t, torch.from_numpy(g).to(dev), dy, dparams, dcoords], desc)
    got = dparams.cpu().numpy().astype(np.float64)
    want, want_c = oracle.tcnn_grid_backward64(pts, g, dy.cpu().numpy(), offs, L, F, cfg["N_min"], b, int(offs[-1]))
    absw, _ = oracle.tcnn_grid_backward64(pts, np.abs(g), dy.cpu().numpy(), offs, L, F, cfg["N_min"], b,
                                          int(offs[-1]))
    assert (got[absw == 0] == 0).all()
    assert (np.abs(got - want) <= 1e-5 * absw + 1e-30).all()
    np.testing.assert_allclose(dcoords.cpu().numpy(), want_c, rtol=1e-5, atol=1e-5 * float(np.abs(want_c).max()))


def test_tcnn_encoder_module_autograd():
    """TCNNHashGridEncoder (encoders.py:259-305): positions mapped from [-bound, bound],
    (encodings [N, L*F], 0); gradients reach the table and the positions."""
    dev = torch.device("cuda:0")
    enc_mod = TCNNHashGridEncoder(L=8, T=2 ** 14, F=2, N_min=8, N_max=128, device=dev,
                                  generator=torch.Generator().manual_seed(0))
    assert enc_mod.table.shape == (int(enc_mod.offsets_host[-1]), 2)
    pos = (torch.rand(700, 3, device=dev) * 1.6 - 0.8).requires_grad_(True)
    out, tv = enc_mod(pos, 1.0)
    assert out.shape == (700, 16) and tv == 0
    g = torch.randn_like(out)
    out.backward(g)
    assert enc_mod.table.grad is not None and pos.grad is not None
    # direct call of the jaxtcnn-shaped function gives the same encodings
    p01 = ((pos.detach() + 1.0) / 2.0).t().contiguous()
    e2 = hashgrid_encode(8, 2, 8, enc_mod.b, enc_mod.offsets, p01, enc_mod.table.detach())
    assert torch.equal(e2.t(), out.detach())
    # d(loss)/d(pos) = dL/dcoords / (2 bound)
    dc = torch.empty(3, 700, device=dev)
    dp = torch.empty_like(enc_mod.table)
    dy = torch.empty(16, 700, 3, device=dev)
    e3 = torch.empty(16, 700, device=dev)
    desc = _lib.tcnn_hashgrid_descriptor(700, 8, 2, 8, enc_mod.b)
    _lib.call("tcnn_hashgrid_encode", [enc_mod.offsets, p01, enc_mod.table.detach(), e3, dy], desc)
    _lib.call("tcnn_hashgrid_encode_backward", [enc_mod.offsets, p01, g.t().contiguous(), dy, dp, dc], desc)
    torch.testing.assert_close(pos.grad, dc.t() / 2.0, rtol=1e-6, atol=1e-7)
    torch.testing.assert_close(enc_mod.table.grad, dp, rtol=1e-5, atol=1e-6)  # atomics: any order


def test_tcnn_no_points_zeroes_the_table_gradient():
    dev = torch.device("cuda:0")
    offs = torch.tensor([0, 64, 576], dtype=torch.int32, device=dev).view(torch.uint32)
    dp = torch.full((576, 2), 3.0, device=dev)
    z = torch.zeros(1, device=dev)
    _lib.call("tcnn_hashgrid_encode_backward", [offs, z, z, z, dp, z], _lib.tcnn_hashgrid_descriptor(0, 2, 2, 4, 2.0))
    torch.cuda.synchronize()
    assert (dp == 0).all()
