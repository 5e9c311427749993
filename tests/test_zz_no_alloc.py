"""nerf_mlp_backward_no_alloc (fixed device scratch, no allocation) against
nerf_mlp_backward (caller workspace), called through the C ABI.

Both run the same kernels with the same number of workgroups, and the weight-
gradient rows are summed in a fixed order, so the outputs must agree bit for bit.
Two calls in a row, the second on a smaller input, show that the scratch's
leftovers from the first call do not leak into the second (every workgroup
writes its whole row)."""
import numpy as np
import pytest
import torch

from jaxngp_amd import _lib
from jaxngp_amd.nerfs import N_PARAMS, glorot_uniform_, mlp_workspace_blocks

pytestmark = pytest.mark.gpu


def _inputs(dev, S, seed, level_major, f16):
    rng = np.random.default_rng(seed)
    enc = rng.normal(0, 0.5, (S, 32)).astype(np.float32)
    if level_major:
        enc = np.ascontiguousarray(enc.reshape(S, 16, 2).transpose(1, 0, 2))
    d = rng.normal(size=(S, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    params = glorot_uniform_(torch.empty(N_PARAMS), torch.Generator().manual_seed(seed))
    g = rng.normal(size=(S, 4)).astype(np.float32)
    T = lambda a: torch.from_numpy(a).to(dev)
    p = params.to(dev)
    return T(enc), T(d), (p.half() if f16 else p), T(g)


def _backward(name, enc, d, params, g, blocks, layout, dtype):
    S = d.shape[0]
    denc = torch.full_like(enc, float("nan"))
    dparams = torch.full((N_PARAMS,), float("nan"), device=enc.device)
    bufs = [enc, d, params, g, denc, dparams]
    if name == "nerf_mlp_backward":
        # garbage in the caller's workspace: every row must be overwritten
        bufs.append(torch.full((blocks, N_PARAMS), float("nan"), device=enc.device))
    _lib.call(name, bufs, _lib.nerf_mlp_descriptor(S, blocks, layout, dtype))
    torch.cuda.synchronize()
    return denc.cpu(), dparams.cpu()


@pytest.mark.parametrize("level_major,f16", [(False, False), (True, True)])
def test_no_alloc_matches_workspace_entry_point(dev, level_major, f16):
    layout = _lib.ENC_LEVEL_MAJOR if level_major else _lib.ENC_POINT_MAJOR
    dtype = _lib.DTYPE_F16 if f16 else _lib.DTYPE_F32
    for S, seed in ((70001, 1), (3000, 2)):  # 512 workgroups, then 47: the second call is smaller
        blocks = mlp_workspace_blocks(S)
        print(f"no-alloc MLP backward: S={S} blocks={blocks} level_major={level_major} f16={f16}", flush=True)
        enc, d, params, g = _inputs(dev, S, seed, level_major, f16)
        want = _backward("nerf_mlp_backward", enc, d, params, g, blocks, layout, dtype)
        got = _backward("nerf_mlp_backward_no_alloc", enc, d, params, g, blocks, layout, dtype)
        assert torch.isfinite(want[0]).all() and torch.isfinite(want[1]).all()
        assert torch.equal(got[0], want[0]), "dL_dencodings"
        assert torch.equal(got[1], want[1]), "dL_dparams"


def test_no_alloc_rejects_more_blocks_than_its_scratch(dev):
    enc, d, params, g = _inputs(dev, 64, 3, False, False)
    denc, dparams = torch.empty_like(enc), torch.empty(N_PARAMS, device=dev)
    with pytest.raises(_lib.NGPError, match="exceeds the fixed scratch"):
        _lib.call("nerf_mlp_backward_no_alloc", [enc, d, params, g, denc, dparams],
                  _lib.nerf_mlp_descriptor(64, 513))
    assert _lib.last_error() == (0, "")
