"""GPU parity of occupancy_grid_update (utils/types.py:1149-1225) against the oracle.

The oracle restates the kernel's draws (tests/test_oracle_ogrid.py pins them), so the
chosen cells and their points are known exactly: which cells change is compared
exactly, the decay (f32 multiply by 0.95f) and the max bit for bit, and the new
densities against float64 (ogrid_density64) within 1e-4 relative (f32 encoder and
MFMA products, K <= 64).  At the training grid size (128^3 cells) the fused kernel
is compared with the unfused chain on the same points (hashgrid_encode then
nerf_mlp_forward), which must give the same densities bit for bit."""
import numpy as np
import pytest
import torch

import oracle
from jaxngp_amd import _lib
from jaxngp_amd.encoders import HashGridEncoder
from jaxngp_amd.nerfs import N_PARAMS, glorot_uniform_, nerf_mlp

pytestmark = pytest.mark.gpu

SEED, STEP = 0x1234_5678_9ABC_DEF1, 77
REL = 1e-4


def _field(dev, T, seed, table_dtype=torch.float32):
    enc = HashGridEncoder(L=16, T=T, F=2, N_min=16, N_max=512, generator=torch.Generator().manual_seed(seed))
    rng = np.random.default_rng(seed)
    lv = oracle.hashgrid_levels(16, T, 16, 512)
    table = rng.uniform(-0.6, 0.6, (lv["n_rows"], 2)).astype(np.float32)
    with torch.no_grad():
        enc.table.data = torch.from_numpy(table).to(table_dtype)
    enc = enc.to(dev)
    params = glorot_uniform_(torch.empty(N_PARAMS), torch.Generator().manual_seed(seed + 1)).to(dev)
    return enc, table, params.cpu().numpy().astype(np.float64), params, lv


def _grid(rng, G, K, alive_frac, occ_frac):
    n = K * G ** 3
    alive = rng.random(n) < alive_frac
    occ = alive & (rng.random(n) < occ_frac)
    density = np.where(alive, 0.0, -1.0).astype(np.float32)
    bits = np.packbits(occ.reshape(-1, 8), axis=-1, bitorder="little").reshape(-1)
    return alive, occ, density, bits


def _run(dev, enc, params_dev, G, K, cas, alive, bits, density, update_all, bound, f16_weights=False):
    G3 = G ** 3
    alive_c = np.flatnonzero(alive[cas * G3:(cas + 1) * G3]).astype(np.int32)
    dens = torch.from_numpy(density.copy()).to(dev)
    prefix = torch.full((G3 // 64 + 1,), -1, dtype=torch.int32, device=dev)
    w = params_dev.half() if f16_weights else params_dev
    desc = _lib.occupancy_update_descriptor(enc.descriptor(0, bound), G, K, cas, alive_c.size, update_all,
                                            _lib.DTYPE_F16 if f16_weights else _lib.DTYPE_F32, SEED, STEP)
    print(f"occupancy_grid_update G={G} K={K} cascade={cas} alive={alive_c.size} update_all={update_all} "
          f"table={enc.table.dtype} f16_weights={f16_weights}", flush=True)
    _lib.call("occupancy_grid_update", [torch.from_numpy(alive_c).to(dev), torch.from_numpy(bits).to(dev),
                                        enc.table.detach(), w, dens, prefix], desc)
    torch.cuda.synchronize()
    return alive_c, dens.cpu().numpy(), prefix.cpu().numpy()


def _expected(alive_c, occ_c, G, cas, update_all, bound, table, params, lv, T, old):
    """Chosen cells, the float64 density max per chosen cell, and the decayed grid."""
    cells, pts = oracle.ogrid_choice(alive_c, occ_c, G, cas, update_all, SEED, STEP, bound)
    sig = oracle.ogrid_density64(pts, table, params, bound, lv, T)
    best = {}
    for c, s in zip(cells.tolist(), sig.tolist()):
        best[c] = max(best.get(c, -np.inf), s)
    decayed = old.copy()
    decayed[alive_c] = old[alive_c] * oracle.OGRID_DECAY
    return best, decayed


@pytest.mark.parametrize("update_all", [1, 0])
def test_update_matches_oracle(dev, update_all):
    rng = np.random.default_rng(3)
    G, K, T, bound = 32, 1, 1 << 16, 1.0
    enc, table, params, params_dev, lv = _field(dev, T, 5)
    alive, occ, density, bits = _grid(rng, G, K, 0.8, 0.05)
    alive_c, got, prefix = _run(dev, enc, params_dev, G, K, 0, alive, bits, density, update_all, bound)
    best, decayed = _expected(alive_c, occ, G, 0, update_all, bound, table, params, lv, T, density)
    changed = np.flatnonzero(got != decayed)
    assert set(changed.tolist()) == set(best)  # all densities start at 0: chosen cells rise above it
    want = np.array([best[c] for c in changed.tolist()])
    assert (np.abs(got[changed] - want) <= REL * want).all(), np.abs(got[changed] / want - 1).max()
    assert (got[~alive] == -1).all()
    if not update_all:
        W = G ** 3 // 64
        words = np.packbits(occ.reshape(-1, 8), axis=-1, bitorder="little").reshape(-1).view("<u8")
        pop = np.array([bin(int(x)).count("1") for x in words], np.int64)
        np.testing.assert_array_equal(prefix[:W + 1], np.concatenate([[0], np.cumsum(pop)]))


def test_decay_and_max(dev):
    """Random old densities: cells no slot chose are exactly 0.95f * old; chosen ones
    are max(decayed, density) (either value where the two lie within the tolerance)."""
    rng = np.random.default_rng(4)
    G, K, T, bound = 32, 1, 1 << 16, 1.0
    enc, table, params, params_dev, lv = _field(dev, T, 6)
    alive, occ, density, bits = _grid(rng, G, K, 0.9, 0.1)
    probe = oracle.ogrid_density64(rng.uniform(-1, 1, (4096, 3)).astype(np.float32), table, params, bound, lv, T)
    density = np.where(alive, rng.uniform(0, 2 * np.median(probe), alive.size), -1.0).astype(np.float32)
    alive_c, got, _ = _run(dev, enc, params_dev, G, K, 0, alive, bits, density, 0, bound)
    best, decayed = _expected(alive_c, occ, G, 0, 0, bound, table, params, lv, T, density)
    chosen = np.zeros(alive.size, bool)
    chosen[list(best)] = True
    np.testing.assert_array_equal(got[~chosen], decayed[~chosen])
    idx = np.flatnonzero(chosen)
    want = np.array([best[c] for c in idx.tolist()])
    d = decayed[idx].astype(np.float64)
    clear_decay = d > want * (1 + REL)
    clear_new = want > d * (1 + REL)
    assert clear_decay.sum() > 100 and clear_new.sum() > 100
    np.testing.assert_array_equal(got[idx[clear_decay]], decayed[idx[clear_decay]])
    assert (np.abs(got[idx[clear_new]] - want[clear_new]) <= REL * want[clear_new]).all()
    amb = ~(clear_decay | clear_new)
    assert (np.abs(got[idx[amb]] - np.maximum(d[amb], want[amb])) <= 2 * REL * want[amb]).all()


def test_second_cascade_leaves_the_first_alone(dev):
    rng = np.random.default_rng(5)
    G, K, T, bound = 16, 2, 1 << 14, 2.0
    enc, table, params, params_dev, lv = _field(dev, T, 7)
    alive, occ, density, bits = _grid(rng, G, K, 0.7, 0.2)
    G3 = G ** 3
    for update_all in (1, 0):
        alive_c, got, _ = _run(dev, enc, params_dev, G, K, 1, alive, bits, density, update_all, bound)
        np.testing.assert_array_equal(got[:G3], density[:G3])
        best, decayed = _expected(alive_c, occ[G3:], G, 1, update_all, bound, table, params, lv, T, density[G3:])
        changed = np.flatnonzero(got[G3:] != decayed)
        assert set(changed.tolist()) == set(best)
        want = np.array([best[c] for c in changed.tolist()])
        assert (np.abs(got[G3:][changed] - want) <= REL * want).all()


def test_no_occupied_cell_and_no_alive_cell(dev):
    rng = np.random.default_rng(6)
    G, K, T, bound = 16, 1, 1 << 14, 1.0
    enc, table, params, params_dev, lv = _field(dev, T, 8)
    alive, occ, density, bits = _grid(rng, G, K, 0.5, 0.0)
    alive_c, got, _ = _run(dev, enc, params_dev, G, K, 0, alive, bits, density, 0, bound)
    best, decayed = _expected(alive_c, occ, G, 0, 0, bound, table, params, lv, T, density)
    assert set(np.flatnonzero(got != decayed).tolist()) == set(best)
    none = np.zeros_like(alive)
    _, got0, _ = _run(dev, enc, params_dev, G, K, 0, none, bits, density, 0, bound)
    np.testing.assert_array_equal(got0, density)


def test_f16_table_and_weights(dev):
    """The precision path chooses the same cells; its densities (f16 table rows and
    MLP weights, f32 accumulation) stay within 3e-2 of float64 for 99% of the cells."""
    rng = np.random.default_rng(7)
    G, K, T, bound = 32, 1, 1 << 16, 1.0
    enc, table, params, params_dev, lv = _field(dev, T, 9, table_dtype=torch.float16)
    alive, occ, density, bits = _grid(rng, G, K, 0.8, 0.05)
    alive_c, got, _ = _run(dev, enc, params_dev, G, K, 0, alive, bits, density, 0, bound, f16_weights=True)
    t16 = table.astype(np.float16).astype(np.float32)
    p16 = params.astype(np.float16).astype(np.float64)
    best, decayed = _expected(alive_c, occ, G, 0, 0, bound, t16, p16, lv, T, density)
    changed = np.flatnonzero(got != decayed)
    assert set(changed.tolist()) == set(best)
    want = np.array([best[c] for c in changed.tolist()])
    assert np.mean(np.abs(got[changed] / want - 1) <= 3e-2) > 0.99


def test_full_grid_equals_unfused_chain(dev):
    """128^3 cells, all alive, every cell updated: the fused densities equal
    hashgrid_encode + nerf_mlp_forward on the same points bit for bit."""
    G, K, T, bound = 128, 1, 1 << 19, 1.0
    enc, table, params, params_dev, lv = _field(dev, T, 10)
    G3 = G ** 3
    alive = np.ones(G3, bool)
    density = np.zeros(G3, np.float32)
    bits = np.full(G3 // 8, 255, np.uint8)
    alive_c, got, _ = _run(dev, enc, params_dev, G, K, 0, alive, bits, density, 1, bound)
    cells, pts = oracle.ogrid_choice(alive_c, np.ones(G3, bool), G, 0, True, SEED, STEP, bound)
    with torch.no_grad():
        x = torch.from_numpy(pts).to(dev)
        e, _ = enc(x, bound)
        dirs = torch.zeros(G3, 3, device=dev)
        sig = nerf_mlp(e, dirs, params_dev)[:, 0].cpu().numpy()
    np.testing.assert_array_equal(got[cells], np.maximum(np.float32(0), sig))
