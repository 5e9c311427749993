"""Pins of the occupancy-grid update oracle (utils/types.py:1149-1225):
the random draws against the published splitmix64 reference sequence, the cell
choice against its definition (every alive cell; or halves uniform over the alive
and the occupied cells), the points against the cell geometry (each point lies in
its own cell, within half a cell width of the centre), and the density against the
pinned float64 NeRF MLP."""
import numpy as np
import pytest

import oracle


def test_mix64_is_splitmix64():
    """splitmix64's output k is mix64(seed + (k + 1) * 0x9e3779b97f4a7c15); the values
    are those of the reference implementation (xoshiro.di.unimi.it/splitmix64.c) for
    seed 1477776061723855037, as published with rand_xoshiro's SplitMix64 test."""
    g, s = 0x9E3779B97F4A7C15, 1477776061723855037
    states = np.array([(s + (k + 1) * g) % 2 ** 64 for k in range(6)], np.uint64)
    want = [1985237415132408290, 2979275885539914483, 13511426838097143398, 8488337342461049707,
            15141737807933549159, 17093170987380407015]
    assert [int(x) for x in oracle.mix64(states)] == want


def test_below_is_a_multiply_shift_index():
    r = np.array([0, (1 << 32) - 1, 1 << 63, 0xFFFFFFFF_FFFFFFFF], np.uint64)
    assert oracle.below(r, 10).tolist() == [0, 0, 5, 9]
    rng = np.random.default_rng(0)
    x = oracle.below(rng.integers(0, 2 ** 63, 200000, dtype=np.int64).astype(np.uint64) << np.uint64(1), 7)
    counts = np.bincount(x.astype(np.int64), minlength=7)
    assert counts.size == 7 and abs(counts - 200000 / 7).max() < 5 * np.sqrt(200000 / 7)


@pytest.mark.parametrize("G,cascade,bound", [(16, 0, 1.0), (32, 1, 2.0), (8, 2, 3.0)])
def test_points_lie_in_their_cells(G, cascade, bound):
    alive = np.arange(G ** 3)
    cells, pts = oracle.ogrid_choice(alive, np.zeros(G ** 3, bool), G, cascade, True, 7, 3, bound)
    np.testing.assert_array_equal(cells, alive)
    mip = min(bound, 2.0 ** cascade)
    hcw = mip / G
    v = oracle.morton3d_invert(cells.astype(np.uint32)).astype(np.float64)
    centre = (2 * v + 1) * hcw - mip  # (v / (G-1) * 2 - 1) * (mip - hcw) is this centre
    off = pts.astype(np.float64) - centre
    assert (np.abs(off) <= hcw * (1 + 1e-5)).all()
    inside = np.floor((pts.astype(np.float64) + mip) / (2 * hcw))
    assert (inside == v).mean() > 0.999  # all but f32 ties on a cell face
    # the offsets cover the cell: both signs on every axis, spread over the width
    assert (off.min(0) < -0.9 * hcw).all() and (off.max(0) > 0.9 * hcw).all()


def test_sampled_choice_halves():
    rng = np.random.default_rng(1)
    G = 16
    alive = np.sort(rng.choice(G ** 3, 3000, replace=False))
    occupied = np.zeros(G ** 3, bool)
    occ_cells = rng.choice(alive, 40, replace=False)
    occupied[occ_cells] = True
    cells, pts = oracle.ogrid_choice(alive, occupied, G, 0, False, 11, 300, 1.0)
    M = max(1, 3000 // 2)
    h = max(1, M // 2)
    assert cells.shape == (2 * h,) and pts.shape == (2 * h, 3)
    assert np.isin(cells[:h], alive).all()
    assert np.isin(cells[h:], occ_cells).all()
    # uniform over the occupied cells: every one drawn, counts within 5 sigma
    c = np.bincount(np.searchsorted(np.sort(occ_cells), cells[h:]), minlength=40)
    assert (c > 0).all() and abs(c - h / 40).max() < 5 * np.sqrt(h / 40)
    # the first half spreads over the alive cells: distinct cells as expected of h
    # uniform draws among 3000
    expect = 3000 * (1 - (1 - 1 / 3000) ** h)
    assert abs(np.unique(cells[:h]).size - expect) < 0.1 * expect
    # no occupied cell: the second half is uniform over the alive cells too
    cells0, _ = oracle.ogrid_choice(alive, np.zeros(G ** 3, bool), G, 0, False, 11, 300, 1.0)
    np.testing.assert_array_equal(cells0[:h], cells[:h])
    assert np.isin(cells0[h:], alive).all() and not np.isin(cells0[h:], occ_cells).all()
    # one alive cell: one slot per half
    one, _ = oracle.ogrid_choice(alive[:1], occupied, G, 0, False, 11, 300, 1.0)
    assert one.shape == (2,) and one[0] == alive[0]


def test_draws_depend_on_step_cascade_and_seed():
    alive = np.arange(4096)
    a = oracle.ogrid_choice(alive, np.zeros(4096, bool), 16, 0, False, 5, 17, 1.0)[0]
    for args in ((5, 18, 0), (6, 17, 0), (5, 17, 1)):
        b = oracle.ogrid_choice(alive, np.zeros(4096, bool), 16, args[2], False, args[0], args[1], 1.0)[0]
        assert (a != b).mean() > 0.9


def test_density_is_the_mlp_density_output():
    """The update's density equals the first output of the pinned float64 NeRF
    forward (density = trunc_exp(density_mlp(enc)[..., 0]), nerfs.py:66-69), whatever
    the direction."""
    rng = np.random.default_rng(2)
    lv = oracle.hashgrid_levels(16, 1 << 12, 16, 256)
    table = rng.uniform(-1, 1, (lv["n_rows"], 2)).astype(np.float32)
    params = rng.normal(0, 0.3, 9408)
    pts = rng.uniform(-1, 1, (500, 3)).astype(np.float32)
    dens = oracle.ogrid_density64(pts, table, params, 1.0, lv, 1 << 12)
    dirs = rng.normal(size=(500, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    want = oracle.mlp_forward64(oracle.hashgrid_forward64(pts, table, 1.0, lv, 1 << 12), dirs, params)[0][:, 0]
    np.testing.assert_allclose(dens, want, rtol=1e-12)
    assert dens.std() > 0.1 * dens.mean()
