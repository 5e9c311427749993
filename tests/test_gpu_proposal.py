"""GPU parity of the learned-proposal kernels against the float64 definitions in
the oracle (proposal_mlp_forward64 / backward64 / proposal_loss64, pinned by finite
differences in test_oracle_pins.py).

Tolerances (f32 arithmetic against float64):
  points:            bit-exact against the same f32 formula in numpy
  MLP sigma:         |a - b| <= 1e-5 * |b|
  MLP gradients:     |a - b| <= 1e-5 * max|b| + 1e-6 * |b|  (sums over up to 2^17 points)
  loss:              relative 1e-4;  dL/dsigma_p: |a - b| <= 1e-4 * max|b|
  sampling:          counts and fallback flags exact; sample distances within
                     2e-5 of the segment length; ds = forward differences exactly
                     (within 1e-6 of the segment); the last sample's ds within 1e-3
                     relative for 99% of rays (its bin can differ at a CDF boundary)
"""
import numpy as np
import pytest
import torch

import oracle
from jaxngp_amd import _lib
from tests.helpers import blob_bitfield
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


def test_proposal_sample(dev):
    rng = np.random.default_rng(11)
    n, N = 2000, 48
    o = rng.normal(size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    t0 = rng.uniform(0, 2, n).astype(np.float32)
    t1 = (t0 + rng.uniform(0.5, 3, n)).astype(np.float32)
    t1[::97] = t0[::97]  # empty segments
    # densities: a few bumps per ray, some rays nearly transparent (fallback)
    bins = np.arange(64)[None, :]
    c = rng.uniform(0, 64, (n, 1))
    sig = (rng.uniform(5, 400, (n, 1)) * np.exp(-0.5 * ((bins - c) / rng.uniform(1, 8, (n, 1))) ** 2)).astype(np.float32)
    sig[::7] *= 0.01
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    start = torch.empty(n, dtype=torch.int32, device=dev).view(torch.uint32)
    cnt = torch.empty(n, dtype=torch.int32, device=dev).view(torch.uint32)
    fb = torch.empty(n, dtype=torch.bool, device=dev)
    xyz = torch.empty(n, N, 3, device=dev)
    ds = torch.empty(n, N, device=dev)
    z = torch.empty(n, N, device=dev)
    full = torch.full((128 ** 3 // 8,), 255, dtype=torch.uint8, device=dev)
    _lib.call("proposal_sample", [T(o), T(d), T(t0), T(t1), T(sig), full, start, cnt, fb, xyz, ds, z],
              _lib.proposal_sample_descriptor(n, N, 0.9))
    want_c, want_fb, want_t, want_ds = oracle.proposal_sample64(t0, t1, sig, N, 0.9)
    # masses within 1e-6 of the threshold could round either way: none here
    got_c = cnt.view(torch.int32).cpu().numpy()
    np.testing.assert_array_equal(got_c, want_c)
    np.testing.assert_array_equal(fb.cpu().numpy(), want_fb)
    np.testing.assert_array_equal(start.view(torch.int32).cpu().numpy(), np.arange(n) * N)
    ok = want_c > 0
    assert ok.sum() > n // 2 and want_fb.sum() > 100
    zt = z.cpu().numpy()
    span = (t1 - t0)[:, None]
    assert (np.abs(zt[ok] - want_t[ok]) <= 2e-5 * span[ok] + 1e-6).all(), np.abs(zt[ok] - want_t[ok]).max()
    assert (zt[~ok] == 0).all() and (ds.cpu().numpy()[~ok] == 0).all()
    np.testing.assert_allclose(xyz.cpu().numpy()[ok], (o[:, None, :] + zt[:, :, None] * d[:, None, :])[ok],
                               rtol=0, atol=1e-5)
    # ds: forward differences of the sample positions (and the last sample's bin spacing)
    dsg = ds.cpu().numpy()
    assert (np.abs(dsg[ok][:, :-1] - np.diff(zt[ok], axis=1)) <= 1e-6 * span[ok]).all()
    assert (np.abs(dsg[ok][:, -1] - want_ds[ok][:, -1]) <= 1e-3 * np.abs(want_ds[ok][:, -1]) + 1e-6).mean() > 0.99


def test_proposal_sample_support_mode(dev):
    rng = np.random.default_rng(12)
    n, N = 1500, 48
    o = rng.normal(size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    t0 = rng.uniform(0, 2, n).astype(np.float32)
    t1 = (t0 + rng.uniform(0.5, 3, n)).astype(np.float32)
    bins = np.arange(64)[None, :]
    c = rng.uniform(0, 64, (n, 1))
    sig = (rng.uniform(5, 400, (n, 1)) * np.exp(-0.5 * ((bins - c) / rng.uniform(1, 8, (n, 1))) ** 2)).astype(np.float32)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    start = torch.empty(n, dtype=torch.int32, device=dev).view(torch.uint32)
    cnt = torch.empty(n, dtype=torch.int32, device=dev).view(torch.uint32)
    fb = torch.empty(n, dtype=torch.bool, device=dev)
    xyz = torch.empty(n, N, 3, device=dev)
    ds = torch.empty(n, N, device=dev)
    z = torch.empty(n, N, device=dev)
    full = torch.full((128 ** 3 // 8,), 255, dtype=torch.uint8, device=dev)
    _lib.call("proposal_sample", [T(o), T(d), T(t0), T(t1), T(sig), full, start, cnt, fb, xyz, ds, z],
              _lib.proposal_sample_descriptor(n, N, 0.9, _lib.PROPOSAL_SUPPORT, 1e-3))
    want_c, want_fb, want_t, want_ds = oracle.proposal_sample_support64(t0, t1, sig, N, 0.9, 1e-3)
    np.testing.assert_array_equal(cnt.view(torch.int32).cpu().numpy(), want_c)
    np.testing.assert_array_equal(fb.cpu().numpy(), want_fb)
    ok = want_c > 0
    span = (t1 - t0)[:, None]
    zt, dsg = z.cpu().numpy(), ds.cpu().numpy()
    # the support's end bins can differ where a cumulative share sits within f32
    # rounding of eps or 1 - eps: allow that for 1% of the rays
    close = (np.abs(zt - want_t) <= 2e-5 * span + 1e-6).all(axis=1) & \
            (np.abs(dsg - want_ds) <= 2e-5 * span + 1e-7).all(axis=1)
    assert close[ok].mean() > 0.99, close[ok].mean()

    # occupancy: with a grid whose cells are empty for x < 0, the samples in those
    # cells are dropped and the others kept in order (grid x = floor((x + 1) / 2 * 128))
    from tests.helpers import blob_bitfield
    xs, ys, zs_ = np.meshgrid(np.arange(128), np.arange(128), np.arange(128), indexing="ij")
    m = oracle.morton3d(np.stack([xs.ravel(), ys.ravel(), zs_.ravel()], 1).astype(np.uint32))
    flat = np.zeros(128 ** 3, bool)
    flat[m] = (xs >= 64).ravel()
    half = T(np.packbits(flat.reshape(-1, 8), axis=-1, bitorder="little").reshape(-1))
    cnt2 = torch.empty_like(cnt)
    xyz2, ds2, z2 = torch.empty_like(xyz), torch.empty_like(ds), torch.empty_like(z)
    _lib.call("proposal_sample", [T(o), T(d), T(t0), T(t1), T(sig), half, start, cnt2, fb, xyz2, ds2, z2],
              _lib.proposal_sample_descriptor(n, N, 0.9, _lib.PROPOSAL_SUPPORT, 1e-3))
    x = xyz.cpu().numpy()[..., 0]
    ix = np.floor(((x / np.float32(1.0) + np.float32(1.0)) * np.float32(0.5)) * np.float32(128))
    c2 = cnt2.view(torch.int32).cpu().numpy()
    zz2, dd2 = z2.cpu().numpy(), ds2.cpu().numpy()
    for r in np.nonzero(ok)[0][:300]:
        keep = ix[r] >= 64
        assert c2[r] == keep.sum()
        np.testing.assert_array_equal(zz2[r, :c2[r]], zt[r][keep])
        np.testing.assert_array_equal(dd2[r, :c2[r]], dsg[r][keep])
        assert (zz2[r, c2[r]:] == 0).all() and (dd2[r, c2[r]:] == 0).all()


def test_proposal_render_samples_fused(dev):
    """The fused kernel equals proposal_points + hashgrid_encode + proposal_mlp_forward
    + proposal_sample bit for bit (same arithmetic in the same order)."""
    from jaxngp_amd.proposal import ProposalNet
    rng = np.random.default_rng(13)
    n, N = 3000, 48
    prop = ProposalNet(1.0, dev, seed=5, table_dtype=torch.float16)
    with torch.no_grad():
        prop.position_encoder.table.uniform_(-1.0, 1.0)
        prop.mlp_params.mul_(3.0)
    prop.mlp_half = prop.mlp_params.detach().half()
    o = (rng.normal(size=(n, 3)) * 0.3 - np.array([0, 0, 3.0])).astype(np.float32)
    d = rng.normal(size=(n, 3)) * 0.2 + np.array([0, 0, 1.0])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    from jaxngp_amd.renderer import make_near_far_from_bound
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ot, dt = T(o), T(d)
    t0, t1 = make_near_far_from_bound(1.0, ot, dt)
    bits, _ = blob_bitfield(rng)
    bt = T(bits)
    outs = []
    for fused in (False, True):
        start = torch.empty(n, dtype=torch.int32, device=dev).view(torch.uint32)
        cnt = torch.empty_like(start)
        fb = torch.empty(n, dtype=torch.bool, device=dev)
        xyz, ds, z = torch.empty(n, N, 3, device=dev), torch.empty(n, N, device=dev), torch.empty(n, N, device=dev)
        sd = _lib.proposal_sample_descriptor(n, N, 0.5, _lib.PROPOSAL_SUPPORT, 1e-3)
        if fused:
            w, pd = prop.weights()
            _lib.call("proposal_render_samples", [ot, dt, t0, t1, prop.position_encoder.table.detach(), w, bt, start,
                                                  cnt, fb, xyz, ds, z],
                      _lib.proposal_render_descriptor(prop.position_encoder.descriptor(0, 1.0), sd, pd))
        else:
            sig = prop.bin_density(ot, dt, t0, t1)
            _lib.call("proposal_sample", [ot, dt, t0, t1, sig, bt, start, cnt, fb, xyz, ds, z], sd)
        outs.append([a.cpu() for a in (start.view(torch.int32), cnt.view(torch.int32), fb, xyz, ds, z)])
    for a, b in zip(*outs):
        assert torch.equal(a, b)
    assert outs[0][1].sum() > 0 and outs[0][2].sum() > 0


def _train_sample_inputs(dev, seed, n, table_dtype):
    from jaxngp_amd.proposal import ProposalNet
    from jaxngp_amd.renderer import make_near_far_from_bound
    rng = np.random.default_rng(seed)
    prop = ProposalNet(1.0, dev, seed=seed, table_dtype=table_dtype)
    with torch.no_grad():
        prop.position_encoder.table.uniform_(-1.0, 1.0)
        prop.mlp_params.mul_(3.0)
    if table_dtype == torch.float16:
        prop.mlp_half = prop.mlp_params.detach().half()
    o = (rng.normal(size=(n, 3)) * 0.3 - np.array([0, 0, 3.0])).astype(np.float32)
    d = rng.normal(size=(n, 3)) * 0.25 + np.array([0, 0, 1.0])
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    d[:7] = [0, 1, 0]  # rays that miss the box: t_end <= t_start
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    ot, dt = T(o), T(d)
    t0, t1 = make_near_far_from_bound(1.0, ot, dt)
    bits, _ = blob_bitfield(rng, n_blobs=40)
    return rng, prop, ot, dt, t0, t1, T(bits)


def _train_samples(prop, ot, dt, t0, t1, noise, bt, N, S):
    dev = ot.device
    n = ot.shape[0]
    u32 = torch.uint32
    out = dict(next_loc=torch.empty(1, dtype=u32, device=dev), exceeded=torch.empty(1, dtype=u32, device=dev),
               valid=torch.empty(n, dtype=torch.bool, device=dev), ns=torch.empty(n, dtype=u32, device=dev),
               start=torch.empty(n, dtype=u32, device=dev), idcs=torch.full((S,), 7, dtype=u32, device=dev),
               xyzs=torch.full((S, 3), 7.0, device=dev), dirs=torch.full((S, 3), 7.0, device=dev),
               dss=torch.full((S,), 7.0, device=dev), zs=torch.full((S,), 7.0, device=dev))
    w, pd = prop.weights()
    rdesc = _lib.proposal_render_descriptor(prop.position_encoder.descriptor(0, 1.0),
                                            _lib.proposal_sample_descriptor(n, N, 0.0), pd)
    stage = torch.full((5 * max(n, 1) * 64,), 7.0, device=dev)  # scratch
    _lib.call("proposal_train_samples", [ot, dt, t0, t1, noise, prop.position_encoder.table.detach(), w, bt] +
              list(out.values()) + [stage], _lib.proposal_train_descriptor(rdesc, S))
    torch.cuda.synchronize()
    return {k: v.cpu() for k, v in out.items()}


@pytest.mark.parametrize("table_dtype,fill", [(torch.float32, 0.6), (torch.float16, 1.1)])
def test_proposal_train_samples_are_packed_render_samples(dev, table_dtype, fill):
    """With the jitter at 1/2 for every ray and no mass threshold, the training
    sampler places exactly proposal_render_samples' kept samples, packed in ray order
    in march_rays' layout: start = exclusive prefix of the counts, a ray is valid iff
    t_end > t_start and its samples fit, idcs = the ray, dirs = its direction, the
    tail zero, next_sample_write_location = every ray's count and
    number_of_exceeded_samples = those of the rays that did not fit.  The slots hold
    `fill` times the rays' samples (0.6: the later rays do not fit)."""
    n, N = 3000, 48
    rng, prop, ot, dt, t0, t1, bt = _train_sample_inputs(dev, 21, n, table_dtype)
    start = torch.empty(n, dtype=torch.uint32, device=dev)
    cnt, fb = torch.empty_like(start), torch.empty(n, dtype=torch.bool, device=dev)
    xyz, ds, z = torch.empty(n, N, 3, device=dev), torch.empty(n, N, device=dev), torch.empty(n, N, device=dev)
    w, pd = prop.weights()
    _lib.call("proposal_render_samples", [ot, dt, t0, t1, prop.position_encoder.table.detach(), w, bt, start, cnt, fb,
                                          xyz, ds, z],
              _lib.proposal_render_descriptor(prop.position_encoder.descriptor(0, 1.0),
                                              _lib.proposal_sample_descriptor(n, N, 0.0), pd))
    c = cnt.view(torch.int32).cpu().numpy().astype(np.int64)
    S = int(fill * c.sum())
    print(f"proposal_train_samples: rays={n} S={S} table={table_dtype}", flush=True)
    got = _train_samples(prop, ot, dt, t0, t1, torch.full((n,), 0.5, device=dev), bt, N, S)
    seg = (t1 > t0).cpu().numpy()
    assert (~seg).sum() >= 7 and c[seg].sum() > 4 * n
    pre = np.concatenate([[0], np.cumsum(c)[:-1]])
    valid = seg & (pre < S) & (pre + c <= S)
    assert valid.sum() > n // 4 and (valid[seg].all() if fill > 1 else not valid[seg].all())
    np.testing.assert_array_equal(got["valid"].numpy(), valid)
    np.testing.assert_array_equal(got["ns"].view(torch.int32).numpy(), np.where(valid, c, 0))
    np.testing.assert_array_equal(got["start"].view(torch.int32).numpy(), np.where(valid & (c > 0), pre, 0))
    assert int(got["next_loc"].view(torch.int32)) == c.sum()
    assert int(got["exceeded"].view(torch.int32)) == c[~valid].sum()
    m = int(c[valid].sum())
    rays = np.repeat(np.arange(n), np.where(valid, c, 0))
    slot = np.concatenate([np.arange(k) for k in np.where(valid, c, 0)])
    np.testing.assert_array_equal(got["idcs"].view(torch.int32).numpy()[:m], rays)
    np.testing.assert_array_equal(got["xyzs"].numpy()[:m], xyz.cpu().numpy()[rays, slot])
    np.testing.assert_array_equal(got["dirs"].numpy()[:m], dt.cpu().numpy()[rays])
    np.testing.assert_array_equal(got["dss"].numpy()[:m], ds.cpu().numpy()[rays, slot])
    np.testing.assert_array_equal(got["zs"].numpy()[:m], z.cpu().numpy()[rays, slot])
    for k in ("idcs", "xyzs", "dirs", "dss", "zs"):
        assert (got[k][m:] == 0).all(), k


def test_proposal_train_samples_stratified(dev):
    """Per-ray jitter: sample j sits at the (j + noise) / n quantile of the ray's
    predicted mass (the oracle's float64 placement with the bin densities of the
    unfused chain), before empty cells are dropped (all cells occupied here)."""
    n, N, S = 2000, 48, 1 << 17
    rng, prop, ot, dt, t0, t1, _ = _train_sample_inputs(dev, 22, n, torch.float32)
    full = torch.full((128 ** 3 // 8,), 255, dtype=torch.uint8, device=dev)
    noise = rng.uniform(0, 1, n).astype(np.float32)
    got = _train_samples(prop, ot, dt, t0, t1, torch.from_numpy(noise).to(dev), full, N, S)
    sig = prop.bin_density(ot, dt, t0, t1).cpu().numpy()
    T0, T1 = t0.cpu().numpy(), t1.cpu().numpy()
    want_c, _, want_t, want_ds = oracle.proposal_sample64(T0, T1, sig, N, 0.0, jitter=noise)
    ns = got["ns"].view(torch.int32).numpy()
    valid = got["valid"].numpy()
    np.testing.assert_array_equal(ns[valid], want_c[valid])
    assert valid.sum() > n // 2
    st = got["start"].view(torch.int32).numpy()
    zt = got["zs"].numpy()
    # sample j sits at the (j + noise) / n quantile: the float64 CDF of the ray's
    # predicted mass at the GPU's position equals that quantile (a quantile on a
    # plateau of the CDF, where bins hold almost no mass, may land anywhere on it,
    # so positions are compared through the CDF), and positions agree to 1e-4 of the
    # segment for 98% of the rays
    rays = np.nonzero(valid & (ns > 0))[0]
    assert (ns[rays] == N).all()  # every cell occupied: every sample kept
    # The kernel's bin weights come from 1 - exp(-sigma D) in f32, which is off by up
    # to ~1.2e-7 absolute (two ulps of exp near 1): a relative error of 1.2e-7 / a_b in
    # each bin's weight, so the CDF may move by up to sum_b p_b 1.2e-7 / a_b (large
    # for nearly transparent rays); f32 prefix sums add ~1e-5; and the position itself
    # is an f32 (t0 + x D, two roundings of up to half an ulp of t each), which the
    # CDF's slope p_b / D turns into up to p_b ulp(t) / D.  A misplaced stratum would
    # be 1/48 off.
    worst, close = 0.0, 0
    for r in rays:
        D = (T1[r] - T0[r]) / 64
        a = 1.0 - np.exp(-sig[r].astype(np.float64) * D)
        Tb = np.concatenate([[1.0], np.cumprod(1.0 - a)[:-1]])
        p = Tb * a / (Tb * a).sum()
        C = np.concatenate([[0.0], np.cumsum(p)])
        x = (zt[st[r]:st[r] + N].astype(np.float64) - T0[r]) / D
        b = np.clip(np.floor(x).astype(np.int64), 0, 63)
        F = C[b] + p[b] * (x - b)
        u = (np.arange(N) + np.float64(noise[r])) / N
        zr = zt[st[r]:st[r] + N]
        tol = 1e-5 + 2.0 * (p * 1.2e-7 / np.maximum(a, 1e-12)).sum() + p[b] * np.spacing(zr).astype(np.float64) / D
        worst = max(worst, (np.abs(F - u) / tol).max())
        close += np.abs(zt[st[r]:st[r] + N] - want_t[r]).max() <= 1e-4 * (T1[r] - T0[r]) + 1e-6
    assert worst <= 1.0, worst
    assert close > 0.98 * len(rays)
    # jitter moves samples: not the render placement
    rend = oracle.proposal_sample64(T0, T1, sig, N, 0.0)[2]
    assert np.abs(want_t[valid] - rend[valid]).max() > 1e-3


def test_proposal_train_samples_with_uniform_samples(dev):
    """n_uniform U > 0: each ray's samples are its n_per_ray quantile samples (as with
    U = 0) and U stratified uniform ones at t0 + (i + noise) (t1 - t0) / U, merged in
    ascending order; ds is the distance to the next sample, the last one's
    min(t1 - t, (t1 - t0) / U); every cell occupied, so every sample is kept."""
    n, N, U, S = 1500, 48, 16, 1 << 17
    rng, prop, ot, dt, t0, t1, _ = _train_sample_inputs(dev, 23, n, torch.float32)
    full = torch.full((128 ** 3 // 8,), 255, dtype=torch.uint8, device=dev)
    noise = rng.uniform(0, 1, n).astype(np.float32)
    nz = torch.from_numpy(noise).to(dev)
    base = _train_samples(prop, ot, dt, t0, t1, nz, full, N, S)
    w, pd = prop.weights()
    out = {k: torch.empty_like(v, device=dev) for k, v in base.items()}
    rdesc = _lib.proposal_render_descriptor(prop.position_encoder.descriptor(0, 1.0),
                                            _lib.proposal_sample_descriptor(n, N, 0.0), pd)
    stage = torch.empty(5 * n * 64, device=dev)
    _lib.call("proposal_train_samples", [ot, dt, t0, t1, nz, prop.position_encoder.table.detach(), w, full] +
              list(out.values()) + [stage], _lib.proposal_train_descriptor(rdesc, S, U))
    got = {k: v.cpu() for k, v in out.items()}
    T0, T1 = t0.cpu().numpy(), t1.cpu().numpy()
    seg = T1 > T0
    ns_b, ns = base["ns"].view(torch.int32).numpy(), got["ns"].view(torch.int32).numpy()
    st_b, st = base["start"].view(torch.int32).numpy(), got["start"].view(torch.int32).numpy()
    valid = got["valid"].numpy()
    assert valid.sum() > n // 2 and (ns[valid & seg] == N + U).all()
    zb, z, ds = base["zs"].numpy(), got["zs"].numpy(), got["dss"].numpy()
    f = np.float32
    for r in np.nonzero(valid & seg)[0][:400]:
        span = f((T1[r] - T0[r]) / f(U))
        uni = T0[r] + (np.arange(U, dtype=f) + noise[r]) * span
        want = np.sort(np.concatenate([zb[st_b[r]:st_b[r] + ns_b[r]], uni.astype(f)]))
        zr = z[st[r]:st[r] + ns[r]]
        np.testing.assert_array_equal(zr, want)
        dr = ds[st[r]:st[r] + ns[r]]
        np.testing.assert_array_equal(dr[:-1], np.diff(zr).astype(f))
        assert dr[-1] == min(T1[r] - zr[-1], span)
    for k in ("idcs", "xyzs", "dirs", "dss", "zs"):
        m = int(ns[valid].sum())
        assert (got[k][m:] == 0).all(), k
