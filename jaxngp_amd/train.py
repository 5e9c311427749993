"""Training step and occupancy grid on the gfx950 kernels.

Mirrors the reference's training loop (app/nerf/train.py:28-96) and step
(app/nerf/_utils.py:80-170) for the NeRF-synthetic style setup of the workload:
bound 1 (one cascade), 128^3 occupancy grid, diagonal_n_steps 1024, perturbed
marching, random background, Huber(0.1) loss over valid rays, Adam(1e-2, 0.9, 0.99,
eps 1e-15, eps_root 1e-15) with the staircase decay, the weight-decay term added
after the learning-rate scaling as optax.chain does there, occupancy-grid updates
every min(16, step//16 + 1) steps (all cells before step 256).

A step is the sequence of C-ABI targets below on preallocated device buffers, all
enqueued on the current stream with no host synchronisation (encodings and their
gradients use the level-major [16, S, 2] layout between the encoder and the MLP):
  train_make_rays -> march_rays -> hashgrid_encode -> nerf_mlp_forward ->
  integrate_rays -> train_loss_grad -> integrate_rays_backward ->
  nerf_mlp_backward -> hashgrid_encode_backward -> adam_update (table, MLP)
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .nerfs import N_PARAMS, NeRF, mlp_workspace_blocks
from .scene import SceneData

SQRT3 = 3 ** 0.5
# distinguishes the occupancy update's random draws from the ray generation's,
# which share the trainer's seed and step
OGRID_SEED_TAG = 0x6F63637570616E63


@dataclass
class TrainConfig:
    total_samples: int = 1 << 18
    n_rays: int | None = None  # rays per batch; None: the reference's n_rays = total_samples
    lr: float = 1e-2
    diagonal_n_steps: int = 1024
    grid_res: int = 128
    perturb: bool = True
    random_bg: bool = True
    bg: tuple = (1.0, 1.0, 1.0)
    near: float = 0.3
    seed: int = 1_000_000_007  # utils/args.py:18-27
    # "fp32": the reference's precision.  "fp16": f16 hash table and f16 MLP weights
    # on MFMA (f32 accumulation), f32 master weights, f32 transmittance, dynamic
    # loss scaling from 2^15 (halve on overflow, double after 2000 finite steps)
    precision: str = "fp32"
    loss_scale_init: float = 2.0 ** 15
    loss_scale_growth_interval: int = 2000
    # learned proposal trained online on the marcher's sample weights of the first
    # `proposal_rays` rays of every batch (the valid rays of a batch come first)
    proposal: bool = False
    proposal_rays: int = 1024
    proposal_every: int = 4  # proposal update on every n-th step
    # train the field on samples the proposal places (proposal_train_samples:
    # `proposal_samples_per_ray` per ray at stratified quantiles of its predicted
    # mass, minus empty occupancy cells) instead of the occupancy-grid marcher's;
    # the proposal then learns from the field's weights at its own samples.  Needs
    # `proposal`.  n_rays defaults to total_samples / 32 in this mode (about 44 kept
    # samples a ray fill the slots from the first ~6,000; the rays that do not fit are
    # dropped, and each offered ray costs one proposal evaluation)
    proposal_guided: bool = False
    proposal_samples_per_ray: int = 32
    # guided mode: samples per ray added at stratified uniform positions, so the
    # field also sees (and clears) what the proposal predicts empty
    proposal_uniform_samples: int = 32
    # guided mode: the occupancy-grid marcher places the samples of the first
    # `proposal_guided_warmup` steps and of every `proposal_guided_plain_every`-th
    # step after them (0: none)
    proposal_guided_warmup: int = 0
    proposal_guided_plain_every: int = 0
    # sky-sphere background model trained with the field, for scenes marked "bg"
    # (app/nerf/_utils.py:128-133; random_bg is then ignored, as there)
    background_model: bool = False
    # entries per hash-grid level (2^19 in make_nerf_ngp; the precision workload
    # also trains 2^22)
    log2_hashmap_size: int = 19
    # occupancy threshold cap: None takes the reference's 0.01 * diagonal_n_steps /
    # (2 min(bound, 1) sqrt 3) (types.py:1227-1239), the density at which one step
    # of the smallest size has opacity 1%; in guided mode the same formula with the
    # samples per ray of the guided sampler (proposal_samples_per_ray +
    # proposal_uniform_samples) in place of diagonal_n_steps.  The grid thresholds
    # at the smaller of the cap and the mean density
    density_threshold: float | None = None
    # without prefetch_march: run the optimizer update on a side stream, overlapping
    # the next step's ray generation and march (False: on the main stream, in order)
    overlap_update: bool = True
    # make the next step's rays and samples (train_make_rays + march_rays or the guided
    # sampler) on a stream of their own while this step computes, into a second set
    # of ray and sample arrays, whenever nothing this step does changes what they read
    # (no grid update after it; in guided mode, no proposal update in it).  The
    # samples are the same either way (False: made at the start of each step)
    prefetch_march: bool = True
    # FP32: the table's optimizer step taken inside the table gradient's last pass
    # (hashgrid_encode_backward_adam; same results, the gradient never goes through
    # memory) when the encoder allows it.  Off by default: measured, the step was no
    # faster (0.784 against 0.780 ms): the separate update streams its 341 MB at
    # about 6 TB/s, while in the reducer's tail (one workgroup per CU, held by its
    # 128 KiB of LDS) the same rows move at about 5.5 TB/s, which eats the 98 MB saved
    fused_table_update: bool = False


def lr_at(step: int, lr: float) -> float:
    """optax.exponential_decay(lr, 10000, 1/3, staircase, transition_begin=10000, end=lr/100)
    (app/nerf/_utils.py:20-27)."""
    if step <= 10000:
        return lr
    return max(lr * (1 / 3) ** ((step - 10000) // 10000), lr / 100)


class OccupancyGrid:
    """Occupancy density grid (utils/types.py:93-145): f32 density, bool mask and u8
    bitfield over K cascades of G^3 cells in Morton order, plus the alive cells."""

    def __init__(self, K: int, G: int, device):
        self.K, self.G = K, G
        n = K * G ** 3
        self.density = torch.zeros(n, dtype=torch.float32, device=device)
        self.occ_mask = torch.zeros(n, dtype=torch.bool, device=device)
        self.bitfield = torch.full((n // 8,), 255, dtype=torch.uint8, device=device)
        self.alive = torch.arange(n, dtype=torch.int64, device=device)
        self.alive_offsets = [i * G ** 3 for i in range(K + 1)]
        # per cascade: its alive cells' Morton indices within the cascade (u32, ascending)
        self.alive_cas = [torch.arange(G ** 3, dtype=torch.int32, device=device) for _ in range(K)]
        # occupied-cell prefix counts of the sampled update (u32[G^3 / 64 + 1])
        self.prefix = torch.zeros(G ** 3 // 64 + 1, dtype=torch.int32, device=device)

    def mean_density_up_to_cascade(self, cas: int) -> torch.Tensor:
        return self.density[self.alive[: self.alive_offsets[cas]]].mean()


_PF_STREAMS: dict = {}


def _prefetch_stream(dev) -> torch.cuda.Stream:
    """The prefetch stream of a device (default priority), created once per process:
    HIP maps streams onto a few hardware queues in creation order, and a stream created
    anew for every trainer sometimes shared the main stream's queue, which serialises
    the prefetch with the step (measured: 0.77 or 0.85 ms per step from run to run)."""
    key = torch.device(dev).index
    if key is None:
        key = torch.cuda.current_device()
    if key not in _PF_STREAMS:
        _PF_STREAMS[key] = torch.cuda.Stream(device=dev)
    return _PF_STREAMS[key]


def density_threshold_from_min_step_size(diagonal_n_steps: int, bound: float) -> float:
    """utils/types.py:1367-1369."""
    return 0.01 * diagonal_n_steps / (2 * min(bound, 1) * SQRT3)


class Trainer:
    def __init__(self, data: SceneData, cfg: TrainConfig = TrainConfig(), device=None, table_dtype=torch.float32):
        self.cfg = cfg
        self.data = data
        self.dev = device or data.rgba.device
        # The optimizer runs on a side stream, overlapping the next step's ray
        # generation and march (which read no parameters); see step().
        self._adam_pending = False
        self.bound = float(data.scene.bound)
        self.K = max(1, int(1 + math.ceil(math.log2(self.bound))))  # SceneMeta.cascades, types.py:858-860
        self.portion = 0.0 if self.bound <= 1 else 1 / 256  # stepsize_portion (types.py:874-886), bound in (1, 4)
        if cfg.precision not in ("fp32", "fp16"):
            raise ValueError(f"precision must be 'fp32' or 'fp16', got {cfg.precision!r}")
        self.fp16 = cfg.precision == "fp16"
        if self.fp16:
            table_dtype = torch.float16
        self._model = NeRF(bound=self.bound, device=self.dev, seed=cfg.seed % 2 ** 31, table_dtype=table_dtype,
                           log2_T=cfg.log2_hashmap_size)
        self.grid = OccupancyGrid(self.K, cfg.grid_res, self.dev)
        self.step_count = 0
        S = cfg.total_samples
        if cfg.proposal_guided and not cfg.proposal:
            raise ValueError("proposal_guided needs proposal=True")
        n = cfg.n_rays or (S // 32 if cfg.proposal_guided else S)
        self.n_rays = n
        f32, u32 = torch.float32, torch.uint32
        z = lambda *shape, dtype=f32: torch.zeros(*shape, dtype=dtype, device=self.dev)
        # the step's counters share one 16-byte tensor, so a caller can log them with one copy
        self.counters = z(4, dtype=u32)
        cnt = self.counters
        self.b = dict(
            o=z(n, 3), d=z(n, 3), t0=z(n), t1=z(n), noise=z(n), bgs=z(n, 3), gt=z(n, 3), pix=z(n, dtype=u32),
            next_loc=cnt[0:1], exceeded=cnt[1:2], valid=z(n, dtype=torch.bool), ns=z(n, dtype=u32),
            start=z(n, dtype=u32), idcs=z(S, dtype=u32), xyzs=z(S, 3), dirs=z(S, 3), dss=z(S), zs=z(S),
            enc=z(16, S, 2), drgbs=z(S, 4), measured=cnt[3:4], rgbd=z(n, 4), op=z(n), n_valid=cnt[2:3],
            loss=z(1), g_rgbd=z(n, 4), dbgs=z(n, 3), dz=z(S), ddrgbs=z(S, 4), denc=z(16, S, 2), dparams=z(N_PARAMS),
            ws=z(mlp_workspace_blocks(S), N_PARAMS), dtable=z(*self.model.position_encoder.table.shape),
        )
        if cfg.proposal_guided:
            # the guided sampler's staging workspace (f32[5, n, 64])
            self.b["pstage"] = torch.empty(5 * n * 64, dtype=f32, device=self.dev)
        enc_mod = self.model.position_encoder
        self.b["hg_ws"] = torch.empty(
            _lib.hashgrid_backward_workspace_bytes(enc_mod.descriptor(S, self.bound, _lib.ENC_LEVEL_MAJOR)),
            dtype=torch.uint8, device=self.dev)
        self.b["counters"] = cnt
        # two sets of the ray and sample arrays (the other buffers are shared): step s
        # uses set s % 2, so the next step's can be filled while this one computes
        self._sets = [self.b, self._second_set(self.b)]
        self._cur = 0
        table = self.model.position_encoder.table
        self.table_master = table.detach().float().clone()
        self.opt = dict(tm=torch.zeros_like(self.table_master), tv=torch.zeros_like(self.table_master),
                        mm=torch.zeros(N_PARAMS, device=self.dev), mv=torch.zeros(N_PARAMS, device=self.dev))
        if self.fp16:
            self.model.mlp_half = self.model.mlp_params.detach().half()
            self.loss_state = _lib.loss_scale_state(self.dev, cfg.loss_scale_init)
        self.prop = None
        self.prop_t = 0
        if cfg.proposal:
            from .proposal import BINS, ProposalNet
            self.prop = ProposalNet(self.bound, self.dev, seed=cfg.seed % 2 ** 31 + 1, table_dtype=table_dtype)
            R = min(cfg.proposal_rays, n)
            P = R * BINS
            pt = self.prop.position_encoder.table
            self.prop_master = pt.detach().float().clone()
            self.bp = dict(pts=z(P, 3), enc=z(5, P, 2), sig=z(P), dsig=z(P), denc=z(5, P, 2),
                           dparams=z(_lib.PROPOSAL_N_PARAMS), ws=z(max(1, min(512, (P + 255) // 256)),
                                                                   _lib.PROPOSAL_N_PARAMS),
                           dtable=z(*pt.shape), loss=z(1),
                           tm=torch.zeros_like(self.prop_master), tv=torch.zeros_like(self.prop_master),
                           mm=z(_lib.PROPOSAL_N_PARAMS), mv=z(_lib.PROPOSAL_N_PARAMS))
            self.bp["hg_ws"] = torch.empty(_lib.hashgrid_backward_workspace_bytes(
                self.prop.position_encoder.descriptor(P, self.bound, _lib.ENC_LEVEL_MAJOR)), dtype=torch.uint8,
                device=self.dev)
            if self.fp16:
                self.prop.mlp_half = self.prop.mlp_params.detach().half()
        self._bg_model = None
        if cfg.background_model:
            from .background import make_skysphere_background_model_ngp
            self._bg_model = make_skysphere_background_model_ngp(self.bound, device=self.dev,
                                                                 seed=cfg.seed % 2 ** 31 + 2)
            nb = self._bg_model.params.numel()
            sky = _lib.sky_sphere_bg_descriptor(n, 0, 1, self._bg_model.radius)
            self.bb = dict(dbg=z(n, 3), dparams=z(nb), mm=z(nb), mv=z(nb),
                           ws=torch.empty(_lib.sky_sphere_bg_workspace_bytes(sky), dtype=torch.uint8, device=self.dev))
        self.main_stream = torch.cuda.current_stream(self.dev)
        # With the prefetch the next step's sampler already runs beside this step, and an
        # update on a side stream would only add two cross-stream waits (about 30 us a
        # step, measured) to the critical path: the update then runs in order.
        side = cfg.overlap_update and not cfg.prefetch_march
        self.side_stream = torch.cuda.Stream(self.dev) if side else self.main_stream  # default priority
        self._bwd_done = torch.cuda.Event()
        self._adam_done = torch.cuda.Event()
        self.pf_stream = _prefetch_stream(self.dev) if cfg.prefetch_march else None
        self._pf_ready = [torch.cuda.Event(), torch.cuda.Event()]  # a set's prefetched samples are written
        self._pf_gate = torch.cuda.Event()  # the main stream's work enqueued before a prefetch
        self._pf_before = "mlp_bwd"  # the call of the step before which the prefetch is issued (measured best)
        self._pf_pending = [None, None]  # per set: (bitfield, step) of a prefetch not yet consumed
        self._calls = None
        self.mark_untrained_density_grid()

    # the arrays train_make_rays and the sampler write, and the step's counters
    RAY_KEYS = ("o", "d", "t0", "t1", "noise", "bgs", "gt", "pix", "valid", "ns", "start", "idcs", "xyzs", "dirs",
                "dss", "zs")

    def _second_set(self, b: dict) -> dict:
        b2 = dict(b)
        for key in self.RAY_KEYS:
            b2[key] = torch.zeros_like(b[key])
        if "pstage" in b:  # the guided sampler's scratch, per set
            b2["pstage"] = torch.empty_like(b["pstage"])
        cnt = torch.zeros_like(b["counters"])
        b2.update(counters=cnt, next_loc=cnt[0:1], exceeded=cnt[1:2], n_valid=cnt[2:3], measured=cnt[3:4])
        return b2

    # -- parameters and the optimizer's side stream ----------------------------------------

    def sync(self) -> None:
        """Makes work enqueued on the main stream from now on wait for the last step's
        optimizer update (which runs on the side stream)."""
        if self._adam_pending:
            self.main_stream.wait_event(self._adam_done)

    def close(self) -> None:
        """Orders the main stream after the pending optimizer update.  The parameters,
        gradients and moments it reads and writes were allocated on the main stream, so
        once the trainer is dropped the caching allocator may hand their blocks to new
        main-stream tensors; those must not be written while the side stream still
        works on them.  Called by __del__; safe to call more than once."""
        if getattr(self, "_adam_pending", False):
            self.main_stream.wait_event(self._adam_done)
            self._adam_pending = False
        for i, pending in enumerate(getattr(self, "_pf_pending", ())):
            if pending is not None:
                self.main_stream.wait_event(self._pf_ready[i])
                self._pf_pending[i] = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def model(self) -> NeRF:
        """The field; reading it orders the main stream after the last update."""
        self.sync()
        return self._model

    @property
    def bg_model(self):
        """The background model (or None); as `model`."""
        self.sync()
        return self._bg_model

    # -- occupancy grid -------------------------------------------------------------------

    def mark_untrained_density_grid(self):
        """Cells seen by no training camera are untrainable (types.py:1241-1362)."""
        G, K = self.cfg.grid_res, self.K
        G3 = G ** 3
        dev = self.dev
        cam = self.data.camera
        all_idx = torch.arange(K * G3, device=dev)
        alive = torch.zeros(K * G3, dtype=torch.bool, device=dev)
        from .volrend import morton3d_invert
        base = morton3d_invert(torch.arange(G3, device=dev, dtype=torch.int32).view(torch.uint32))
        base = base.view(torch.int32).float()
        offs = torch.tensor([[0, 0, 0], [0, 0, 1], [0, 1, 0], [0, 1, 1], [1, 0, 0], [1, 0, 1], [1, 1, 0], [1, 1, 1]],
                            dtype=torch.float32, device=dev)
        for cas in range(K):
            mip = min(2 ** cas, self.bound)
            xyz = (base / G - 0.5) * 2 * mip
            verts = xyz[:, None, :] + (2 * mip / G) * offs[None]  # [G3, 8, 3]
            seen = torch.zeros(G3, dtype=torch.bool, device=dev)
            for v in range(self.data.n_views):
                pose = torch.as_tensor(self.data.poses[v], dtype=torch.float32, device=dev)
                p_cam = (verts - pose[:, 3]) @ pose[:, :3]  # R^T (p - t)
                front = p_cam[..., 2] < 0
                u = p_cam[..., 0] / (-p_cam[..., 2]) * cam.fx + cam.cx
                vv = -p_cam[..., 1] / (-p_cam[..., 2]) * cam.fy + cam.cy
                inside = (u >= 0) & (u < cam.width) & (vv >= 0) & (vv < cam.height)
                seen |= (front & inside).any(-1)
                if bool(seen.all()):
                    break
            alive[cas * G3:(cas + 1) * G3] = seen
        g = self.grid
        g.density = torch.where(alive, g.density, torch.full_like(g.density, -1.0))
        thr = -0.5 if self.step_count == 0 else min(
            self.density_threshold_cap, float(g.mean_density_up_to_cascade(1)))
        from .volrend import packbits
        g.occ_mask, g.bitfield = packbits(thr, g.density)
        g.alive = all_idx[alive]
        counts = [int(alive[c * G3:(c + 1) * G3].sum()) for c in range(K)]
        g.alive_offsets = np.cumsum([0] + counts).tolist()
        g.alive_cas = [(g.alive[g.alive_offsets[c]:g.alive_offsets[c + 1]] - c * G3).to(torch.int32).contiguous()
                       for c in range(K)]

    def update_ogrid_density(self, cas: int, update_all: bool, max_inference: int | None = None):
        """types.py:1149-1225 for cascade `cas`, in one fused launch
        (occupancy_grid_update): decay the alive cells by 0.95, then max each chosen
        cell's density with the field's density at a random point of the cell (every
        alive cell when `update_all`, else half uniform over the alive cells and half
        uniform over the occupied ones).  `max_inference` (the reference's chunk size
        for its field evaluations) is accepted and not needed: the kernel evaluates
        every chosen cell in one launch."""
        g = self.grid
        alive = g.alive_cas[cas]
        if alive.numel() == 0:
            return
        m = self.model  # orders the main stream after the last optimizer update
        enc = m.position_encoder
        w16 = self.fp16
        params = m.mlp_half if w16 else m.mlp_params.detach()
        desc = _lib.occupancy_update_descriptor(
            enc.descriptor(0, self.bound), g.G, g.K, cas, alive.numel(), update_all,
            _lib.DTYPE_F16 if w16 else _lib.DTYPE_F32, self.cfg.seed ^ OGRID_SEED_TAG, self.step_count)
        _lib.call("occupancy_grid_update", [alive, g.bitfield, enc.table.detach(), params, g.density, g.prefix], desc)

    def threshold_ogrid(self):
        """types.py:1227-1239."""
        g = self.grid
        # clamp against a Python scalar: no host-to-device copy, so no host sync
        thr = torch.clamp(g.mean_density_up_to_cascade(1), max=self.density_threshold_cap)
        from .volrend import packbits
        g.occ_mask, g.bitfield = packbits(thr, g.density)

    @property
    def density_threshold_cap(self) -> float:
        return self.threshold_cap_of(self.cfg, self.bound)

    @staticmethod
    def threshold_cap_of(c: "TrainConfig", bound: float = 1.0) -> float:
        """The occupancy threshold cap configuration `c` trains with (see
        TrainConfig.density_threshold)."""
        if c.density_threshold is not None:
            return float(c.density_threshold)
        if c.proposal_guided:
            return density_threshold_from_min_step_size(c.proposal_samples_per_ray + c.proposal_uniform_samples, bound)
        return density_threshold_from_min_step_size(c.diagonal_n_steps, bound)

    @property
    def proposal_render_settings(self) -> dict:
        return self.proposal_render_settings_of(self.cfg)

    @staticmethod
    def proposal_render_settings_of(c: "TrainConfig") -> dict:
        """render_image_proposal keywords for the proposal configuration `c` trains.  A proposal
        trained on the guided samples (which include stratified uniform ones, so it
        also learns where it predicted nothing) renders with 32 samples spread over
        the support of its predicted mass and falls back below predicted mass 0.5;
        measured on the 8-blob scene at 5,000 steps, 20 held-out 800x800 views: every
        view within 0.01 dB of the same field rendered by the occupancy-grid marcher,
        at 0.32 of that render's frame time.  A proposal trained beside the marcher
        keeps the renderer's defaults (48 samples at the quantiles, fallback below
        0.98): with the guided settings it lost 1.8 dB on the same views."""
        if c.proposal_guided:
            return {"n_per_ray": 32, "mass_threshold": 0.5, "placement": _lib.PROPOSAL_SUPPORT}
        return {"n_per_ray": 48, "mass_threshold": 0.98, "placement": _lib.PROPOSAL_QUANTILES}

    @property
    def update_ogrid_interval(self) -> int:
        return min(16, self.step_count // 16 + 1)

    def maybe_update_grid(self):
        """train.py:75-85 with the cadence of types.py:1379-1392."""
        s = self.step_count
        if s > 0 and s % self.update_ogrid_interval == 0:
            for cas in range(self.K):
                self.update_ogrid_density(cas, update_all=s < 256)
            self.threshold_ogrid()

    # -- one step -----------------------------------------------------------------------

    def _prepare(self, b: dict) -> dict:
        """Bound calls of one training step on buffer set `b` (buffers and static
        descriptors fixed)."""
        c, m = self.cfg, self.model
        enc_mod = m.position_encoder
        table = enc_mod.table.detach()
        params = m.mlp_params.detach()
        S, n = c.total_samples, self.n_rays
        data = self.data
        LM = _lib.ENC_LEVEL_MAJOR
        hdesc = enc_mod.descriptor(S, self.bound, LM)
        P = _lib.Prepared
        o = self.opt
        mirror = table if table.dtype == torch.float16 else None
        tab_master = table if mirror is None else self.table_master
        if self.fp16:
            mlp_w, pd, flags = m.mlp_half, _lib.DTYPE_F16, _lib.LOSS_SCALED
            state = [self.loss_state]
        else:
            mlp_w, pd, flags, state = params, _lib.DTYPE_F32, 0, []
        none = torch.empty(0, device=self.dev)  # a NULL mirror slot
        fused = (c.fused_table_update and mirror is None and not self.fp16 and
                 _lib.hashgrid_backward_adam_supported(hdesc))
        side = self.side_stream.cuda_stream  # the optimizer's stream (see step)
        march_out = [b["next_loc"], b["exceeded"], b["valid"], b["ns"], b["start"], b["idcs"], b["xyzs"], b["dirs"],
                     b["dss"], b["zs"]]
        if c.proposal_guided:
            # samples placed by the proposal (its parameters are updated on the main
            # stream, so this launch is ordered after their update)
            ppe = self.prop.position_encoder
            pw, ppd = self.prop.weights()
            rdesc = _lib.proposal_render_descriptor(
                ppe.descriptor(0, self.bound),
                _lib.proposal_sample_descriptor(n, c.proposal_samples_per_ray, 0.0, _lib.PROPOSAL_QUANTILES, 1e-3,
                                                c.diagonal_n_steps, self.K, c.grid_res, self.bound, self.portion), ppd)
            guided = P("proposal_train_samples", [b["o"], b["d"], b["t0"], b["t1"], b["noise"], ppe.table.detach(),
                                                  pw, self.grid.bitfield] + march_out + [b["pstage"]],
                       _lib.proposal_train_descriptor(rdesc, S, c.proposal_uniform_samples))
        march = P("march_rays", [b["o"], b["d"], b["t0"], b["t1"], b["noise"], self.grid.bitfield] + march_out,
                  _lib.marching_descriptor(n, S, c.diagonal_n_steps, self.K, c.grid_res, self.bound, self.portion))
        calls = dict(
            rays=P("train_make_rays", [data.transforms, data.rgba, b["o"], b["d"], b["t0"], b["t1"], b["noise"],
                                       b["bgs"], b["gt"], b["pix"]], b""),
            march=march,
            guided=guided if c.proposal_guided else None,
            encode=P("hashgrid_encode", [b["xyzs"], table, b["enc"]], hdesc),
            mlp=P("nerf_mlp_forward", [b["enc"], b["dirs"], mlp_w, b["drgbs"]], _lib.nerf_mlp_descriptor(S, 0, LM, pd)),
            integrate=P("integrate_rays", [b["start"], b["ns"], b["bgs"], b["dss"], b["zs"], b["drgbs"], b["measured"],
                                           b["rgbd"], b["op"]], _lib.integrating_descriptor(n, S)),
            loss=P("train_loss_grad", [b["rgbd"], b["gt"], b["valid"], b["n_valid"], b["g_rgbd"], b["loss"]] + state,
                   _lib.loss_descriptor(n, 0.1, flags)),
            integrate_bwd=P("integrate_rays_backward", [b["start"], b["ns"], b["bgs"], b["dss"], b["zs"], b["drgbs"],
                                                        b["rgbd"], b["op"], b["g_rgbd"], b["dbgs"], b["dz"],
                                                        b["ddrgbs"]],
                            _lib.integrating_backward_descriptor(n, S, c.near)),
            mlp_bwd=P("nerf_mlp_backward", [b["enc"], b["dirs"], mlp_w, b["ddrgbs"], b["denc"], b["dparams"],
                                            b["ws"]], _lib.nerf_mlp_descriptor(S, b["ws"].shape[0], LM, pd)),
            encode_bwd=(P("hashgrid_encode_backward_adam", [b["xyzs"], b["denc"], table, o["tm"], o["tv"],
                                                            b["hg_ws"]], hdesc) if fused else
                        P("hashgrid_encode_backward", [b["xyzs"], b["denc"], b["dtable"], b["hg_ws"]], hdesc)),
            fused_table=fused,
            adam_table=P("adam_update", [tab_master, b["dtable"], o["tm"], o["tv"]] +
                         [mirror if mirror is not None else none] + state, b"", stream=side),
            adam_mlp=P("adam_update", [params, b["dparams"], o["mm"], o["mv"]] +
                       [m.mlp_half if self.fp16 else none] + state, b"", stream=side),
            mirror=mirror is not None,
            flags=flags,
        )
        if self.prop is not None:
            bp, pr = self.bp, self.prop
            R = min(c.proposal_rays, n)
            Pn = R * _lib.PROPOSAL_BINS
            ppe = pr.position_encoder
            ptab = ppe.table.detach()
            pdesc = ppe.descriptor(Pn, self.bound, LM)
            pw, ppd = pr.weights()
            pmirror = ptab if ptab.dtype == torch.float16 else None
            calls.update(
                p_points=P("proposal_points", [b["o"][:R], b["d"][:R], b["t0"][:R], b["t1"][:R], bp["pts"]],
                           _lib.proposal_rays_descriptor(R)),
                p_encode=P("hashgrid_encode", [bp["pts"], ptab, bp["enc"]], pdesc),
                p_mlp=P("proposal_mlp_forward", [bp["enc"], pw, bp["sig"]], _lib.proposal_mlp_descriptor(Pn, 0, LM, ppd)),
                p_loss=P("proposal_loss", [b["start"][:R], b["ns"][:R], b["dss"], b["zs"], b["drgbs"], b["t0"][:R],
                                           b["t1"][:R], bp["sig"], b["valid"][:R], b["n_valid"], bp["dsig"],
                                           bp["loss"]], _lib.proposal_loss_descriptor(R)),
                p_mlp_bwd=P("proposal_mlp_backward", [bp["enc"], pw, bp["dsig"], bp["denc"], bp["dparams"], bp["ws"]],
                            _lib.proposal_mlp_descriptor(Pn, bp["ws"].shape[0], LM, ppd)),
                p_encode_bwd=P("hashgrid_encode_backward", [bp["pts"], bp["denc"], bp["dtable"], bp["hg_ws"]], pdesc),
                p_adam_table=P("adam_update", [ptab if pmirror is None else self.prop_master, bp["dtable"], bp["tm"],
                                               bp["tv"], pmirror if pmirror is not None else none], b""),
                p_adam_mlp=P("adam_update", [pr.mlp_params.detach(), bp["dparams"], bp["mm"], bp["mv"],
                                             pr.mlp_half if pr.mlp_half is not None else none], b""),
            )
        if self._bg_model is not None:
            bb, bgp = self.bb, self._bg_model.params.detach()
            sky = _lib.sky_sphere_bg_descriptor(n, 0, 1, self._bg_model.radius)
            tb = _lib.train_bg_descriptor(n)
            calls.update(
                bg_fwd=P("sky_sphere_bg", [b["o"], b["d"], none, bgp, b["bgs"], bb["ws"]], sky),
                bg_blend=P("train_gt_blend", [data.rgba, b["pix"], b["bgs"], b["gt"]], tb),
                bg_grad=P("train_bg_grad", [data.rgba, b["pix"], b["g_rgbd"], b["dbgs"], bb["dbg"]], tb),
                bg_bwd=P("sky_sphere_bg_backward", [b["o"], b["d"], none, bgp, bb["dbg"], bb["dparams"], none, bb["ws"]],
                         sky),
                adam_bg=P("adam_update", [bgp, bb["dparams"], bb["mm"], bb["mv"], none] + state, b"", stream=side),
            )
            if self.fp16:
                calls["check_bg"] = P("loss_scale_check", [bb["dparams"], self.loss_state],
                                      _lib.loss_scale_check_descriptor(bgp.numel()), stream=side)
        if self.fp16:
            calls["check_table"] = P("loss_scale_check", [b["dtable"], self.loss_state],
                                     _lib.loss_scale_check_descriptor(b["dtable"].numel()), stream=side)
            calls["check_mlp"] = P("loss_scale_check", [b["dparams"], self.loss_state],
                                   _lib.loss_scale_check_descriptor(N_PARAMS), stream=side)
            calls["scale_update"] = P("loss_scale_update", [self.loss_state],
                                      _lib.loss_scale_update_descriptor(c.loss_scale_growth_interval), stream=side)
        return calls

    def step(self) -> None:
        """One training step (app/nerf/_utils.py:80-170), then the grid cadence.
        Kernel order: train_make_rays -> march_rays (or the guided sampler) -> (wait for
        the previous update) -> hashgrid_encode -> nerf_mlp_forward -> integrate_rays ->
        train_loss_grad -> integrate_rays_backward -> nerf_mlp_backward ->
        hashgrid_encode_backward, then adam_update (table, MLP) on the side stream.
        With `prefetch_march`, the next step's train_make_rays and sampler are issued on
        the prefetch stream into the other buffer set before this step's MLP backward, so
        the latency-bound marcher overlaps the MLP backward, the table gradient and the
        update (issued before the encoder, it slowed the encoder and MLP forward it ran
        beside by about 40 us); a step whose samples were prefetched starts with its
        encoder.  With a proposal, its update (every proposal_every-th step) runs just
        before that point, on this step's samples and the forward's outputs, so a guided
        next step reads the updated proposal.  A prefetch is made only when nothing
        before the next step changes what it reads (see _prefetchable), and is dropped
        if the grid's bitfield was replaced meanwhile.  With a background model,
        sky_sphere_bg + train_gt_blend follow the wait and train_bg_grad +
        sky_sphere_bg_backward follow integrate_rays_backward.  Readers of the
        parameters outside the step go through `model`, `bg_model` or `sync()`."""
        c = self.cfg
        if self._calls is None:
            # (re)bind: a prefetch made before (e.g. before a checkpoint restore) is dropped
            self.close()
            self._calls = [self._prepare(b) for b in self._sets]
        cur = self._cur
        self.b = self._sets[cur]
        self.counters = self.b["counters"]
        k = self._calls[cur]
        s0 = self.step_count
        bits = self.grid.bitfield
        pending = self._pf_pending[cur]
        if pending is not None:
            self.main_stream.wait_event(self._pf_ready[cur])
            self._pf_pending[cur] = None
        if pending is None or pending[0] is not bits or pending[1] != s0:
            self._rays_and_samples(k, s0, None)
        nxt = 1 - cur
        prefetch = self.pf_stream is not None and self._prefetchable(s0)
        # rays and samples read no field parameters: they overlap the last step's update
        # on the side stream; everything after waits for it
        self.sync()
        bg = self._bg_model is not None
        order = ("encode", "mlp", "integrate", "loss", "integrate_bwd", "mlp_bwd", "encode_bwd")
        if bg:
            # per-ray backgrounds from the model, the ground truth blended over them, and
            # their gradient after the integrator's (app/nerf/_utils.py:128-150)
            order = ("bg_fwd", "bg_blend") + order[:5] + ("bg_grad", "bg_bwd") + order[5:]
        train_prop = self.prop is not None and s0 % c.proposal_every == 0
        t = s0 + 1  # the optimizer's step count after this step
        lr = lr_at(t - 1, c.lr)
        b1, b2 = 0.9, 0.99
        bp1, bp2 = float(np.float32(b1) ** t), float(np.float32(b2) ** t)
        n_table = k["adam_table"].keep[0].numel()
        for name in order:
            if name == self._pf_before and train_prop:
                # the proposal's update, on this step's samples and the field's weights at
                # them (the forward's outputs), before the next step's sampler reads it
                self._train_proposal(k, s0 + 1)
            if prefetch and name == self._pf_before:
                # the next step's rays and samples, on the prefetch stream into the other
                # set, after everything enqueued on the main stream so far (the last step
                # that used that set, the grid update that may have followed it, and this
                # step's kernels up to here): the latency-bound marcher then runs beside
                # the rest of this step
                pf = self.pf_stream
                self._pf_gate.record(self.main_stream)
                pf.wait_event(self._pf_gate)
                with torch.cuda.stream(pf):
                    self._rays_and_samples(self._calls[nxt], s0 + 1, pf.cuda_stream)
                self._pf_ready[nxt].record(pf)
                self._pf_pending[nxt] = (bits, s0 + 1)
            if name == "encode_bwd" and k["fused_table"]:
                # the table's optimizer step inside its gradient's last pass
                k[name](_lib.hashgrid_adam_descriptor(
                    k[name].opaque, _lib.adam_descriptor(n_table, lr, b1, b2, 1e-15, 1e-15, bp1, bp2, 0.0)))
            else:
                k[name]()
        self.step_count += 1
        # the update on the side stream, once the gradients exist; the next step's
        # rays and march run meanwhile
        self._bwd_done.record(self.main_stream)
        self.side_stream.wait_event(self._bwd_done)
        with torch.cuda.stream(self.side_stream):
            if self.fp16:
                k["check_table"]()
                k["check_mlp"]()
                if bg:
                    k["check_bg"]()
            if not k["fused_table"]:
                k["adam_table"](_lib.adam_descriptor(n_table, lr, b1, b2, 1e-15, 1e-15, bp1, bp2, 0.0,
                                                     mirror_dtype=_lib.DTYPE_F16 if k["mirror"] else _lib.DTYPE_F32,
                                                     flags=k["flags"]))
            k["adam_mlp"](_lib.adam_descriptor(N_PARAMS, lr, b1, b2, 1e-15, 1e-15, bp1, bp2, 1e-6,
                                               mirror_dtype=_lib.DTYPE_F16 if self.fp16 else _lib.DTYPE_F32,
                                               flags=k["flags"]))
            if bg:
                # params["bg"] is optimised like the NeRF MLP: the "network" Adam and
                # weight decay 1e-6 (app/nerf/_utils.py:28-76)
                k["adam_bg"](_lib.adam_descriptor(k["adam_bg"].keep[0].numel(), lr, b1, b2, 1e-15, 1e-15, bp1, bp2,
                                                  1e-6, flags=k["flags"]))
            if self.fp16:
                k["scale_update"]()
        self._adam_done.record(self.side_stream)
        self._adam_pending = True
        self._cur = nxt
        _lib.check()
        self.maybe_update_grid()

    def _train_proposal(self, k: dict, t: int) -> None:
        """One update of the proposal (its t-th step of training overall is the
        trainer's step t): bin densities of the first proposal_rays rays, the
        histogram-bound loss against the field's weights at this step's samples, the
        backward and Adam."""
        c = self.cfg
        for name in ("p_points", "p_encode", "p_mlp", "p_loss", "p_mlp_bwd", "p_encode_bwd"):
            k[name]()
        b1, b2 = 0.9, 0.99
        lr = lr_at(t - 1, c.lr)
        pm16 = self.prop.position_encoder.table.dtype == torch.float16
        self.prop_t += 1  # the proposal's own update count for Adam's bias corrections
        q1, q2 = float(np.float32(b1) ** self.prop_t), float(np.float32(b2) ** self.prop_t)
        k["p_adam_table"](_lib.adam_descriptor(k["p_adam_table"].keep[0].numel(), lr, b1, b2, 1e-15, 1e-15, q1, q2,
                                               0.0, mirror_dtype=_lib.DTYPE_F16 if pm16 else _lib.DTYPE_F32))
        k["p_adam_mlp"](_lib.adam_descriptor(_lib.PROPOSAL_N_PARAMS, lr, b1, b2, 1e-15, 1e-15, q1, q2, 1e-6,
                                             mirror_dtype=_lib.DTYPE_F16 if self.prop.mlp_half is not None
                                             else _lib.DTYPE_F32))

    def _guided_at(self, s: int) -> bool:
        """Whether step s takes the guided sampler's samples (else the marcher's)."""
        c = self.cfg
        return c.proposal_guided and s >= c.proposal_guided_warmup and not (
            c.proposal_guided_plain_every > 0 and (s - c.proposal_guided_warmup) % c.proposal_guided_plain_every == 0)

    def _grid_update_after(self, s: int) -> bool:
        """Whether maybe_update_grid runs an update at the end of step s."""
        t = s + 1
        return t % min(16, t // 16 + 1) == 0

    def _prefetchable(self, s: int) -> bool:
        """Step s + 1's samples can be made during step s: the grid does not change
        after step s.  (The proposal a guided step reads is updated in step s before the
        prefetch is issued.)"""
        return not self._grid_update_after(s)

    def _rays_and_samples(self, k: dict, s: int, stream) -> None:
        """train_make_rays and the sampler of step s on set k's buffers (on `stream`,
        or the calls' own)."""
        c, cam, data = self.cfg, self.data.camera, self.data
        guided = self._guided_at(s)
        march = k["guided"] if guided else k["march"]
        march.set_buffer(7 if guided else 5, self.grid.bitfield)
        bg = self._bg_model is not None
        flags = (_lib.TRAIN_PERTURB if c.perturb else 0) | (_lib.TRAIN_RANDOM_BG if c.random_bg and not bg else 0)
        k["rays"](_lib.train_rays_descriptor(self.n_rays, data.n_views, cam.width, cam.height, cam.fx, cam.fy, cam.cx,
                                             cam.cy, self.bound, c.seed, s, flags, c.bg), stream=stream)
        march(stream=stream)

    # -- metrics (these synchronise) -------------------------------------------------------

    def metrics(self) -> dict:
        self.sync()  # the loss-scale state is written by the update
        b = self.b
        u = lambda t: int(t.view(torch.int32)[0].item())
        out = dict(loss=float(b["loss"][0].item()), n_valid_rays=u(b["n_valid"]),
                   measured_batch_size_before_compaction=u(b["next_loc"]) - u(b["exceeded"]),
                   measured_batch_size=u(b["measured"]))
        if self.fp16:
            out["loss_scale"] = float(self.loss_state[:4].view(torch.float32)[0].item())
        if self.prop is not None:
            out["proposal_loss"] = float(self.bp["loss"][0].item())
        return out
