"""Throughput of the NeRF training step on one MI355X.

Workload (plain FP32, occupancy-grid marching): the synthetic 8-blob scene, 100
training views of 800x800 rendered in float64, batches of 2^18 sample slots with
the reference's n_rays = total_samples convention, 128^3 occupancy grid,
diagonal_n_steps 1024, perturbed marching, random background, hash grid
L=16/F=2/T=2^19, NeRF MLPs 32->64->16 and (16+16)->64->64->3, Adam lr 1e-2.

A step is one full training step (ray generation, marching, encoding, MLP,
integration, loss, backward through all of them, Adam on the hash table and the
MLP, and the occupancy-grid update on its cadence).  `value` is marched samples
per second (the samples that went through the MLP:
measured_batch_size_before_compaction, summed over the timed steps on the device)
divided by the timed wall time.  The grid-update cadence makes steps after 256
the steady state, so warm-up defaults past it.

Prints one JSON line.  Only one GPU is supported.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# algorithmic bytes / flops per unit (hash encoder per point, marcher per sample
# and per ray, integrator per sample and per ray, MLP flops per sample)
ENC_BYTES_PER_POINT = 12 + 16 * 8 * 2 * 4 + 16 * 2 * 4  # 1164
MARCH_BYTES_PER_SAMPLE, MARCH_BYTES_PER_RAY = 36, 45
INT_FWD_BYTES_PER_SAMPLE, INT_FWD_BYTES_PER_RAY = 24, 40
INT_BWD_BYTES_PER_SAMPLE, INT_BWD_BYTES_PER_RAY = 44, 68
MLP_FWD_FLOP = 2 * (32 * 64 + 64 * 16 + 32 * 64 + 64 * 64 + 64 * 3)  # 18816
# the table gradient with the optimizer step fused: per table parameter, params, m
# and v read and written once (f32)
ADAM_BYTES_PER_PARAM = 6 * 4
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3


def cpu_baseline(seconds: float = 12.0) -> dict:
    """The oracle's CPU training step on a bounded sample (2^14 sample slots)."""
    import oracle
    from tests.helpers import march_inputs

    rng = np.random.default_rng(0)
    S = 1 << 14
    n = S
    o, d, t0, t1, noise, bits, _ = march_inputs(rng, n)
    bgs = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    gt = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    lv = oracle.hashgrid_levels(16, 1 << 19, 16, 2048)
    table = rng.uniform(-1e-4, 1e-4, (lv["n_rows"], 2)).astype(np.float32)
    params = rng.uniform(-0.2, 0.2, 9408)
    samples, steps = 0, 0
    t_start = time.perf_counter()
    while True:
        samples += oracle.cpu_train_step(o, d, t0, t1, noise, bits, bgs, gt, table, params, lv, 1 << 19, S)
        steps += 1
        el = time.perf_counter() - t_start
        if el > seconds:
            break
    cores = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", cores))
    return {"value": samples / el, "unit": "samples/s", "cores": min(cores, threads), "kind": "port",
            "sample": f"{steps} CPU training steps of 2^14 sample slots (n_rays = 2^14) on a random-blob "
                      f"occupancy grid: oracle C marcher/encoder/integrator (OpenMP) + numpy f32 MLP"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_launches(fn, reps: int, stream) -> float:
    """Median ms of `fn()` over `reps` launches, each bracketed by HIP events on
    `stream` (the stream the targets are launched on)."""
    import torch
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for _ in range(3):
        fn()
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    t = sorted(a.elapsed_time(b) for a, b in ev)
    return t[len(t) // 2]


def bench_encoder(dev, reps: int = 50) -> dict:
    """The hash encoder alone on its own workload: L=16, F=2, T=2^19, N_min=16,
    N_max=2048, f32 table U(-1e-4, 1e-4) (seed 1), 2^18 points xyz ~ U[0,1)^3
    (seed 0) passed as pos = 2 xyz - 1 with bound 1 to the drop-in point-major
    encoder, cotangent N(0,1) [2^18, 32] (seed 2): forward and backward (table
    gradient) per launch, against 1,164 algorithmic bytes per point each."""
    import torch
    from jaxngp_amd import _lib
    from jaxngp_amd.encoders import HashGridEncoder
    N = 1 << 18
    enc = HashGridEncoder(L=16, T=1 << 19, F=2, N_min=16, N_max=2048,
                          generator=torch.Generator(device="cpu").manual_seed(1)).to(dev)
    xyz = torch.from_numpy(np.random.default_rng(0).uniform(0, 1, (N, 3)).astype(np.float32)).to(dev)
    pos = (2 * xyz - 1).contiguous()
    g = torch.from_numpy(np.random.default_rng(2).normal(size=(N, 32)).astype(np.float32)).to(dev)
    table = enc.table.detach()
    out = torch.empty(N, 32, device=dev)
    dtable = torch.empty_like(table)
    desc = enc.descriptor(N, 1.0)
    ws = torch.empty(_lib.hashgrid_backward_workspace_bytes(desc), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    fwd = _lib.Prepared("hashgrid_encode", [pos, table, out], desc)
    bwd = _lib.Prepared("hashgrid_encode_backward", [pos, g, dtable, ws], desc)
    f_ms = time_launches(fwd, reps, stream)
    b_ms = time_launches(bwd, reps, stream)
    _lib.check()
    algo = N * ENC_BYTES_PER_POINT
    res = {"points": N, "table": "f32 [6098120, 2]", "layout": "point-major [N, 32]",
           "forward_ms": f_ms, "backward_ms": b_ms,
           "forward_GBps": algo / (f_ms * 1e-3) / 1e9, "backward_GBps": algo / (b_ms * 1e-3) / 1e9,
           "algorithmic_bytes_per_launch": algo}
    res["forward_frac"] = res["forward_GBps"] / HBM_PEAK_GBS
    res["backward_frac"] = res["backward_GBps"] / HBM_PEAK_GBS
    log(f"encoder alone: forward {f_ms * 1e3:.1f} us ({res['forward_GBps']:.0f} GB/s), backward "
        f"{b_ms * 1e3:.1f} us ({res['backward_GBps']:.0f} GB/s)")
    return res


def bench_integrate(dev, reps: int = 50) -> dict:
    """integrate_rays forward and backward alone: 4096 rays of S / 4096 samples each
    (S = 2^18 and 2^20), densities softplus(N(0,2)), rgb and bg U[0,1], ds =
    2 sqrt(3) / 1024, z = 3 + k ds, near 0.3 (seed 4), against 24 / 44 algorithmic
    bytes per sample and 40 / 68 per ray."""
    import torch
    from jaxngp_amd import _lib
    out = {}
    rng = np.random.default_rng(4)
    stream = torch.cuda.current_stream()
    n = 4096
    for S in (1 << 18, 1 << 20):
        per = S // n
        ns = np.full(n, per, np.uint32)
        start = (np.arange(n, dtype=np.uint64) * per).astype(np.uint32)
        ds = np.float32(2 * 3 ** 0.5 / 1024)
        dss = np.full(S, ds, np.float32)
        z = (3 + np.tile(np.arange(per, dtype=np.float32), n) * ds).astype(np.float32)
        sig = np.log1p(np.exp(rng.normal(0, 2, S))).astype(np.float32)
        drgbs = np.concatenate([sig[:, None], rng.uniform(0, 1, (S, 3)).astype(np.float32)], 1)
        bgs = rng.uniform(0, 1, (n, 3)).astype(np.float32)
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        st, nst, bg, dsd, zd, dr = T(start), T(ns), T(bgs), T(dss), T(z), T(drgbs)
        meas = torch.empty(1, dtype=torch.uint32, device=dev)
        rgbd, op = torch.empty(n, 4, device=dev), torch.empty(n, device=dev)
        g = T(rng.normal(size=(n, 4)).astype(np.float32))
        dbg, dz, dd = torch.empty(n, 3, device=dev), torch.empty(S, device=dev), torch.empty(S, 4, device=dev)
        fwd = _lib.Prepared("integrate_rays", [st, nst, bg, dsd, zd, dr, meas, rgbd, op],
                            _lib.integrating_descriptor(n, S))
        bwd = _lib.Prepared("integrate_rays_backward", [st, nst, bg, dsd, zd, dr, rgbd, op, g, dbg, dz, dd],
                            _lib.integrating_backward_descriptor(n, S, 0.3))
        f_ms = time_launches(fwd, reps, stream)
        b_ms = time_launches(bwd, reps, stream)
        _lib.check()
        fa = S * INT_FWD_BYTES_PER_SAMPLE + n * INT_FWD_BYTES_PER_RAY
        ba = S * INT_BWD_BYTES_PER_SAMPLE + n * INT_BWD_BYTES_PER_RAY
        key = f"samples_2^{S.bit_length() - 1}"
        out[key] = {"rays": n, "samples_per_ray": per, "forward_ms": f_ms, "backward_ms": b_ms,
                    "fwd_bwd_ms": f_ms + b_ms, "forward_frac": fa / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "backward_frac": ba / (b_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        log(f"integrate alone, S={S}: forward {f_ms * 1e3:.1f} us, backward {b_ms * 1e3:.1f} us")
    return out


def steady_state(tr, steps: int = 200) -> dict:
    """Plain timing (no events) of `steps` training steps of an already trained
    field: the grid update runs every 16 steps there (after step 256)."""
    import torch
    counts = torch.zeros(steps, 4, dtype=torch.int32, device=tr.dev)
    first = tr.step_count
    torch.cuda.synchronize()
    gc.collect()
    gc.disable()
    t0 = time.perf_counter()
    for i in range(steps):
        tr.step()
        counts[i].copy_(tr.counters.view(torch.int32))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    gc.enable()
    c = counts.long()
    n_samples = int((c[:, 0] - c[:, 1]).sum().item())
    n_valid = int(c[:, 2].sum().item())
    res = {"after_steps": first, "steps": steps, "ms_per_step": el / steps * 1e3, "value": n_samples / el,
           "unit": "samples/s", "rays_per_s": n_valid / el, "samples_per_step": n_samples / steps}
    log(f"[{tr.cfg.precision}] steady state after {first} steps: {res['ms_per_step']:.3f} ms/step, "
        f"{res['value'] / 1e6:.1f} M samples/s")
    return res


CALIBRATION_STEPS = 30
ROOFLINE_CANDIDATES = ("hashgrid_encode", "hashgrid_encode_backward", "hashgrid_encode_backward_adam", "march_rays",
                       "integrate_rays", "integrate_rays_backward", "nerf_mlp_forward", "nerf_mlp_backward")


def run_training(data, cfg, steps: int, warmup: int, dev) -> dict:
    """Trains W warm-up steps, runs an untimed calibration pass, then times K steps.

    Launch durations come from HIP events recorded around the bound calls on the
    stream each is launched on.  Every recorded event also costs step time on this
    GPU, so the calibration pass (CALIBRATION_STEPS steps right before the timed
    region) times every target and picks the one with the longest launch, and the
    timed region records events around that target only: its launch time (the
    roofline's) is measured live in the timed region, the others come from the
    calibration pass."""
    import torch
    from jaxngp_amd.train import Trainer

    tr = Trainer(data, cfg, device=dev)
    for i in range(warmup):
        tr.step()
        if i % 100 == 0:
            torch.cuda.synchronize()
            log(f"[{cfg.precision}] warmup step {i}: {tr.metrics()}")
    torch.cuda.synchronize()
    targets = ["march_rays", "hashgrid_encode", "nerf_mlp_forward", "integrate_rays", "integrate_rays_backward",
               "nerf_mlp_backward", "hashgrid_encode_backward", "hashgrid_encode_backward_adam", "adam_update"]

    class Timed:
        def __init__(self, call, events):
            self.call, self.events = call, events

        def __getattr__(self, a):
            return getattr(self.call, a)

        def __call__(self, opaque=None, stream=None):
            # the events go on the stream the call is launched on: the trainer passes
            # a stream only inside that stream's context
            cur = torch.cuda.current_stream()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(cur)
            self.call(opaque, stream)
            b.record(cur)
            self.events.setdefault(self.call.name, []).append((a, b))

    plain = tr._calls  # one dict of bound calls per buffer set

    def instrument(names, events):
        tr._calls = [{k: (Timed(v, events) if hasattr(v, "name") and v.name in names else v) for k, v in d.items()}
                     for d in plain]

    def per_launch(events):
        return {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in events.items() if v}

    calib: dict = {}
    instrument(targets, calib)
    for _ in range(CALIBRATION_STEPS):
        tr.step()
    torch.cuda.synchronize()
    per = per_launch(calib)
    # the roofline kernel is the longest on the step's critical path: a prefetched
    # sampler runs beside the step's compute (its events span that overlap)
    off_path = {"march_rays"} if tr.pf_stream is not None else set()
    dom = max((k for k in per if k in ROOFLINE_CANDIDATES and k not in off_path), key=lambda k: per[k])
    live: dict = {}
    instrument([dom], live)
    # per-step counters (next_loc, exceeded, n_valid, measured) logged on the device
    # with one 16-byte copy per step, summed after the timed region
    counts = torch.zeros(steps, 4, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    gc.collect()
    gc.disable()  # no collector pause inside the timed region (the host must keep ahead of the GPU)
    t_start = time.perf_counter()
    for i in range(steps):
        tr.step()
        counts[i].copy_(tr.counters.view(torch.int32))  # the set the step used
    t_host = time.perf_counter() - t_start
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    gc.enable()
    tr._calls = plain
    c = counts.long()
    n_samples = int((c[:, 0] - c[:, 1]).sum().item())
    n_valid = int(c[:, 2].sum().item())
    ms = elapsed / steps * 1e3
    dom_ms = per_launch(live)[dom]
    log(f"[{cfg.precision}] timed {steps} steps: {ms:.3f} ms/step ({t_host / steps * 1e3:.3f} ms/step to "
        f"enqueue), {n_samples / steps:.0f} samples/step, {n_valid / steps:.0f} valid rays/step, "
        f"{dom} {dom_ms:.4f} ms/launch (live), metrics {tr.metrics()}")
    return dict(trainer=tr, per=per, dom=dom, dom_ms=dom_ms, elapsed=elapsed, samples=n_samples, valid=n_valid,
                ms=ms)


POOL = 1 << 18  # ray slots of the inference renderer (the reference uses 8192; same image)


def train_to(tr, total_steps: int) -> None:
    """Untimed training until the trainer has taken `total_steps` steps in all, so
    that the frames are rendered from a field trained as long as the PSNR figures
    are quoted for."""
    import torch

    while tr.step_count < total_steps:
        tr.step()
    torch.cuda.synchronize()


# Proposal render settings the bench reports: (samples per ray, fallback mass threshold,
# placement).  The first is the renderer's default; Trainer.proposal_render_settings names
# the one each field is rendered at (the last two for a proposal-guided field).
PROPOSAL_RENDER_SETTINGS = ((48, 0.98, "quantiles"), (48, 0.9, "support"), (32, 0.4, "support"),
                            (32, 0.5, "support"))


def bench_inference(tr, data, n_frames: int, dev, proposal: bool = False, empty_mass: float = 0.0) -> dict:
    """ms per 800x800 frame of the trained field from held-out seeded poses, with the
    reference's ray-pool renderer (march_steps_cap 8, white background) on a pool of
    2^18 slots, median over the timed frames after three warm-up renders; PSNR of the
    u8 render against the float64 ground truth blended on white.  With `proposal`,
    the frames are rendered with `render_image_proposal` at each of
    PROPOSAL_RENDER_SETTINGS (occupancy-grid fallback below the predicted mass
    threshold, on its own pool of 2^16 slots) and the same field is also rendered the
    plain way, for the PSNR difference; the top-level figures are those of the
    trainer's own setting (Trainer.proposal_render_settings), `settings` holds every
    setting's."""
    import torch
    from jaxngp_amd import _lib
    from jaxngp_amd import scene as sc
    from jaxngp_amd.renderer import render_image_inference, render_image_proposal

    poses = sc.look_at_poses(n_frames, seed=2, golden=False)
    gt = sc.make_views(data.scene, data.camera, poses, dev)
    c = tr.cfg
    kw = dict(K=tr.K, G=c.grid_res, diagonal_n_steps=c.diagonal_n_steps, stepsize_portion=tr.portion,
              march_steps_cap=8, pool=POOL)
    pose_t = [torch.as_tensor(p, dtype=torch.float32, device=dev) for p in poses]
    refs = [sc.blend_u8(gt.rgba[i], (1.0, 1.0, 1.0)) for i in range(n_frames)]
    plain = lambda pose, st=None: render_image_inference(tr.model, data.camera, pose, tr.grid.bitfield, **kw)
    placements = {"quantiles": _lib.PROPOSAL_QUANTILES, "support": _lib.PROPOSAL_SUPPORT}

    def frames(render):
        for i in range(3):
            render(pose_t[i % n_frames], {})
        times, psnrs, fallback = [], [], []
        for i in range(n_frames):
            st = {}
            torch.cuda.synchronize()
            t = time.perf_counter()
            _, rgb, _ = render(pose_t[i], st)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t) * 1e3)
            psnrs.append(sc.psnr_u8(rgb, refs[i]))
            if st:
                fallback.append(st["fallback_rays"] / max(1, st["fallback_rays"] + st["proposal_rays"]))
        return sorted(times)[len(times) // 2], psnrs, fallback

    base = {"frames": n_frames, "resolution": "800x800", "pool": POOL, "march_steps_cap": 8, "background": "white",
            "train_steps": tr.step_count}
    ms, plain_psnrs, _ = frames(plain)
    if not proposal:
        res = dict(ms_per_frame=ms, psnr_db=float(np.mean(plain_psnrs)), **base)
        log(f"inference: {ms:.2f} ms/frame (median of {n_frames}), PSNR {res['psnr_db']:.2f} dB after "
            f"{tr.step_count} steps")
        return res
    settings = []
    for n_per_ray, thr, placement in PROPOSAL_RENDER_SETTINGS:
        render = lambda pose, st: render_image_proposal(tr.model, tr.prop, data.camera, pose, tr.grid.bitfield,
                                                        n_per_ray=n_per_ray, mass_threshold=thr,
                                                        placement=placements[placement], stats=st,
                                                        empty_mass=empty_mass, **kw)
        pms, psnrs, fallback = frames(render)
        settings.append({"samples_per_ray": n_per_ray, "mass_threshold": thr, "placement": placement,
                         "empty_mass": empty_mass, "ms_per_frame": pms, "psnr_db": float(np.mean(psnrs)),
                         "delta_psnr_db": float(np.mean(psnrs) - np.mean(plain_psnrs)),
                         "worst_view_delta_psnr_db": float(min(a - b for a, b in zip(psnrs, plain_psnrs))),
                         "fallback_ray_fraction": float(np.mean(fallback))})
        log(f"inference (proposal, {n_per_ray} samples, threshold {thr}, {placement}): {pms:.2f} ms/frame, "
            f"PSNR {settings[-1]['psnr_db']:.2f} dB, delta {settings[-1]['delta_psnr_db']:+.3f} dB "
            f"(worst view {settings[-1]['worst_view_delta_psnr_db']:+.3f}), fallback rays "
            f"{settings[-1]['fallback_ray_fraction']:.1%}")
    own = tr.proposal_render_settings
    own = (own["n_per_ray"], own["mass_threshold"], {v: k for k, v in placements.items()}[own["placement"]])
    res = dict(settings[PROPOSAL_RENDER_SETTINGS.index(own)], **base)
    res.update(psnr_plain_render_db=float(np.mean(plain_psnrs)), ms_per_frame_plain_render=ms, settings=settings)
    return res


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--views", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--frames", type=int, default=20, help="timed 800x800 inference frames (held-out poses)")
    ap.add_argument("--render-steps", type=int, default=5000,
                    help="train (untimed, after the timed steps) to this many steps in all before rendering frames")
    ap.add_argument("--no-fp16", action="store_true", help="skip the f16-weight precision path")
    ap.add_argument("--no-proposal", action="store_true", help="skip the learned-proposal path")
    ap.add_argument("--no-t22", action="store_true", help="skip the precision path at 2^22 entries per level")
    ap.add_argument("--no-marcher-proposal", action="store_true",
                    help="skip the proposal trained beside the marcher-trained field")
    ap.add_argument("--no-legs", action="store_true", help="skip the encoder-alone and integrate-alone timings")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy 5 times (the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 64 << 20  # floats: 256 MiB
        src = torch.ones(n, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from jaxngp_amd import scene as sc
    from jaxngp_amd.train import TrainConfig

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    blobs = sc.BlobScene.random(seed=3)
    cam = sc.Camera()
    poses = sc.look_at_poses(args.views, seed=1)
    data = sc.make_views(blobs, cam, poses, dev, progress=lambda v: log(f"rendering view {v}"))
    torch.cuda.synchronize()
    log(f"scene: {args.views} views {cam.width}x{cam.height} in {time.perf_counter() - t0:.1f}s")

    res = run_training(data, TrainConfig(), args.steps, args.warmup, dev)
    tr, per, dom, dom_ms, elapsed, n_samples, n_valid, ms = (
        res[k] for k in ("trainer", "per", "dom", "dom_ms", "elapsed", "samples", "valid", "ms"))
    loss_fp32 = tr.metrics()["loss"]  # after warmup + timed steps, as the other paths' losses
    cfg = tr.cfg
    S = cfg.total_samples
    n_rays = tr.n_rays
    spp = n_samples / args.steps
    algo = {
        "hashgrid_encode": ("hbm", S * ENC_BYTES_PER_POINT),
        "hashgrid_encode_backward": ("hbm", S * ENC_BYTES_PER_POINT),
        "hashgrid_encode_backward_adam": ("hbm", S * ENC_BYTES_PER_POINT +
                                          tr.model.position_encoder.table.numel() * ADAM_BYTES_PER_PARAM),
        "march_rays": ("hbm", spp * MARCH_BYTES_PER_SAMPLE + n_rays * MARCH_BYTES_PER_RAY),
        "integrate_rays": ("hbm", S * INT_FWD_BYTES_PER_SAMPLE + n_rays * INT_FWD_BYTES_PER_RAY),
        "integrate_rays_backward": ("hbm", S * INT_BWD_BYTES_PER_SAMPLE + n_rays * INT_BWD_BYTES_PER_RAY),
        "nerf_mlp_forward": ("mfma", S * MLP_FWD_FLOP),
        "nerf_mlp_backward": ("mfma", S * MLP_FWD_FLOP * 3),
    }
    assert set(algo) == set(ROOFLINE_CANDIDATES)
    bound, work = algo[dom]
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on this
    # workload (copy-calibrated), and the L2 hit rate from TCC_HIT_sum / TCC_MISS_sum,
    # committed beside the kernel-time summary
    from jaxngp_amd.profiling import read_l2_hit_rate, read_traffic
    prof = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    traffic = read_traffic(os.path.join(prof, "training_step_fp32_hbm_traffic.txt")).get(dom)
    l2_hit_rate = read_l2_hit_rate(os.path.join(prof, "training_step_fp32_l2_hit_rate.txt")).get(dom)
    if bound == "hbm":
        achieved = work / (dom_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "algorithmic_bytes_per_launch": work, "ms_per_launch": dom_ms, "l2_hit_rate": l2_hit_rate}
    else:
        achieved = work / (dom_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_MFMA_PEAK_TFLOPS, "traffic": traffic, "kernel": dom,
                "algorithmic_flops_per_launch": work, "ms_per_launch": dom_ms, "l2_hit_rate": l2_hit_rate}

    out = {
        "metric": "train samples/s & ms per 800² frame: proposal+FP16 vs plain, ΔPSNR≤0.1dB",
        "value": n_samples / elapsed,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain FP32 occupancy-grid training step, 8-blob procedural scene, "
                               f"{args.views} views 800x800, 2^18 sample slots, n_rays = 2^18",
                   "total_samples": S, "n_rays": n_rays, "grid": "128^3 x 1 cascade", "hashgrid": "L16 F2 T2^19",
                   "path": "plain-fp32"},
        "rays_per_s": n_valid / elapsed,
        "samples_per_step": spp,
        "kernel_ms_per_launch": {k: v for k, v in per.items()},
        "kernel_times": f"ms per launch from HIP events: the {dom} launch time (roofline) live in the timed "
                        f"region, the others from a {CALIBRATION_STEPS}-step untimed calibration pass before it",
        "roofline": roof,
    }
    out["loss"] = loss_fp32
    out["loss_at_step"] = tr.step_count
    if not args.no_legs:
        out["encoder_alone"] = bench_encoder(dev)
        out["integrate_alone"] = bench_integrate(dev)
    if args.frames > 0:
        train_to(tr, args.render_steps)
        out["steady_state"] = steady_state(tr)
        out["inference"] = bench_inference(tr, data, args.frames, dev)
    if not args.no_proposal:
        # the learned-proposal path on the precision path.  `proposal_fp16`: the field
        # trains on proposal-placed samples (32 at the proposal's quantiles + 32
        # stratified uniform per ray, occupancy threshold for that sample spacing) and
        # renders with 48 proposal samples per ray (occupancy-grid fallback);
        # `proposal_fp16_marcher_trained`: the marcher-trained FP16 field with the same
        # proposal beside it; `plain_fp32_at_guided_threshold`: the plain FP32 path with
        # the guided path's occupancy threshold, the control for the threshold
        from jaxngp_amd.train import Trainer
        runs = [("proposal_fp16", TrainConfig(precision="fp16", proposal=True, proposal_guided=True))]
        if not args.no_marcher_proposal:
            runs.append(("proposal_fp16_marcher_trained", TrainConfig(precision="fp16", proposal=True)))
        guided_cap = Trainer.threshold_cap_of(runs[0][1], float(data.scene.bound))
        runs.append(("plain_fp32_at_guided_threshold", TrainConfig(density_threshold=guided_cap)))
        for key, pcfg in runs:
            rp = run_training(data, pcfg, args.steps, args.warmup, dev)
            tp = rp["trainer"]
            m = tp.metrics()
            out[key] = {"value": rp["samples"] / rp["elapsed"], "unit": "samples/s",
                        "ms_per_step": rp["ms"], "rays_per_s": rp["valid"] / rp["elapsed"],
                        "n_rays": tp.n_rays, "samples_per_step": rp["samples"] / args.steps,
                        "density_threshold_cap": tp.density_threshold_cap,
                        "loss": m["loss"], "loss_at_step": tp.step_count}
            if pcfg.proposal:
                out[key]["proposal_loss"] = m["proposal_loss"]
            if pcfg.proposal_guided:
                out[key].update(guided_samples_per_ray=pcfg.proposal_samples_per_ray,
                                uniform_samples_per_ray=pcfg.proposal_uniform_samples)
            if args.frames > 0:
                train_to(tp, args.render_steps)
                out[key]["steady_state"] = steady_state(tp)
                out[key]["inference"] = bench_inference(tp, data, args.frames, dev, proposal=pcfg.proposal)
                if "inference" in out:
                    out[key]["inference"]["delta_psnr_vs_plain_fp32_db"] = (
                        out[key]["inference"]["psnr_db"] - out["inference"]["psnr_db"])
            del rp, tp
        if args.frames > 0 and "steady_state" in out:
            # the target's ratios, steady state (after the render_steps of training)
            pf, pl = out["proposal_fp16"], out
            ctl = out["plain_fp32_at_guided_threshold"]
            fr = pf["inference"]
            out["target"] = {
                "train_samples_per_s_ratio": pf["steady_state"]["value"] / pl["steady_state"]["value"],
                "train_rays_per_s_ratio": pf["steady_state"]["rays_per_s"] / pl["steady_state"]["rays_per_s"],
                "render_setting": {k: fr[k] for k in ("samples_per_ray", "mass_threshold", "placement")},
                "frame_ms_ratio": fr["ms_per_frame"] / pl["inference"]["ms_per_frame"],
                "frame_ms_ratio_renderer_default": fr["settings"][0]["ms_per_frame"] / pl["inference"]["ms_per_frame"],
                "delta_psnr_db": fr["psnr_db"] - pl["inference"]["psnr_db"],
                "same_field_delta_psnr_db": fr["delta_psnr_db"],
                "same_field_worst_view_delta_psnr_db": fr["worst_view_delta_psnr_db"],
                "vs_plain_at_guided_threshold": {
                    "train_samples_per_s_ratio": pf["steady_state"]["value"] / ctl["steady_state"]["value"],
                    "train_rays_per_s_ratio": pf["steady_state"]["rays_per_s"] / ctl["steady_state"]["rays_per_s"],
                    "frame_ms_ratio": fr["ms_per_frame"] / ctl["inference"]["ms_per_frame"],
                    "delta_psnr_db": fr["psnr_db"] - ctl["inference"]["psnr_db"]},
                "note": "plain = the reference configuration (occupancy threshold cap 0.01 * 1024 / (2 sqrt 3)); "
                        "the guided path's cap follows the same formula with its 64 samples per ray",
            }
    if not args.no_fp16:
        # the precision path on the same scene: f16 table and MLP weights, loss scaling
        r16 = run_training(data, TrainConfig(precision="fp16"), args.steps, args.warmup, dev)
        t16 = r16["trainer"]
        out["fp16"] = {"value": r16["samples"] / r16["elapsed"], "unit": "samples/s", "ms_per_step": r16["ms"],
                       "rays_per_s": r16["valid"] / r16["elapsed"], "kernel_ms_per_launch": r16["per"],
                       "loss": t16.metrics()["loss"], "loss_fp32": loss_fp32, "loss_at_step": t16.step_count}
        if args.frames > 0:
            train_to(t16, args.render_steps)
            out["fp16"]["steady_state"] = steady_state(t16)
            out["fp16"]["inference"] = bench_inference(t16, data, args.frames, dev)
            if "inference" in out:
                out["fp16"]["delta_psnr_vs_fp32_db"] = out["fp16"]["inference"]["psnr_db"] - out["inference"]["psnr_db"]
        del r16, t16
        if not args.no_t22:
            # the precision path with 2^22 entries per level (39,474,024 rows, f16)
            r22 = run_training(data, TrainConfig(precision="fp16", log2_hashmap_size=22), args.steps, args.warmup,
                               dev)
            t22 = r22["trainer"]
            rec = {"value": r22["samples"] / r22["elapsed"], "unit": "samples/s", "ms_per_step": r22["ms"],
                   "table_rows": t22.model.position_encoder.levels.n_rows, "loss": t22.metrics()["loss"],
                   "loss_at_step": t22.step_count}
            if args.frames > 0:
                train_to(t22, args.render_steps)
                rec["inference"] = bench_inference(t22, data, args.frames, dev)
                if "inference" in out:
                    rec["delta_psnr_vs_fp32_db"] = rec["inference"]["psnr_db"] - out["inference"]["psnr_db"]
            out["fp16_T2^22"] = rec
            del r22, t22
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
        out["cpu_baseline"]["cpu_model"] = cpu_model()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
