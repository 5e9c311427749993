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


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--views", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from jaxngp_amd import scene as sc
    from jaxngp_amd.train import TrainConfig, Trainer

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    blobs = sc.BlobScene.random(seed=3)
    cam = sc.Camera()
    poses = sc.look_at_poses(args.views, seed=1)
    data = sc.make_views(blobs, cam, poses, dev, progress=lambda v: log(f"rendering view {v}"))
    torch.cuda.synchronize()
    log(f"scene: {args.views} views {cam.width}x{cam.height} in {time.perf_counter() - t0:.1f}s")

    cfg = TrainConfig()
    tr = Trainer(data, cfg, device=dev)
    for i in range(args.warmup):
        tr.step()
        if i % 100 == 0:
            torch.cuda.synchronize()
            log(f"warmup step {i}: {tr.metrics()}")
    torch.cuda.synchronize()

    # per-target HIP events on the launch stream
    timed = ["march_rays", "hashgrid_encode", "nerf_mlp_forward", "integrate_rays", "integrate_rays_backward",
             "nerf_mlp_backward", "hashgrid_encode_backward", "adam_update"]
    from jaxngp_amd import _lib
    events: dict = {k: [] for k in timed}
    orig_call = _lib.call

    def timed_call(name, tensors, opaque, stream=None):
        if name in events:
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            orig_call(name, tensors, opaque, stream)
            b.record()
            events[name].append((a, b))
        else:
            orig_call(name, tensors, opaque, stream)

    samples = torch.zeros((), dtype=torch.int64, device=dev)
    rays = torch.zeros((), dtype=torch.int64, device=dev)
    _lib.call = timed_call
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        tr.step()
        b = tr.b
        samples += b["next_loc"].view(torch.int32).long()[0] - b["exceeded"].view(torch.int32).long()[0]
        rays += b["n_valid"].view(torch.int32).long()[0]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    _lib.call = orig_call
    n_samples = int(samples.item())
    n_valid = int(rays.item())
    ms = elapsed / args.steps * 1e3
    log(f"timed {args.steps} steps: {ms:.3f} ms/step, {n_samples / args.steps:.0f} samples/step, "
        f"{n_valid / args.steps:.0f} valid rays/step, metrics {tr.metrics()}")

    per = {}
    for k, lst in events.items():
        if lst:
            per[k] = sum(a.elapsed_time(b) for a, b in lst) / len(lst)  # ms per launch
    S = cfg.total_samples
    n_rays = tr.n_rays
    spp = n_samples / args.steps
    algo = {
        "hashgrid_encode": ("hbm", S * ENC_BYTES_PER_POINT),
        "hashgrid_encode_backward": ("hbm", S * ENC_BYTES_PER_POINT),
        "march_rays": ("hbm", spp * MARCH_BYTES_PER_SAMPLE + n_rays * MARCH_BYTES_PER_RAY),
        "integrate_rays": ("hbm", S * INT_FWD_BYTES_PER_SAMPLE + n_rays * INT_FWD_BYTES_PER_RAY),
        "integrate_rays_backward": ("hbm", S * INT_BWD_BYTES_PER_SAMPLE + n_rays * INT_BWD_BYTES_PER_RAY),
        "nerf_mlp_forward": ("mfma", S * MLP_FWD_FLOP),
        "nerf_mlp_backward": ("mfma", S * MLP_FWD_FLOP * 3),
    }
    dom = max((k for k in per if k in algo), key=lambda k: per[k])
    bound, work = algo[dom]
    if bound == "hbm":
        achieved = work / (per[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                "algorithmic_bytes_per_launch": work, "ms_per_launch": per[dom]}
    else:
        achieved = work / (per[dom] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP32_MFMA_PEAK_TFLOPS, "traffic": None, "kernel": dom,
                "algorithmic_flops_per_launch": work, "ms_per_launch": per[dom]}

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
        "kernel_ms_per_step": {k: v for k, v in per.items()},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
