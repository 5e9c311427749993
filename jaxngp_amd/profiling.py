"""Summaries of rocprofv3 output for the kernels of this library.

    python -m jaxngp_amd.profiling stats  KERNEL_STATS_CSV [TOP]
    python -m jaxngp_amd.profiling traffic FETCH_DIR WRITE_DIR COPY_FETCH_DIR COPY_WRITE_DIR COPY_BYTES
    python -m jaxngp_amd.profiling l2 HIT_MISS_DIR
    python -m jaxngp_amd.profiling counters PMC_DIR
    python -m jaxngp_amd.profiling gaps KERNEL_TRACE_CSV [STEP_KERNEL] [LAST_STEPS]

`stats` prints the kernels of a `--kernel-trace --stats` run by total time.
`traffic` turns two `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of a workload, and
the same two counters read on a plain device copy of COPY_BYTES bytes, into HBM
bytes per launch of each target: every counter is scaled so that the copy reads and
writes exactly COPY_BYTES (the factor is printed), summed over the target's kernels
and divided by the target's launches.
`l2` turns one `--pmc TCC_HIT_sum TCC_MISS_sum` run into each target's L2 hit rate,
hits / (hits + misses) summed over the target's kernels.
`counters` prints, for each kernel of the targets, every counter of a `--pmc` run
averaged per dispatch.
`gaps` reads a `--kernel-trace` CSV and reports, per training step, the time some
kernel ran and the GPU's idle rest.
"""
from __future__ import annotations

import collections
import csv
import glob
import re
import sys

# the kernels each target launches, and one that runs exactly once per launch
TARGET_KERNELS = {
    "march_rays": (["march_count_kernel", "march_scan_kernel", "march_emit_kernel", "march_fill_kernel",
                    "march_finish_kernel", "march_empty_kernel"], "march_scan_kernel"),
    "hashgrid_encode": (["hashgrid_fwd_lm_kernel", "hashgrid_fwd_kernel"], None),
    "hashgrid_encode_backward": (["bwd_bucket_kernel", "bwd_slice_parts_kernel", "bwd_bucket_plan_kernel",
                                  "bwd_reduce_runs_kernel", "bwd_count_kernel", "scan_sums_kernel",
                                  "scan_apply_kernel", "plan_parts_kernel", "bwd_scatter_kernel", "bwd_reduce_kernel",
                                  "bwd_combine_kernel", "zero_table_kernel"], "bwd_combine_kernel"),
    # the same kernels with the table's optimizer step in the reduce and combine passes
    "hashgrid_encode_backward_adam": (["bwd_bucket_kernel", "bwd_slice_parts_kernel", "bwd_bucket_plan_kernel",
                                       "bwd_reduce_runs_kernel", "bwd_combine_kernel", "adam_zero_grad_kernel"],
                                      "bwd_combine_kernel"),
    "nerf_mlp_forward": (["mlp_forward_kernel", "mlp_forward_h_kernel"], None),
    "nerf_mlp_backward": (["mlp_backward_kernel", "mlp_backward_h_kernel", "sum_rows_kernel"], "sum_rows_kernel"),
    "integrate_rays": (["integrate_kernel"], "integrate_kernel"),
    "integrate_rays_backward": (["integrate_backward_kernel"], "integrate_backward_kernel"),
    "occupancy_grid_update": (["ogrid_update_kernel", "ogrid_decay_kernel", "ogrid_prefix_kernel"],
                              "ogrid_update_kernel"),
    "proposal_train_samples": (["proposal_train_kernel", "proposal_train_scan_kernel", "proposal_train_write_kernel",
                                "proposal_train_tail_kernel"], "proposal_train_scan_kernel"),
}


# kernels outside the targets above whose counters are also of interest
EXTRA_KERNELS = ("mlp_forward_h_kernel", "mlp_backward_h_kernel", "field_chunks_kernel", "proposal_render_kernel")


def _short(name: str) -> str:
    if name.startswith("_Z"):
        # a mangled name the tool left as is: its first unqualified identifier after the
        # namespaces (length-prefixed), e.g. _ZN3ngp12_GLOBAL__N_121mlp_backward_h_kernelE...
        m = re.search(r"_GLOBAL__N_1(\d+)", name) or re.search(r"^_ZN3ngp(\d+)", name)
        if m:
            return name[m.end():m.end() + int(m.group(1))]
    n = name.replace("(anonymous namespace)::", "").replace("void ", "").replace("ngp::", "")
    return n if n.startswith("at::") else n.split("(")[0]


def _base(name: str) -> str:
    return _short(name).split("<")[0]


def stats(path: str, top: int = 40) -> str:
    rows = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    out = [f"{'kernel':64s} {'calls':>7s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'%time':>6s}"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        out.append(f"{_short(r['Name'])[:64]:64s} {r['Calls']:>7s} {float(r['AverageNs']) / 1e3:9.1f} "
                   f"{float(r['MinNs']) / 1e3:9.1f} {float(r['MaxNs']) / 1e3:9.1f} "
                   f"{float(r['TotalDurationNs']) / tot * 100:6.1f}")
    return "\n".join(out)


def _counter(d: str, name: str | None = None) -> dict:
    """kernel base name -> [dispatches, summed counter value], over the rows of
    counter `name` (all rows when None: a run of one counter)"""
    agg = collections.defaultdict(lambda: [0, 0.0])
    for f in glob.glob(d + "/**/*counter_collection*.csv", recursive=True):
        for r in csv.DictReader(open(f)):
            if name is not None and r["Counter_Name"] != name:
                continue
            a = agg[_base(r["Kernel_Name"])]
            a[0] += 1
            a[1] += float(r["Counter_Value"] or 0)
    return agg


def traffic(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_bytes) -> str:
    f, w = _counter(fetch_dir), _counter(write_dir)
    cf, cw = _counter(copy_fetch_dir), _counter(copy_write_dir)
    pick = lambda c: max(c.items(), key=lambda kv: kv[1][1])  # the copy kernel: the largest counter total
    (cname, (cn, cfv)), (_, (cwn, cwv)) = pick(cf), pick(cw)
    kf = copy_bytes / (cfv / cn * 1024.0)
    kw = copy_bytes / (cwv / cwn * 1024.0)
    out = [f"copy calibration: kernel {cname}, {copy_bytes} bytes; FETCH_SIZE reads {cfv / cn:.1f} KiB -> scale "
           f"{kf:.4f}; WRITE_SIZE reads {cwv / cwn:.1f} KiB -> scale {kw:.4f}", ""]
    out.append(f"{'target':26s} {'launches':>8s} {'fetch_MB':>10s} {'write_MB':>10s} {'traffic_MB':>11s}  kernels")
    for tgt, (kernels, once) in TARGET_KERNELS.items():
        present = [k for k in kernels if k in f and k in w]
        if not present:
            continue
        launches = f[once][0] if once in f else max(f[k][0] for k in present)
        fb = sum(f[k][1] for k in present) * 1024.0 * kf / launches
        wb = sum(w[k][1] for k in present) * 1024.0 * kw / launches
        out.append(f"{tgt:26s} {launches:8d} {fb / 1e6:10.1f} {wb / 1e6:10.1f} {(fb + wb) / 1e6:11.1f}  "
                   f"{','.join(present)}")
        out.append(f"traffic_bytes_per_launch {tgt} {fb + wb:.0f}")
    return "\n".join(out)


def l2(hit_miss_dir: str) -> str:
    h, m = _counter(hit_miss_dir, "TCC_HIT_sum"), _counter(hit_miss_dir, "TCC_MISS_sum")
    out = [f"{'target':26s} {'hits_M':>10s} {'misses_M':>10s} {'hit_rate':>9s}  kernels"]
    for tgt, (kernels, _) in TARGET_KERNELS.items():
        present = [k for k in kernels if k in h and k in m]
        if not present:
            continue
        hs, ms = sum(h[k][1] for k in present), sum(m[k][1] for k in present)
        rate = hs / (hs + ms) if hs + ms else float("nan")
        out.append(f"{tgt:26s} {hs / 1e6:10.2f} {ms / 1e6:10.2f} {rate:9.3f}  {','.join(present)}")
        out.append(f"l2_hit_rate {tgt} {rate:.4f}")
    out.append("")
    out.append(f"{'kernel':40s} {'hit_rate':>9s}")
    for k in sorted(set(h) & set(m)):
        if any(k in ks for ks, _ in TARGET_KERNELS.values()):
            tot = h[k][1] + m[k][1]
            out.append(f"{k:40s} {h[k][1] / tot if tot else float('nan'):9.3f}")
    return "\n".join(out)


def _read_keyed(path: str, key: str) -> dict:
    res = {}
    try:
        for line in open(path):
            p = line.split()
            if len(p) == 3 and p[0] == key:
                res[p[1]] = float(p[2])
    except OSError:
        pass
    return res


def counters(pmc_dir: str) -> str:
    rows = collections.defaultdict(lambda: collections.defaultdict(lambda: [0, 0.0]))
    for f in glob.glob(pmc_dir + "/**/*counter_collection*.csv", recursive=True):
        for r in csv.DictReader(open(f)):
            a = rows[_base(r["Kernel_Name"])][r["Counter_Name"]]
            a[0] += 1
            a[1] += float(r["Counter_Value"] or 0)
    ours = {k for ks, _ in TARGET_KERNELS.values() for k in ks} | set(EXTRA_KERNELS)
    names = sorted({c for k in rows if k in ours for c in rows[k]})
    out = [f"{'kernel (average per dispatch)':32s} {'dispatches':>10s} " + " ".join(f"{c:>26s}" for c in names)]
    for k in sorted(k for k in rows if k in ours):
        n = max(v[0] for v in rows[k].values())
        out.append(f"{k:32s} {n:10d} " + " ".join(
            f"{rows[k][c][1] / rows[k][c][0]:26.4g}" if rows[k][c][0] else f"{'-':>26s}" for c in names))
    return "\n".join(out)


def gaps(trace_csv: str, step_kernel: str = "make_rays_kernel", last: int = 0) -> str:
    """Per training step (the span from one `step_kernel` dispatch to the next): the
    span, the time some kernel was running (the union of the dispatch intervals, so
    overlapping streams count once) and the idle rest, as medians over the steps
    (the last `last` steps only, when given), plus the dispatches per step."""
    rows = list(csv.DictReader(open(trace_csv)))
    iv = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), _base(r["Kernel_Name"])) for r in rows)
    starts = [a for a, _, k in iv if k == step_kernel]
    steps = []
    for s0, s1 in zip(starts, starts[1:]):
        busy, cur_a, cur_b, n = 0, None, None, 0
        for a, b, _ in iv:
            if b <= s0 or a >= s1:
                continue
            a, b = max(a, s0), min(b, s1)
            n += 1
            if cur_b is None or a > cur_b:
                if cur_b is not None:
                    busy += cur_b - cur_a
                cur_a, cur_b = a, b
            else:
                cur_b = max(cur_b, b)
        if cur_b is not None:
            busy += cur_b - cur_a
        steps.append((s1 - s0, busy, n))
    if last:
        steps = steps[-last:]
    med = lambda xs: sorted(xs)[len(xs) // 2]
    span, busy, n = med([x[0] for x in steps]), med([x[1] for x in steps]), med([x[2] for x in steps])
    return (f"steps {len(steps)}: median span {span / 1e3:.1f} us, kernels running {busy / 1e3:.1f} us, "
            f"idle {(span - busy) / 1e3:.1f} us ({(span - busy) / span:.1%}), {n} dispatches per step")


def read_traffic(path: str) -> dict:
    """{target: HBM bytes per launch} from a summary written by `traffic`."""
    return _read_keyed(path, "traffic_bytes_per_launch")


def read_l2_hit_rate(path: str) -> dict:
    """{target: L2 hit rate} from a summary written by `l2`."""
    return _read_keyed(path, "l2_hit_rate")


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        print(stats(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 40))
    elif sys.argv[1] == "traffic":
        print(traffic(*sys.argv[2:6], float(sys.argv[6])))
    elif sys.argv[1] == "l2":
        print(l2(sys.argv[2]))
    elif sys.argv[1] == "counters":
        print(counters(sys.argv[2]))
    elif sys.argv[1] == "gaps":
        print(gaps(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else "make_rays_kernel",
                   int(sys.argv[4]) if len(sys.argv) > 4 else 0))
    else:
        raise SystemExit(__doc__)
