"""The rocprofv3 summaries bench.py and the committed profiles rely on, on
synthetic counter files in rocprofv3's CSV layout (one row per dispatch and
counter)."""
import csv

import pytest

from jaxngp_amd import profiling

FIELDS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(d, rows):
    d.mkdir(parents=True, exist_ok=True)
    with open(d / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, FIELDS)
        w.writeheader()
        for i, (k, c, v) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": c, "Counter_Value": v})


def test_traffic_scales_by_the_copy_and_divides_by_launches(tmp_path):
    k = "void ngp::(anonymous namespace)::integrate_kernel(ngp::Desc, float const*)"
    _write(tmp_path / "f", [(k, "FETCH_SIZE", 100.0), (k, "FETCH_SIZE", 300.0)])
    _write(tmp_path / "w", [(k, "WRITE_SIZE", 50.0), (k, "WRITE_SIZE", 50.0)])
    copy = "__amd_rocclr_copyBuffer"
    _write(tmp_path / "cf", [(copy, "FETCH_SIZE", 512.0)])  # half of the 1 MiB the copy reads
    _write(tmp_path / "cw", [(copy, "WRITE_SIZE", 1024.0)])
    out = profiling.traffic(*(str(tmp_path / n) for n in ("f", "w", "cf", "cw")), 1048576.0)
    (tmp_path / "t.txt").write_text(out)
    # per launch: fetch (400 / 2) KiB x 2, write (100 / 2) KiB x 1
    assert profiling.read_traffic(str(tmp_path / "t.txt")) == {"integrate_rays": (400 + 50) * 1024.0}


def test_l2_hit_rate_per_target_and_kernel(tmp_path):
    a = "void ngp::(anonymous namespace)::bwd_scatter_kernel<2, false>(ngp::HashGridDescriptor)"
    b = "void ngp::(anonymous namespace)::bwd_reduce_kernel<2, false>(ngp::HashGridDescriptor)"
    _write(tmp_path / "hm", [(a, "TCC_HIT_sum", 30.0), (a, "TCC_MISS_sum", 10.0),
                             (b, "TCC_HIT_sum", 10.0), (b, "TCC_MISS_sum", 50.0),
                             ("other_kernel", "TCC_HIT_sum", 1.0), ("other_kernel", "TCC_MISS_sum", 0.0)])
    out = profiling.l2(str(tmp_path / "hm"))
    rates = {p[1]: float(p[2]) for p in (line.split() for line in out.splitlines()) if p and p[0] == "l2_hit_rate"}
    assert rates == {"hashgrid_encode_backward": pytest.approx(40 / 100)}
    (tmp_path / "l2.txt").write_text(out)
    assert profiling.read_l2_hit_rate(str(tmp_path / "l2.txt")) == {"hashgrid_encode_backward": 0.4}
    per_kernel = dict(line.split() for line in out.splitlines()[-2:])
    assert float(per_kernel["bwd_scatter_kernel"]) == pytest.approx(0.75)
    assert float(per_kernel["bwd_reduce_kernel"]) == pytest.approx(1 / 6, abs=1e-3)


def test_read_traffic_of_a_missing_file_is_empty(tmp_path):
    assert profiling.read_traffic(str(tmp_path / "none.txt")) == {}


def test_counters_average_per_dispatch(tmp_path):
    k = "void ngp::(anonymous namespace)::mlp_backward_kernel(ngp::MlpDescriptor)"
    _write(tmp_path / "c", [(k, "SQ_BUSY_CYCLES", 100.0), (k, "SQ_VALU_MFMA_BUSY_CYCLES", 30.0),
                            (k, "SQ_BUSY_CYCLES", 300.0), (k, "SQ_VALU_MFMA_BUSY_CYCLES", 50.0),
                            ("at::native::foo", "SQ_BUSY_CYCLES", 7.0)])
    lines = profiling.counters(str(tmp_path / "c")).splitlines()
    assert lines[0].split()[-2:] == ["SQ_BUSY_CYCLES", "SQ_VALU_MFMA_BUSY_CYCLES"]
    assert lines[1].split() == ["mlp_backward_kernel", "2", "200", "40"]
    assert len(lines) == 2  # kernels outside the library are left out


def test_gaps_counts_overlap_once_and_idle_time(tmp_path):
    """Per-step idle time from a kernel trace: overlapping dispatches (two streams)
    count once, and the idle rest is the span minus that."""
    import csv
    from jaxngp_amd import profiling
    rows, t = [], 0
    for _ in range(4):
        rows.append(("void ngp::(anonymous namespace)::make_rays_kernel(int)", t, t + 10_000))
        rows.append(("march_emit_kernel", t + 15_000, t + 65_000))
        rows.append(("adam_kernel<float>", t + 40_000, t + 80_000))  # overlaps the march by 25 us
        t += 100_000
    f = tmp_path / "trace.csv"
    with open(f, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writerows(rows)
    out = profiling.gaps(str(f))
    # span 100 us; busy 10 + (15..80 = 65) = 75 us; idle 25 us
    assert "median span 100.0 us" in out and "kernels running 75.0 us" in out and "idle 25.0 us" in out
    assert "3 dispatches per step" in out


def test_mangled_kernel_names_are_shortened():
    from jaxngp_amd.profiling import _base
    assert _base("_ZN3ngp12_GLOBAL__N_121mlp_backward_h_kernelEjbPKfS2_PKDF16_S2_P") == "mlp_backward_h_kernel"
    assert _base("_ZN3ngp12_GLOBAL__N_119ogrid_update_kernelIDF16_Lb1EEEv18HashGri") == "ogrid_update_kernel"
    assert _base("void ngp::(anonymous namespace)::bwd_bucket_kernel<2, false>(HashGridDescriptor)") == "bwd_bucket_kernel"
