"""The offline MLUPS / roofline reporter on synthetic rocprofv3 CSV files."""
import csv

import pytest

from multicomponent_color_gradient_amd import perf

COUNTER_COLS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
                "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
                "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def write_counters(path, rows):
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(COUNTER_COLS)
        for disp, name, counter, value in rows:
            w.writerow([disp, disp, "Agent 1", 1, 1, 1, 1024, 1, name, 256, 0, 0, 32, 0, 32, counter, value, 0, 1])


def synthetic(tmp_path):
    # setup copy (small), probe copies (1 GiB read back as 0.5 GiB FETCH / 1 GiB WRITE),
    # two dispatches of the kernel, a later big copy (adjoint) that must be ignored
    gib_kib = (1 << 30) / 1024
    fetch = [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 12.0),
             (2, "void mcg::k_fused<false, 0>(...)", "FETCH_SIZE", 3.0e6),
             (3, "void mcg::k_fused<false, 0>(...)", "FETCH_SIZE", 3.2e6),
             (4, "__amd_rocclr_copyBuffer", "FETCH_SIZE", gib_kib / 2),
             (5, "__amd_rocclr_copyBuffer", "FETCH_SIZE", gib_kib / 2),
             (6, "__amd_rocclr_copyBuffer", "FETCH_SIZE", gib_kib / 2),
             (7, "__amd_rocclr_copyBuffer", "FETCH_SIZE", 3 * gib_kib)]
    write = [(d, n, "WRITE_SIZE", v * (2.0 if "copy" in n and v > 1e5 else 1.0)) for d, n, _, v in fetch]
    write[1] = (2, "void mcg::k_fused<false, 0>(...)", "WRITE_SIZE", 7.0e6)
    write[2] = (3, "void mcg::k_fused<false, 0>(...)", "WRITE_SIZE", 7.0e6)
    fp, wp = tmp_path / "fetch.csv", tmp_path / "write.csv"
    write_counters(fp, fetch)
    write_counters(wp, write)
    return str(fp), str(wp)


def test_copy_probe_scale_and_traffic(tmp_path):
    fp, wp = synthetic(tmp_path)
    f, w = perf.read_counters(fp), perf.read_counters(wp)
    assert perf.copy_probe_scale(f, "FETCH_SIZE") == pytest.approx(2048.0)  # bytes per KiB unit: counter reads half
    assert perf.copy_probe_scale(w, "WRITE_SIZE") == pytest.approx(1024.0)
    t = perf.traffic(f, w, "k_fused", nodes=1000)
    assert t["dispatches"] == 2
    assert t["read_bytes"] == pytest.approx(3.1e6 * 2048)
    assert t["write_bytes"] == pytest.approx(7.0e6 * 1024)
    assert t["bytes_per_launch"] == pytest.approx(t["read_bytes"] + t["write_bytes"])
    assert t["read_per_node"] == pytest.approx(t["read_bytes"] / 1000)
    assert t["fetch_scale"] == pytest.approx(2.0) and t["write_scale"] == pytest.approx(1.0)


def test_missing_probe_is_an_error(tmp_path):
    p = tmp_path / "c.csv"
    write_counters(p, [(1, "void k(...)", "FETCH_SIZE", 5.0)])
    with pytest.raises(ValueError):
        perf.copy_probe_scale(perf.read_counters(str(p)), "FETCH_SIZE")


def test_kernel_stats_table(tmp_path):
    p = tmp_path / "run_kernel_stats.csv"
    with open(p, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"])
        w.writerow(["void mcg::k_fused<false, 0>(mcg::DevParams)", 22, 89953944, 4088815.6, 82.49, 4030021, 4147606, 1])
    rows = perf.read_kernel_stats(str(p))
    assert rows[0]["calls"] == 22 and rows[0]["avg_ns"] == pytest.approx(4088815.6)
    table = perf.stats_table(rows)
    assert "| 22 | 4088.8 | 4030.0 | 4147.6 | 82.49 |" in table


def test_mlups_and_roofline():
    # the reference's definition: steps * cells / t; 432 B per update at 4.1 ms for 32.9 M updates
    assert perf.mlups(200, 256 * 512 * 256, 0.8) == pytest.approx(200 * 33554432 / 0.8 / 1e6)
    r = perf.roofline(432, 254 * 510 * 254, 4.1e-3)
    assert r["achieved"] == pytest.approx(432 * 254 * 510 * 254 / 4.1e-3 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)


def test_cli(tmp_path, capsys):
    fp, wp = synthetic(tmp_path)
    assert perf.main(["traffic", fp, wp, "--kernel", "k_fused", "--nodes", "10"]) == 0
    assert "bytes_per_launch" in capsys.readouterr().out


def test_trace_groups_launches_by_grid(tmp_path):
    cols = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
            "Correlation_Id", "Start_Timestamp", "End_Timestamp", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
            "Accum_VGPR_Count", "SGPR_Count", "Workgroup_Size_X", "Workgroup_Size_Y", "Workgroup_Size_Z",
            "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"]
    p = tmp_path / "run_kernel_trace.csv"
    with open(p, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(cols)
        # k on a big grid twice (1000 ns, 3000 ns), then on a small grid once, then another kernel
        for disp, name, t0, t1, gx in [(1, "k", 0, 1000, 4096), (2, "k", 5000, 8000, 4096), (3, "k", 9000, 9100, 64),
                                       (4, "q", 9200, 9300, 64)]:
            w.writerow(["KERNEL_DISPATCH", "Agent 1", 1, 0, 1, disp, 1, name, disp, t0, t1, 0, 0, 32, 0, 32, 64, 4, 1,
                        gx, 4, 2])
    rows = perf.trace_stats(str(p))
    assert [(r["name"], r["grid"], r["calls"]) for r in rows] == [("k", 32768, 2), ("k", 512, 1), ("q", 512, 1)]
    assert rows[0]["avg_ns"] == 2000 and rows[0]["min_ns"] == 1000 and rows[0]["max_ns"] == 3000
    assert "| `k` | 32768 | 2 | 2.0 | 1.0 | 3.0 |" in perf.trace_table(rows)


def sized(tmp_path):
    # setup copy, three probe copies of 1 GiB (128-byte reads, 64-byte writes), two step dispatches
    gib = 1 << 30
    rd, wr = [], []
    for disp, name, n32, n64, n128, w64, w32 in [
            (1, "__amd_rocclr_copyBuffer", 0, 24, 1, 30, 2),
            (2, "void mcg::k_fused<false, 0>(...)", 10, 300, 74_000_000, 112_000_000, 7000),
            (3, "void mcg::k_fused<false, 0>(...)", 10, 340, 74_100_000, 112_100_000, 6000),
            (4, "__amd_rocclr_copyBuffer", 0, 200, gib // 128, gib // 64, 0),
            (5, "__amd_rocclr_copyBuffer", 0, 202, gib // 128, gib // 64, 0),
            (6, "__amd_rocclr_copyBuffer", 0, 318, gib // 128, gib // 64, 0)]:
        rd += [(disp, name, "TCC_EA0_RDREQ_32B_sum", n32), (disp, name, "TCC_EA0_RDREQ_64B_sum", n64),
               (disp, name, "TCC_EA0_RDREQ_128B_sum", n128), (disp, name, "TCC_EA0_RDREQ_sum", n32 + n64 + n128)]
        wr += [(disp, name, "TCC_EA0_WRREQ_64B_sum", w64), (disp, name, "TCC_EA0_WRREQ_sum", w64 + w32)]
    rp, wp = tmp_path / "rd.csv", tmp_path / "wr.csv"
    write_counters(rp, rd)
    write_counters(wp, wr)
    return str(rp), str(wp)


def test_request_size_traffic(tmp_path):
    rp, wp = sized(tmp_path)
    t = perf.request_traffic(perf.read_counters(rp), perf.read_counters(wp), "k_fused", nodes=100)
    assert t["dispatches"] == 2
    r2 = 32 * 10 + 64 * 300 + 128 * 74_000_000
    r3 = 32 * 10 + 64 * 340 + 128 * 74_100_000
    assert t["read_bytes"] == pytest.approx((r2 + r3) / 2)
    assert t["write_bytes"] == pytest.approx((64 * 112_000_000 + 32 * 7000 + 64 * 112_100_000 + 32 * 6000) / 2)
    assert t["probe_copy"]["copies"] == 3
    # the 64-byte reads of the probe copies are the only excess over 1 GiB
    assert t["probe_copy"]["read_ratio"] == pytest.approx(1 + 64 * 240 / (1 << 30))
    assert t["probe_copy"]["write_ratio"] == pytest.approx(1.0)
    assert t["read_per_node"] == pytest.approx(t["read_bytes"] / 100)


def test_request_traffic_needs_the_probe(tmp_path):
    rp, wp = sized(tmp_path)
    rows = [r for r in csv.DictReader(open(rp)) if r["Dispatch_Id"] not in ("4", "5", "6")]
    write_counters(rp, [(int(r["Dispatch_Id"]), r["Kernel_Name"], r["Counter_Name"], float(r["Counter_Value"]))
                        for r in rows])
    with pytest.raises(ValueError):
        perf.request_traffic(perf.read_counters(rp), perf.read_counters(wp), "k_fused")


def test_counters_of_one_grid(tmp_path):
    p = tmp_path / "c.csv"
    with open(p, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(COUNTER_COLS)
        for disp, grid, name, value in [(1, 1024, "k", 5.0), (2, 2048, "k", 7.0), (3, 64, "__amd_rocclr_copyBuffer", 1.0)]:
            w.writerow([disp, disp, "Agent 1", 1, 1, 1, grid, 1, name, 256, 0, 0, 32, 0, 32, "X", value, 0, 1])
    c = perf.read_counters(str(p), grid=2048)
    assert perf.per_dispatch(c, "k", "X") == [7.0]
    # the probe's copies are kept whatever their grid
    assert perf.per_dispatch(c, "copyBuffer", "X") == [1.0]
    assert perf.per_dispatch(perf.read_counters(str(p)), "k", "X") == [5.0, 7.0]
