"""The rocprofv3 summaries (kites_amd/profiling.py) on small synthetic CSV
files laid out as rocprofv3 writes them: the roofline traffic scaling by the
copy probe, and the per-iteration counter table."""
import csv
import json
import os

from kites_amd import profiling as P

STEP = "void kites::detail::step_kernel<true, true, false, P>(P)"
COPY = "__amd_rocclr_copyBuffer"
FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id",
          "Grid_Size", "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size",
          "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count", "SGPR_Count", "Counter_Name",
          "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def _counters(path, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, FIELDS)
        w.writeheader()
        for disp, name, counter, value in rows:
            w.writerow({k: 0 for k in FIELDS} | {"Dispatch_Id": disp, "Kernel_Name": name,
                                                 "Counter_Name": counter, "Counter_Value": value})


def _trace(path, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for disp, name, start, end in rows:
            w.writerow({"Dispatch_Id": disp, "Kernel_Name": name, "Start_Timestamp": start,
                        "End_Timestamp": end})


def test_counter_table(tmp_path):
    # two passes over the same program: 3 step_kernel dispatches and a copy
    p1 = tmp_path / "p1" / "run" / "a_counter_collection.csv"
    p2 = tmp_path / "p2" / "run" / "b_counter_collection.csv"
    _counters(str(p1), [(1, STEP, "SQ_WAVE_CYCLES", 1000), (1, STEP, "SQ_WAIT_ANY", 900),
                        (2, COPY, "SQ_WAVE_CYCLES", 5), (3, STEP, "SQ_WAVE_CYCLES", 400),
                        (3, STEP, "SQ_WAIT_ANY", 100), (4, STEP, "SQ_WAVE_CYCLES", 10),
                        (4, STEP, "SQ_WAIT_ANY", 0)])
    _counters(str(p2), [(7, STEP, "SQ_INSTS_VALU", 11), (9, STEP, "SQ_INSTS_VALU", 22),
                        (12, STEP, "SQ_INSTS_VALU", 33)])
    kt = tmp_path / "kt" / "k_kernel_trace.csv"
    _trace(str(kt), [(1, STEP, 1000, 3000), (2, COPY, 3000, 4000), (3, STEP, 5000, 5500),
                     (4, STEP, 6000, 6100)])
    out = tmp_path / "c.md"
    P.main(["counters", str(out), str(tmp_path / "p1"), str(tmp_path / "p2"), "--ktrace",
            str(tmp_path / "kt"), "--labels", "a", "--keep", "0,2", "--command", "cmd one"])
    text = out.read_text()
    assert "cmd one" in text
    rows = [r for r in text.splitlines() if r.startswith("| ") and r[2].isdigit()]
    assert len(rows) == 2                       # dispatches 0 and 2 kept
    first = [c.strip() for c in rows[0].strip("|").split("|")]
    last = [c.strip() for c in rows[1].strip("|").split("|")]
    # dispatch, run, us, WAVE_CYCLES, WAIT_ANY, INSTS_VALU, wait share
    assert first == ["0", "a", "2.0", "1e+03", "900", "11", "0.90"]
    assert last == ["2", "a", "0.1", "10", "0", "33", "0.00"]


def _request_runs(tmp_path, copy_bubble):
    """Read- and write-request passes over a program with a copy probe of
    4 MiB each way and two step_kernel dispatches."""
    mib = 4
    rd = tmp_path / "rd" / "r_counter_collection.csv"
    wr = tmp_path / "wr" / "w_counter_collection.csv"
    lines = mib * (1 << 20) // 128          # the copy's 128-byte read requests
    _counters(str(rd), [(1, COPY, "TCC_EA0_RDREQ_sum", lines), (1, COPY, "TCC_EA0_RDREQ_32B_sum", 0),
                        (1, COPY, "TCC_BUBBLE_sum", lines if copy_bubble else 0),
                        (2, STEP, "TCC_EA0_RDREQ_sum", 100), (2, STEP, "TCC_EA0_RDREQ_32B_sum", 40),
                        (2, STEP, "TCC_BUBBLE_sum", 10),
                        (3, STEP, "TCC_EA0_RDREQ_sum", 300), (3, STEP, "TCC_EA0_RDREQ_32B_sum", 60),
                        (3, STEP, "TCC_BUBBLE_sum", 30)])
    _counters(str(wr), [(1, COPY, "TCC_EA0_WRREQ_sum", mib * (1 << 20) // 64),
                        (1, COPY, "TCC_EA0_WRREQ_64B_sum", mib * (1 << 20) // 64),
                        (2, STEP, "TCC_EA0_WRREQ_sum", 10), (2, STEP, "TCC_EA0_WRREQ_64B_sum", 4),
                        (3, STEP, "TCC_EA0_WRREQ_sum", 30), (3, STEP, "TCC_EA0_WRREQ_64B_sum", 6)])
    stats = tmp_path / "kt" / "bench_kernel_stats.csv"
    stats.parent.mkdir()
    stats.write_text("Name,Calls,TotalDurationNs,AverageNs,Percentage\n"
                     f'"{STEP}",2,4000,2000,50.0\n')
    bench = tmp_path / "bench.json"
    bench.write_text(json.dumps({
        "config": {"workload": "w"},
        "roofline": {"kernel": "step_kernel", "bytes_per_launch": 10000, "launch_ms": 0.002,
                     "achieved": 500.0, "peak": 8000.0, "frac": 0.0625}}) + "\n")
    out = tmp_path / "out"
    out.mkdir()
    P.main([str(tmp_path / "kt"), str(tmp_path / "rd"), str(tmp_path / "wr"), str(bench),
            str(out), "--copy-mib", str(mib), "--command", "cmd a"])
    return json.loads((out / "traffic.json").read_text()), out


def test_roofline_traffic_by_request_size(tmp_path):
    """Read bytes follow the request sizes (32 B counted apart, 128 B from
    TCC_BUBBLE, the rest 64 B), write bytes 32 or 64 B per request; the
    copy probe's counters resolve to the bytes it moved."""
    t, out = _request_runs(tmp_path, copy_bubble=True)
    assert t["non_32B_reads_taken_as_128B"] is False
    assert t["copy_probe"]["read_bytes"] == t["copy_probe"]["known_bytes"] == 4 << 20
    assert t["copy_probe"]["write_bytes"] == 4 << 20
    k = t["kernels"]["step_kernel"]
    # per dispatch: reads (200 all, 50 of 32 B, 20 of 128 B), writes (20 all, 5 of 64 B)
    rd = 32 * 50 + 128 * 20 + 64 * (200 - 50 - 20)
    wr = 32 * 15 + 64 * 5
    assert k["read_bytes_per_launch"] == rd and k["write_bytes_per_launch"] == wr
    assert k["traffic_bytes_per_launch"] == rd + wr
    assert k["algorithmic_bytes_per_launch"] == 10000 and k["avg_ns_kernel_trace"] == 2000
    summary = (out / "step_kernel_rmat22_sssp1.md").read_text()
    assert "cmd a" in summary and f"| traffic / algorithmic | {(rd + wr) / 10000:.2f} |" in summary


def test_roofline_traffic_when_the_128B_counter_misses(tmp_path):
    """A copy probe whose 128-byte reads the 128-byte counter does not see
    switches every read request that is not 32 B to 128 B (upper bound)."""
    t, _ = _request_runs(tmp_path, copy_bubble=False)
    assert t["non_32B_reads_taken_as_128B"] is True
    assert t["copy_probe"]["read_bytes_if_128B"] == 4 << 20
    k = t["kernels"]["step_kernel"]
    assert k["read_bytes_per_launch"] == 32 * 50 + 128 * 150
