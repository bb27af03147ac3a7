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


def test_roofline_traffic_scaling(tmp_path):
    """FETCH_SIZE and WRITE_SIZE per dispatch are scaled by the copy probe's
    known bytes over what the counter read for it."""
    mib = 4
    kib = mib * 1024
    fetch = tmp_path / "fetch" / "f_counter_collection.csv"
    write = tmp_path / "write" / "w_counter_collection.csv"
    stats = tmp_path / "kt" / "bench_kernel_stats.csv"
    stats.parent.mkdir()
    stats.write_text("Name,Calls,TotalDurationNs,AverageNs,Percentage\n"
                     f'"{STEP}",2,4000,2000,50.0\n')
    _counters(str(fetch), [(1, COPY, "FETCH_SIZE", kib / 2), (2, STEP, "FETCH_SIZE", 100),
                           (3, STEP, "FETCH_SIZE", 300)])
    _counters(str(write), [(1, COPY, "WRITE_SIZE", kib), (2, STEP, "WRITE_SIZE", 10),
                           (3, STEP, "WRITE_SIZE", 30)])
    bench = tmp_path / "bench.json"
    bench.write_text(json.dumps({
        "config": {"workload": "w"},
        "roofline": {"kernel": "step_kernel", "bytes_per_launch": 1000, "launch_ms": 0.002,
                     "achieved": 500.0, "peak": 8000.0, "frac": 0.0625}}) + "\n")
    out = tmp_path / "out"
    out.mkdir()
    P.main([str(tmp_path / "kt"), str(tmp_path / "fetch"), str(tmp_path / "write"), str(bench),
            str(out), "--copy-mib", str(mib)])
    t = json.loads((out / "traffic.json").read_text())
    assert abs(t["fetch_factor"] - 2.0) < 1e-9 and abs(t["write_factor"] - 1.0) < 1e-9
    k = t["kernels"]["step_kernel"]
    # mean of (100, 300) KiB x 2 + mean of (10, 30) KiB x 1, in bytes
    assert k["traffic_bytes_per_launch"] == int((200 * 2 + 20) * 1024)
    assert k["algorithmic_bytes_per_launch"] == 1000 and k["avg_ns_kernel_trace"] == 2000
    summary = (out / "step_kernel_rmat22_sssp1.md").read_text()
    assert "| traffic / algorithmic | 430.08 |" in summary
