"""The rocprofv3 summary helpers on small synthetic outputs (no GPU): the traffic report's copy-probe scaling and
the line bench.py reads back."""
import csv
import sqlite3

from libopensps_amd import profiling as P

FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size", "Kernel_Id",
          "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count",
          "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def _csv(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, FIELDS)
        w.writeheader()
        for i, (name, grid, val) in enumerate(rows, 1):
            w.writerow({k: 0 for k in FIELDS} | {"Dispatch_Id": i, "Kernel_Name": name, "Grid_Size": grid,
                                                   "Counter_Name": counter, "Counter_Value": val})


def test_traffic_report(tmp_path):
    k = "void k_chain<float, true>(ChainArgs)"
    c = "__amd_rocclr_copyBuffer"
    # a smaller copy of the same grid earlier in the run must not count as the probe
    _csv(tmp_path / "f.csv", "FETCH_SIZE", [(c, 131072, 1000.0), (k, 100, 5000.0), (k, 100, 5002.0),
                                            (c, 131072, 131072.0), (c, 131072, 131072.0)])
    _csv(tmp_path / "w.csv", "WRITE_SIZE", [(c, 131072, 2000.0), (k, 100, 7000.0), (k, 100, 7000.0),
                                            (c, 131072, 262144.0), (c, 131072, 262144.0)])
    t = P.traffic(tmp_path / "f.csv", tmp_path / "w.csv", "void k_chain<float, true>", 256 << 20)
    assert abs(t["fetch_factor"] - 2.0) < 1e-12 and abs(t["write_factor"] - 1.0) < 1e-12
    assert abs(t["bytes"] - (2.0 * 5001.0 + 7000.0) * 1024) < 1e-6
    out = tmp_path / "r.md"
    P.main(["traffic", "--fetch", str(tmp_path / "f.csv"), "--write", str(tmp_path / "w.csv"), "--kernel",
            "void k_chain<float, true>", "--out", str(out), "--command", "cmd"])
    kern, b = P.read_traffic(out)
    assert kern == "void k_chain<float, true>" and b == round((2.0 * 5001.0 + 7000.0) * 1024)
    assert P.read_traffic(tmp_path / "missing.md") is None


def test_stats_report(tmp_path):
    db = tmp_path / "r.db"
    con = sqlite3.connect(db)
    con.execute("create table top_kernels (name text, total_calls int, total_duration real, average real, "
                "percentage real)")
    con.executemany("insert into top_kernels values (?,?,?,?,?)", [("a", 2, 10.0, 5.0, 25.0), ("b", 1, 30.0, 30.0, 75.0)])
    con.commit()
    con.close()
    assert [r[0] for r in P.kernel_stats(db)] == ["b", "a"]
    out = tmp_path / "s.md"
    P.main(["stats", "--db", str(db), "--out", str(out), "--command", "cmd"])
    assert "| `b` | 1 | 30.0 | 30.0 | 75.0 |" in out.read_text()
