"""Summaries of rocprofv3 output for the batched composite-population kernels.

Two reports, written as short Markdown files:

* ``stats``: the per-kernel table of a ``rocprofv3 --kernel-trace --stats`` run (its SQLite results file):
  calls, total and average duration; with ``--by-launch`` one row per kernel and launch shape (a kernel that
  serves several workloads of the run, such as the products, then shows each apart).
* ``counters``: SQ counters per kernel from several ``--pmc`` passes over the same program (CSV output): each
  counter summed over its rows per dispatch, averaged over the kernel's dispatches, and per wavefront; with the
  ratios the text quotes (wait and issue shares, MFMA busy share).
* ``traffic``: the HBM bytes per launch of one kernel from two counter runs, ``--pmc FETCH_SIZE`` and
  ``--pmc WRITE_SIZE`` (CSV output, KiB per dispatch). Each counter is scaled by what it reads for a plain
  device-to-device copy of known size in the same run (the last two ``__amd_rocclr_copyBuffer`` dispatches,
  which ``bench.py --copy-probe`` issues after everything else): factor = bytes the copy moves / bytes the counter reports.

The traffic report ends in a line ``traffic_bytes_per_launch: <int>`` that ``bench.py`` reads for the roofline's
``traffic`` field when the kernel matches.

    python -m libopensps_amd.profiling stats --db RUN_results.db --out profiles/X.md --command "..."
    python -m libopensps_amd.profiling traffic --fetch F.csv --write W.csv --kernel 'k_chain<float, true, false>' \\
        --copy-bytes 268435456 --out profiles/Y.md --command "..."
"""
from __future__ import annotations

import argparse
import csv
import sqlite3
import statistics

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def kernel_stats(db_path):
    """[(name, calls, total_us, avg_us, percent)] from the results database of a --kernel-trace --stats run."""
    con = sqlite3.connect(db_path)
    try:
        rows = con.execute("select name, total_calls, total_duration, average, percentage from top_kernels "
                           "order by total_duration desc").fetchall()
    finally:
        con.close()
    return rows


def kernel_stats_by_launch(db_path):
    """[(name, grid (x, y, z), workgroup, calls, total_us, avg_us, vgprs, lds bytes)] per kernel and launch shape
    (one workload a shape) from the dispatch records of a --kernel-trace run; durations in ns in the database."""
    con = sqlite3.connect(db_path)
    try:
        rows = con.execute("select name, grid_x, grid_y, grid_z, workgroup_x, count(*), sum(duration), avg(duration), "
                           "max(vgpr_count + accum_vgpr_count), max(lds_size) from kernels "
                           "group by name, grid_x, grid_y, grid_z, workgroup_x order by sum(duration) desc").fetchall()
    finally:
        con.close()
    return [(r[0], (r[1], r[2], r[3]), r[4], r[5], r[6] / 1e3, r[7] / 1e3, r[8], r[9]) for r in rows]


def pmc_values(csv_path, counter):
    """[(dispatch id, kernel name, grid size, value)] of one counter from a --pmc CSV."""
    out = []
    with open(csv_path, newline="") as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                out.append((int(r["Dispatch_Id"]), r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"])))
    return sorted(out)


def _copy_factor(vals, copy_bytes, nprobe=2):
    """bytes moved / counter bytes for the probe: the last nprobe copy dispatches of the run."""
    copies = [v for v in vals if v[1].startswith(COPY_KERNEL)]
    if len(copies) < nprobe:
        raise ValueError("no copy probe in the counter run (run bench.py with --copy-probe)")
    kib = [v[3] for v in copies[-nprobe:]]
    return copy_bytes / (1024.0 * statistics.median(kib)), len(kib)


def traffic(fetch_csv, write_csv, kernel, copy_bytes):
    """HBM bytes per launch of kernel (exact name prefix) from the FETCH_SIZE and WRITE_SIZE runs."""
    out = {"kernel": kernel, "copy_bytes": copy_bytes}
    for key, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        vals = pmc_values(path, counter)
        fac, ncopy = _copy_factor(vals, copy_bytes)
        kib = [v[3] for v in vals if v[1].startswith(kernel)]
        if not kib:
            raise ValueError(f"no dispatch of {kernel} in {path}")
        out[key + "_factor"] = fac
        out[key + "_copy_dispatches"] = ncopy
        out[key + "_kib_median"] = statistics.median(kib)
        out[key + "_dispatches"] = len(kib)
        out[key + "_bytes"] = fac * 1024.0 * statistics.median(kib)
    out["bytes"] = out["fetch_bytes"] + out["write_bytes"]
    return out


def counters(csv_paths, kernels):
    """{kernel: {counter: mean over dispatches of the per-dispatch sum}} for kernels whose name contains one of the
    given substrings, over the --pmc CSV files of several passes."""
    acc = {}
    for path in csv_paths:
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                name = r["Kernel_Name"]
                if not any(k in name for k in kernels):
                    continue
                d = acc.setdefault(name, {}).setdefault(r["Counter_Name"], {})
                key = (path, r["Dispatch_Id"])
                d[key] = d.get(key, 0.0) + float(r["Counter_Value"])
    return {k: {c: statistics.mean(v.values()) for c, v in cs.items()} for k, cs in acc.items()}


RATIOS = [("SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES", "SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES", 1.0),
          ("SQ_WAIT_ANY / SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES", 1.0),
          ("SQ_ACTIVE_INST_VALU / SQ_BUSY_CU_CYCLES", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CU_CYCLES", 1.0),
          ("SQ_VALU_MFMA_BUSY_CYCLES / (4 SQ_BUSY_CU_CYCLES)", "SQ_VALU_MFMA_BUSY_CYCLES", "SQ_BUSY_CU_CYCLES", 0.25)]


def _write_counters(args):
    res = counters(args.csv, args.kernel)
    lines = [f"# {args.title}", "", "Commands (one pass a counter group):", "", "```", args.command, "```", "",
             "Each counter: summed over its rows of a dispatch, averaged over the kernel's dispatches; per wavefront: "
             "divided by SQ_WAVES.", ""]
    for name in sorted(res, key=lambda k: -res[k].get("SQ_WAVES", 0)):
        c = res[name]
        waves = c.get("SQ_WAVES", 0.0)
        lines += [f"## `{name[:100]}`", "", "| counter | per dispatch | per wavefront |", "|---|---|---|"]
        for cn in sorted(c):
            pw = f"{c[cn] / waves:.1f}" if waves else ""
            lines.append(f"| {cn} | {c[cn]:.4g} | {pw} |")
        lines += ["", "| ratio | value |", "|---|---|"]
        for label, a, b, f in RATIOS:
            if a in c and b in c and c[b] > 0:
                lines.append(f"| {label} | {f * c[a] / c[b]:.3f} |")
        lines.append("")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")


def _write_stats(args):
    lines = [f"# {args.title}", "", "Command:", "", "```", args.command, "```", ""]
    if args.by_launch:
        rows = kernel_stats_by_launch(args.db)
        grand = sum(r[4] for r in rows) or 1.0
        lines += ["Per kernel and launch shape (grid in work-items; each shape is one workload of the run).", "",
                  "| kernel | grid | workgroup | calls | total (us) | average (us) | % | VGPRs | LDS (B) |",
                  "|---|---|---|---|---|---|---|---|---|"]
        for name, grid, wg, calls, tot, avg, vg, lds in rows[:args.top]:
            g = "x".join(str(v) for v in grid if v != 1) or "1"
            lines.append(f"| `{name[:80]}` | {g} | {wg} | {calls} | {tot:.1f} | {avg:.1f} | {100 * tot / grand:.1f} | "
                         f"{vg} | {lds} |")
    else:
        lines += ["| kernel | calls | total (us) | average (us) | % |", "|---|---|---|---|---|"]
        for name, calls, tot, avg, pct in kernel_stats(args.db)[:args.top]:
            lines.append(f"| `{name[:90]}` | {calls} | {tot:.1f} | {avg:.1f} | {pct:.1f} |")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")


def _write_traffic(args):
    t = traffic(args.fetch, args.write, args.kernel, args.copy_bytes)
    lines = [f"# {args.title}", "", "Commands (one counter per run):", "", "```", args.command, "```", "",
             f"Kernel `{t['kernel']}`: median over {t['fetch_dispatches']} dispatches.", "",
             "| counter | KiB per dispatch | copy-probe factor | bytes per launch |", "|---|---|---|---|",
             f"| FETCH_SIZE | {t['fetch_kib_median']:.1f} | {t['fetch_factor']:.4f} | {t['fetch_bytes']:.4e} |",
             f"| WRITE_SIZE | {t['write_kib_median']:.1f} | {t['write_factor']:.4f} | {t['write_bytes']:.4e} |", "",
             f"Copy probe: a {args.copy_bytes} B device-to-device copy (`{COPY_KERNEL}`); factor = bytes moved / "
             "counter bytes.", "",
             f"traffic_kernel: {t['kernel']}",
             f"traffic_bytes_per_launch: {int(round(t['bytes']))}"]
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_traffic(path):
    """(kernel, bytes per launch) from a traffic report, or None."""
    try:
        with open(path) as f:
            kv = dict(line.strip().split(": ", 1) for line in f if line.startswith("traffic_"))
        return kv["traffic_kernel"], int(kv["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("--db", required=True)
    s.add_argument("--out", required=True)
    s.add_argument("--title", default="rocprofv3 kernel statistics")
    s.add_argument("--command", required=True)
    s.add_argument("--top", type=int, default=12)
    s.add_argument("--by-launch", action="store_true", help="one row per kernel and launch shape")
    q = sub.add_parser("counters")
    q.add_argument("--csv", nargs="+", required=True, help="counter_collection.csv of each pass")
    q.add_argument("--kernel", nargs="+", required=True, help="substrings of the kernel names to report")
    q.add_argument("--out", required=True)
    q.add_argument("--title", default="SQ counters per kernel")
    q.add_argument("--command", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, default=256 << 20)
    t.add_argument("--out", required=True)
    t.add_argument("--title", default="HBM traffic per launch")
    t.add_argument("--command", required=True)
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        _write_stats(args)
    elif args.cmd == "counters":
        _write_counters(args)
    else:
        _write_traffic(args)


if __name__ == "__main__":
    main()
