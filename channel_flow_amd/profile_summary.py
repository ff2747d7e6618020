"""Summarise rocprofv3 output as markdown.

    python -m channel_flow_amd.profile_summary <run_results.db> [--title T] [--command C]
    python -m channel_flow_amd.profile_summary --counters <fetch.csv> <write.csv>
        --calib-fetch <copy_fetch.csv> --calib-write <copy_write.csv> --calib-bytes B
        [--json out.json]

The first form reads a kernel-trace database (rocpd SQLite) and prints
per-kernel call counts, total and average durations, and the register/LDS
footprint of each kernel symbol.

The second form reads two `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` counter CSVs
(KiB per dispatch) of the program under study and the same two counters of a
plain device-to-device copy of B bytes.  Each counter is scaled by
B / (counter of the copy), so that the copy reads and writes exactly B bytes,
and the per-kernel average HBM bytes per dispatch are printed (and written as
JSON when --json is given).
"""
from __future__ import annotations

import argparse
import re
import sqlite3
import sys


def short_name(name: str) -> str:
    name = re.sub(r"\(.*\)$", "", name)
    name = name.replace("chf::", "")
    # HIP_vector_type<double, 2u> -> double2, so template arguments fit
    name = re.sub(r"HIP_vector_type<(\w+), (\d)u>", r"\g<1>\g<2>", name)
    return name[:90]


def summarise(db: str):
    con = sqlite3.connect(db)
    cur = con.cursor()
    rows = cur.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration), "
        "max(lds_size), max(vgpr_count), max(accum_vgpr_count), max(sgpr_count), "
        "max(grid_x), max(grid_y), max(grid_z), max(workgroup_x) "
        "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    return rows, total


def read_counter(csv_path: str, counter: str):
    """Per-kernel list of `counter` values (one per dispatch) from a counter CSV."""
    import csv
    per = {}
    with open(csv_path, newline="") as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            per.setdefault(short_name(row["Kernel_Name"]), []).append(float(row["Counter_Value"]))
    return per


def copy_factor(csv_path: str, counter: str, nbytes: float) -> float:
    """Bytes per counter unit, from the dispatch with the largest count (the copy)."""
    per = read_counter(csv_path, counter)
    top = max(max(v) for v in per.values())
    return nbytes / top


def counters_main(a) -> int:
    import json
    ff = copy_factor(a.calib_fetch, "FETCH_SIZE", a.calib_bytes)
    fw = copy_factor(a.calib_write, "WRITE_SIZE", a.calib_bytes)
    fetch = read_counter(a.counters[0], "FETCH_SIZE")
    write = read_counter(a.counters[1], "WRITE_SIZE")
    out = [f"# {a.title}", ""]
    if a.command:
        out += ["Commands:", "", "```", a.command, "```", ""]
    out += [f"Calibration: a {a.calib_bytes / 2**20:.0f} MiB device copy counts FETCH_SIZE = "
            f"{a.calib_bytes / ff:.0f} KiB and WRITE_SIZE = {a.calib_bytes / fw:.0f} KiB; "
            f"scale factors {ff / 1024:.4f} (fetch) and {fw / 1024:.4f} (write) bytes per counted byte.",
            "The fetch factor is exact for kernels whose reads are full 128-byte lines like the "
            "copy's; kernels with narrower reads are over-estimated by up to that factor.", "",
            "| kernel | dispatches | fetch MB/dispatch | write MB/dispatch | total MB/dispatch |",
            "|---|---|---|---|---|"]
    res = {}
    for k in sorted(set(fetch) | set(write), key=lambda k: -sum(fetch.get(k, [0]))):
        fv, wv = fetch.get(k, []), write.get(k, [])
        fb = ff * sum(fv) / len(fv) if fv else None
        wb = fw * sum(wv) / len(wv) if wv else None
        res[k] = dict(fetch_bytes=fb, write_bytes=wb, dispatches=max(len(fv), len(wv)))
        tot = (fb or 0.0) + (wb or 0.0)
        out.append(f"| `{k}` | {max(len(fv), len(wv))} | {(fb or 0) / 1e6:.1f} | {(wb or 0) / 1e6:.1f} | "
                   f"{tot / 1e6:.1f} |")
    print("\n".join(out))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(dict(fetch_factor=ff / 1024, write_factor=fw / 1024, calib_bytes=a.calib_bytes,
                           kernels=res), f, indent=1)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--counters", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    ap.add_argument("--calib-fetch")
    ap.add_argument("--calib-write")
    ap.add_argument("--calib-bytes", type=float)
    ap.add_argument("--json")
    ap.add_argument("db", nargs="?")
    ap.add_argument("--title", default="rocprofv3 kernel trace")
    ap.add_argument("--command", default="")
    a = ap.parse_args(argv)
    if a.counters:
        return counters_main(a)
    rows, total = summarise(a.db)
    out = [f"# {a.title}", ""]
    if a.command:
        out += ["Command:", "", "```", a.command, "```", ""]
    out += ["| kernel | calls | total ms | avg ms | min ms | max ms | % | LDS B | VGPR | AGPR | SGPR | grid (threads) | wg |",
            "|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for (name, n, tot, avg, mn, mx, lds, vg, ag, sg, gx, gy, gz, wg) in rows:
        out.append(f"| `{short_name(name)}` | {n} | {tot / 1e6:.3f} | {avg / 1e6:.4f} | {mn / 1e6:.4f} | "
                   f"{mx / 1e6:.4f} | {100.0 * tot / total:.1f} | {lds} | {vg} | {ag} | {sg} | "
                   f"{gx}x{gy}x{gz} | {wg} |")
    out.append("")
    out.append(f"Total kernel time: {total / 1e6:.3f} ms")
    print("\n".join(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
