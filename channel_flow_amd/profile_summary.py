"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) as markdown.

    python -m channel_flow_amd.profile_summary <run_results.db> [--title T] [--command C]

Prints per-kernel call counts, total and average durations, and the
register/LDS footprint of each kernel symbol.
"""
from __future__ import annotations

import argparse
import re
import sqlite3
import sys


def short_name(name: str) -> str:
    name = re.sub(r"\(.*\)$", "", name)
    name = name.replace("chf::", "")
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="rocprofv3 kernel trace")
    ap.add_argument("--command", default="")
    a = ap.parse_args(argv)
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
