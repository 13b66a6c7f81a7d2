"""Per-kernel summary of a rocprofv3 kernel trace, split by launch shape.

    python3 -m enjaparticles_amd.rocprof_summary RESULTS.db [--top N] [--match REGEX] [--region NAME]
    python3 -m enjaparticles_amd.rocprof_summary COUNTERS.csv [--match REGEX] [--skip N]

rocprofv3 --stats averages every launch of a kernel name together; a run that
steps systems of different sizes (the plain 1M-particle step, the batched
scenes' sequential comparison, the adjoint's 65,536-particle rollout ...)
launches the same kernel with different grids.  This prints, per kernel name
and grid size, the launch count and the average / min / max duration (us),
read from the rocpd SQLite database that rocprofv3 writes (-d DIR -o NAME
gives DIR/NAME_results.db).  Given the counter_collection.csv of a --pmc run
(--output-format csv), it prints the mean of each counter per kernel and
launch shape instead.

--region NAME keeps only the dispatches that ran inside the marker ranges
called NAME (a run under rocprofv3 --marker-trace; bench.py marks its timed
region "timed"), so the averages are those of the timed steps alone.
--skip N drops the first N dispatches of every kernel and launch shape of a
counter run (the warm-up steps before a timed region).
"""
from __future__ import annotations

import argparse
import csv
import json
import re
import sqlite3
from collections import defaultdict


def short_name(name: str) -> str:
    """Kernel name without its argument list (anonymous namespaces dropped)."""
    return name.replace("(anonymous namespace)::", "").split("(")[0]


def regions(db_path: str, name: str):
    """(start, end) of every marker range called `name`: rocprofv3 records a
    roctx range under the API's name with the message in its extdata."""
    db = sqlite3.connect(db_path)
    out = []
    for start, end, rname, extdata in db.execute("select start, end, name, extdata from regions order by start"):
        message = None
        try:
            message = json.loads(extdata).get("message") if extdata else None
        except (ValueError, AttributeError):
            pass
        if rname == name or message == name:
            out.append((start, end))
    return out


def summarize(db_path: str, top: int = 40, match: str | None = None, region: str | None = None):
    db = sqlite3.connect(db_path)
    cur = db.cursor()
    where = ""
    if region:
        spans = regions(db_path, region)
        if not spans:
            raise SystemExit(f"no marker range called {region!r} in {db_path}")
        where = "where " + " or ".join(f"(d.start >= {a} and d.end <= {b})" for a, b in spans)
    rows = cur.execute(
        "select s.display_name, d.grid_size_x, d.workgroup_size_x, count(*), avg(d.end - d.start), "
        "min(d.end - d.start), max(d.end - d.start), sum(d.end - d.start) "
        "from rocpd_kernel_dispatch d join rocpd_info_kernel_symbol s on d.kernel_id = s.id " + where +
        " group by s.display_name, d.grid_size_x, d.workgroup_size_x order by sum(d.end - d.start) desc").fetchall()
    total = sum(r[7] for r in rows) or 1
    out = []
    for name, grid, wg, n, avg, mn, mx, tot in rows:
        short = short_name(name)
        if match and not re.search(match, short):
            continue
        out.append((short, grid, wg, n, avg / 1e3, mn / 1e3, mx / 1e3, 100.0 * tot / total))
    return out[:top]


def counters(csv_path: str, match: str | None = None, skip: int = 0):
    acc = defaultdict(lambda: [0.0, 0])
    rows = defaultdict(list)
    with open(csv_path, newline="") as f:
        for row in csv.DictReader(f):
            name = short_name(row["Kernel_Name"])
            if match and not re.search(match, name):
                continue
            key = (name, int(row["Grid_Size"]), int(row["Workgroup_Size"]), row["Counter_Name"])
            rows[key].append((int(row.get("Dispatch_Id", 0) or 0), float(row["Counter_Value"])))
    for key, vals in rows.items():
        vals.sort()
        for _, v in vals[skip:]:
            a = acc[key]
            a[0] += v
            a[1] += 1
    return sorted(((k[0], k[1], k[2], k[3], v[1], v[0] / v[1]) for k, v in acc.items()),
                  key=lambda r: (r[0], r[1], r[2], r[3]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--top", type=int, default=40)
    ap.add_argument("--match", default=None)
    ap.add_argument("--region", default=None, help="only dispatches inside the marker ranges of this name")
    ap.add_argument("--skip", type=int, default=0, help="counters: drop the first N dispatches per kernel and shape")
    a = ap.parse_args()
    if a.db.endswith(".csv"):
        print(f"{'kernel':58s} {'grid':>9s} {'wg':>4s} {'counter':>28s} {'calls':>6s} {'mean':>16s}")
        for r in counters(a.db, a.match, a.skip):
            print(f"{r[0][:58]:58s} {r[1]:9d} {r[2]:4d} {r[3]:>28s} {r[4]:6d} {r[5]:16.2f}")
        return
    print(f"{'kernel':58s} {'grid':>9s} {'wg':>4s} {'calls':>7s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s} {'%':>6s}")
    for r in summarize(a.db, a.top, a.match, a.region):
        print(f"{r[0][:58]:58s} {r[1]:9d} {r[2]:4d} {r[3]:7d} {r[4]:10.2f} {r[5]:10.2f} {r[6]:10.2f} {r[7]:6.2f}")


if __name__ == "__main__":
    main()
