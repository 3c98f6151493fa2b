"""Summaries of rocprofv3 result databases (rocpd SQLite) as short text tables.

    python -m quip_amd.prof_summary kernels <results.db> [--match SUBSTR]
    python -m quip_amd.prof_summary counters <results.db> [--match SUBSTR]

`kernels` prints calls, total and average duration per kernel (the --stats
view); `counters` prints the average value per dispatch of each collected PMC
counter (e.g. FETCH_SIZE / WRITE_SIZE in KiB) per kernel.
"""
import argparse
import re
import sqlite3


def short(name):
    name = re.sub(r"\(.*", "", name)
    return name.replace("void ", "").replace("gapx::", "")


def kernels(db, match=None):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(duration), avg(duration) from kernels group by name "
                       "order by sum(duration) desc").fetchall()
    tot = sum(r[2] for r in rows)
    out = ["%-58s %7s %12s %10s %6s" % ("kernel", "calls", "total_us", "avg_us", "pct")]
    for n, c, s, a in rows:
        if match and match not in n:
            continue
        out.append("%-58s %7d %12.1f %10.3f %6.2f" % (short(n)[:58], c, s / 1e3, a / 1e3, 100.0 * s / tot))
    return "\n".join(out)


def counters(db, match=None):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, count(*), avg(value), avg(duration) from counters_collection "
                       "group by kernel_name, counter_name order by kernel_name").fetchall()
    out = ["%-58s %-12s %7s %14s %10s" % ("kernel", "counter", "calls", "avg_value", "avg_us")]
    for n, cn, c, v, d in rows:
        if match and match not in n:
            continue
        out.append("%-58s %-12s %7d %14.1f %10.3f" % (short(n)[:58], cn, c, v, (d or 0) / 1e3))
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["kernels", "counters"])
    ap.add_argument("db")
    ap.add_argument("--match")
    a = ap.parse_args()
    print(kernels(a.db, a.match) if a.what == "kernels" else counters(a.db, a.match))


if __name__ == "__main__":
    main()
