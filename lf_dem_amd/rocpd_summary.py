"""Per-kernel summary (calls, total, average, share) of a rocprofv3 SQLite
result (``rocprofv3 --kernel-trace -d DIR -o NAME -- cmd`` writes
DIR/NAME_results.db).  Used to produce the text summaries kept in profiles/
when rocprofv3's own CSV stats are not written (it can stop with a fault at
process exit after cooperative launches, with the database already complete).

usage: python -m lf_dem_amd.rocpd_summary RESULTS.db [--top N]
       python -m lf_dem_amd.rocpd_summary --pmc COUNTERS.csv --kernel REGEX
(the second form: mean / min / max of each counter over the dispatches of the
kernels matching REGEX in a ``rocprofv3 --pmc ... -f csv`` counter file)
"""
import argparse
import csv
import re
import sqlite3


def summarize(path, top=30):
    con = sqlite3.connect(path)
    rows = con.execute("SELECT name, COUNT(*), SUM(end - start), AVG(end - start), MIN(end - start), "
                       "MAX(end - start) FROM kernels GROUP BY name ORDER BY SUM(end - start) DESC").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = ["%-72s %7s %12s %12s %12s %12s %7s" % ("kernel", "calls", "total_us", "avg_us", "min_us", "max_us",
                                                  "pct")]
    for name, calls, tot, avg, mn, mx in rows[:top]:
        out.append("%-72s %7d %12.1f %12.2f %12.2f %12.2f %6.2f%%" % (name[:72], calls, tot / 1e3, avg / 1e3,
                                                                       mn / 1e3, mx / 1e3, 100.0 * tot / total))
    out.append("total kernel time: %.1f us over %d dispatches" % (total / 1e3, sum(r[1] for r in rows)))
    return "\n".join(out)


def summarize_pmc(path, kernel):
    """Per counter: dispatches, mean, min, max over the kernels matching the regex."""
    pat = re.compile(kernel)
    vals = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            name = row.get("Kernel_Name") or row.get("Kernel-Name") or ""
            if not pat.search(name):
                continue
            key = (row.get("Dispatch_Id"), row.get("Counter_Name"))
            vals.setdefault(row.get("Counter_Name"), {}).setdefault(key, 0.0)
            vals[row.get("Counter_Name")][key] += float(row.get("Counter_Value") or 0)
    out = []
    for cname, per in sorted(vals.items()):
        v = list(per.values())
        out.append("%s  %s  %d dispatches  mean %.1f  min %.1f  max %.1f" % (
            kernel, cname, len(v), sum(v) / len(v), min(v), max(v)))
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db", nargs="?")
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--pmc")
    ap.add_argument("--kernel", default=".")
    a = ap.parse_args()
    if a.pmc:
        print(summarize_pmc(a.pmc, a.kernel))
    else:
        print(summarize(a.db, a.top))


if __name__ == "__main__":
    main()
