"""Per-kernel summary (calls, total, average, share) of a rocprofv3 SQLite
result (``rocprofv3 --kernel-trace -d DIR -o NAME -- cmd`` writes
DIR/NAME_results.db).  Used to produce the text summaries kept in profiles/
when rocprofv3's own CSV stats are not written (it can stop with a fault at
process exit after cooperative launches, with the database already complete).

usage: python -m lf_dem_amd.rocpd_summary RESULTS.db [--top N]
"""
import argparse
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--top", type=int, default=30)
    a = ap.parse_args()
    print(summarize(a.db, a.top))


if __name__ == "__main__":
    main()
