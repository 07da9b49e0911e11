"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) as a text
table: per kernel, dispatch count, total / average / min / max duration.

usage: python rocpd_summary.py <results.db> [title]
"""
import sqlite3
import sys


def main():
    db = sys.argv[1]
    title = sys.argv[2] if len(sys.argv) > 2 else db
    c = sqlite3.connect(db)
    rows = list(c.execute(
        "select name, count(*), sum(end-start), avg(end-start), min(end-start), max(end-start) "
        "from kernels group by name order by sum(end-start) desc"))
    total = sum(r[2] for r in rows) or 1
    print(f"# {title}")
    print(f"{'kernel':64s} {'calls':>6s} {'total ms':>10s} {'avg us':>10s} {'min us':>10s} {'max us':>10s} {'%':>6s}")
    for name, n, tot, avg, mn, mx in rows:
        print(f"{name[:64]:64s} {n:6d} {tot / 1e6:10.3f} {avg / 1e3:10.1f} {mn / 1e3:10.1f} {mx / 1e3:10.1f} "
              f"{100.0 * tot / total:6.1f}")


if __name__ == "__main__":
    main()
