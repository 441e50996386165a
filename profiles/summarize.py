"""Summarise a rocprofv3 output database (``*_results.db``) into the
per-kernel table kept in this directory: calls, total and average
duration (microseconds) and share of GPU time.

    python profiles/summarize.py <run_results.db> [title]
"""
import sqlite3
import sys


def main():
    db = sqlite3.connect(sys.argv[1])
    title = sys.argv[2] if len(sys.argv) > 2 else sys.argv[1]
    rows = db.execute("select name, total_calls, total_duration, average, percentage "
                      "from top_kernels").fetchall()
    print(f"# {title}\n")
    print("| kernel | calls | total us | avg us | % time |")
    print("|---|---|---|---|---|")
    for name, calls, tot, avg, pct in rows:
        short = name if len(name) < 110 else name[:107] + "..."
        print(f"| `{short}` | {calls} | {tot:.1f} | {avg:.1f} | {pct:.2f} |")


if __name__ == "__main__":
    main()
