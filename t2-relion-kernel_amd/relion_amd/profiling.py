"""Summarise a rocprofv3 ``--stats --output-format csv`` kernel table
(``*_kernel_stats.csv``) into the markdown kept in this directory.

    python t2-relion-kernel_amd/relion_amd/profiling.py <run_kernel_stats.csv> [title]
"""
import csv
import sys


def main():
    rows = list(csv.DictReader(open(sys.argv[1])))
    title = sys.argv[2] if len(sys.argv) > 2 else sys.argv[1]
    print(f"# {title}\n")
    print("| kernel | calls | avg us | min us | max us | % time |")
    print("|---|---|---|---|---|---|")
    for r in rows:
        name = r["Name"]
        short = name if len(name) < 100 else name[:97] + "..."
        print(f"| `{short}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
              f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
              f"{float(r['Percentage']):.2f} |")


if __name__ == "__main__":
    main()
