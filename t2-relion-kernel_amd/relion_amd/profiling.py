"""Summaries of rocprofv3 CSV output as the short markdown kept with the
sources (kernel timings, HBM traffic per launch).

    python -m relion_amd.profiling stats   <run_kernel_stats.csv> [title]
    python -m relion_amd.profiling trace   <run_kernel_trace.csv> <kernel substring> [title]
    python -m relion_amd.profiling traffic <fetch.csv> <write.csv> <copy_fetch.csv>
                                           <copy_write.csv> <kernel substring> <copy bytes> [title]
    python -m relion_amd.profiling pmc     <counter_collection.csv> [...] <kernel substring>
                                           [last N dispatches] [title]

``trace`` groups the dispatches of the matching kernels by launch grid, so a
workload's launches can be told apart from other sizes of the same kernel.
``traffic`` scales each counter by what it reads for a plain device copy of
``copy bytes`` bytes (FETCH_SIZE counts the copy's reads, WRITE_SIZE its
writes), and writes one machine-readable line
``hbm_bytes_per_launch: <int>`` for the benchmark.
``pmc`` reads counter CSVs (one pass each) and reports, per matching kernel,
averages over its last N dispatches: the MFMA busy fraction
sum(SQ_VALU_MFMA_BUSY_CYCLES) / (GPU cycles x SIMDs) (rocprofv3's own
definition), the GPU clock implied by GRBM_GUI_ACTIVE (summed over the 8
XCDs) over the dispatch time, cycles per MFMA, the LDS bank-conflict share
of LDS-active cycles and the share of wave time spent in s_waitcnt.
"""
import collections
import csv
import sys

TRAFFIC_KEY = "hbm_bytes_per_launch:"


def _short(name, n=100):
    return name if len(name) < n else name[:n - 3] + "..."


def stats(path, title=None):
    rows = list(csv.DictReader(open(path)))
    out = [f"# {title or path}\n", "| kernel | calls | avg us | min us | max us | % time |",
           "|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.2f} |")
    return "\n".join(out)


def trace_groups(path, match):
    """{(kernel, grid threads, block threads): [duration ns, ...]}"""
    groups = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if match not in r["Kernel_Name"]:
            continue
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        block = int(r["Workgroup_Size_X"]) * int(r["Workgroup_Size_Y"]) * int(
            r["Workgroup_Size_Z"])
        groups[(r["Kernel_Name"], grid, block)].append(
            int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return groups


def trace(path, match, title=None):
    g = trace_groups(path, match)
    out = [f"# {title or path}\n",
           "| kernel | workgroups x threads | calls | avg us | min us | max us |",
           "|---|---|---|---|---|---|"]
    for (name, grid, block), d in sorted(g.items(), key=lambda kv: -len(kv[1])):
        out.append(f"| `{_short(name)}` | {grid // block} x {block} | {len(d)} | "
                   f"{sum(d) / len(d) / 1e3:.1f} | {min(d) / 1e3:.1f} | {max(d) / 1e3:.1f} |")
    return "\n".join(out)


def _per_dispatch(path, match):
    per = collections.defaultdict(float)
    names = {}
    for r in csv.DictReader(open(path)):
        if match in r["Kernel_Name"]:
            per[r["Dispatch_Id"]] += float(r["Counter_Value"])
            names[r["Dispatch_Id"]] = r["Kernel_Name"]
    return per, names


def traffic(fetch, write, copy_fetch, copy_write, match, copy_bytes, title=None):
    kib = 1024.0
    f, names = _per_dispatch(fetch, match)
    w, _ = _per_dispatch(write, match)
    cf, _ = _per_dispatch(copy_fetch, "copyBuffer")
    cw, _ = _per_dispatch(copy_write, "copyBuffer")
    f_scale = copy_bytes / (max(cf.values()) * kib)
    w_scale = copy_bytes / (max(cw.values()) * kib)
    fb = sum(f.values()) / len(f) * kib * f_scale
    wb = sum(w.values()) / len(w) * kib * w_scale
    out = [f"# {title or 'HBM traffic'}\n",
           f"kernel: `{_short(next(iter(names.values())), 200)}`, {len(f)} dispatch(es)\n",
           "| counter | KiB per dispatch (raw) | copy calibration (true / counted) | bytes per launch |",
           "|---|---|---|---|",
           f"| FETCH_SIZE | {sum(f.values()) / len(f):.4g} | {f_scale:.3f} | {fb:.4g} |",
           f"| WRITE_SIZE | {sum(w.values()) / len(w):.4g} | {w_scale:.3f} | {wb:.4g} |",
           "",
           f"calibration: `hipMemcpy`-style device copy of {copy_bytes} bytes "
           f"(FETCH_SIZE read {max(cf.values()):.4g} KiB, WRITE_SIZE {max(cw.values()):.4g} KiB)",
           "",
           f"{TRAFFIC_KEY} {int(round(fb + wb))}"]
    return "\n".join(out)


def _pmc_rows(paths, match):
    """{(kernel, dispatch): {counter: value, 'ns': duration}} over several
    counter-collection CSVs (dispatch ids differ per pass, so each pass's
    dispatches of a kernel are matched by their order)."""
    per_pass = []
    for path in paths:
        d = collections.OrderedDict()
        for r in csv.DictReader(open(path)):
            if match not in r["Kernel_Name"]:
                continue
            key = (r["Kernel_Name"], r["Dispatch_Id"])
            e = d.setdefault(key, {"ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"])})
            e[r["Counter_Name"]] = e.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        by_kernel = collections.defaultdict(list)
        for (name, _), e in d.items():
            by_kernel[name].append(e)
        per_pass.append(by_kernel)
    return per_pass


def pmc(paths, match, last=3, title=None, simds=1024, xcds=8):
    passes = _pmc_rows(paths, match)
    names = sorted({n for p in passes for n in p})
    out = [f"# {title or 'counters'}\n",
           "| kernel | dispatches | avg us | clock GHz | MFMA busy | cycles/MFMA | MFMA/VALU co-issue | "
           "LDS bank-conflict share | s_waitcnt share of wave time |",
           "|---|---|---|---|---|---|---|---|---|"]
    for name in names:
        merged = []
        for p in passes:
            rows = p.get(name, [])[-last:]
            for k, e in enumerate(rows):
                if k >= len(merged):
                    merged.append(dict(e))
                else:
                    merged[k].update({c: v for c, v in e.items() if c != "ns"})
                    merged[k]["ns"] = (merged[k]["ns"] + e["ns"]) / 2
        if not merged:
            continue
        avg = lambda c: sum(e.get(c, float("nan")) for e in merged) / len(merged)
        ns = avg("ns")
        gpu_cycles = avg("GRBM_GUI_ACTIVE") / xcds
        busy = avg("SQ_VALU_MFMA_BUSY_CYCLES") / simds / gpu_cycles
        per = avg("SQ_VALU_MFMA_BUSY_CYCLES") / avg("SQ_INSTS_MFMA")
        coex = avg("SQ_VALU_MFMA_COEXEC_CYCLES") / avg("SQ_VALU_MFMA_BUSY_CYCLES")
        bank = avg("SQ_LDS_BANK_CONFLICT") / avg("SQ_LDS_IDX_ACTIVE")
        wait = avg("SQ_WAIT_INST_ANY") / avg("SQ_WAVE_CYCLES")
        fmt = lambda x, f: "n/a" if x != x else f.format(x)
        out.append(f"| `{_short(name)}` | {len(merged)} | {ns / 1e3:.1f} | "
                   f"{fmt(gpu_cycles / ns, '{:.2f}')} | {fmt(busy, '{:.1%}')} | {fmt(per, '{:.1f}')} | "
                   f"{fmt(coex, '{:.1%}')} | {fmt(bank, '{:.1%}')} | {fmt(wait, '{:.1%}')} |")
    return "\n".join(out)


def read_traffic(path):
    """The ``hbm_bytes_per_launch`` value of a traffic summary, or None."""
    try:
        for line in open(path):
            if line.startswith(TRAFFIC_KEY):
                return int(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def main(argv):
    cmd = argv[1]
    if cmd == "stats":
        print(stats(argv[2], argv[3] if len(argv) > 3 else None))
    elif cmd == "trace":
        print(trace(argv[2], argv[3], argv[4] if len(argv) > 4 else None))
    elif cmd == "pmc":
        paths = [a for a in argv[2:] if a.endswith(".csv")]
        rest = [a for a in argv[2:] if not a.endswith(".csv")]
        print(pmc(paths, rest[0], int(rest[1]) if len(rest) > 1 else 3,
                  rest[2] if len(rest) > 2 else None))
    elif cmd == "traffic":
        print(traffic(argv[2], argv[3], argv[4], argv[5], argv[6], int(argv[7]),
                      argv[8] if len(argv) > 8 else None))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv)
