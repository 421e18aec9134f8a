"""Summaries of rocprofv3 runs of bench.py (kernel time and HBM traffic).

Traffic per launch comes from two counter runs of the same command,
`rocprofv3 --pmc FETCH_SIZE` and `rocprofv3 --pmc WRITE_SIZE` (KiB per
dispatch).  Each counter is scaled by its reading for a plain device copy of
known size that bench.py issues (torch clone of the n-word input: 8n bytes
read, 8n written), so the result is in bytes as the copy defines them.

    python -m goldilocks_amd.profiling --workload ntt_2^24 --copy-bytes 134217728 \\
        --stats <kernel_stats.csv> --fetch <counter_collection.csv> \\
        --write <counter_collection.csv> --json profiles/traffic.json --md out.md
"""
import argparse
import collections
import csv
import json
import os


def per_dispatch(path, counter):
    """{dispatch id: (kernel name, value)} for one counter (summed over dimensions)."""
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            d = int(r["Dispatch_Id"])
            name, v = out.get(d, (r["Kernel_Name"], 0.0))
            out[d] = (name, v + float(r["Counter_Value"]))
    return out


def calibrated(path, counter, copy_bytes):
    """Per-kernel average bytes per dispatch, scaled so that the largest copy
    dispatch reads (or writes) copy_bytes."""
    disp = per_dispatch(path, counter)
    copies = [v for name, v in disp.values() if "copyBuffer" in name]
    if not copies or max(copies) <= 0:
        raise SystemExit(f"{path}: no copy dispatch to calibrate {counter}")
    scale = copy_bytes / max(copies)  # bytes per counter unit
    acc = collections.defaultdict(list)
    for name, v in disp.values():
        acc[name].append(v * scale)
    return scale, {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def short(name):
    name = name.replace("(gl::PassArgs)", "").replace("void ", "").replace("gl::", "")
    return name.strip()


def kernel_stats(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]) / 1e3,
                         float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    return sorted(rows, key=lambda x: -x[2])


def _union(intervals):
    """Sorted, merged [start, end) intervals."""
    out = []
    for a, b in sorted(intervals):
        if out and a <= out[-1][1]:
            out[-1][1] = max(out[-1][1], b)
        else:
            out.append([a, b])
    return out


def _overlap(a, b):
    """Total length of the intersection of two merged interval lists."""
    i = j = 0
    both = 0
    while i < len(a) and j < len(b):
        lo, hi = max(a[i][0], b[j][0]), min(a[i][1], b[j][1])
        if lo < hi:
            both += hi - lo
        if a[i][1] < b[j][1]:
            i += 1
        else:
            j += 1
    return both


def copy_activity(copies_csv, kernels_csv=None):
    """Host <-> device copy activity per stream: DMA copies from a rocprofv3
    --memory-copy-trace CSV (with their direction) and, from a --kernel-trace
    CSV, the runtime's blit kernels (__amd_rocclr_copyBuffer), which carry
    some copies instead of the DMA engines.  Returns {stream: {"dma": {dir:
    n}, "blits": n, "busy": merged intervals}}."""
    acts = collections.defaultdict(lambda: {"dma": collections.Counter(), "blits": 0, "iv": []})
    with open(copies_csv) as f:
        for r in csv.DictReader(f):
            a = acts[r["Stream_Id"]]
            a["dma"][r["Direction"].replace("MEMORY_COPY_", "")] += 1
            a["iv"].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    if kernels_csv:
        with open(kernels_csv) as f:
            for r in csv.DictReader(f):
                if "copyBuffer" in r["Kernel_Name"]:
                    a = acts[r["Stream_Id"]]
                    a["blits"] += 1
                    a["iv"].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    return {s: {"dma": dict(a["dma"]), "blits": a["blits"], "busy": _union(a["iv"])} for s, a in acts.items()}


def write_stats_md(f, args, stats):
    f.write(f"# rocprofv3 summary: {args.workload}\n\n")
    if args.command:
        f.write("Commands (repo root, one MI355X):\n\n")
        for c in args.command.split(";;"):
            f.write(f"    {c.strip()}\n")
        f.write("\n")
    f.write("Kernel time (`--kernel-trace --stats`, kernel_stats.csv):\n\n")
    f.write("| kernel | calls | total us | average us | % |\n|---|---|---|---|---|\n")
    for name, calls, tot, avg, pct in stats[:12]:
        f.write(f"| `{short(name)}` | {calls} | {tot:.1f} | {avg:.2f} | {pct:.1f} |\n")


def counters_md(f, path, kernel_filter):
    """Per-kernel averages of every counter in a --pmc run, plus the VALU
    issue fraction: SQ_INSTS_VALU wave instructions x 4 cycles (wave64 on a
    16-lane SIMD) over 1024 SIMDs, against SQ_BUSY_CYCLES / 32 (summed over
    the 32 shader engines of 8 XCDs)."""
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    with open(path) as fh:
        for r in csv.DictReader(fh):
            if kernel_filter not in r["Kernel_Name"]:
                continue
            k = short(r["Kernel_Name"])
            per[k][r["Counter_Name"]] += float(r["Counter_Value"])
            disp[k].add(int(r["Dispatch_Id"]))
    for k, c in per.items():
        nd = len(disp[k])
        names = sorted(c)
        f.write(f"\n`{k}`, {nd} dispatches, per-dispatch averages:\n\n")
        f.write("| " + " | ".join(names) + " |\n|" + "---|" * len(names) + "\n")
        f.write("| " + " | ".join(f"{c[n] / nd:.0f}" for n in names) + " |\n")
        if "SQ_INSTS_VALU" in c and "SQ_BUSY_CYCLES" in c:
            valu = c["SQ_INSTS_VALU"] / nd * 4 / 1024
            busy = c["SQ_BUSY_CYCLES"] / nd / 32
            f.write(f"\nVALU issue: {valu:.0f} cycles per SIMD over {busy:.0f} busy cycles = "
                    f"**{valu / busy:.2f}** of peak\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", required=True)
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--stats", default=None)
    ap.add_argument("--fetch", default=None)
    ap.add_argument("--write", default=None)
    ap.add_argument("--algorithmic-bytes", type=int, default=None,
                    help="algorithmic bytes per launch of the pass kernels (16 n for an NTT pass)")
    ap.add_argument("--json", default=None)
    ap.add_argument("--md", required=True)
    ap.add_argument("--command", default="")
    ap.add_argument("--counters", default=None, help="counter_collection.csv of a --pmc run")
    ap.add_argument("--kernel", default="ntt_pass", help="kernel-name filter for --counters")
    ap.add_argument("--copies", default=None, help="memory_copy_trace.csv of a --memory-copy-trace run")
    ap.add_argument("--kernel-trace", default=None, help="kernel_trace.csv of the same run (blit copies)")
    args = ap.parse_args()
    if args.copies:
        acts = copy_activity(args.copies, args.kernel_trace)
        with open(args.md, "w") as f:
            if args.stats:
                write_stats_md(f, args, kernel_stats(args.stats))
            f.write("\nHost <-> device copies per stream (DMA copies from `--memory-copy-trace`, "
                    "blit kernels from `--kernel-trace`):\n\n")
            f.write("| stream | DMA copies | blit kernels | copy time in flight ms |\n|---|---|---|---|\n")
            for st, a in sorted(acts.items(), key=lambda x: int(x[0])):
                dma = ", ".join(f"{n} {d}" for d, n in sorted(a["dma"].items())) or "-"
                busy = sum(b - a_ for a_, b in a["busy"]) / 1e6
                f.write(f"| {st} | {dma} | {a['blits']} | {busy:.3f} |\n")
            pairs = []
            keys = sorted(acts, key=int)
            for x in range(len(keys)):
                for y in range(x + 1, len(keys)):
                    both = _overlap(acts[keys[x]]["busy"], acts[keys[y]]["busy"])
                    if both:
                        pairs.append((keys[x], keys[y], both))
            if pairs:
                f.write("\nCopies of two streams in flight at once:\n\n| streams | ms |\n|---|---|\n")
                for a_, b, both in pairs:
                    f.write(f"| {a_} and {b} | {both / 1e6:.3f} |\n")
        return
    if args.counters:
        with open(args.md, "w") as f:
            f.write(f"# rocprofv3 counters: {args.workload}\n\n")
            if args.command:
                f.write("Command (repo root, one MI355X):\n\n")
                for c in args.command.split(";;"):
                    f.write(f"    {c.strip()}\n")
            counters_md(f, args.counters, args.kernel)
        return
    stats = kernel_stats(args.stats)
    if not (args.fetch and args.write):
        with open(args.md, "w") as f:
            write_stats_md(f, args, stats)
        return
    fs, fetch = calibrated(args.fetch, "FETCH_SIZE", args.copy_bytes)
    ws, write = calibrated(args.write, "WRITE_SIZE", args.copy_bytes)
    data = {}
    if os.path.exists(args.json):
        with open(args.json) as f:
            data = json.load(f)
    rec = {}
    for name in fetch:
        if name not in write or "copyBuffer" in name:
            continue
        fb, fn = fetch[name]
        wb, wn = write[name]
        rec[short(name)] = {"bytes_per_launch": round(fb + wb), "fetch_bytes": round(fb),
                            "write_bytes": round(wb), "dispatches": [fn, wn]}
    data[args.workload] = rec
    data.setdefault("_calibration", {})[args.workload] = {
        "FETCH_SIZE_bytes_per_unit": fs, "WRITE_SIZE_bytes_per_unit": ws, "copy_bytes": args.copy_bytes}
    with open(args.json, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    with open(args.md, "w") as f:
        write_stats_md(f, args, stats)
        f.write(f"\nHBM traffic per launch (`--pmc FETCH_SIZE`, then `--pmc WRITE_SIZE`; one counter unit = "
                f"{fs:.1f} B read / {ws:.1f} B written, from the {args.copy_bytes}-byte copy):\n\n")
        f.write("| kernel | launches | fetch MiB | write MiB | total MiB |")
        if args.algorithmic_bytes:
            f.write(" / algorithmic |")
        f.write("\n|---|---|---|---|---|" + ("---|" if args.algorithmic_bytes else "") + "\n")
        for name, r in sorted(rec.items(), key=lambda kv: -kv[1]["bytes_per_launch"]):
            line = (f"| `{name}` | {r['dispatches'][0]} | {r['fetch_bytes'] / 2**20:.1f} | "
                    f"{r['write_bytes'] / 2**20:.1f} | {r['bytes_per_launch'] / 2**20:.1f} |")
            if args.algorithmic_bytes:
                line += f" {r['bytes_per_launch'] / args.algorithmic_bytes:.3f} |"
            f.write(line + "\n")
    print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    main()
