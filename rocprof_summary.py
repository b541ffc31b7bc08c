"""Short text summaries of rocprofv3 output for the committed profiles (measurement tooling beside bench.py; not part of the library).

    python rocprof_summary.py stats KERNEL_STATS_CSV [--top N]
    python rocprof_summary.py window KERNEL_TRACE_CSV --last N \
        EPILOGUE [EPILOGUE ...]
    python rocprof_summary.py traffic --fetch CSV --write CSV \\
        --copy-fetch CSV --copy-write CSV
    python rocprof_summary.py counters COUNTER_CSV EPILOGUE [EPILOGUE ...]

``stats`` turns a ``--kernel-trace --stats`` kernel_stats.csv into a
markdown table.  ``window`` reads a ``--kernel-trace`` kernel_trace.csv and
gives, per CSR pass epilogue, the durations of its last N dispatches (in
start order): bench.py's own kernel timing launches each pass last, so
these are the launches its roofline figure averages.  ``traffic`` reads four ``--pmc`` counter_collection.csv
files (FETCH_SIZE and WRITE_SIZE of the benchmark, and of the 1 GiB plain
copy of ``bench.py --copy-calibration``), scales each counter by what it
reads for the copy of known size, and prints a table plus a JSON block of
HBM bytes per launch per CSR pass epilogue, which bench.py reads back.
``counters`` gives the mean of every counter of one ``--pmc`` run per CSR
pass epilogue (L2 hits and misses, L1→L2 read latency, wave cycles).
"""
from __future__ import annotations

import argparse
import collections
import csv
import json


def pass_key(name: str) -> str | None:
    """The epilogue of a csr_pass / csr_pass_persistent /
    csr_pass_xcd_segments instantiation (its fifth template argument), e.g.
    'EpiBatchPrimalY<8, false>', or 'k_batch_slice_dualz<8, false>' (or _nt) for the
    combine launch of the sliced A pass; None for other kernels."""
    for combine in ("k_batch_slice_dualz<", "k_batch_slice_dualz_nt<"):
        j = name.find(combine)
        if j >= 0:  # the second launch of the sliced A pass: its own name
            return name[j:name.index(">", j) + 1]
    for head in ("csr_pass_persistent<", "csr_pass_xcd_segments<", "csr_pass<"):
        i = name.find(head)
        if i >= 0:
            break
    else:
        return None
    i += len(head)
    depth, args, cur = 0, [], ""
    for ch in name[i:]:
        if ch == "<":
            depth += 1
        elif ch == ">":
            if depth == 0:
                args.append(cur.strip())
                break
            depth -= 1
        elif ch == "," and depth == 0:
            args.append(cur.strip())
            cur = ""
            continue
        cur += ch
    return args[4].replace("pdfp::", "") if len(args) > 4 else None


def short_name(name: str) -> str:
    key = pass_key(name)
    if key is None:
        return name.split("(")[0].replace("pdfp::", "")[:70]
    head = name.split("(")[0].replace("pdfp::", "").replace("void ", "")
    return head[:110]


def stats(path: str, top: int) -> str:
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| kernel | calls | mean us | min us | max us | % time |",
           "|---|---:|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append(f"| `{short_name(r['Name'])}` | {r['Calls']} | "
                   f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
                   f"{float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def window(path: str, keys: list, last: int) -> str:
    runs = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        key = pass_key(r["Kernel_Name"])
        if key in keys:
            runs[key].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                              short_name(r["Kernel_Name"])))
    out = [f"| epilogue | kernel | last dispatches | mean us | min us | max us |",
           "|---|---|---:|---:|---:|---:|"]
    for key in keys:
        d = sorted(runs.get(key, []))[-last:]
        if not d:
            continue
        us = [(e - s) / 1e3 for s, e, _ in d]
        names = sorted({n for _, _, n in d})
        out.append(f"| `{key}` | `{'`, `'.join(names)}` | {len(d)} | "
                   f"{sum(us) / len(us):.2f} | {min(us):.2f} | {max(us):.2f} |")
    return "\n".join(out)


def _per_kernel(path: str) -> dict:
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return d


def _copy_kib(path: str) -> float:
    """Mean counter value (KiB) over the large copy dispatches."""
    vals = [v for k, vs in _per_kernel(path).items() if "copyBuffer" in k or "copy" in k.lower()
            for v in vs]
    big = max(vals)
    sel = [v for v in vals if v >= 0.5 * big]
    return sum(sel) / len(sel)


def traffic(fetch: str, write: str, copy_fetch: str, copy_write: str,
            copy_bytes: int = 1 << 30) -> str:
    ff = copy_bytes / (_copy_kib(copy_fetch) * 1024.0)
    fw = copy_bytes / (_copy_kib(copy_write) * 1024.0)
    F, W = _per_kernel(fetch), _per_kernel(write)

    # Per epilogue, the one instantiation present in both counter runs with
    # the most dispatches: with the schedules pinned (bench.py --schedule,
    # --batch-schedule) that is the benchmarked kernel, and the fetch and
    # write figures then come from the same kernel.
    best: dict = {}
    for name in set(F) & set(W):
        key = pass_key(name)
        if key is None:
            continue
        n = len(F[name]) + len(W[name])
        if key not in best or n > len(F[best[key]]) + len(W[best[key]]):
            best[key] = name
    bf = bw = best
    recs = {}
    lines = [f"copy calibration: FETCH_SIZE x {ff:.3f}, WRITE_SIZE x {fw:.3f} "
             f"(1 GiB copy / counter reading)", "",
             "| epilogue | dispatches (fetch / write run) | FETCH KiB | WRITE KiB | "
             "HBM MB per launch |",
             "|---|---:|---:|---:|---:|"]
    for key in sorted(best, key=lambda k: -len(F[bf[k]])):
        nf, nw = bf[key], bw[key]
        fk = sum(F[nf]) / len(F[nf])
        wk = sum(W[nw]) / len(W[nw])
        t = fk * 1024.0 * ff + wk * 1024.0 * fw
        recs[key] = {"kernel_fetch_run": short_name(nf), "kernel_write_run": short_name(nw),
                     "dispatches": [len(F[nf]), len(W[nw])],
                     "fetch_kib": round(fk, 1), "write_kib": round(wk, 1),
                     "traffic_bytes": round(t)}
        lines.append(f"| `{key}` | {len(F[nf])} / {len(W[nw])} | {fk:.0f} | {wk:.0f} | "
                     f"{t / 1e6:.2f} |")
    lines += ["", "```json", json.dumps(recs, indent=1), "```"]
    return "\n".join(lines)


def counters(path: str, keys: list) -> str:
    """Mean of every counter in one --pmc counter_collection.csv per CSR
    pass epilogue (the instantiation with the most dispatches), plus the
    figures derived from them when their counters are present."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        key = pass_key(r["Kernel_Name"])
        if key in keys:
            per[(key, r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    best: dict = {}
    for (key, name), cs in per.items():
        n = max(len(v) for v in cs.values())
        if key not in best or n > best[key][1]:
            best[key] = (name, n)
    out = []
    for key in keys:
        if key not in best:
            continue
        name, n = best[key]
        mean = {c: sum(v) / len(v) for c, v in per[(key, name)].items()}
        out += [f"`{key}`: `{short_name(name)}`, {n} dispatches", "",
                "| counter | mean per dispatch |", "|---|---:|"]
        out += [f"| {c} | {v:.1f} |" for c, v in sorted(mean.items())]
        derived = []
        if "TCC_HIT_sum" in mean and "TCC_MISS_sum" in mean:
            h, m = mean["TCC_HIT_sum"], mean["TCC_MISS_sum"]
            derived.append(f"L2 hit rate {h / max(h + m, 1.0):.3f}")
        if "TCP_TCC_READ_REQ_LATENCY_sum" in mean and "TCP_TCC_READ_REQ_sum" in mean:
            derived.append("mean L1→L2 read latency "
                           f"{mean['TCP_TCC_READ_REQ_LATENCY_sum'] / max(mean['TCP_TCC_READ_REQ_sum'], 1.0):.0f} cycles")
        if "SQ_WAIT_INST_ANY" in mean and "SQ_WAVE_CYCLES" in mean:
            derived.append("SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES "
                           f"{mean['SQ_WAIT_INST_ANY'] / max(mean['SQ_WAVE_CYCLES'], 1.0):.2f}")
        if derived:
            out += ["", "derived: " + "; ".join(derived)]
        out.append("")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=15)
    w = sub.add_parser("window")
    w.add_argument("csv")
    w.add_argument("keys", nargs="+")
    w.add_argument("--last", type=int, required=True)
    t = sub.add_parser("traffic")
    for a in ("--fetch", "--write", "--copy-fetch", "--copy-write"):
        t.add_argument(a, required=True)
    c = sub.add_parser("counters")
    c.add_argument("csv")
    c.add_argument("keys", nargs="+")
    a = ap.parse_args()
    if a.cmd == "stats":
        print(stats(a.csv, a.top))
    elif a.cmd == "window":
        print(window(a.csv, a.keys, a.last))
    elif a.cmd == "counters":
        print(counters(a.csv, a.keys))
    else:
        print(traffic(a.fetch, a.write, a.copy_fetch, a.copy_write))


if __name__ == "__main__":
    main()
