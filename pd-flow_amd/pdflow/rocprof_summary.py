"""Summaries of rocprofv3 output for profiles/ (measurement tooling; runs on the CSV files only).

``traffic``: HBM bytes per launch of the batched iteration kernel from two ``rocprofv3 --pmc``
runs of the same bench command (FETCH_SIZE, then WRITE_SIZE), each scaled by what the counter read
for the plain copy kernel of known size that ``bench.py --copy-calibration BYTES`` runs first
(every byte read once and written once).  The launches summarised are the last ``--launches``
dispatches of kb_pd_tb, the eager solves whose HIP-event timing gives bench.py's roofline.  The
last line, ``traffic: bytes_per_launch=... workload=... iterate_source_sha16=...``, is the one
bench.py reads back.

``trace``: the average duration of those launches in a ``--kernel-trace`` CSV of the same command,
to set beside bench.py's HIP-event average.

``l2``: the L2 hit rate of those launches from one ``--pmc TCC_HIT_sum TCC_MISS_sum`` run of the
same command.

``counters``: per-launch averages of every counter in one or more ``--pmc`` runs (the SQ
instruction and wave-state counters) over the same launches, by kernel instance and grid.

``kernels``: every dispatch of a ``--kernel-trace`` CSV grouped by kernel and grid, with totals
divided by ``--per`` (the number of solves or pairs traced), and the average of the last
``--iteration-launches`` dispatches of the single-pair iteration kernels.

    python -m pdflow.rocprof_summary traffic FETCH.csv WRITE.csv --copy-bytes 1073741824 \\
        --launches 117 --workload KEY
    python -m pdflow.rocprof_summary trace TRACE.csv --launches 117
    python -m pdflow.rocprof_summary kernels TRACE.csv --per 43 --iteration-launches 940
"""
import argparse
import collections
import csv
import hashlib
import os

# the sources of the iteration kernels (the kernels and the per-pixel arithmetic they inline)
ITERATE_SRCS = [os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "csrc", f)
                for f in ("pdflow_iterate.hip", "pdflow_pixel.hpp")]


def _rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def _sorted_dispatches(rows):
    return sorted(rows, key=lambda r: int(r["Dispatch_Id"]) if "Dispatch_Id" in r else int(r["Start_Timestamp"]))


def _factor(rows, copy_bytes):
    """bytes per counter KiB: the copy kernel moves copy_bytes each way"""
    vals = [float(r["Counter_Value"]) for r in rows if "k_stream_copy" in r["Kernel_Name"]]
    if not vals:
        raise SystemExit("no calibration copy in the counter file")
    return copy_bytes / (sum(vals) / len(vals) * 1024.0), vals


def traffic(a):
    fetch, write = _sorted_dispatches(_rows(a.fetch)), _sorted_dispatches(_rows(a.write))
    ff, fv = _factor(fetch, a.copy_bytes)
    wf, wv = _factor(write, a.copy_bytes)
    tb_f = [r for r in fetch if "kb_pd_tb" in r["Kernel_Name"]][-a.launches:]
    tb_w = [r for r in write if "kb_pd_tb" in r["Kernel_Name"]][-a.launches:]
    if len(tb_f) != a.launches or len(tb_w) != a.launches:
        raise SystemExit("fewer kb_pd_tb dispatches than --launches")
    kinds = collections.OrderedDict()
    for rf, rw in zip(tb_f, tb_w):
        if rf["Kernel_Name"] != rw["Kernel_Name"] or rf["Grid_Size"] != rw["Grid_Size"]:
            raise SystemExit("the two runs dispatched different launches")
        k = (rf["Kernel_Name"].split("kb_pd_tb")[1].split("(")[0], int(rf["Grid_Size"]))
        kinds.setdefault(k, []).append((float(rf["Counter_Value"]) * 1024 * ff, float(rw["Counter_Value"]) * 1024 * wf))
    print(f"Calibration: k_stream_copy of {a.copy_bytes} bytes, {len(fv)} dispatches: FETCH_SIZE "
          f"{sum(fv) / len(fv):,.0f} KiB -> factor {ff:.4f}; WRITE_SIZE {sum(wv) / len(wv):,.0f} KiB -> "
          f"factor {wf:.4f} (bytes = counter KiB x 1024 x factor)")
    print()
    print("| instance | grid (threads) | launches | read GB | write GB | HBM GB per launch |")
    print("|---|---|---|---|---|---|")
    tot = 0.0
    for (name, grid), v in kinds.items():
        r = sum(x for x, _ in v) / len(v)
        w = sum(y for _, y in v) / len(v)
        tot += (r + w) * len(v)
        print(f"| kb_pd_tb{name} | {grid} | {len(v)} | {r / 1e9:.3f} | {w / 1e9:.3f} | {(r + w) / 1e9:.3f} |")
    per = tot / a.launches
    print()
    print(f"Average over the {a.launches} launches: {per / 1e9:.4f} GB of HBM traffic per launch.")
    h = hashlib.sha256()
    for path in ITERATE_SRCS:
        with open(path, "rb") as f:
            h.update(f.read())
    sha = h.hexdigest()[:16]
    print()
    print(f"traffic: bytes_per_launch={per:.6e} workload={a.workload} iterate_source_sha16={sha}")


def l2(a):
    """L2 hit rate of the last --launches kb_pd_tb dispatches from one --pmc TCC_HIT_sum TCC_MISS_sum run"""
    rows = _sorted_dispatches(_rows(a.counters))
    by = collections.defaultdict(dict)
    for r in rows:
        by[int(r["Dispatch_Id"])][r["Counter_Name"]] = float(r["Counter_Value"])
        by[int(r["Dispatch_Id"])]["name"] = r["Kernel_Name"]
        by[int(r["Dispatch_Id"])]["grid"] = int(r["Grid_Size"])
    tb = [by[d] for d in sorted(by) if "kb_pd_tb" in by[d]["name"]][-a.launches:]
    if len(tb) != a.launches:
        raise SystemExit("fewer kb_pd_tb dispatches than --launches")
    kinds = collections.OrderedDict()
    for d in tb:
        k = (d["name"].split("kb_pd_tb")[1].split("(")[0], d["grid"])
        kinds.setdefault(k, []).append((d["TCC_HIT_sum"], d["TCC_MISS_sum"]))
    print("| instance | grid (threads) | launches | L2 hits | L2 misses | hit rate |")
    print("|---|---|---|---|---|---|")
    th = tm = 0.0
    for (name, grid), v in kinds.items():
        h, m = sum(x for x, _ in v), sum(y for _, y in v)
        th, tm = th + h, tm + m
        print(f"| kb_pd_tb{name} | {grid} | {len(v)} | {h / len(v):,.0f} | {m / len(v):,.0f} | {h / (h + m):.3f} |")
    print()
    print(f"L2 hit rate over the {a.launches} launches: {th / (th + tm):.3f}")


def counters(a):
    """per-launch averages of every counter of one or more --pmc runs over the last --launches
    kb_pd_tb dispatches of each, by kernel instance and grid"""
    table = collections.OrderedDict()
    names = []
    for path in a.csv:
        rows = _sorted_dispatches(_rows(path))
        by = collections.OrderedDict()
        for r in rows:
            d = by.setdefault(int(r["Dispatch_Id"]), {"name": r["Kernel_Name"], "grid": int(r["Grid_Size"])})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        tb = [d for d in by.values() if "kb_pd_tb" in d["name"]][-a.launches:]
        if len(tb) != a.launches:
            raise SystemExit(f"{path}: fewer kb_pd_tb dispatches than --launches")
        for d in tb:
            k = (d["name"].split("kb_pd_tb")[1].split("(")[0], d["grid"])
            for c, v in d.items():
                if c in ("name", "grid"):
                    continue
                if c not in names:
                    names.append(c)
                table.setdefault(k, {}).setdefault(c, []).append(v)
    print("| instance | grid (threads) | " + " | ".join(names) + " |")
    print("|---|---|" + "---|" * len(names))
    for (name, grid), cs in table.items():
        cells = [f"{sum(cs[c]) / len(cs[c]):.4g}" if c in cs else "" for c in names]
        print(f"| kb_pd_tb{name} | {grid} | " + " | ".join(cells) + " |")


def trace(a):
    rows = sorted(_rows(a.trace), key=lambda r: int(r["Start_Timestamp"]))
    tb = [r for r in rows if "kb_pd_tb" in r["Kernel_Name"]][-a.launches:]
    dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in tb]
    ov = sum(1 for x, y in zip(tb, tb[1:]) if int(y["Start_Timestamp"]) < int(x["End_Timestamp"]))
    print(f"last {len(tb)} kb_pd_tb dispatches: average {sum(dur) / len(dur):.1f} us, {ov} overlapping pairs")
    by = collections.OrderedDict()
    for r, d in zip(tb, dur):
        k = (r["Kernel_Name"].split("kb_pd_tb")[1].split("(")[0], int(r.get("Grid_Size_X") or r.get("Grid_Size")))
        by.setdefault(k, []).append(d)
    for (name, grid), v in by.items():
        print(f"  kb_pd_tb{name} grid {grid}: {len(v)} launches, average {sum(v) / len(v):.1f} us")


ITERATION_KERNELS = ("k_pd_tb", "k_pd_level", "k_pd_fused")


def _short(name):
    """kernel name without return type, namespaces and parameter list"""
    name = name[5:] if name.startswith("void ") else name
    if name.endswith(")"):                 # drop the parameter list: the '(' matching the last ')'
        depth = 0
        for i in range(len(name) - 1, -1, -1):
            depth += {")": 1, "(": -1}.get(name[i], 0)
            if depth == 0:
                name = name[:i]
                break
    for ns in ("pdf::cnn::(anonymous namespace)::", "pdf::(anonymous namespace)::", "pdf::"):
        name = name.replace(ns, "")
    return name


def kernels(a):
    rows = sorted(_rows(a.trace), key=lambda r: int(r["Start_Timestamp"]))
    groups = collections.OrderedDict()
    for r in rows:
        k = (_short(r["Kernel_Name"]), int(r.get("Grid_Size_X") or r.get("Grid_Size")))
        groups.setdefault(k, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    print(f"{len(rows)} dispatches; per unit = total / {a.per}")
    print()
    print("| kernel | grid x (threads) | dispatches | average us | per unit ms |")
    print("|---|---|---|---|---|")
    tot = 0.0
    for (name, grid), v in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        tot += sum(v)
        if sum(v) / a.per / 1e3 < a.min_ms:
            continue
        print(f"| {name} | {grid} | {len(v)} | {sum(v) / len(v):.2f} | {sum(v) / a.per / 1e3:.4f} |")
    print()
    print(f"All dispatches: {tot / a.per / 1e3:.4f} ms per unit (rows under {a.min_ms} ms per unit omitted above).")
    if a.iteration_launches:
        it = [r for r in rows if _short(r["Kernel_Name"]).startswith(ITERATION_KERNELS)][-a.iteration_launches:]
        dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in it]
        print()
        print(f"last {len(it)} iteration-kernel dispatches ({', '.join(ITERATION_KERNELS)}): average "
              f"{sum(dur) / len(dur):.2f} us")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    h = sub.add_parser("l2")
    h.add_argument("counters")
    h.add_argument("--launches", type=int, required=True)
    g = sub.add_parser("kernels")
    g.add_argument("trace")
    g.add_argument("--per", type=float, default=1.0, help="divide totals by this count (solves, pairs)")
    g.add_argument("--min-ms", type=float, default=0.0)
    g.add_argument("--iteration-launches", type=int, default=0)
    t = sub.add_parser("traffic")
    t.add_argument("fetch")
    t.add_argument("write")
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--launches", type=int, required=True)
    t.add_argument("--workload", required=True)
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    c.add_argument("--launches", type=int, required=True)
    r = sub.add_parser("trace")
    r.add_argument("trace")
    r.add_argument("--launches", type=int, required=True)
    a = ap.parse_args(argv)
    {"traffic": traffic, "trace": trace, "kernels": kernels, "l2": l2, "counters": counters}[a.cmd](a)


if __name__ == "__main__":
    main()
