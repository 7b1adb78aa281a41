"""Summaries of rocprofv3 CSV output for the contraction kernels.

    python -m qcmaquis_amd.profiling KERNEL_TRACE_OR_COUNTER_CSV [--last N] [--scale COUNTER=F ...]
    python -m qcmaquis_amd.profiling CSV --around REGEX [--before B] [--after A] [--scale ...]

Groups the dispatches of a ``*_kernel_trace.csv`` or ``*_counter_collection.csv`` by kernel name,
keeps the last N dispatches of each kernel (the timed region of a benchmark is the end of its
run: a workload set-up that launches the same kernels comes first) and prints a markdown table of
their count, mean / min / max duration and, for counter files, the mean counter value (KiB for
FETCH_SIZE / WRITE_SIZE) multiplied by the given calibration factor.  With --around, lists the
dispatches in launch order from B before to A after the last dispatch whose name matches REGEX
(one H*psi's launch sequence, for instance).
"""
import argparse
import collections
import csv
import re
import sys


def short_name(name):
    """kernel name without return type, namespaces and argument list; template arguments kept"""
    head = name.strip()
    if head.endswith(")"):  # drop the trailing argument list
        depth = 0
        for i in range(len(head) - 1, -1, -1):
            depth += {")": 1, "(": -1}.get(head[i], 0)
            if depth == 0:
                head = head[:i]
                break
    head = head.replace("(anonymous namespace)::", "")
    if head.startswith("void "):
        head = head[5:]
    depth, cut = 0, 0
    for i, ch in enumerate(head):  # last "::" outside template brackets
        depth += {"<": 1, ">": -1}.get(ch, 0)
        if depth == 0 and head[i:i + 2] == "::":
            cut = i + 2
    return head[cut:]


def load(path):
    """dispatch rows grouped by kernel name (and counter, for counter files with several)"""
    rows = list(csv.DictReader(open(path)))
    out = collections.OrderedDict()
    for r in rows:
        key = (r["Kernel_Name"], r.get("Counter_Name", ""))
        out.setdefault(key, []).append(r)
    return out


def summarize(path, last=0, scale=None, match=None):
    scale = scale or {}
    table = []
    for (name, _), rows in load(path).items():
        if match and not re.search(match, name):
            continue
        sel = rows[-last:] if last else rows
        dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6 for r in sel]
        ent = {"kernel": short_name(name), "dispatches": len(sel), "of": len(rows),
               "mean_ms": sum(dur) / len(dur), "min_ms": min(dur), "max_ms": max(dur)}
        if "Counter_Name" in sel[0]:
            cn = sel[0]["Counter_Name"]
            v = [float(r["Counter_Value"]) for r in sel]
            f = scale.get(cn, 1.0)
            ent["counter"] = cn
            ent["mean_kib"] = sum(v) / len(v)
            ent["scaled_gb"] = ent["mean_kib"] * 1024 * f / 1e9
        table.append(ent)
    return table


def around(path, regex, before=3, after=4, scale=None):
    """the dispatches in launch order around the last one matching regex"""
    scale = scale or {}
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    hits = [i for i, r in enumerate(rows) if re.search(regex, r["Kernel_Name"])]
    if not hits:
        return []
    i = hits[-1]
    out = []
    for r in rows[max(0, i - before):i + after + 1]:
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-6
        ent = {"kernel": short_name(r["Kernel_Name"]), "dispatches": 1, "of": 1, "mean_ms": dur, "min_ms": dur,
               "max_ms": dur}
        if "Counter_Name" in r:
            ent["counter"] = r["Counter_Name"]
            ent["mean_kib"] = float(r["Counter_Value"])
            ent["scaled_gb"] = ent["mean_kib"] * 1024 * scale.get(r["Counter_Name"], 1.0) / 1e9
        out.append(ent)
    return out


def markdown(table):
    has_c = any("counter" in t for t in table)
    head = "| kernel | dispatches (of) | mean ms | min ms | max ms |" + (" counter | mean KiB | scaled GB |" if has_c else "")
    lines = [head, "|" + "---|" * (head.count("|") - 1)]
    for t in table:
        s = "| %s | %d (%d) | %.4f | %.4f | %.4f |" % (t["kernel"], t["dispatches"], t["of"], t["mean_ms"], t["min_ms"],
                                                    t["max_ms"])
        if has_c:
            s += " %s | %.1f | %.4f |" % (t.get("counter", ""), t.get("mean_kib", 0), t.get("scaled_gb", 0))
        lines.append(s)
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("csv")
    ap.add_argument("--last", type=int, default=0, help="keep the last N dispatches of each kernel (0 = all)")
    ap.add_argument("--match", default=None, help="regex on the kernel name")
    ap.add_argument("--scale", nargs="*", default=[], help="COUNTER=factor calibration, e.g. FETCH_SIZE=2.0")
    ap.add_argument("--around", default=None, help="regex: list the launch sequence around its last dispatch")
    ap.add_argument("--before", type=int, default=3)
    ap.add_argument("--after", type=int, default=4)
    a = ap.parse_args(argv)
    scale = {k: float(v) for k, v in (s.split("=") for s in a.scale)}
    if a.around:
        print(markdown(around(a.csv, a.around, a.before, a.after, scale)))
    else:
        print(markdown(summarize(a.csv, a.last, scale, a.match)))


if __name__ == "__main__":
    sys.exit(main())
