#!/usr/bin/env python3
"""Turn rocprofv3 CSV output into the short summaries kept under profiles/.

  prof_summary.py stats <kernel-trace dir> <out.md> [command text]
      per-kernel calls / total / average time from run_kernel_stats.csv
  prof_summary.py traffic <fetch dir> <write dir> <copy fetch dir> <copy write dir> <copy MiB> <out.json> [command]
      HBM bytes per dispatch of every kernel from FETCH_SIZE / WRITE_SIZE
      (KiB per dispatch), each counter scaled so that a plain copy of
      <copy MiB> read + <copy MiB> written reads as exactly that

Kernel names are shortened: namespaces and argument lists are dropped, and a
numeric template argument (lanes per row) is folded into the base name.
"""
import csv
import json
import os
import re
import sys


def short_name(name: str) -> str:
    n = name
    if "rocprim::" in n:  # library kernels: the algorithm step and the key / value types
        step = re.search(r"detail::(radix_sort_onesweep_iteration|radix_sort_onesweep_global_offsets"
                         r"|radix_sort_block_sort|merge_sort_block_merge_impl|merge_\w+?_config_selector"
                         r"|transform_impl)", n)
        types = re.search(r"_config<[^,<>]*default_config, ([\w: ]+), ([\w: ]+)>", n)
        out = "rocprim " + (step.group(1) if step else "kernel")
        return out + (f"<{types.group(1)}, {types.group(2)}>" if types else "")
    if n.startswith("void "):
        n = n[5:]
    depth, cut = 0, len(n)
    for i, ch in enumerate(n):  # drop the argument list (first '(' outside template brackets)
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and not n[i:].startswith("(anonymous"):
            cut = i
            break
    n = n[:cut]
    n = n.replace("(anonymous namespace)::", "").replace("gp::", "")
    n = re.sub(r"<\d+>$", "", n)
    n = re.sub(r"<(true|false)>$", "", n)
    return n


def read_csv(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def stats(trace_dir, out, command=""):
    rows = read_csv(os.path.join(trace_dir, "run_kernel_stats.csv"))
    agg = {}
    for r in rows:
        k = short_name(r["Name"])
        a = agg.setdefault(k, [0, 0.0])
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"])
    total = sum(v[1] for v in agg.values())
    lines = ["# Kernel time, rocprofv3 --kernel-trace --stats", ""]
    if command:
        lines += ["Command:", "", "```", command, "```", ""]
    lines += [f"Total kernel time {total / 1e6:.2f} ms.", "",
              "| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for k, (calls, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:40]:
        lines.append(f"| `{k}` | {calls} | {ns / 1e6:.2f} | {ns / calls / 1e3:.1f} | {100 * ns / total:.1f} |")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def per_kernel(counter_dir, counter):
    rows = read_csv(os.path.join(counter_dir, "run_counter_collection.csv"))
    acc = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        k = short_name(r["Kernel_Name"])
        a = acc.setdefault(k, [0, 0.0])
        a[0] += 1
        a[1] += float(r["Counter_Value"])
    return {k: (n, v / n) for k, (n, v) in acc.items()}


def traffic(fetch_dir, write_dir, cfetch_dir, cwrite_dir, copy_mib, out, command=""):
    copy_bytes = float(copy_mib) * (1 << 20)
    cf = per_kernel(cfetch_dir, "FETCH_SIZE")
    cw = per_kernel(cwrite_dir, "WRITE_SIZE")
    # the copy kernel is the one with the largest reading
    fetch_kib = max(v for _, v in cf.values())
    write_kib = max(v for _, v in cw.values())
    f_scale = copy_bytes / (fetch_kib * 1024.0)
    w_scale = copy_bytes / (write_kib * 1024.0)
    fetch = per_kernel(fetch_dir, "FETCH_SIZE")
    write = per_kernel(write_dir, "WRITE_SIZE")
    kernels = {}
    for k in sorted(set(fetch) | set(write)):
        fk = fetch.get(k, (0, 0.0))[1]
        wk = write.get(k, (0, 0.0))[1]
        kernels[k] = {"dispatches": max(fetch.get(k, (0, 0))[0], write.get(k, (0, 0))[0]),
                      "fetch_kib_raw": round(fk, 1), "write_kib_raw": round(wk, 1),
                      "bytes_per_launch": round(fk * 1024.0 * f_scale + wk * 1024.0 * w_scale)}
    # keep the 40 kernels that move the most bytes in total
    top = sorted(kernels, key=lambda k: -kernels[k]["bytes_per_launch"] * kernels[k]["dispatches"])[:40]
    doc = {"command": command, "copy_calibration_bytes": copy_bytes, "fetch_scale": f_scale,
           "write_scale": w_scale, "kernels": {k: kernels[k] for k in top}}
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2], sys.argv[3], sys.argv[4] if len(sys.argv) > 4 else "")
    elif sys.argv[1] == "traffic":
        traffic(*sys.argv[2:8], command=sys.argv[8] if len(sys.argv) > 8 else "")
    else:
        raise SystemExit(__doc__)
