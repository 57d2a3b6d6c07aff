"""Summaries of rocprofv3 CSV output (kernel trace, counter collection) as
short markdown, grouped by kernel and grid size (one grid size per
multigrid level).

  python -m flow_structure_interaction_fem_ibm_amd.profile_summary trace <kernel_trace.csv> [--last-solve]
      [--match SUBSTRING ...] [--by-level]
  python -m flow_structure_interaction_fem_ibm_amd.profile_summary pmc <counter_collection.csv> [...more csv]
  python -m flow_structure_interaction_fem_ibm_amd.profile_summary traffic <fetch.csv> <write.csv> \
      --copy-mib 512 --out profiles/traffic.json
"""
from __future__ import annotations

import argparse
import collections
import csv
import json


def _short(name):
    n = name.split("(")[0].replace("void ", "")
    return n


def trace(path, last_solve=False, top=30, match=None):
    """Per-kernel table of a kernel trace; last_solve keeps the kernels from
    the last solve's first kernel on, match (substrings) keeps the kernels
    whose name holds one of them."""
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    if last_solve:
        starts = [i for i, r in enumerate(rows) if "k_mask_rhs" in r["Kernel_Name"]]
        if starts:
            rows = rows[starts[-1]:]
    if match:
        rows = [r for r in rows if any(m in r["Kernel_Name"] for m in match)]
    agg = collections.defaultdict(lambda: [0, 0.0])
    for r in rows:
        key = (_short(r["Kernel_Name"]), f'{r["Grid_Size_X"]}x{r["Grid_Size_Y"]}x{r["Grid_Size_Z"]}',
               f'{r["Workgroup_Size_X"]}x{r["Workgroup_Size_Y"]}')
        agg[key][0] += 1
        agg[key][1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
    busy = sum(v[1] for v in agg.values())
    wall = (int(rows[-1]["End_Timestamp"]) - int(rows[0]["Start_Timestamp"])) * 1e-3 if rows else 0.0
    out = [f"kernels: {len(rows)}, summed kernel time {busy / 1e3:.3f} ms, first start to last end {wall / 1e3:.3f} ms",
           "", "| kernel | grid (threads) | block | calls | total us | avg us |", "|---|---|---|---|---|---|"]
    for (k, g, b), (n, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"| `{k}` | {g} | {b} | {n} | {t:.1f} | {t / n:.2f} |")
    return "\n".join(out)


_DOWN = ("k_resrestrict2", "k_restrict<")   # a transfer to the next coarser level
_UP = ("k_prolong2", "k_prolong<")           # a transfer to the next finer level


def by_level(path, last_solve=True):
    """Device time per multigrid level of the last solve, from the launch
    order: a restriction ends the work of its fine level, a prolongation
    starts the work of its fine level again; the transfer kernels are counted
    on their fine level, the dense coarsest solve on the coarsest, and the
    outer-iteration kernels (Krylov, outer residual, ghost rows, reductions)
    on level 0."""
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    if last_solve:
        starts = [i for i, r in enumerate(rows) if "k_mask_rhs" in r["Kernel_Name"]]
        if starts:
            rows = rows[starts[-1]:]
    depth = 0
    lev = collections.defaultdict(lambda: [0, 0.0])
    for r in rows:
        n = _short(r["Kernel_Name"])
        t = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        if any(n.startswith("pp::" + k) or n.startswith(k) for k in _UP):
            depth -= 1
            key = depth
        elif any(n.startswith("pp::" + k) or n.startswith(k) for k in _DOWN):
            key = depth
            depth += 1
        else:
            key = depth
        lev[key][0] += 1
        lev[key][1] += t
    total = sum(v[1] for v in lev.values())
    out = ["| level | launches | total us | share |", "|---|---|---|---|"]
    for k in sorted(lev):
        out.append(f"| {k} | {lev[k][0]} | {lev[k][1]:.1f} | {lev[k][1] / total:.1%} |")
    below = sum(v[1] for k, v in lev.items() if k > 0)
    out.append(f"\nlevels below the finest: {below / 1e3:.3f} ms of {total / 1e3:.3f} ms summed kernel time")
    return "\n".join(out)


def pmc(paths):
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    cnt = collections.defaultdict(collections.Counter)
    for p in paths:
        for r in csv.DictReader(open(p)):
            key = (_short(r["Kernel_Name"]), r.get("Grid_Size", ""))
            agg[key][r["Counter_Name"]] += float(r["Counter_Value"])
            cnt[key][r["Counter_Name"]] += 1
    names = sorted({c for v in agg.values() for c in v})
    out = ["| kernel | grid (threads) | dispatches | " + " | ".join(f"{c} / dispatch" for c in names) + " |",
           "|---|---|---|" + "---|" * len(names)]
    res = {}
    for key, v in sorted(agg.items(), key=lambda kv: -max(kv[1].values())):
        n = max(cnt[key].values())
        per = {c: v[c] / max(cnt[key][c], 1) for c in names if c in v}
        res[f"{key[0]}|{key[1]}"] = per
        out.append(f"| `{key[0]}` | {key[1]} | {n} | " + " | ".join(
            f"{per[c]:.4g}" if c in per else "-" for c in names) + " |")
    return "\n".join(out), res


# finest-level (256x192x192) kernels of the FP32 V-cycle (the line passes
# on the colour-split copy): class -> (name prefix, grid threads)
FINEST = {"yline": ("pp::k_yzline_split<8, 1>", 589824), "zline": ("pp::k_yzline_split<8, 2>", 589824),
          "xline": ("pp::k_xline<float, 32, 8, true>", 589824), "residual": ("pp::k_residual4<float>", 2359296)}
ALGORITHMIC = {"yline": 6.5, "zline": 6.5, "xline": 6.5, "residual": 13.0}  # bytes per cell per launch


def traffic(fetch_csv, write_csv, copy_mib, cells=256 * 192 * 192):
    """HBM bytes per launch of the finest-level kernels from one FETCH_SIZE and
    one WRITE_SIZE run, each scaled by what the counter reads for the
    device-to-device copy of copy_mib MiB in the same run."""
    def per_dispatch(path, counter):
        out = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                out[(_short(r["Kernel_Name"]), int(r.get("Grid_Size", 0) or 0))].append(float(r["Counter_Value"]))
        return out
    fe, wr = per_dispatch(fetch_csv, "FETCH_SIZE"), per_dispatch(write_csv, "WRITE_SIZE")
    copy_kib = copy_mib * 1024.0

    def copy_value(d):
        vals = [v for (k, g), vs in d.items() if "copyBuffer" in k for v in vs]
        big = [v for v in vals if v > 0.25 * copy_kib]
        return sum(big) / len(big)
    ff = copy_kib / copy_value(fe)
    wf = copy_kib / copy_value(wr)
    res = {"workload": "256x192x192", "fetch_factor": ff, "write_factor": wf, "copy_mib": copy_mib, "kernels": {}}
    for cls, (prefix, grid) in FINEST.items():
        fk = [v for (k, g), vs in fe.items() if k.startswith(prefix) and g == grid for v in vs]
        wk = [v for (k, g), vs in wr.items() if k.startswith(prefix) and g == grid for v in vs]
        if not fk or not wk:
            continue
        fkib, wkib = sum(fk) / len(fk), sum(wk) / len(wk)
        name = next(k for (k, g) in fe if k.startswith(prefix) and g == grid)
        res["kernels"][cls] = {"kernel": name, "grid_threads": grid, "dispatches": len(fk),
                               "fetch_kib": fkib, "write_kib": wkib,
                               "hbm_bytes_per_launch": (fkib * ff + wkib * wf) * 1024.0,
                               "hbm_bytes_unscaled": (fkib + wkib) * 1024.0,
                               "algorithmic_bytes_per_launch": ALGORITHMIC[cls] * cells}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=("trace", "pmc", "traffic"))
    ap.add_argument("paths", nargs="+")
    ap.add_argument("--last-solve", action="store_true")
    ap.add_argument("--match", nargs="*", default=None, help="trace: keep kernels whose name holds one of these")
    ap.add_argument("--by-level", action="store_true", help="trace: also the device time per multigrid level")
    ap.add_argument("--json", default=None, help="pmc: also write the per-dispatch figures as JSON")
    ap.add_argument("--copy-mib", type=int, default=512, help="traffic: size of the calibration copy")
    ap.add_argument("--out", default=None, help="traffic: output JSON path")
    args = ap.parse_args()
    if args.kind == "traffic":
        res = traffic(args.paths[0], args.paths[1], args.copy_mib)
        text = json.dumps(res, indent=1)
        if args.out:
            open(args.out, "w").write(text + "\n")
        print(text)
        return
    if args.kind == "trace":
        print(trace(args.paths[0], args.last_solve, match=args.match))
        if args.by_level:
            print()
            print(by_level(args.paths[0], args.last_solve))
    else:
        text, res = pmc(args.paths)
        print(text)
        if args.json:
            json.dump(res, open(args.json, "w"), indent=1)


if __name__ == "__main__":
    main()
