#!/usr/bin/env python3
"""Summarise rocprofv3 output of bench.py runs into the short text files kept in profiles/.

  python profile_summary.py stats  <kernel_stats.csv> [--top 12]
  python profile_summary.py traffic <fetch_counter_collection.csv>
         <write_counter_collection.csv> --kernel <substring> --copy-bytes <bytes>
         [--atoms N] [--command "..."]
  python profile_summary.py steady <kernel_trace.csv> [--last 4]
  python profile_summary.py counters <counter_collection.csv> [--last 4] [--simds 1024]
         [--xcds 8]
  python profile_summary.py waits <counter_collection.csv> [--last 4]

`stats` prints the per-kernel table of a `rocprofv3 --kernel-trace --stats` run (calls,
average / total duration, share of GPU time).  `traffic` reads one `--pmc FETCH_SIZE` and one
`--pmc WRITE_SIZE` run, takes the per-dispatch averages (KiB) of the kernel whose name
contains --kernel, and scales each by the factor that makes the plain copy kernel of the same
run (bench.py --copy-check) read and write its known --copy-bytes: HBM bytes per launch =
counter * copy_bytes / counter(copy).  It prints a JSON object that bench.py can report as
roofline.traffic.  `steady` averages the last --last dispatches of every kernel of a kernel
trace (the steady evaluations after the autotuner's trial launches).  `counters` does the same
for a `--pmc` run of SQ_WAVES, SQ_INSTS_VALU, SQ_INSTS_LDS, SQ_INSTS_SALU, SQ_INSTS_SMEM and
GRBM_GUI_ACTIVE: the SQ counters are totals over the device (wave-instructions), and
GRBM_GUI_ACTIVE is summed over the XCDs, so one XCD's busy cycles are GRBM / --xcds.  It
reports per dispatch the instructions per wave and the issue rates per SIMD-cycle (VALU) and
per CU-cycle (LDS, SALU, SMEM); "VALU x4" is the fraction of SIMD cycles that the VALU
instructions would fill at 4 cycles per wave64 instruction.  `waits` prints every other SQ
counter of a run that also collected SQ_WAVES and SQ_WAVE_CYCLES as a fraction of the wave-cycles
(e.g. SQ_WAIT_ANY: waiting on a dependency or barrier, SQ_WAIT_INST_ANY: ready but not issued,
SQ_ACTIVE_INST_ANY: issuing).
"""
import argparse
import collections
import csv
import json
import sys


def stats(path, top):
    rows = list(csv.DictReader(open(path)))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print(f"{'kernel':70s} {'calls':>6s} {'avg us':>10s} {'total ms':>10s} {'share':>6s}")
    for r in rows[:top]:
        name = r["Name"][:70]
        t = float(r["TotalDurationNs"])
        print(f"{name:70s} {int(r['Calls']):6d} {float(r['AverageNs']) / 1e3:10.1f} "
              f"{t / 1e6:10.3f} {100 * t / total:5.1f}%")


def per_kernel(path, counter):
    """{kernel name: (mean counter value per dispatch, dispatch count, largest dispatch)}"""
    acc = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v), max(v)) for k, v in acc.items()}


def find(d, sub):
    """The kernel whose name contains `sub`; among several instantiations the one dispatched
    most often (the autotuner's choice runs every step, the other candidates only while
    being timed)."""
    hits = [k for k in d if sub in k]
    if not hits:
        raise SystemExit(f"no kernel matching {sub!r}")
    return max(hits, key=lambda k: d[k][1])


def traffic(fetch_csv, write_csv, kernel, copy_bytes, atoms, command):
    fetch = per_kernel(fetch_csv, "FETCH_SIZE")
    write = per_kernel(write_csv, "WRITE_SIZE")
    # the largest device-to-device copy dispatch is the calibration copy (bench.py
    # --copy-check): it reads and writes copy_bytes
    ck = "__amd_rocclr_copyBuffer"
    if ck not in fetch or ck not in write:
        raise SystemExit("no copyBuffer dispatch: run the profiled command with --copy-check")
    f_scale = copy_bytes / (fetch[ck][2] * 1024.0)
    w_scale = copy_bytes / (write[ck][2] * 1024.0)
    k = find(fetch, kernel)
    print(f"{'kernel':64s} {'calls':>5s} {'fetch MB':>10s} {'write MB':>10s}", file=sys.stderr)
    for name in sorted(fetch, key=lambda n: -fetch[n][0]):
        w = write.get(name, (0.0, 0))[0]
        print(f"{name[:64]:64s} {fetch[name][1]:5d} {fetch[name][0] * 1024 * f_scale / 1e6:10.1f} "
              f"{w * 1024 * w_scale / 1e6:10.1f}", file=sys.stderr)
    out = {
        "kernel": k[:120],
        "fetch_kib": fetch[k][0],
        "write_kib": write[k][0],
        "copy_kernel": ck,
        "copy_fetch_kib": fetch[ck][2],
        "copy_write_kib": write[ck][2],
        "copy_bytes": copy_bytes,
        "fetch_scale": f_scale,
        "write_scale": w_scale,
    }
    out["hbm_bytes"] = out["fetch_kib"] * 1024.0 * f_scale + out["write_kib"] * 1024.0 * w_scale
    out["atoms"] = atoms
    out["hbm_bytes_per_atom"] = out["hbm_bytes"] / atoms if atoms else None
    out["command"] = command
    print(json.dumps(out, indent=1))


def steady(path, last):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    acc = collections.defaultdict(list)
    for r in rows:
        acc[r["Kernel_Name"]].append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"]), int(r["Scratch_Size"]),
             int(r["VGPR_Count"]), int(r["LDS_Block_Size"])))
    print(f"{'kernel':72s} {'calls':>5s} {'steady us':>10s} {'VGPR':>5s} {'scratch':>7s} "
          f"{'LDS':>7s}")
    for name, v in sorted(acc.items(), key=lambda kv: -sum(x[0] for x in kv[1][-last:])):
        tail = v[-last:]
        us = sum(x[0] for x in tail) / len(tail) / 1e3
        if us < 20:
            continue
        print(f"{name[:72]:72s} {len(v):5d} {us:10.1f} {tail[-1][2]:5d} {tail[-1][1]:7d} "
              f"{tail[-1][3]:7d}")


def counters(path, last, simds, xcds):
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        per[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    cus = simds // 4
    print(f"{'kernel':72s} {'VALU/wave':>9s} {'LDS/wave':>8s} {'SALU/wave':>9s} "
          f"{'SMEM/wave':>9s} {'us@2.4':>7s} {'VALU x4':>7s} {'LDS/CUc':>7s} {'SALU/CUc':>8s}")
    rows = []
    for name, d in per.items():
        if "GRBM_GUI_ACTIVE" not in d or "SQ_WAVES" not in d:
            continue
        m = {c: sum(v[-last:]) / len(v[-last:]) for c, v in d.items()}
        cyc = m["GRBM_GUI_ACTIVE"] / xcds
        if cyc < 20e3:
            continue
        w = max(m["SQ_WAVES"], 1.0)
        rows.append((cyc, name, m, w))
    for cyc, name, m, w in sorted(rows, key=lambda r: -r[0]):
        print(f"{name[:72]:72s} {m.get('SQ_INSTS_VALU', 0) / w:9.0f} "
              f"{m.get('SQ_INSTS_LDS', 0) / w:8.0f} {m.get('SQ_INSTS_SALU', 0) / w:9.0f} "
              f"{m.get('SQ_INSTS_SMEM', 0) / w:9.0f} {cyc / 2400.0:7.0f} "
              f"{4.0 * m.get('SQ_INSTS_VALU', 0) / (simds * cyc):7.2f} "
              f"{m.get('SQ_INSTS_LDS', 0) / (cus * cyc):7.3f} "
              f"{m.get('SQ_INSTS_SALU', 0) / (cus * cyc):8.3f}")


def waits(path, last):
    """Per kernel: every SQ counter of the run except SQ_WAVES / SQ_WAVE_CYCLES as a fraction
    of SQ_WAVE_CYCLES (wave-cycles), means of the last `last` dispatches."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        per[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    rows = []
    for name, d in per.items():
        if "SQ_WAVE_CYCLES" not in d or "SQ_WAVES" not in d:
            continue
        m = {c: sum(v[-last:]) / len(v[-last:]) for c, v in d.items()}
        if m["SQ_WAVE_CYCLES"] <= 0 or m["SQ_WAVE_CYCLES"] / max(m["SQ_WAVES"], 1.0) < 5e3:
            continue
        rows.append((name, m))
    cols = sorted({c for _, m in rows for c in m
                   if c.startswith("SQ_") and c not in ("SQ_WAVES", "SQ_WAVE_CYCLES")})
    print(f"{'kernel':72s} {'cyc/wave':>9s} " + " ".join(f"{c[3:][:14]:>14s}" for c in cols))
    for name, m in sorted(rows, key=lambda r: -r[1]["SQ_WAVE_CYCLES"]):
        wc = m["SQ_WAVE_CYCLES"]
        print(f"{name[:72]:72s} {wc / max(m['SQ_WAVES'], 1.0):9.0f} " +
              " ".join(f"{m.get(c, 0.0) / wc:14.3f}" for c in cols))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=12)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--atoms", type=int, default=0)
    t.add_argument("--command", default="", help="the profiled command, recorded in the output")
    d = sub.add_parser("steady")
    d.add_argument("csv")
    d.add_argument("--last", type=int, default=4)
    c = sub.add_parser("counters")
    c.add_argument("csv")
    c.add_argument("--last", type=int, default=4)
    c.add_argument("--simds", type=int, default=1024)
    c.add_argument("--xcds", type=int, default=8)
    w = sub.add_parser("waits")
    w.add_argument("csv")
    w.add_argument("--last", type=int, default=4)
    a = ap.parse_args()
    if a.cmd == "waits":
        waits(a.csv, a.last)
    elif a.cmd == "stats":
        stats(a.csv, a.top)
    elif a.cmd == "steady":
        steady(a.csv, a.last)
    elif a.cmd == "counters":
        counters(a.csv, a.last, a.simds, a.xcds)
    else:
        traffic(a.fetch_csv, a.write_csv, a.kernel, a.copy_bytes, a.atoms, a.command)


if __name__ == "__main__":
    sys.exit(main())
