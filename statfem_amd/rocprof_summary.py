"""Summaries of rocprofv3 CSV output for this library's kernels.

    python -m statfem_amd.rocprof_summary stats  <kernel_stats.csv> [rows]
    python -m statfem_amd.rocprof_summary traffic <fetch counter csv> <write counter csv> [kernel substring]
    python -m statfem_amd.rocprof_summary mfma <counter csv> <kernel trace csv> [kernel substring]

`traffic` reports mean FETCH_SIZE / WRITE_SIZE (KiB per dispatch) per kernel
and grid size, and converts them to bytes with the factors that the same
counters give for the largest device-to-device copy in the run
(a copy of B bytes reads B and writes B)."""
import collections
import csv
import re
import sys


def short(name):
    name = re.sub(r"sfh::\(anonymous namespace\)::", "", name)
    return re.sub(r"\(.*", "", name).replace("void ", "").strip()


def stats(path, top=20):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = [f"{'kernel':44s} {'calls':>7s} {'avg us':>10s} {'total ms':>10s} {'%':>6s}"]
    for r in rows[:top]:
        out.append(f"{short(r['Name'])[:44]:44s} {int(r['Calls']):7d} {float(r['AverageNs']) / 1e3:10.2f} "
                   f"{float(r['TotalDurationNs']) / 1e6:10.2f} {float(r['Percentage']):6.2f}")
    return "\n".join(out)


def _counter(path, cname):
    agg = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != cname:
            continue
        agg[(short(r["Kernel_Name"]), int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    return agg


def calibration(agg):
    """KiB reported for the largest copy dispatch and its grid size."""
    best = None
    for (k, g), v in agg.items():
        if "copyBuffer" in k:
            m = max(v)
            if best is None or m > best:
                best = m
    return best


def traffic(fetch_csv, write_csv, copy_bytes=1 << 30, match=""):
    f, w = _counter(fetch_csv, "FETCH_SIZE"), _counter(write_csv, "WRITE_SIZE")
    cf, cw = calibration(f), calibration(w)
    sf = copy_bytes / (cf * 1024.0) if cf else None
    sw = copy_bytes / (cw * 1024.0) if cw else None
    out = {"fetch_scale": sf, "write_scale": sw, "kernels": {}}
    for key in sorted(set(f) | set(w)):
        if match and match not in key[0]:
            continue
        fv, wv = f.get(key, []), w.get(key, [])
        fk = sum(fv) / len(fv) if fv else 0.0
        wk = sum(wv) / len(wv) if wv else 0.0
        out["kernels"][f"{key[0]} grid={key[1]}"] = {
            "dispatches": max(len(fv), len(wv)), "fetch_KiB": fk, "write_KiB": wk,
            "bytes": (fk * 1024 * (sf or 1.0)) + (wk * 1024 * (sw or 1.0))}
    return out


def mfma(counter_csv, trace_csv, match="", simds=1024, xcds=8, peak_tflops=78.6):
    """Per kernel (name, grid): FP64 MFMA flops (SQ_INSTS_VALU_MFMA_MOPS_F64 x
    512) per dispatch, their rate over the dispatch's duration from the kernel
    trace of the same run, and the MFMA busy fraction
    SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / XCDs x SIMDs): the busy
    cycles are summed over every SIMD, GRBM_GUI_ACTIVE over the 8 XCDs, each
    of which counts the same elapsed cycles."""
    dur = {}
    for r in csv.DictReader(open(trace_csv)):
        dur[r["Dispatch_Id"]] = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    for r in csv.DictReader(open(counter_csv)):
        per[(short(r["Kernel_Name"]), int(r["Grid_Size"]), r["Dispatch_Id"])][r["Counter_Name"]] += \
            float(r["Counter_Value"])
    agg = collections.defaultdict(list)
    for (k, g, d), c in per.items():
        if match and match not in k:
            continue
        flops = c.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0) * 512
        ns = dur.get(d)
        busy, act = c.get("SQ_VALU_MFMA_BUSY_CYCLES"), c.get("GRBM_GUI_ACTIVE")
        agg[(k, g)].append((flops, ns, busy / (act / xcds * simds) if busy is not None and act else None))
    out = {}
    for (k, g), v in agg.items():
        fl = sum(x[0] for x in v) / len(v)
        ts = [x[1] for x in v if x[1]]
        us = [x[2] for x in v if x[2] is not None]
        t = sum(ts) / len(ts) if ts else None
        out[f"{k} grid={g}"] = {"dispatches": len(v), "gflop": fl / 1e9, "us": t / 1e3 if t else None,
                               "tflops": fl / t / 1e3 if t else None,
                               "frac": fl / t / 1e3 / peak_tflops if t else None,
                               "mfma_busy": sum(us) / len(us) if us else None}
    return out


def main(argv):
    if argv[0] == "stats":
        print(stats(argv[1], top=int(argv[2]) if len(argv) > 2 else 20))
    elif argv[0] == "traffic":
        t = traffic(argv[1], argv[2], match=argv[3] if len(argv) > 3 else "")
        print(f"calibration: FETCH_SIZE x {t['fetch_scale']:.3f}, WRITE_SIZE x {t['write_scale']:.3f} "
              f"(1 GiB device-to-device copy)")
        print(f"{'kernel':60s} {'n':>6s} {'fetch KiB':>12s} {'write KiB':>12s} {'bytes/launch':>14s}")
        for k, v in t["kernels"].items():
            if v["bytes"] < 1 << 20:
                continue
            print(f"{k[:60]:60s} {v['dispatches']:6d} {v['fetch_KiB']:12.1f} {v['write_KiB']:12.1f} "
                  f"{v['bytes']:14.4e}")
    elif argv[0] == "mfma":
        t = mfma(argv[1], argv[2], match=argv[3] if len(argv) > 3 else "")
        print(f"{'kernel':48s} {'n':>5s} {'GFLOP':>9s} {'us':>9s} {'TFLOP/s':>9s} {'of peak':>8s} {'busy':>7s}")
        for k, v in sorted(t.items(), key=lambda kv: -(kv[1]["gflop"] or 0)):
            if v["gflop"] <= 0:
                continue
            f = lambda x, fmt: format(x, fmt) if x is not None else "-"
            print(f"{k[:48]:48s} {v['dispatches']:5d} {v['gflop']:9.3f} {f(v['us'], '9.1f')} "
                  f"{f(v['tflops'], '9.2f')} {f(v['frac'], '8.3f')} {f(v['mfma_busy'], '7.3f')}")
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
