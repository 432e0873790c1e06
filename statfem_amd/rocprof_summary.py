"""Summaries of rocprofv3 CSV output for this library's kernels.

    python -m statfem_amd.rocprof_summary stats  <kernel_stats.csv> [rows]
    python -m statfem_amd.rocprof_summary traffic <fetch counter csv> <write counter csv> [kernel substring]

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
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
