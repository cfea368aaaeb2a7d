"""Summaries of rocprofv3 runs of bench.py for profiles/.

    python -m magic_amd.profiling stats  <kernel_stats.csv> <title> <command>
    python -m magic_amd.profiling traffic <fetch_counter_collection.csv>
                                          <write_counter_collection.csv>
                                          <copy_mib> <out.json> <command...>

`stats` prints a markdown table of the kernel-trace summary.  `traffic`
averages FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel and scales
each by what the same counter reads for the device-to-device copy of
<copy_mib> MiB that `bench.py --calib-copy-mib` runs in the same process
(the largest copy dispatch), giving HBM bytes per launch; it writes them as
JSON (read by bench.py for `roofline.traffic`) and prints a markdown table.
"""
import csv
import json
import re
import sys
from collections import defaultdict


def short(name):
    n = name.replace("mhx::(anonymous namespace)::", "").replace("void ", "")
    n = re.sub(r"\(.*", "", n)
    return n.strip()


def stats(path, title, cmd):
    rows = list(csv.DictReader(open(path)))
    out = ["# %s" % title, "", "Command (after "
           "`cd /tmp && export TMPDIR=/tmp`):", "", "    %s" % cmd, "",
           "rocprofv3 kernel-trace summary (durations in microseconds):", "",
           "| kernel | calls | total us | average us | % |", "|---|---|---|---|---|"]
    for r in rows:
        out.append("| %s | %s | %.1f | %.1f | %.1f |" % (
            short(r["Name"])[:80], r["Calls"], float(r["TotalDurationNs"]) / 1e3,
            float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    return "\n".join(out)


def per_kernel(path, counter):
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        acc[(short(r["Kernel_Name"]), int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    return acc


def traffic(fetch_csv, write_csv, copy_mib, out_json, cmd):
    f = per_kernel(fetch_csv, "FETCH_SIZE")
    w = per_kernel(write_csv, "WRITE_SIZE")
    copy_kib = copy_mib * 1024.0
    # calibration: the largest copy dispatch (the copy of copy_mib MiB)
    cf = max(v for (k, g), vs in f.items() if "copyBuffer" in k or "copy" in k.lower() for v in vs)
    cw = max(v for (k, g), vs in w.items() if "copyBuffer" in k or "copy" in k.lower() for v in vs)
    ff, wf = copy_kib / cf, copy_kib / cw
    res = {"copy_mib": copy_mib, "fetch_factor": ff, "write_factor": wf, "kernels": {}}
    for (k, g), vs in f.items():
        if "mhx" not in k and not k.startswith("k_"):
            continue
        wv = w.get((k, g), [0.0])
        fk = sum(vs) / len(vs)
        wk = sum(wv) / len(wv)
        name = k + " grid=%d" % g
        res["kernels"][name] = {"fetch_kib": fk, "write_kib": wk,
                                "bytes": (fk * ff + wk * wf) * 1024.0, "dispatches": len(vs)}
    json.dump(res, open(out_json, "w"), indent=1, sort_keys=True)
    out = ["HBM traffic per launch from rocprofv3 PMC (FETCH_SIZE and WRITE_SIZE runs,",
           "KiB per dispatch), scaled by the %d MiB copy of the same run: fetch x%.3f,"
           " write x%.3f." % (copy_mib, ff, wf), "", "Commands:", ""]
    out += ["    %s" % c for c in cmd.split(" ;; ")]
    out += ["", "| kernel (grid) | FETCH KiB | WRITE KiB | HBM MB per launch |", "|---|---|---|---|"]
    for name, v in sorted(res["kernels"].items(), key=lambda kv: -kv[1]["bytes"]):
        out.append("| %s | %.0f | %.0f | %.1f |" % (name, v["fetch_kib"], v["write_kib"],
                                                    v["bytes"] / 1e6))
    return "\n".join(out)


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        print(stats(sys.argv[2], sys.argv[3], sys.argv[4]))
    elif sys.argv[1] == "traffic":
        print(traffic(sys.argv[2], sys.argv[3], int(sys.argv[4]), sys.argv[5],
                      " ".join(sys.argv[6:])))
    else:
        sys.exit(__doc__)
