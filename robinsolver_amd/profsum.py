"""Summaries of rocprofv3 output for this library's kernels.

  python -m robinsolver_amd.profsum stats <kernel_trace.csv> "<title>" "<command>"
      per-kernel table (calls, total, average) as markdown
  python -m robinsolver_amd.profsum traffic <fetch_dir> <write_dir> <kernel> <category> "<command>"
      HBM bytes per launch of one kernel from a FETCH_SIZE run and a WRITE_SIZE
      run (rocprofv3 --pmc ... --output-format csv), as JSON
  python -m robinsolver_amd.profsum mfma <pmc_dir> <after_kernel> <until_kernel|-> "<title>" "<command>"
      matrix-core utilisation per kernel, as markdown, from a run with
      --pmc MfmaUtil SQ_INSTS_VALU_MFMA_MOPS_F64 SQ_INSTS_VALU_MFMA_MOPS_F32
      over a window of dispatches (from the last <after_kernel> to the next
      <until_kernel>): MfmaUtil weighted by dispatch duration, and the MFMA
      flops (MOPS * 512, rocprofv3's MfmaFlops)

FETCH_SIZE and WRITE_SIZE are reported in KiB per dispatch.  Each is scaled by
what it reads for a plain device copy of known size in the same process (the
largest __amd_rocclr_copyBuffer dispatch, a copy of `copy_bytes` bytes):
scale = copy_bytes / (counter KiB * 1024).  For the leaf kernel the solve runs
it alternately as the up pass and the down pass (up first); `--select odd`
takes the down-pass dispatches, `--last N` the last N of those (the timed
steps at the end of a bench run).
"""
import csv
import json
import os
import sys
from collections import defaultdict


def short_name(name):
    s = name.replace("(anonymous namespace)::", "").replace("void ", "").replace("rs::", "")
    head = s.split("(")[0]
    return head.strip()


def _rows(path):
    with open(path) as fh:
        return list(csv.DictReader(fh))


def kernel_stats_md(trace_csv, title, command):
    rows = _rows(trace_csv)
    agg = defaultdict(lambda: [0, 0.0])
    for r in rows:
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        k = short_name(r["Kernel_Name"])
        agg[k][0] += 1
        agg[k][1] += d
    tot = sum(v[1] for v in agg.values())
    out = ["# %s" % title, "", "Command (MI355X):", "", "```", command, "```", "",
           "Kernel durations from `rocprofv3 --kernel-trace` (microseconds):", "",
           "| kernel | calls | total us | avg us | % |", "|---|---:|---:|---:|---:|"]
    for k, (n, t) in sorted(agg.items(), key=lambda x: -x[1][1])[:30]:
        out.append("| %s | %d | %.1f | %.1f | %.2f |" % (k.replace("|", "/"), n, t, t / n, 100.0 * t / tot))
    return "\n".join(out) + "\n"


def _counter(path_dir, counter):
    f = os.path.join(path_dir, "run_counter_collection.csv")
    rows = [r for r in _rows(f) if r["Counter_Name"] == counter]
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return rows


def traffic(fetch_dir, write_dir, kernel, category, command, copy_bytes=1 << 28, select="odd", last=0):
    res = {"category": category, "kernel": kernel, "command": command, "select": select, "last": last}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _counter(d, counter)
        cal = max((r for r in rows if "copyBuffer" in r["Kernel_Name"]), key=lambda r: float(r["Counter_Value"]))
        scale = copy_bytes / (float(cal["Counter_Value"]) * 1024.0)
        ks = [r for r in rows if kernel in short_name(r["Kernel_Name"]) or kernel in r["Kernel_Name"]]
        if select == "odd":
            ks = ks[1::2]
        elif select == "even":
            ks = ks[0::2]
        if last:
            ks = ks[-last:]
        vals = [float(r["Counter_Value"]) * 1024.0 * scale for r in ks]
        res[counter.lower() + "_bytes_per_launch"] = sum(vals) / len(vals)
        res[counter.lower() + "_scale"] = scale
        res[counter.lower() + "_dispatches"] = len(vals)
    res["hbm_bytes_per_launch"] = res["fetch_size_bytes_per_launch"] + res["write_size_bytes_per_launch"]
    return res


def mfma_md(pmc_dir, after_kernel, until_kernel, title, command):
    """Per kernel: rocprofv3's MfmaUtil (per dispatch, weighted by the dispatch
    durations) and the MFMA flops of the MOPS counters, over the dispatches from
    the last <after_kernel> up to the first <until_kernel> after it ("-": to the end)."""
    rows = _rows(os.path.join(pmc_dir, "run_counter_collection.csv"))
    per = defaultdict(dict)
    names, dur = {}, {}
    for r in rows:
        d = int(r["Dispatch_Id"])
        per[d][r["Counter_Name"]] = float(r["Counter_Value"])
        names[d] = short_name(r["Kernel_Name"])
        dur[d] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
    ids = sorted(per)
    first = max((d for d in ids if after_kernel in names[d]), default=ids[0])
    last = ids[-1]
    if until_kernel != "-":
        last = min((d for d in ids if d > first and until_kernel in names[d]), default=ids[-1])
    agg = defaultdict(lambda: defaultdict(float))
    for d in ids:
        if d < first or d > last:
            continue
        a = agg[names[d]]
        c = per[d]
        a["n"] += 1
        a["t"] += dur[d]
        a["ut"] += c.get("MfmaUtil", 0.0) * dur[d]
        a["f64"] += c.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0) * 512
        a["f32"] += c.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0.0) * 512
    out = ["# %s" % title, "", "Command (MI355X):", "", "```", command, "```", "",
           "Per kernel, over the dispatches from the last `%s`%s: rocprofv3's `MfmaUtil` (matrix-core busy"
           " cycles over GPU busy cycles times the SIMD count) averaged with the dispatch durations as"
           " weights, and the MFMA floating-point operations of the `SQ_INSTS_VALU_MFMA_MOPS_*` counters"
           " (x 512; complex products run as Gauss's three real products, 6 real flops per complex"
           " multiply-add). Durations are those of the counting run." % (
               after_kernel, "" if until_kernel == "-" else " to the next `%s`" % until_kernel), "",
           "| kernel | dispatches | total us | MfmaUtil % | MFMA TFLOP/s | FP64 MFMA GFLOP | FP32 MFMA GFLOP |",
           "|---|---:|---:|---:|---:|---:|---:|"]
    tot = defaultdict(float)
    for k, a in sorted(agg.items(), key=lambda x: -x[1]["t"])[:20]:
        out.append("| %s | %d | %.1f | %.1f | %.1f | %.1f | %.1f |" % (
            k.replace("|", "/"), a["n"], a["t"] * 1e6, a["ut"] / a["t"] if a["t"] else 0.0,
            (a["f64"] + a["f32"]) / a["t"] / 1e12 if a["t"] else 0.0, a["f64"] / 1e9, a["f32"] / 1e9))
    for a in agg.values():
        for k, v in a.items():
            tot[k] += v
    out += ["", "All of them: %.1f us, MfmaUtil %.1f%%, FP64 MFMA %.1f GFLOP, FP32 MFMA %.1f GFLOP." % (
        tot["t"] * 1e6, tot["ut"] / tot["t"] if tot["t"] else 0.0, tot["f64"] / 1e9, tot["f32"] / 1e9)]
    return "\n".join(out) + "\n"


def main(argv):
    if argv[1] == "stats":
        sys.stdout.write(kernel_stats_md(argv[2], argv[3], argv[4]))
    elif argv[1] == "traffic":
        sel, last = "odd", 0
        if "--select" in argv:
            sel = argv[argv.index("--select") + 1]
        if "--last" in argv:
            last = int(argv[argv.index("--last") + 1])
        print(json.dumps(traffic(argv[2], argv[3], argv[4], argv[5], argv[6], select=sel, last=last), indent=1))
    elif argv[1] == "mfma":
        sys.stdout.write(mfma_md(argv[2], argv[3], argv[4], argv[5], argv[6]))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv)
