"""Short text summaries of rocprofv3 CSV output (``-f csv``).

    python -m quip_amd.prof_summary kernels  <run_kernel_trace.csv> [--match S] [--top N]
    python -m quip_amd.prof_summary counters <run_counter_collection.csv> [--match S]
    python -m quip_amd.prof_summary traffic  <fetch_counter.csv> <write_counter.csv> --kernel S [--grid G]
                                            [--probe-bytes B]
    python -m quip_amd.prof_summary stages   <kernel_trace.csv> <fetch.csv> <write.csv> <sq.csv>
                                            [--match A|B|...] [--probe-bytes B]
    python -m quip_amd.prof_summary steps    <run_kernel_trace.csv> --start S [--contains T] [--which K]

kernels:  calls, total and average duration per (kernel, grid size), the
          --stats view split by launch shape.
counters: average counter value per dispatch per (kernel, grid size, counter).
stages:   per kernel shape the achieved HBM GB/s and FP64 VALU TFLOP/s against the
          MI355X peaks, from a kernel trace, FETCH_SIZE and WRITE_SIZE runs and a run
          of SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64, SQ_INSTS_VALU, SQ_INSTS_LDS, SQ_WAVES.
steps:    the kernels of one evaluation in launch order: a step begins at each launch of
          the kernel named S; of the complete steps that launch a kernel named T, the
          K-th from the end (default 2, a late timed step) is listed with its start,
          duration and grid, then the sum of durations and first start to next step.
traffic:  HBM bytes per launch of one kernel shape from a FETCH_SIZE run and a
          WRITE_SIZE run (KiB per dispatch), scaled by what the same counters
          read for the device-to-device copy kernel (__amd_rocclr_copyBuffer)
          of a probe of B bytes (B read + B written): factor = B / counter.
          Without --grid, --kernel is a name prefix and the (kernel, grid) shape
          with the most dispatches in the FETCH_SIZE run is taken (the shape an
          autotuner settled on after timing its candidates once each).
"""
import argparse
import csv
import re
from collections import defaultdict


def short(name):
    name = re.sub(r"\(.*", "", name)
    return name.replace("void ", "").replace("gapx::", "")


def _grid(row):
    if "Grid_Size_X" in row:
        return "%sx%sx%s" % (row["Grid_Size_X"], row["Grid_Size_Y"], row["Grid_Size_Z"])
    return row.get("Grid_Size", "?")


def kernels(path, match=None, top=40):
    agg = defaultdict(lambda: [0, 0.0])
    for r in csv.DictReader(open(path)):
        k = (short(r["Kernel_Name"]), _grid(r))
        agg[k][0] += 1
        agg[k][1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    tot = sum(v[1] for v in agg.values())
    out = ["%-44s %-18s %6s %11s %10s %6s" % ("kernel", "grid (threads)", "calls", "total_us", "avg_us", "pct")]
    rows = sorted(agg.items(), key=lambda kv: -kv[1][1])
    n = 0
    for (name, grid), (c, s) in rows:
        if match and match not in name:
            continue
        out.append("%-44s %-18s %6d %11.1f %10.3f %6.2f" % (name[:44], grid, c, s, s / c, 100.0 * s / tot))
        n += 1
        if n >= top:
            break
    return "\n".join(out)


def _counter_avgs(path):
    agg = defaultdict(list)
    for r in csv.DictReader(open(path)):
        agg[(short(r["Kernel_Name"]), _grid(r), r["Counter_Name"])].append(float(r["Counter_Value"]))
    return {k: (len(v), sum(v) / len(v)) for k, v in agg.items()}


def counters(path, match=None):
    out = ["%-44s %-12s %-12s %6s %14s" % ("kernel", "grid", "counter", "calls", "avg_value")]
    for (name, grid, cn), (c, v) in sorted(_counter_avgs(path).items()):
        if match and match not in name:
            continue
        out.append("%-44s %-12s %-12s %6d %14.1f" % (name[:44], grid, cn, c, v))
    return "\n".join(out)


def busy(path, match=None):
    """MFMA pipe occupancy and per-wave instruction counts of each (kernel, grid) from a
    run with SQ_VALU_MFMA_BUSY_CYCLES, SQ_BUSY_CYCLES, SQ_INSTS_MFMA, SQ_WAVES,
    SQ_WAIT_INST_ANY, SQ_WAVE_CYCLES and SQ_INSTS_LDS.  The MFMA busy counter sums over
    the 1024 SIMDs and the busy counter over the 32 shader engines, so busy =
    MFMA_BUSY / (32 BUSY)."""
    per = defaultdict(dict)
    for (name, grid, cn), (_, v) in _counter_avgs(path).items():
        per[(name, grid)][cn] = v
    out = []
    for (name, grid), a in sorted(per.items()):
        if match and match not in name:
            continue
        w = a["SQ_WAVES"]
        out.append("## %s, grid %s (%d waves)" % (name, grid, int(w)))
        out.append("  MFMA busy %.1f%%, MFMA / wave %.0f, LDS / wave %.0f, waiting %.0f%% of wave cycles"
                   % (100 * a["SQ_VALU_MFMA_BUSY_CYCLES"] / (32 * a["SQ_BUSY_CYCLES"]), a["SQ_INSTS_MFMA"] / w,
                      a["SQ_INSTS_LDS"] / w, 100 * a["SQ_WAIT_INST_ANY"] / a["SQ_WAVE_CYCLES"]))
        out.append("  " + ", ".join("%s %d" % (k, round(v)) for k, v in sorted(a.items())))
        out.append("")
    return "\n".join(out)


def traffic(fetch_csv, write_csv, kernel, grid, probe_bytes):
    f, w = _counter_avgs(fetch_csv), _counter_avgs(write_csv)

    def pick(d, name, cn, g=None):
        vals = [(k, v) for k, v in d.items() if k[0] == name and k[2] == cn and (g is None or k[1] == g)]
        return max(vals, key=lambda kv: kv[1][1]) if vals else None

    # the probe's dispatch is the copy kernel shape with the largest counter value
    pf, pw = pick(f, "__amd_rocclr_copyBuffer", "FETCH_SIZE"), pick(w, "__amd_rocclr_copyBuffer", "WRITE_SIZE")
    if grid is None:   # the most-dispatched shape whose name starts with `kernel`
        cand = [(v[0], k) for k, v in f.items() if k[0].startswith(kernel) and k[2] == "FETCH_SIZE"]
        if not cand:
            raise SystemExit("no FETCH_SIZE dispatches of %s" % kernel)
        _, (kernel, grid, _) = max(cand)
    kf, kw = pick(f, kernel, "FETCH_SIZE", grid), pick(w, kernel, "WRITE_SIZE", grid)
    ff = probe_bytes / pf[1][1]
    fw = probe_bytes / pw[1][1]
    lines = ["copy probe: %d bytes read + written; FETCH_SIZE %.1f, WRITE_SIZE %.1f per dispatch"
             % (probe_bytes, pf[1][1], pw[1][1]),
             "scale: %.2f bytes per FETCH_SIZE unit, %.2f bytes per WRITE_SIZE unit" % (ff, fw),
             "%s grid %s: FETCH_SIZE %.1f (%d dispatches), WRITE_SIZE %.1f (%d dispatches)"
             % (kernel, grid, kf[1][1], kf[1][0], kw[1][1], kw[1][0]),
             "fetch_bytes_per_launch %.0f" % (kf[1][1] * ff),
             "write_bytes_per_launch %.0f" % (kw[1][1] * fw),
             "hbm_bytes_per_launch %.0f" % (kf[1][1] * ff + kw[1][1] * fw)]
    return "\n".join(lines)


FP64_PEAK = 78.6e12     # MI355X FP64 vector and matrix peak, FLOP/s
HBM_PEAK = 8.0e12       # MI355X HBM3E, bytes/s


def stages(trace_csv, fetch_csv, write_csv, sq_csv, match, probe_bytes):
    """Per kernel shape: average duration (kernel trace), HBM bytes per launch
    (FETCH_SIZE + WRITE_SIZE scaled by the copy probe), FP64 VALU FLOPs per launch
    (64 lanes x (ADD + MUL + TRANS + 2 FMA) wave instructions, full wavefronts assumed)
    and what they give against the 8 TB/s and 78.6 TFLOP/s peaks; VALU and LDS
    instructions per wavefront."""
    dur = defaultdict(list)
    for r in csv.DictReader(open(trace_csv)):   # keyed by total grid threads, as the counter files are
        g = str(int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"]))
        dur[(short(r["Kernel_Name"]), g)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    f, w, q = _counter_avgs(fetch_csv), _counter_avgs(write_csv), _counter_avgs(sq_csv)
    pf = max((v for k, v in f.items() if k[0] == "__amd_rocclr_copyBuffer" and k[2] == "FETCH_SIZE"),
             key=lambda v: v[1])
    pw = max((v for k, v in w.items() if k[0] == "__amd_rocclr_copyBuffer" and k[2] == "WRITE_SIZE"),
             key=lambda v: v[1])
    ff, fw = probe_bytes / pf[1], probe_bytes / pw[1]
    out = ["copy probe: %d bytes read + written; %.2f bytes per FETCH_SIZE unit, %.2f per WRITE_SIZE unit"
           % (probe_bytes, ff, fw),
           "%-34s %-12s %6s %9s %10s %8s %6s %10s %8s %6s %9s %8s" % (
               "kernel", "grid", "calls", "avg_us", "hbm_MB", "GB/s", "hbm%", "fp64_GF", "TF/s", "valu%",
               "valu/wave", "lds/wave")]
    shapes = sorted({(k[0], k[1]) for k in f} | {(k[0], k[1]) for k in q})
    for name, grid in shapes:
        if match and not any(m_ in name for m_ in match.split("|")):
            continue
        d = dur.get((name, grid))
        if not d:
            continue
        avg = sum(d) / len(d)

        def cv(src, cn):
            v = src.get((name, grid, cn))
            return v[1] if v else 0.0
        hbm = cv(f, "FETCH_SIZE") * ff + cv(w, "WRITE_SIZE") * fw
        flops = 64.0 * (cv(q, "SQ_INSTS_VALU_ADD_F64") + cv(q, "SQ_INSTS_VALU_MUL_F64") +
                        cv(q, "SQ_INSTS_VALU_TRANS_F64") + 2.0 * cv(q, "SQ_INSTS_VALU_FMA_F64"))
        waves = max(cv(q, "SQ_WAVES"), 1.0)
        t = avg * 1e-6
        out.append("%-34s %-12s %6d %9.1f %10.3f %8.0f %6.1f %10.4f %8.2f %6.1f %9.0f %8.0f" % (
            name[:34], grid, len(d), avg, hbm / 1e6, hbm / t / 1e9, 100.0 * hbm / t / HBM_PEAK, flops / 1e9,
            flops / t / 1e12, 100.0 * flops / t / FP64_PEAK, cv(q, "SQ_INSTS_VALU") / waves,
            cv(q, "SQ_INSTS_LDS") / waves))
    return "\n".join(out)


def steps(trace_csv, start, contains, which):
    rows = list(csv.DictReader(open(trace_csv)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    idx = [i for i, r in enumerate(rows) if start in r["Kernel_Name"]]
    spans = [(a, b) for a, b in zip(idx, idx[1:])
             if not contains or any(contains in r["Kernel_Name"] for r in rows[a:b])]
    if len(spans) < which:
        return "fewer than %d complete steps" % which
    a, b = spans[-which]
    t0 = int(rows[a]["Start_Timestamp"])
    out, tot = [], 0
    for r in rows[a:b]:
        s0, e0 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        tot += e0 - s0
        g = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        out.append("  %-44s %10d  start %9.1f us  %8.1f us" % (short(r["Kernel_Name"])[:44], g, (s0 - t0) / 1e3,
                                                                (e0 - s0) / 1e3))
    out.append("  sum of kernel durations %.1f us; first start to next step %.1f us (%d steps seen)"
               % (tot / 1e3, (int(rows[b]["Start_Timestamp"]) - t0) / 1e3, len(spans)))
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["kernels", "counters", "busy", "traffic", "stages", "steps"])
    ap.add_argument("files", nargs="+")
    ap.add_argument("--match")
    ap.add_argument("--top", type=int, default=40)
    ap.add_argument("--kernel")
    ap.add_argument("--grid")
    ap.add_argument("--probe-bytes", type=int, default=256 << 20)
    ap.add_argument("--start")
    ap.add_argument("--contains")
    ap.add_argument("--which", type=int, default=2)
    a = ap.parse_args()
    if a.what == "kernels":
        print(kernels(a.files[0], a.match, a.top))
    elif a.what == "counters":
        print(counters(a.files[0], a.match))
    elif a.what == "busy":
        print(busy(a.files[0], a.match))
    elif a.what == "traffic":
        print(traffic(a.files[0], a.files[1], a.kernel, a.grid, a.probe_bytes))
    elif a.what == "steps":
        print(steps(a.files[0], a.start, a.contains, a.which))
    else:
        print(stages(a.files[0], a.files[1], a.files[2], a.files[3], a.match, a.probe_bytes))


if __name__ == "__main__":
    main()
