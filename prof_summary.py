"""Summaries of rocprofv3 CSV output for the committed measurement files.

Kernel trace (``--kernel-trace --stats``)::

    python prof_summary.py trace DIR/NAME_kernel_trace.csv DIR/NAME_kernel_stats.csv \
        --title TITLE --command "rocprofv3 ..." [--notes FILE] [--group NAME=REGEX ...] > SUMMARY.md

HBM traffic (one ``--pmc FETCH_SIZE`` and one ``--pmc WRITE_SIZE`` run of the workload, and the same two
of the plain copy kernel of known size for the counter calibration)::

    python prof_summary.py traffic FETCH.csv WRITE.csv COPY_FETCH.csv COPY_WRITE.csv \
        --copy-bytes BYTES --op NAME=REGEX=LAUNCHES_PER_OP ... [--bench BENCH.json] \
        [--workload-key JSON] [--json OUT.json] --commands FILE > SUMMARY.md

FETCH_SIZE and WRITE_SIZE are in KiB per dispatch; each is scaled by (true bytes of the copy) / (counter
of the copy), so the figures are bytes as the plain copy kernel counts them.  On gfx950 FETCH_SIZE
counts a 128-byte read request as 64 bytes (a copy of 128-byte or longer runs reads 0.5x, one of 32- or
64-byte runs 1.0x the true bytes), so one scale factor fits only kernels with one request size; with
``--fetch-requests`` the fetch runs are ``--pmc TCC_EA0_RDREQ_32B TCC_EA0_RDREQ_64B TCC_EA0_RDREQ_128B
TCC_EA0_RDREQ`` instead and each request counts its own size.

Shader counters (``--pmc`` runs of the workload, each with its own counter set, beside its kernel trace)::

    python prof_summary.py counters DIR/NAME_kernel_trace.csv PMC1.csv PMC2.csv ... --kernel REGEX ... \
        --title TITLE --commands FILE > SUMMARY.md

On gfx950 the SQ cycle counters (SQ_WAVE_CYCLES, SQ_BUSY_CYCLES, SQ_ACTIVE_INST_*, SQ_WAIT_*) count
quad-cycles, and GRBM_GUI_ACTIVE counts cycles summed over the 8 XCDs: the kernel's cycles under the
counters are GRBM_GUI_ACTIVE / 8, the mean resident waves per SIMD are 4 SQ_WAVE_CYCLES / (cycles x
1024 SIMDs), and the share of SIMD cycles issuing vector instructions is 4 SQ_ACTIVE_INST_VALU /
(cycles x 1024).
"""

import argparse
import collections
import csv
import json
import re
import sys


def short_name(name):
    """'void pms::(anonymous namespace)::k_conv_cols<13, true>(...)' -> 'k_conv_cols<13, true>'."""
    m = re.search(r"(k_\w+(<[^()]*>)?)\(", name)
    return m.group(1) if m else name[:80]


def load_counters(path, counter):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"].startswith(counter):
                rows.append((short_name(r["Kernel_Name"]), int(r["Grid_Size"]), float(r["Counter_Value"])))
    return rows


def trace(args):
    per = collections.defaultdict(list)
    grid = {}
    with open(args.kernel_trace) as f:
        for r in csv.DictReader(f):
            n = short_name(r["Kernel_Name"])
            per[n].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0)
            grid[n] = "%s x %s" % (r["Grid_Size_X"], r.get("Workgroup_Size_X", ""))
    out = ["# Kernel trace: %s" % args.title, "", "Command (repo root, on an MI355X):", "", "```",
           args.command, "```", ""]
    if args.notes:
        out += [open(args.notes).read().rstrip(), ""]
    if args.group:
        total = sum(sum(v) for v in per.values())
        out += ["Kernel time by family (busy time = sum of kernel durations, %.1f ms in all):" % (total / 1e3), "",
                "| family | kernels | busy ms | share |", "|---|---|---|---|"]
        left = dict(per)
        for spec in args.group + ["other=."]:
            name, regex = spec.split("=", 1)
            sel = [k for k in left if re.search(regex, k)]
            t = sum(sum(left.pop(k)) for k in sel)
            if sel:
                out.append("| %s | %s | %.2f | %.1f%% |" % (name, ", ".join(sorted(set(re.sub("<.*", "", k) for k in sel))),
                                                             t / 1e3, 100.0 * t / total))
        out.append("")
    out += ["| kernel | grid threads x block | calls | avg us | min us | max us |", "|---|---|---|---|---|---|"]
    for n, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        out.append("| %s | %s | %d | %.1f | %.1f | %.1f |" % (n, grid[n], len(v), sum(v) / len(v), min(v), max(v)))
    out += ["", "rocprofv3 --stats summary (kernel_stats.csv):", "", "| name | calls | avg ns | total ns | % |",
            "|---|---|---|---|---|"]
    with open(args.kernel_stats) as f:
        for r in csv.DictReader(f):
            out.append("| %s | %s | %.1f | %s | %s |" % (short_name(r["Name"]), r["Calls"], float(r["AverageNs"]),
                                                         r["TotalDurationNs"], r["Percentage"]))
    print("\n".join(out))


def load_request_bytes(path):
    """Read bytes per dispatch from a run of TCC_EA0_RDREQ_32B/_64B/_128B: 32, 64 and 128 bytes per
    request (FETCH_SIZE's expression counts a 128-byte request as 64 bytes on gfx950), in KiB."""
    per = collections.defaultdict(dict)
    meta = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            d = r["Dispatch_Id"]
            per[d][r["Counter_Name"]] = float(r["Counter_Value"])
            meta[d] = (short_name(r["Kernel_Name"]), int(r["Grid_Size"]))
    rows = []
    for d, c in per.items():
        b = 32 * c.get("TCC_EA0_RDREQ_32B", 0.0) + 64 * c.get("TCC_EA0_RDREQ_64B", 0.0) + \
            128 * c.get("TCC_EA0_RDREQ_128B", 0.0)
        rows.append((meta[d][0], meta[d][1], b / 1024.0))
    return rows


def traffic(args):
    if args.fetch_requests:
        fetch = load_request_bytes(args.fetch)
        cf = [v for n, g, v in load_request_bytes(args.copy_fetch) if n.startswith("k_copy")]
    else:
        fetch = load_counters(args.fetch, "FETCH_SIZE")
        cf = [v for n, g, v in load_counters(args.copy_fetch, "FETCH_SIZE") if n.startswith("k_copy")]
    write = load_counters(args.write, "WRITE_SIZE")
    cw = [v for n, g, v in load_counters(args.copy_write, "WRITE_SIZE") if n.startswith("k_copy")]
    if not cf or not cw:
        sys.exit("no k_copy dispatches in the calibration runs")
    kib_f, kib_w = sum(cf) / len(cf), sum(cw) / len(cw)
    f_scale = args.copy_bytes / (kib_f * 1024.0)
    w_scale = args.copy_bytes / (kib_w * 1024.0)
    fetch_name = "(from its read requests)" if args.fetch_requests else "FETCH_SIZE ="

    def table(rows):
        t = collections.defaultdict(list)
        for n, g, v in rows:
            t[(n, g)].append(v)
        return t

    tf, tw = table(fetch), table(write)
    out = ["# HBM traffic per launch: %s" % args.title, "",
           "Commands (repo root, on an MI355X), one counter per pass:", "",
           "```", open(args.commands).read().rstrip(), "```", "",
           "Calibration: the plain copy kernel (k_copy) moving %d bytes each way reads %s %.1f KiB and "
           "WRITE_SIZE = %.1f KiB per dispatch, so the scale factors to true bytes are **fetch x%.3f** and "
           "**write x%.3f**." % (args.copy_bytes, fetch_name, kib_f, kib_w, f_scale, w_scale), "",
           "| kernel | grid | dispatches | %s KiB (raw avg) | WRITE KiB (raw avg) | traffic MB (calibrated) |"
           % ("read" if args.fetch_requests else "FETCH"),
           "|---|---|---|---|---|---|"]
    per_kernel = {}
    for key in sorted(set(tf) | set(tw), key=lambda k: -sum(tf.get(k, [0])) - sum(tw.get(k, [0]))):
        f = tf.get(key, [0.0])
        w = tw.get(key, [0.0])
        fb = sum(f) / len(f) * 1024 * f_scale
        wb = sum(w) / len(w) * 1024 * w_scale
        per_kernel[key] = (len(f), sum(f) * 1024 * f_scale, sum(w) * 1024 * w_scale)
        out.append("| %s | %d | %d | %.0f | %.0f | %.1f |" % (key[0], key[1], len(f), sum(f) / len(f), sum(w) / len(w),
                                                             (fb + wb) / 1e6))
    bench_ops = {}
    if args.bench:
        with open(args.bench) as fb_:
            bench_ops = json.loads(fb_.read().strip().splitlines()[-1]).get("ops", {})
    ops = {}
    out += ["", "Per operation (kernels of the operation summed, divided by the operation count):", ""]
    for spec in args.op:
        name, regex, per_op = spec.split("=")
        per_op = int(per_op)
        sel = [k for k in per_kernel if re.search(regex, k[0])]
        n_disp = sum(per_kernel[k][0] for k in sel)
        fbytes = sum(per_kernel[k][1] for k in sel)
        wbytes = sum(per_kernel[k][2] for k in sel)
        n_ops = n_disp / per_op
        rec = {"kernels": sorted(set(k[0] for k in sel)), "launches_per_op": per_op,
               "fetch_bytes": fbytes / n_ops, "write_bytes": wbytes / n_ops,
               "traffic_bytes": (fbytes + wbytes) / n_ops}
        line = "* %s (%s; %d launches per operation): fetch %.1f MB + write %.1f MB = **%.1f MB** per operation" % (
            name, ", ".join(rec["kernels"]), per_op, rec["fetch_bytes"] / 1e6, rec["write_bytes"] / 1e6,
            rec["traffic_bytes"] / 1e6)
        if name in bench_ops and "alg_bytes" in bench_ops[name]:
            alg = bench_ops[name]["alg_bytes"]
            rec["alg_bytes"] = alg
            line += " vs %.1f MB algorithmic (%.2fx)" % (alg / 1e6, rec["traffic_bytes"] / alg)
        out.append(line)
        ops[name] = rec
    print("\n".join(out))
    if args.json:
        doc = {"workload": json.loads(args.workload_key) if args.workload_key else None,
               "calibration": {"copy_bytes": args.copy_bytes, "fetch_scale": f_scale, "write_scale": w_scale,
                               "fetch_counter": "TCC_EA0_RDREQ_32B/64B/128B" if args.fetch_requests else "FETCH_SIZE"},
               "ops": ops}
        with open(args.json, "w") as f:
            json.dump(doc, f, indent=1)


SIMDS = 1024  # 256 CUs x 4 SIMDs
VALU_ISSUE = SIMDS * 2.4e9 / 4  # wave64 vector instructions per second without packed FP32


def counters(args):
    dur = collections.defaultdict(list)
    with open(args.kernel_trace) as f:
        for r in csv.DictReader(f):
            dur[short_name(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0)
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    res = {}
    for path in args.pmc:
        with open(path) as f:
            for r in csv.DictReader(f):
                n = short_name(r["Kernel_Name"])
                vals[n][r["Counter_Name"]].append(float(r["Counter_Value"]))
                res[n] = (r["Workgroup_Size"], r["VGPR_Count"], r["Accum_VGPR_Count"], r["SGPR_Count"])
    names = [n for n in vals if any(re.search(k, n) for k in args.kernel)]
    names.sort(key=lambda n: -sum(dur.get(n, [0.0])) )
    out = ["# Shader counters: %s" % args.title, "", "Commands (repo root, on an MI355X):", "", "```",
           open(args.commands).read().rstrip(), "```", "",
           "Per kernel: the trace's average duration, the counters' average per dispatch, and what they imply "
           "(SQ cycle counters in quad-cycles; the kernel's cycles under the counters = GRBM_GUI_ACTIVE / 8).", "",
           "| kernel | block, VGPR | avg us (trace) | VALU wave-instr | VALU floor us | SIMD VALU busy | "
           "waves/SIMD | LDS instr | LDS-wait share | bank-conflict cycles / CU | any-wait share |",
           "|---|---|---|---|---|---|---|---|---|---|---|"]

    def avg(n, c):
        v = vals[n].get(c)
        return sum(v) / len(v) if v else float("nan")

    for n in names:
        d = dur.get(n, [float("nan")])
        us = sum(d) / len(d)
        cyc = avg(n, "GRBM_GUI_ACTIVE") / 8.0
        valu = avg(n, "SQ_INSTS_VALU")
        wave = avg(n, "SQ_WAVE_CYCLES")
        out.append("| %s | %s, %s | %.1f | %.3g | %.1f | %.2f | %.2f | %.3g | %.2f | %.0f | %.2f |" % (
            n, res[n][0], res[n][1], us, valu, valu / VALU_ISSUE * 1e6, 4 * avg(n, "SQ_ACTIVE_INST_VALU") / (cyc * SIMDS),
            4 * wave / (cyc * SIMDS), avg(n, "SQ_INSTS_LDS"), avg(n, "SQ_WAIT_INST_LDS") / wave,
            avg(n, "SQ_LDS_BANK_CONFLICT") / 256.0, avg(n, "SQ_WAIT_ANY") / wave))
    print("\n".join(out))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("kernel_trace")
    t.add_argument("kernel_stats")
    t.add_argument("--title", required=True)
    t.add_argument("--command", required=True)
    t.add_argument("--notes")
    t.add_argument("--group", action="append", default=[],
                   help="NAME=REGEX: sum the kernels whose short name matches into a family (first match wins)")
    r = sub.add_parser("traffic")
    for a in ("fetch", "write", "copy_fetch", "copy_write"):
        r.add_argument(a)
    r.add_argument("--copy-bytes", type=float, required=True)
    r.add_argument("--fetch-requests", action="store_true",
                   help="FETCH and COPY_FETCH are runs of TCC_EA0_RDREQ_32B/_64B/_128B instead of FETCH_SIZE")
    r.add_argument("--op", action="append", default=[])
    r.add_argument("--bench")
    r.add_argument("--title", required=True)
    r.add_argument("--commands", required=True)
    r.add_argument("--workload-key")
    r.add_argument("--json")
    c = sub.add_parser("counters")
    c.add_argument("kernel_trace")
    c.add_argument("pmc", nargs="+")
    c.add_argument("--kernel", action="append", required=True, help="REGEX on the kernel's short name")
    c.add_argument("--title", required=True)
    c.add_argument("--commands", required=True)
    args = ap.parse_args()
    {"trace": trace, "traffic": traffic, "counters": counters}[args.cmd](args)


if __name__ == "__main__":
    main()
