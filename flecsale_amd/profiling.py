"""HBM-traffic measurement for the step kernels with rocprofv3 hardware counters.

Two passes, each under its own rocprofv3 run (the counters do not fit one pass):

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d prof/pmc_fetch -o run -- \\
        python3 -m flecsale_amd.profiling run
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d prof/pmc_write -o run -- \\
        python3 -m flecsale_amd.profiling run
    python3 -m flecsale_amd.profiling summarize prof/pmc_fetch prof/pmc_write --workload muscl2d

`run` first launches a plain device-to-device copy of a known size (512 MiB,
larger than the 256 MiB Infinity Cache) so each counter can be calibrated
against bytes that are known exactly, then a few eager steps of the workload.
`summarize` reads both databases, scales each kernel's per-dispatch counter by
(known copy bytes / counted copy bytes) and writes profiles/<workload>_traffic.json
(HBM bytes per launch, by kernel) and a markdown table.
"""
import argparse
import json
import os
import re
import sqlite3
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COPY_BYTES = 512 << 20   # per direction


def _copy_calibration():
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so")
    n = COPY_BYTES
    a, b = ct.c_void_p(), ct.c_void_p()
    assert hip.hipMalloc(ct.byref(a), ct.c_size_t(n)) == 0
    assert hip.hipMalloc(ct.byref(b), ct.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ct.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    for _ in range(3):   # hipMemcpyDeviceToDevice = 3
        assert hip.hipMemcpy(b, a, ct.c_size_t(n), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def run(workload, steps, size):
    sys.path.insert(0, ROOT)
    import flecsale_amd as fa
    from flecsale_amd import inputs
    print("[profiling] copy calibration", flush=True)
    _copy_calibration()
    if workload == "muscl2d":
        prob = inputs.quadrants_2d(size or 1024, steps=steps, flux=fa.FLUX_HLLC | fa.BOUNDARY_TRANSMISSIVE)
        ctx = fa.EulerContext(prob["coords"], prob["cells"])
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], steps, flux_kind=prob["flux"], use_graph=False)
    elif workload == "muscl3d":
        prob = inputs.kelvin_helmholtz_3d(size or 128, steps=steps)
        ctx = fa.EulerContext(prob["coords"], prob["cells"])
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], steps, flux_kind=prob["flux"], use_graph=False)
    else:
        if workload == "maire3d":
            prob = inputs.maire_bubble_3d(size or 128, steps=steps)
        else:
            prob = inputs.maire_bubble_2d(size or 512, steps=steps)
        ctx = fa.MaireContext(prob["coords"], prob["cells"])
        ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
        ctx.set_time(0.0, prob["dt0"])
        ctx.advance(prob["cfl"], prob["final_time"], steps, use_graph=False)
    ctx.sync()
    print(f"[profiling] {workload}: {steps} steps done", flush=True)
    ctx.close()


def _short(name):
    return re.sub(r"\(.*", "", str(name).replace("fsale::(anonymous namespace)::", "").replace("void ", ""))


def _per_kernel(path, counter=None, reduce="mean"):
    """{kernel short name: (dispatches, mean (or max) counter value per dispatch)} from
    rocprofv3 output: a counter_collection.csv (or a directory holding one),
    or a rocpd SQLite database."""
    acc = {}
    if os.path.isdir(path):
        hits = [os.path.join(dp, f) for dp, _, fs in os.walk(path) for f in fs
                if f.endswith("counter_collection.csv")]
        if not hits:
            raise SystemExit(f"no counter_collection.csv under {path}")
        path = hits[0]
    if path.endswith(".csv"):
        import csv
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                if counter and r.get("Counter_Name") != counter:
                    continue
                name = _short(r.get("Kernel_Name", ""))
                disp = r.get("Dispatch_Id")
                acc.setdefault(name, {}).setdefault(disp, 0.0)
                acc[name][disp] += float(r["Counter_Value"])
    else:
        db = sqlite3.connect(path)
        cur = db.cursor()
        rows = cur.execute("select * from counters_collection").fetchall()
        cols = [d[0] for d in cur.description]
        ic = cols.index("kernel_name" if "kernel_name" in cols else "name")
        iv = cols.index("value")
        idisp = cols.index("dispatch_id") if "dispatch_id" in cols else None
        for r in rows:
            name = _short(r[ic])
            disp = r[idisp] if idisp is not None else None
            acc.setdefault(name, {}).setdefault(disp, 0.0)
            acc[name][disp] += float(r[iv])
    if reduce == "max":
        return {k: (len(v), max(v.values())) for k, v in acc.items()}
    return {k: (len(v), sum(v.values()) / len(v)) for k, v in acc.items()}


def summarize(fetch_db, write_db, workload):
    fetch = _per_kernel(fetch_db)
    write = _per_kernel(write_db)
    # calibration: the largest dispatch of the runtime's copy kernel is one of
    # the 512 MiB device-to-device copies (smaller copy dispatches are uploads)
    fmax = _per_kernel(fetch_db, reduce="max")
    wmax = _per_kernel(write_db, reduce="max")
    copy = [k for k in fetch if "copy" in k.lower()]
    if not copy:
        raise SystemExit("no copy kernel found for calibration")
    ck = max(copy, key=lambda k: fmax[k][1])
    f_scale = COPY_BYTES / (fmax[ck][1] * 1024.0)
    w_scale = COPY_BYTES / (wmax[ck][1] * 1024.0)
    table = {}
    for k in fetch:
        if k not in write or k in copy:
            continue
        table[k] = {"dispatches": fetch[k][0], "fetch_KiB": fetch[k][1], "write_KiB": write[k][1],
                    "hbm_bytes": fetch[k][1] * 1024.0 * f_scale + write[k][1] * 1024.0 * w_scale}
    out = {"copy_kernel": ck, "copy_bytes_per_direction": COPY_BYTES, "fetch_scale": f_scale,
           "write_scale": w_scale, "kernels": table}
    short = {}
    for k, v in table.items():
        m = re.match(r"k_(\w+?)(<|$)", k)
        if m:
            short[m.group(1)] = v["hbm_bytes"]
    out.update(short)
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    w = workload
    lines = [f"# rocprofv3 HBM traffic per launch: {w}", "", "Commands (repository root, one MI355X):", "", "```",
             f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d {fetch_db} -o run -- "
             f"python3 -m flecsale_amd.profiling run --workload {w} --steps 3",
             f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d {write_db} -o run -- "
             f"python3 -m flecsale_amd.profiling run --workload {w} --steps 3",
             f"python3 -m flecsale_amd.profiling summarize {fetch_db} {write_db} "
             f"--workload {w}", "```", "",
             "FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what they read for the runtime's "
             "512 MiB device-to-device copy (scale factors in the first line).", "",
             f"copy kernel {ck}: fetch scale {f_scale:.4f}, write scale {w_scale:.4f}", "",
             "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes (calibrated) |",
             "|---|---|---|---|---|"]
    for k, v in sorted(table.items(), key=lambda kv: -kv[1]["hbm_bytes"]):
        lines.append(f"| {k} | {v['dispatches']} | {v['fetch_KiB']:.0f} | {v['write_KiB']:.0f} | "
                     f"{v['hbm_bytes']:.4e} |")
    with open(os.path.join(ROOT, "profiles", f"{workload}_traffic.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def _window_dispatches(trace_dir, anchor, last):
    """Dispatches of a rocprofv3 kernel_trace.csv inside the time span of the
    last `last` dispatches of the kernel whose name starts with `anchor`, as
    rows shaped like kernel_stats.csv (one per kernel)."""
    import csv
    hits = [os.path.join(dp, f) for dp, _, fs in os.walk(trace_dir) for f in fs if f.endswith("kernel_trace.csv")]
    if not hits:
        raise SystemExit(f"no kernel_trace.csv under {trace_dir}")
    ev = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
          for r in csv.DictReader(open(hits[0], newline=""))]
    ev.sort()
    mine = [e for e in ev if _short(e[2]).startswith(anchor)]
    if not mine:
        raise SystemExit(f"no dispatch of {anchor}")
    t0, t1 = mine[-last:][0][0], mine[-1][1]
    per = {}
    for b, e, name in ev:
        if t0 <= b <= t1:
            per.setdefault(name, []).append(e - b)
    return [{"Name": n, "Calls": len(d), "AverageNs": sum(d) / len(d), "MinNs": min(d), "MaxNs": max(d),
             "TotalDurationNs": sum(d)} for n, d in per.items()]


def trace(trace_dir, out_path, command, last=None, anchor=None):
    """Markdown summary of a rocprofv3 --kernel-trace --stats run: calls,
    average and total time per kernel, over the whole run (kernel_stats.csv)
    or (last, anchor) over the span of the anchor kernel's last `last`
    dispatches (kernel_trace.csv): the benchmark's own timed replay, after its
    launch-shape autotuning."""
    import csv
    if last:
        rows = _window_dispatches(trace_dir, anchor, last)
    else:
        hits = [os.path.join(dp, f) for dp, _, fs in os.walk(trace_dir) for f in fs if f.endswith("kernel_stats.csv")]
        if not hits:
            raise SystemExit(f"no kernel_stats.csv under {trace_dir}")
        rows = list(csv.DictReader(open(hits[0], newline="")))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    scope = ([f"Dispatches within the span of the last {last} `{anchor}` dispatches: the benchmark's profiled "
              "replay of its timed steps, at the launch shapes its autotuner chose (the earlier dispatches "
              "include the autotuner's candidates).", ""] if last else [])
    lines = [f"# rocprofv3 kernel trace: {os.path.basename(out_path).rsplit('.', 1)[0]}", "",
             "Command (from the repository root on one MI355X):", "", "```", command, "```", ""] + scope + [
             "| kernel | calls | avg us | min us | max us | total ms | % |", "|---|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        lines.append(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.3f} | {100 * float(r['TotalDurationNs']) / total:.1f} |")
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("trace_dir")
    t.add_argument("out")
    t.add_argument("--command", required=True,
                   help="the full rocprofv3 command that made the trace, written into the summary")
    t.add_argument("--last", type=int, default=None,
                   help="summarise the span of the anchor kernel's last N dispatches")
    t.add_argument("--anchor", default=None, help="kernel name prefix, e.g. k_eul_fluxupd")
    r = sub.add_parser("run")
    r.add_argument("--workload", default="muscl2d", choices=["muscl2d", "muscl3d", "maire2d", "maire3d"])
    r.add_argument("--steps", type=int, default=3)
    r.add_argument("--size", type=int, default=None)
    s = sub.add_parser("summarize")
    s.add_argument("fetch_db")
    s.add_argument("write_db")
    s.add_argument("--workload", default="muscl2d")
    a = ap.parse_args()
    if a.cmd == "run":
        run(a.workload, a.steps, a.size)
    elif a.cmd == "trace":
        trace(a.trace_dir, a.out, a.command, a.last, a.anchor)
    else:
        summarize(a.fetch_db, a.write_db, a.workload)


if __name__ == "__main__":
    main()
