#!/usr/bin/env python3
"""Benchmark: zone-updates/s of the Eulerian MUSCL step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload muscl2d|muscl3d|maire2d|maire3d]

Default workload (the configuration the metric is quoted on): second-order
MUSCL (LSQ gradients + vertex limiter, HLLC flux) four-quadrant Riemann
problem on a 1024x1024 quad mesh of [0,1]^2 (1,048,576 cells), gamma 1.4,
FP64, 200 steps.  A "step" is one full time step (dt reduction + two
gradient/flux/update half steps) of the whole mesh; a zone-update is one cell
advanced one step.  The state is resident in HBM before the timed region
(inputs uploaded, warm-up steps run, state reset on the device).

Prints ONE JSON line on stdout.  `roofline` is the dominant kernel: its
algorithmic bytes per launch (task-dataflow compulsory bytes: each task reads
every array it consumes once and writes each output once) over its average
launch time, measured with HIP events on the context's stream in a profiled
replay of the same K steps.  `cpu_baseline` times the CPU restatement
(oracle/, OpenMP over the host cores) on a bounded sample of the same
workload.  With the default workload the line also carries
`other_configs`: the MUSCL3D 128^3 (100 steps) and Maire 512^2 (500 steps)
configurations measured the same way in the same run, and `serving`, the
4096-request served stream, batched against one context per request.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0   # MI355X HBM3E peak


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_bytes_euler(nc, nf, nv, dim):
    """Algorithmic bytes per launch of each Eulerian kernel (FP64 8 B, int32 4 B)."""
    if dim == 2:
        # dt: v, a, V, cell->face CSR | n, A ; gradients: 6 vars, centroid, neighbour CSR,
        # cell->vertex CSR, 12 gradients | coords ; fluxes: 6 vars + 12 gradients + centroid |
        # centroid, n, A, face->cell, 4 fluxes ; update: d, v, e in, 7 state out, V, CSR | flux, face->cell
        b = {0: 52 * nc + 24 * nf, 1: 216 * nc + 16 * nv, 2: 160 * nc + 80 * nf, 3: 116 * nc + 40 * nf}
    else:
        b = {0: 68 * nc + 32 * nf, 1: 392 * nc + 24 * nv, 2: 248 * nc + 104 * nf, 3: 140 * nc + 48 * nf}
    # The step loop's fused kernel does the flux and update tasks (and every
    # second launch the next step's dt) without the face flux array, the
    # face->cell map and the update's second read of the state between them:
    # its own compulsory bytes are, per cell, the state (N+4 values), the
    # gradients ((N+4)*N), centroid, volume, cell->face codes (int32) and the
    # new state written (d, v, e, p, T, a); per face, its centroid, normal and
    # area.  The sum of the two tasks' bytes is kept as "task_bytes".
    if dim == 2:
        b[10] = 240 * nc + 40 * nf
    else:
        b[10] = 344 * nc + 56 * nf
    b["task_10"] = b[2] + b[3] + b[0] // 2
    return b


def kernel_bytes_maire(nc, nv, ncn):
    """Algorithmic bytes per launch of each Lagrangian kernel.

    nodal: cell d, a, p, v + cell->vertex | vertex->corner CSR, coords, u_p out | corner npc, Fpc out;
    residual: cell->vertex, dU out | u_p | npc, Fpc; dt: a, min_length, dV, V;
    move: x0, u_p in, x out; update: geometry (cell->vertex, coords, V/centroid/min_length out)
    + state (M, v, e, dU, gamma in; v, e, V, d, p, T, a out); save/restore: x, v, e copies."""
    return {4: 56 * nc + 36 * nv + 36 * ncn, 5: 48 * nc + 16 * nv + 32 * ncn, 6: 32 * nc,
            7: 48 * nv, 8: 184 * nc + 16 * nv, 9: 48 * nc + 32 * nv}


def kernel_bytes_maire3(nc, nv, ncn):
    """The same model for hexes (3 velocity components, 8 corners of 12 doubles'
    worth of corner output per cell): nodal 80 B/cell (d, a, p, v, cell->vertex),
    52 B/vertex (coords, u_p out, corner count), 52 B/corner (vertex->corner
    entry, npc and Fpc out); residual 72 B/cell, 24 B/vertex, 48 B/corner; move
    72 B/vertex; update 232 B/cell + 24 B/vertex."""
    return {4: 80 * nc + 52 * nv + 52 * ncn, 5: 72 * nc + 24 * nv + 48 * ncn, 6: 32 * nc,
            7: 72 * nv, 8: 232 * nc + 24 * nv, 9: 64 * nc + 48 * nv}


def load_traffic(workload):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if present."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_native_oracle():
    """The CPU restatement built for this host's cores (-O3 -march=native),
    compiled here, on the machine that times it."""
    import subprocess
    import oracle as O
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "liboracle_native.so"], check=True)
    O.load(os.path.join(ROOT, "oracle", "liboracle_native.so"))
    return O


def cpu_baseline(prob, kind, budget_s=8.0):
    """The CPU restatement (OpenMP over the host cores, then one core) on a
    bounded number of steps of the same workload."""
    O = load_native_oracle()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = _cpu_sample(O, prob, kind, budget_s, threads)
    one = _cpu_sample(O, prob, kind, budget_s, 1)
    out["one_core"] = {"value": one["value"], "sample": one["sample"]}
    out["cpu_model"] = cpu_model()
    out["build"] = "gcc -O3 -march=native -ffp-contract=off -fopenmp"
    return out


def _cpu_sample(O, prob, kind, budget_s, threads):
    O.set_threads(threads)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    nc = len(prob["cells"])
    if kind == "euler":
        s = O.euler_initial_state(mesh, prob["eos"][0], prob["eos"][1], prob["d"], prob["v"], prob["p"])
        steps, t = 0, 0.0
        t0 = time.perf_counter()
        while steps < 2 or (time.perf_counter() - t0 < budget_s and steps < prob["max_steps"]):
            rc, t, n = O.euler_run(mesh, s, prob["eos"][0], prob["eos"][1], prob["cfl"], prob["final_time"], 1,
                                   flux_kind=prob["flux"], soln_time=t)
            steps += 1
        el = time.perf_counter() - t0
    else:
        s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
        steps, t, dt = 0, 0.0, prob["dt0"]
        t0 = time.perf_counter()
        while steps < 2 or (time.perf_counter() - t0 < budget_s and steps < prob["max_steps"]):
            rc, t, n, dt = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], 1, dt, soln_time=t)
            steps += 1
        el = time.perf_counter() - t0
    return dict(value=nc * steps / el, unit="zone-updates/s", cores=threads, kind="port",
                sample=f"first {steps} steps of the same workload ({nc} cells), {el:.1f} s, "
                       f"{threads} OpenMP thread{'s' if threads > 1 else ''}")


# ------------------------------------------------------------------ serving
SERVE_KINDS = ("euler", "maire")


def serve_requests(inputs, n, steps):
    """The mixed-size stream's problems, built before any timing (the load
    generator's own work stays out of the served loop)."""
    sides = inputs.stream_sides(n)
    return [(SERVE_KINDS[i % 2], inputs.riemann_2d(int(sides[i]), SERVE_KINDS[i % 2], index=i, steps=steps))
            for i in range(n)]


def serve_trace(fa, requests, mode, arrivals=None, capacity=None, window=1024, graphs=True):
    """Serves `requests` (solver kind, problem) through fsale_server in `mode`
    (0 batched, 1 one plain context per request).  arrivals: seconds after
    the start (Poisson trace), or None to keep up to `window` requests queued
    or running (saturated: more than both batches hold at once).  Requests are
    submitted when they arrive, while the previous scheduling round runs."""
    n = len(requests)
    solver = {"euler": fa.SOLVER_EULER, "maire": fa.SOLVER_MAIRE}
    # resident capacity per batch: ~170 average requests (47K cells each)
    cap = capacity or (8 << 20, 17 << 20, 9 << 20, 66 << 20)
    with fa.Server(mode=mode, max_problems=1024, euler=cap, maire=cap, chunk_steps=10, lookahead=64,
                   frac_hbm=0.5, graphs=graphs) as srv:
        t0 = srv.clock()
        ids, cells = [], []
        i = done = pending = 0
        while done < n:
            now = srv.clock() - t0
            while i < n and ((arrivals is not None and arrivals[i] <= now) or
                             (arrivals is None and pending < window)):
                kind, prob = requests[i]
                at = t0 + (arrivals[i] if arrivals is not None else now)
                ids.append(srv.submit(prob, solver[kind], arrival=at))
                cells.append(len(prob["cells"]))
                i += 1
                pending += 1
            c, pending = srv.poll()
            done += c
            if pending == 0 and i < n and arrivals is not None:
                time.sleep(max(0.0, min(0.005, arrivals[i] - (srv.clock() - t0))))
        makespan = srv.clock() - t0
        host = srv.stats() if mode == 0 else None
        lat, zu, errors = [], 0, 0
        for rid, nc in zip(ids, cells):
            a, _, f = srv.times(rid)
            st, _ = srv.result(rid, fields=False)
            lat.append(f - a)
            zu += nc * st.steps
            errors += int(st.error != 0)
    lat = np.array(lat)
    return {"requests": n, "makespan_s": makespan, "zone_updates": zu, "zu_per_s": zu / makespan,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "errors": errors, "host": host}


def serve_stream(fa, inputs, n, steps=100, calib=24, one_saturated=False, graphs=True, modes=("batched", "one_at_a_time"),
                 cpu=False):
    """The mixed-size request stream: one-at-a-time service rate mu from the
    first `calib` requests; a Poisson trace at 2 mu served both ways (latency
    under the same saturated trace); and both ways with the queue kept full
    (throughput capacity).  cpu: also the same trace replayed on the host
    cores (cpu_replay)."""
    from flecsale_amd.inputs import stream_arrivals
    t_gen = time.perf_counter()
    requests = serve_requests(inputs, n, steps)
    log(f"[serve] {n} requests built in {time.perf_counter() - t_gen:.1f}s")
    # register every distinct mesh once (connectivity cache) before any timing,
    # so both serving modes start from the same warm cache
    t_reg = time.perf_counter()
    seen = set()
    for _, prob in requests:
        key = id(prob["cells"])
        if key not in seen:
            seen.add(key)
            fa.connectivity(prob["coords"], prob["cells"])
    log(f"[serve] {len(seen)} distinct meshes registered in {time.perf_counter() - t_reg:.1f}s")
    cal = serve_trace(fa, requests[:calib], 1, graphs=graphs)
    mu = calib / cal["makespan_s"]
    rate = 2 * mu
    arrivals = stream_arrivals(n, rate)
    log(f"[serve] one-at-a-time service rate {mu:.1f} req/s; Poisson trace at {rate:.1f} req/s")
    out = {"requests": n, "steps_per_request": steps, "arrival_rate_per_s": rate,
           "stream": "two-state Riemann problems, side round(exp(U[ln 32, ln 512])), alternating MUSCL "
                     "(HLLE, walls) / Maire (symmetry walls), gamma 1.4 or 1.67"}
    for name, mode in (("batched", 0), ("one_at_a_time", 1)):
        if name not in modes:
            continue
        r = serve_trace(fa, requests, mode, arrivals=arrivals, graphs=graphs)
        log(f"[serve] {name} trace: {r['zu_per_s']:.3e} ZU/s, p50 {r['p50_s']*1e3:.1f} ms, "
            f"p99 {r['p99_s']*1e3:.1f} ms, makespan {r['makespan_s']:.2f} s")
        out[name + "_trace"] = r
        if mode == 1 and not one_saturated:
            # requests arrive at twice its service rate: from the first
            # arrival on it is never idle, so the trace is its saturated run
            continue
        r = serve_trace(fa, requests, mode, graphs=graphs)
        log(f"[serve] {name} saturated: {r['zu_per_s']:.3e} ZU/s, makespan {r['makespan_s']:.2f} s"
            + (f", host {json.dumps({k: round(v, 3) for k, v in r['host'].items()})}" if r["host"] else ""))
        out[name + "_saturated"] = r
    if cpu:
        log("[serve] timing the CPU restatement on the first requests of the stream")
        base = stream_cpu_baseline(inputs, inputs.stream_sides(n), steps)
        out["cpu_baseline"] = {k: v for k, v in base.items() if k != "per_kind"}
        out["cpu_replay"] = cpu_replay(requests, arrivals, steps, base)
        log(f"[serve] cpu replay: {out['cpu_replay']['zu_per_s']:.3e} ZU/s, "
            f"p50 {out['cpu_replay']['p50_s']:.1f} s, p99 {out['cpu_replay']['p99_s']:.1f} s")
    if "one_at_a_time" not in modes:
        return out
    one = out.get("one_at_a_time_saturated", out["one_at_a_time_trace"])
    out["throughput_ratio"] = out["batched_saturated"]["zu_per_s"] / one["zu_per_s"]
    out["throughput_ratio_basis"] = ("batched saturated / one at a time saturated" if one_saturated else
                                     "batched saturated / one at a time under the 2x-service-rate trace")
    return out


def measure(args, fa, inputs, workload, steps=None, size=None, with_cpu=True):
    """One workload on one context: warm-up, the timed K steps (graph replay),
    a profiled replay of the same steps for the per-kernel times, the CPU
    restatement on a bounded sample.  Returns the bench line's fields."""
    t_build = time.perf_counter()
    if workload == "muscl2d":
        n = size or 1024
        K = steps or 200
        prob = inputs.quadrants_2d(n, steps=K, flux=fa.FLUX_HLLC | fa.BOUNDARY_TRANSMISSIVE)
        kind, dim = "euler", 2
        config = {"workload": f"MUSCL2D four-quadrant Riemann problem, {n}x{n} quads of [0,1]^2, "
                              f"LSQ gradients + Barth-Jespersen vertex limiter, HLLC, CFL 0.5, gamma 1.4, "
                              f"transmissive boundaries (with reflecting walls the quadrant problem "
                              f"reaches a negative state at step 6)",
                  "cells": n * n, "steps_per_run": K}
    elif workload == "muscl3d":
        n = size or 128
        K = steps or 100
        prob = inputs.kelvin_helmholtz_3d(n, steps=K)
        kind, dim = "euler", 3
        config = {"workload": f"MUSCL3D Kelvin-Helmholtz, {n}^3 hexes jittered +-10%, HLLE, CFL 1/3, "
                              f"reflecting walls",
                  "cells": n ** 3, "steps_per_run": K}
    elif workload == "maire3d":
        n = size or 128
        K = steps or 100
        prob = inputs.maire_bubble_3d(n, steps=K)
        kind, dim = "maire", 3
        config = {"workload": f"Maire cell-centred Lagrangian, {n}^3 hexes jittered +-15%, two ideal gases "
                              f"(bubble gamma 1.67), left-driven shock, symmetry walls on six sides",
                  "cells": n ** 3, "steps_per_run": K}
    else:
        n = size or 512
        K = steps or 500
        prob = inputs.maire_bubble_2d(n, steps=K)
        kind, dim = "maire", 2
        config = {"workload": f"Maire cell-centred Lagrangian, {n}x{n} quads jittered +-15%, two ideal "
                              f"gases (bubble gamma 1.67), left-driven shock, symmetry walls",
                  "cells": n * n, "steps_per_run": K}
    log(f"[bench] inputs {config['workload']} built in {time.perf_counter() - t_build:.1f}s")
    W = args.warmup

    t0 = time.perf_counter()
    if kind == "euler":
        ctx = fa.EulerContext(prob["coords"], prob["cells"])

        def reset():
            ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])

        def run(k, graph=True):
            ctx.advance(prob["cfl"], prob["final_time"], k, flux_kind=prob["flux"], use_graph=graph)
    else:
        ctx = fa.MaireContext(prob["coords"], prob["cells"])

        def reset():
            ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
            ctx.set_time(0.0, prob["dt0"])

        def run(k, graph=True):
            ctx.advance(prob["cfl"], prob["final_time"], k, use_graph=graph)
    log(f"[bench] context (connectivity + upload + geometry) {time.perf_counter() - t0:.1f}s")

    reset()
    if args.shape is None:
        run(1, graph=False)      # sets the step parameters the tuner runs with
        t0 = time.perf_counter()
        tuned_ms = ctx.autotune(3)
        log(f"[bench] autotuned launch shapes in {time.perf_counter() - t0:.1f}s: "
            f"{json.dumps(ctx.get_launch_shape())} (kernel time {tuned_ms:.3f} ms/step)")
        reset()
    else:
        ctx.launch_shape(args.shape)
    # The step loop replays a captured graph of min(K, 10) steps: the warm-up
    # runs at least that many steps so the timed call replays a graph captured
    # and instantiated before the clock starts.
    W = max(W, min(K, 10)) if W else 0
    if W:
        run(W)
    ctx.sync()
    reset()
    ctx.sync()
    log(f"[bench] warm-up {W} steps done; timing {K} steps")
    t0 = time.perf_counter()
    run(K)
    ctx.sync()
    el = time.perf_counter() - t0
    nc = ctx.nc
    value = nc * K / el
    t_sim, dt_last, steps_done = ctx.time()
    if steps_done != K or ctx.error_cell() != -1:
        raise SystemExit(f"timed run ended after {steps_done} of {K} steps (error cell {ctx.error_cell()})")
    log(f"[bench] {K} steps in {el * 1e3:.2f} ms: {value:.4e} zone-updates/s (t = {t_sim:.6g})")

    # byte model
    if kind == "euler":
        tasks = kernel_bytes_euler(nc, ctx.nf, ctx.nv, dim)
        # task model: one dt, two (gradients, fluxes, update) per step
        bytes_per_step = tasks[0] + 2 * (tasks[1] + tasks[2] + tasks[3])
        # kernels the step loop launches: gradients (the step's first launch
        # also folds the time-step partials into dt), fused fluxes + update
        # (+ the next step's dt partials)
        kb = {1: tasks[1], 10: tasks[10]}
        task_bytes = {10: tasks["task_10"]}
    else:
        kb = (kernel_bytes_maire if dim == 2 else kernel_bytes_maire3)(nc, ctx.nv, nc * ctx.vpc)
        launches_per_step = {4: 2, 5: 2, 6: 1, 7: 2, 8: 2, 9: 2}
        bytes_per_step = sum(kb[k] * launches_per_step[k] for k in kb)
        task_bytes = {}

    roofline = None
    kernels = {}
    if not args.no_profile:
        reset()
        ctx.reset_profile()
        ctx.set_profiling(True)
        run(K, graph=False)
        ctx.sync()
        ctx.set_profiling(False)
        for k in kb:
            tot, nl = ctx.profile(k)
            if k == fa.K_MAIRE_NODAL:
                # the nodal task is two launches: interior vertices and the
                # boundary vertices' symmetry solves
                tot += ctx.profile(fa.K_MAIRE_NODAL_BND)[0]
            if nl:
                avg_s = tot / nl * 1e-3
                kernels[fa.KERNEL_NAMES[k]] = {"avg_us": avg_s * 1e6, "launches": nl,
                                               "alg_bytes": kb[k], "GBps": kb[k] / avg_s / 1e9}
                if k in task_bytes:
                    kernels[fa.KERNEL_NAMES[k]]["task_bytes"] = task_bytes[k]
        dom = max(kernels, key=lambda name: kernels[name]["avg_us"] * kernels[name]["launches"])
        d = kernels[dom]
        traffic = load_traffic(workload)
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(d["GBps"], 1), "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": round(d["GBps"] / PEAK_HBM_GBS, 4),
                    "traffic": next((traffic[k] for k in (dom, dom + "_both", dom + "_brick")
                                     if traffic and traffic.get(k)), None)}
        log("[bench] per-kernel: " + json.dumps({k: {kk: round(vv, 2) for kk, vv in v.items()}
                                                  for k, v in kernels.items()}))

    cpu = None
    if with_cpu and not args.no_cpu_baseline:
        log("[bench] timing the CPU restatement on a bounded sample")
        cpu = cpu_baseline(prob, kind)
        log(f"[bench] cpu baseline {cpu['value']:.4e} zone-updates/s ({cpu['sample']})")

    ctx_shape = ctx.get_launch_shape()
    lds = ctx.lds_staging() if kind == "euler" else None
    ctx.close()
    return {
        "value": value, "steps": K, "warmup": W, "ms_per_step": el * 1e3 / K, "config": config,
        "alg_bytes_per_step": bytes_per_step, "alg_GBps": bytes_per_step * K / el / 1e9,
        "alg_frac_of_peak": round(bytes_per_step * K / el / 1e9 / PEAK_HBM_GBS, 4),
        "roofline": roofline, "cpu_baseline": cpu, "kernels": kernels, "launch_shape": ctx_shape,
        "lds_staging": lds,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="muscl2d",
                    choices=["muscl2d", "muscl3d", "maire2d", "maire3d", "serve"])
    ap.add_argument("--requests", type=int, default=None,
                    help="requests of the served stream (default 4096)")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream measurement")
    ap.add_argument("--size", type=int, default=None, help="override mesh size (cells per side)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="served stream: skip the one-at-a-time comparison (kernel traces of the batched path)")
    ap.add_argument("--eager", action="store_true",
                    help="served stream: launch each step's kernels instead of replaying captured graphs")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="default workload: skip the MUSCL3D 128^3 and Maire 512^2 lines measured beside it")
    ap.add_argument("--shape", type=int, default=None,
                    help="blocks per CU for every kernel (0 = one block per tile); default: autotune")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "flecsale_amd runs on one GPU; multi-GPU sharding is a design "
                                   "(domain decomposition + halo exchange), not a measured path",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import flecsale_amd as fa
    from flecsale_amd import inputs

    if args.workload == "serve":
        main_serve(args, fa, inputs)
        return

    r = measure(args, fa, inputs, args.workload, args.steps, args.size)
    # the other single-GPU configurations, at their own step counts, measured
    # in the same run (not the metric's workload)
    others = None
    if args.workload == "muscl2d" and args.size is None and not args.no_other_configs:
        others = {}
        for w in ("muscl3d", "maire2d"):
            log(f"[bench] other configuration: {w}")
            o = measure(args, fa, inputs, w)
            others[w] = {k: o[k] for k in ("value", "steps", "warmup", "ms_per_step", "config",
                                           "alg_bytes_per_step", "alg_frac_of_peak", "roofline",
                                           "cpu_baseline", "kernels", "lds_staging")}
    serving = None
    if not args.no_serve:
        n_req = args.requests or 4096
        log(f"[bench] served stream sample: {n_req} requests")
        serving = serve_stream(fa, inputs, n_req, graphs=not args.eager, cpu=not args.no_cpu_baseline)
    out = {
        "metric": "zone-updates/s on 1 MI355X; batched p50/p99 latency vs one-at-a-time GPU",
        "value": r["value"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": r["warmup"],
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (deterministic generated mesh and initial data)",
        "config": r["config"],
        "alg_bytes_per_step": r["alg_bytes_per_step"],
        "alg_GBps": r["alg_GBps"],
        "alg_frac_of_peak": r["alg_frac_of_peak"],
        "roofline": r["roofline"],
        "cpu_baseline": r["cpu_baseline"],
        "kernels": r["kernels"],
        "launch_shape": r["launch_shape"],
        "lds_staging": r["lds_staging"],
        "other_configs": others,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


def stream_cpu_baseline(inputs, sides, steps, budget_s=12.0):
    """CPU restatement on the first requests of the stream, one after another
    (zone-updates per second overall and per solver kind)."""
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    zu = k = 0
    per_kind = {kind: [0, 0.0] for kind in SERVE_KINDS}   # zone-updates, seconds
    while k < 2 or (time.perf_counter() - t0 < budget_s and k < len(sides)):
        kind = SERVE_KINDS[k % 2]
        t_req = time.perf_counter()
        prob = inputs.riemann_2d(int(sides[k]), kind, index=k, steps=steps)
        mesh = O.Mesh(prob["coords"], prob["cells"])
        if kind == "euler":
            s = O.euler_initial_state(mesh, prob["eos"][0], prob["eos"][1], prob["d"], prob["v"], prob["p"])
            rc, t, n = O.euler_run(mesh, s, prob["eos"][0], prob["eos"][1], prob["cfl"], prob["final_time"], steps,
                                   flux_kind=prob["flux"])
        else:
            s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
            rc, t, n, dt = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], steps, prob["dt0"])
        zu += len(prob["cells"]) * n
        per_kind[kind][0] += len(prob["cells"]) * n
        per_kind[kind][1] += time.perf_counter() - t_req
        k += 1
    el = time.perf_counter() - t0
    return dict(value=zu / el, unit="zone-updates/s", cores=threads, kind="port",
                sample=f"first {k} requests of the stream served one after another, {el:.1f} s, "
                       f"OMP_NUM_THREADS={threads}",
                per_kind={kind: v[0] / v[1] for kind, v in per_kind.items() if v[1] > 0})


def cpu_replay(requests, arrivals, steps, base):
    """The seeded Poisson trace replayed on the host cores: each request
    served in arrival order, one after another, for its zone-updates at the
    CPU rate measured for its solver kind on the stream's first requests
    (stream_cpu_baseline).  The service times are that model's, not timed
    one by one: the whole stream would keep the host busy for about 20
    minutes."""
    rates = base["per_kind"]
    t_free = 0.0
    lat, zu = [], 0
    for (kind, prob), a in zip(requests, arrivals):
        z = len(prob["cells"]) * steps
        start = max(a, t_free)
        t_free = start + z / rates[kind]
        lat.append(t_free - a)
        zu += z
    span = float(t_free - min(arrivals))
    return {"requests": len(requests), "makespan_s": span, "zone_updates": int(zu), "zu_per_s": zu / span,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(np.mean(lat)), "cores": base["cores"],
            "rates": {k: round(v, 1) for k, v in rates.items()}, "basis": base["sample"]}


def main_serve(args, fa, inputs):
    n = args.requests or 4096
    steps = args.steps or 100
    out = serve_stream(fa, inputs, n, steps, one_saturated=True, graphs=not args.eager,
                       modes=("batched",) if args.batched_only else ("batched", "one_at_a_time"),
                       cpu=not args.no_cpu_baseline)
    cpu = out.pop("cpu_baseline", None)
    b = out["batched_saturated"]
    line = {
        "metric": "zone-updates/s on 1 MI355X; batched p50/p99 latency vs one-at-a-time GPU",
        "value": b["zu_per_s"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded request stream)",
        "config": {"workload": f"served stream of {n} two-state Riemann problems, side log-uniform in "
                               f"[32, 512], alternating MUSCL / Maire, {steps} steps each"},
        "serving": out,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
