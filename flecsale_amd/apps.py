"""The reference's regression executables as drivers of this library: each
reads an Exodus mesh (-m), sets the app's compiled-in problem at the cell
centroids, runs the app's step loop on the GPU and writes the dump the
reference's regression test compares with its .std golden.

    python -m flecsale_amd.apps hydro_2d -m square_32x32.g        # shock_box_2d-solution_rank000000.txt
    python -m flecsale_amd.apps hydro_3d -m sedov_20x20x20.g      # shock_box_3d-solution_rank000000.txt
    python -m flecsale_amd.apps maire_hydro_2d -m sedov_32x32.g   # sedov_2d-solution_rank000000.txt
    python -m flecsale_amd.apps maire_hydro_3d -m sedov_20x20x20.g   # sedov_3d-solution_rank000000.txt
    python -m flecsale_amd.apps hydro_2d -m square_32x32.g -f shock_box_2d.lua   # the problem from a Lua input
    python -m flecsale_amd.numdiff -a 2e-14 <golden>.std shock_box_2d-solution_rank000000.txt

hydro_2d / hydro_3d: apps/hydro/driver.h:237-450 with apps/hydro/{2d,3d}/
inputs.cc; maire_hydro_2d: apps/maire_hydro/driver.h:330-566 with
apps/maire_hydro/2d/inputs_regression.cc (the maire_hydro_2d_regression
target); maire_hydro_3d: the same driver with apps/maire_hydro/3d/inputs.cc
(whose test, apps/maire_hydro/3d/CMakeLists.txt:24-32, runs to completion
without a golden).  The dump name is the app's prefix + "-solution" with the rank
(apps/hydro/driver.h:441-443, apps/hydro/tasks.h:789-794; one rank here).
"""
import argparse
import os
import sys

import numpy as np

import flecsale_amd as fa
from flecsale_amd import inputs, lua_input

APPS = {
    "hydro_2d": ("shock_box_2d", fa.SOLVER_EULER, inputs.shock_box_2d),
    "hydro_3d": ("shock_box_3d", fa.SOLVER_EULER, inputs.shock_box_3d),
    "maire_hydro_2d": ("sedov_2d", fa.SOLVER_MAIRE, inputs.sedov_2d),
    "maire_hydro_3d": ("sedov_3d", fa.SOLVER_MAIRE, inputs.sedov_3d),
}

# inputs_t::output_freq of the compiled-in problems (apps/hydro/2d/inputs.cc:31,
# apps/hydro/3d/inputs.cc:42, apps/maire_hydro/2d/inputs_regression.cc:69,
# apps/maire_hydro/3d/inputs.cc:52)
OUTPUT_FREQ = {"hydro_2d": 1000000, "hydro_3d": 1000000, "maire_hydro_2d": 20, "maire_hydro_3d": 10}

EPSILON = np.finfo(np.float64).eps


def _advance(ctx, solver, prob, nsteps):
    if solver == fa.SOLVER_EULER:
        ctx.advance(prob["cfl"], prob["final_time"], nsteps, flux_kind=prob["flux"])
    else:
        ctx.advance(prob["cfl"], prob["final_time"], nsteps)


def _run_with_output(ctx, solver, prob, freq, prefix):
    """The drivers' step loop with the output task's schedule; returns the
    Exodus files written.  Both write the initial state (iteration 0), then
    after every step whose count is a multiple of output_freq or that is the
    last of max_steps (apps/hydro/driver.h:390-405,
    apps/maire_hydro/driver.h:517-534); hydro once more after the loop when
    the last count was not a multiple (driver.h:424-437), Maire also after the
    step that reaches the final time (driver.h:521), where its loop stops."""
    max_steps = prob["max_steps"]
    if freq <= 0:
        _advance(ctx, solver, prob, max_steps)
        return []
    files = [ctx.output(prefix, 0)]
    done = 0
    while done < max_steps:
        nxt = min((done // freq + 1) * freq, max_steps)
        _advance(ctx, solver, prob, nxt - done)
        t, _, n = ctx.time()
        if n < nxt:   # Maire: the loop stopped at the final time
            if abs(t - prob["final_time"]) < EPSILON:
                files.append(ctx.output(prefix, n))
            done = n
            break
        done = n
        files.append(ctx.output(prefix, done))
    # hydro's post-loop output: the file of the last in-loop output again
    # (the loop always ends at max_steps, which has written it)
    if solver == fa.SOLVER_EULER and done % freq != 0:
        ctx.output(prefix, done)
    return files


def _lua_problem(path, coords, cells, centroids):
    """The Eulerian problem of a Lua input file (inputs_t::load_lua and
    get_initial_conditions, apps/hydro/inputs.h:103-186): the file's step
    parameters and EOS, and its ics(x, t) at every cell centroid at t = 0."""
    h = lua_input.hydro_inputs(path)
    dim = centroids.shape[1]
    d = np.empty(len(centroids))
    p = np.empty(len(centroids))
    v = np.zeros((len(centroids), dim))
    for c, x in enumerate(centroids):
        d[c], vel, p[c] = h["ics"](x, 0.0)
        v[c, :len(vel)] = vel[:dim]
    return h["prefix"], dict(coords=coords, cells=cells, eos=h["eos"], d=d, v=v, p=p, cfl=h["CFL"],
                             final_time=h["final_time"], max_steps=h["max_steps"], flux=0), h["output_freq"]


def run(app, mesh_path, out_dir=".", input_file=None, output=False):
    """Runs `app` on the mesh file (with the problem of a Lua input file when
    given, as the apps' -f option); returns the dump's path, and with
    output=True also the Exodus files of the output task (the reference
    writes them whenever output_freq > 0)."""
    prefix, solver, problem = APPS[app]
    freq = OUTPUT_FREQ[app]
    if input_file and solver != fa.SOLVER_EULER:
        raise ValueError("Lua input files configure the Eulerian apps (hydro_2d, hydro_3d)")
    coords, cells = fa.read_exodus(mesh_path)
    ctx_type = fa.EulerContext if solver == fa.SOLVER_EULER else fa.MaireContext
    with ctx_type(coords, cells) as ctx:
        # initial conditions at the library's cell centroids, as the reference
        # evaluates them at burton's (apps/hydro/tasks.h:46-68)
        centroids = ctx.get(fa.F_CELL_CENTROID)
        if input_file:
            prefix, prob, freq = _lua_problem(input_file, coords, cells, centroids)
        elif app == "maire_hydro_3d":
            # sedov_3d sizes its energy deposit by each cell's volume
            prob = problem(coords, cells, centroids, ctx.get(fa.F_CELL_VOLUME))
        else:
            prob = problem(coords, cells, centroids)
        if solver == fa.SOLVER_EULER:
            ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        else:
            ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
            ctx.set_time(0.0, prob["dt0"])
        files = _run_with_output(ctx, solver, prob, freq if output else 0, os.path.join(out_dir, prefix))
        path = os.path.join(out_dir, f"{prefix}-solution_rank{0:06d}.txt")
        ctx.dump(path)
    return (path, files) if output else path


def main(argv=None):
    ap = argparse.ArgumentParser(description="run one of the reference's hydro apps on the GPU")
    ap.add_argument("app", choices=sorted(APPS))
    ap.add_argument("-m", "--mesh", required=True, help="Exodus-II mesh file")
    ap.add_argument("-f", "--input", default=None, help="Lua input file (hydro apps)")
    ap.add_argument("-o", "--out-dir", default=".")
    ap.add_argument("--no-output", action="store_true", help="skip the output task's Exodus files")
    a = ap.parse_args(argv)
    path, files = run(a.app, a.mesh, a.out_dir, a.input, output=not a.no_output)
    for f in files:
        print(f"Output written to: {f}")
    print(f"Dumping solution to: {path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
