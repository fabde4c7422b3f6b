"""Eulerian (apps/hydro) parity: the HIP path through the C ABI against the CPU
oracle on identical inputs, task by task and over whole runs.

Tolerance (FP64): relative L-inf per field (max|a-b| / max|b|) <= 1e-11 after
one step / task, <= 1e-9 at final time.  The kernels keep the reference's
operation order, so in practice the results are bit-identical; the tests
report the observed difference and assert the stated bound."""
import numpy as np
import pytest

import flecsale_amd as fa
from flecsale_amd import inputs
from conftest import load_golden, load_mesh, rel_linf

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-11
FINAL_TOL = 1e-9


def _random_state(nc, dim, seed):
    rng = np.random.default_rng(seed)
    d = 1.0 + 0.5 * rng.random(nc)
    v = 0.3 * rng.standard_normal((nc, dim))
    p = 1.0 + 0.5 * rng.random(nc)
    # a few sharp jumps so the limiter's branches all fire
    jump = rng.random(nc) < 0.1
    d[jump] *= 4.0
    p[jump] *= 0.2
    return d, v, p


def _oracle_state(O, mesh, eos, d, v, p):
    return O.euler_initial_state(mesh, eos[0], eos[1], d, v, p)


def _cmp_state(ctx, s, tol, what=""):
    got = ctx.state()
    worst = 0.0
    for k, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p), ("T", s.T), ("a", s.a)]:
        err = rel_linf(got[k], want)
        worst = max(worst, err)
        assert err <= tol, f"{what} {k}: {err:.3e}"
    return worst


MESHES = {
    "square_32x32": lambda: load_mesh("square_32x32"),
    "jitter_quads_40x23": lambda: inputs.box_quads(40, 23, jitter=0.2, seed=5),
    "sedov_20x20x20": lambda: load_mesh("sedov_20x20x20"),
    "jitter_hexes_9x7x6": lambda: inputs.box_hexes(9, 7, 6, jitter=0.1, seed=9),
}


@pytest.mark.parametrize("mesh_name", list(MESHES))
def test_tasks_match_oracle(oracle_lib, mesh_name):
    O = oracle_lib
    coords, cells = MESHES[mesh_name]()
    mesh = O.Mesh(coords, cells)
    dim = coords.shape[1]
    eos = (1.4, 1.0)
    d, v, p = _random_state(mesh.nc, dim, 11)
    s = _oracle_state(O, mesh, eos, d, v, p)
    with fa.EulerContext(coords, cells) as ctx:
        # geometry
        geo = mesh.geometry()
        assert rel_linf(ctx.get(fa.F_CELL_CENTROID), geo["cell_centroid"]) == 0.0
        assert rel_linf(ctx.get(fa.F_CELL_VOLUME), geo["cell_volume"]) == 0.0
        assert rel_linf(ctx.get(fa.F_FACE_NORMAL), geo["face_normal"]) == 0.0
        assert rel_linf(ctx.get(fa.F_FACE_AREA), geo["face_area"]) == 0.0
        assert rel_linf(ctx.get(fa.F_FACE_CENTROID), geo["face_centroid"]) == 0.0
        ctx.initial_conditions(eos, d, v, p)
        _cmp_state(ctx, s, 0.0, "ic")
        # evaluate_time_step
        dt = ctx.evaluate_time_step(0.5, 10.0)
        dt_o = O.euler_time_step(mesh, s, 0.5, 10.0, 0.0)
        assert dt == dt_o
        # evaluate_gradients
        ctx.evaluate_gradients()
        g = O.euler_gradients(mesh, s)
        for fid, key in [(fa.F_GRAD_DENSITY, "d"), (fa.F_GRAD_VELOCITY, "v"), (fa.F_GRAD_ENERGY, "e"),
                         (fa.F_GRAD_PRESSURE, "p"), (fa.F_GRAD_SOUND_SPEED, "a")]:
            err = rel_linf(ctx.get(fid), g[key])
            assert err <= STEP_TOL, (key, err)
        # evaluate_fluxes, each flux function
        for kind in (fa.FLUX_HLLE, fa.FLUX_RUSANOV, fa.FLUX_HLLC):
            ctx.evaluate_fluxes(kind)
            F = O.euler_fluxes(mesh, s, g, kind)
            err = rel_linf(ctx.get(fa.F_FLUX), F)
            assert err <= STEP_TOL, (kind, err)
        # apply_update with the HLLE flux
        ctx.evaluate_fluxes(fa.FLUX_HLLE)
        F = O.euler_fluxes(mesh, s, g, fa.FLUX_HLLE)
        ctx.apply_update(0.5)
        rc, bad = O.euler_apply_update(mesh, s, eos[0], eos[1], F, dt_o, 0.5)
        assert rc == 0
        _cmp_state(ctx, s, STEP_TOL, "update")
        t, _, _ = ctx.time()
        assert t == dt_o * 0.5


def test_shock_box_2d_matches_golden_and_oracle(oracle_lib):
    """apps/hydro/2d regression: 20 steps on square_32x32 vs the .std golden and the oracle."""
    O = oracle_lib
    coords, cells = load_mesh("square_32x32")
    mesh = O.Mesh(coords, cells)
    prob = inputs.shock_box_2d(coords, cells, mesh.geometry()["cell_centroid"])
    s = _oracle_state(O, mesh, prob["eos"], prob["d"], prob["v"], prob["p"])
    rc, t_o, n = O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], prob["max_steps"])
    g = load_golden("shock_box_2d_1procs")
    with fa.EulerContext(coords, cells) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], prob["max_steps"], flux_kind=prob["flux"])
        t, dt, steps = ctx.time()
        assert steps == 20 and t == t_o
        assert abs(t - g["time"]) <= 1e-14
        _cmp_state(ctx, s, FINAL_TOL, "shock_box_2d")
        st = ctx.state()
        for key, gk in [("d", "density"), ("e", "internal_energy"), ("p", "pressure"), ("v", "velocity")]:
            assert np.max(np.abs(st[key] - g[gk])) <= 1e-14, key


def test_shock_box_3d_matches_golden_and_oracle(oracle_lib):
    O = oracle_lib
    coords, cells = load_mesh("sedov_20x20x20")
    mesh = O.Mesh(coords, cells)
    prob = inputs.shock_box_3d(coords, cells, mesh.geometry()["cell_centroid"])
    s = _oracle_state(O, mesh, prob["eos"], prob["d"], prob["v"], prob["p"])
    O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], prob["max_steps"])
    g = load_golden("shock_box_3d_1procs")
    with fa.EulerContext(coords, cells) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], prob["max_steps"])
        t, _, steps = ctx.time()
        assert steps == 12 and t == 0.2
        _cmp_state(ctx, s, FINAL_TOL, "shock_box_3d")
        st = ctx.state()
        assert np.max(np.abs(st["d"] - g["density"])) <= 1e-14
        assert np.max(np.abs(st["v"] - g["velocity"])) <= 1e-14


@pytest.mark.parametrize("flux", [fa.FLUX_HLLE, fa.FLUX_RUSANOV, fa.FLUX_HLLC])
def test_quadrants_multi_step(oracle_lib, flux):
    """Four-quadrant problem at 96^2 for 15 steps (the 1024^2 benchmark case, scaled down)."""
    O = oracle_lib
    prob = inputs.quadrants_2d(96, steps=15, flux=flux | fa.BOUNDARY_TRANSMISSIVE)
    flux = prob["flux"]
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = _oracle_state(O, mesh, prob["eos"], prob["d"], prob["v"], prob["p"])
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        for step in range(prob["max_steps"]):
            O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], 1, flux_kind=flux,
                        soln_time=ctx.time()[0])
            ctx.advance(prob["cfl"], prob["final_time"], 1, flux_kind=flux)
            _cmp_state(ctx, s, STEP_TOL, f"step {step}")


def test_kelvin_helmholtz_3d_small(oracle_lib):
    """Jittered-hex KH shear layer at 14^3 for 6 steps."""
    O = oracle_lib
    prob = inputs.kelvin_helmholtz_3d(14, steps=6)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = _oracle_state(O, mesh, prob["eos"], prob["d"], prob["v"], prob["p"])
    O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], prob["max_steps"])
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], prob["max_steps"])
        _cmp_state(ctx, s, FINAL_TOL, "kh3d")


def test_sod_first_order_to_final_time(oracle_lib):
    """Sod tube 100x10, first order, HLLE, CFL 0.5, stepping stops at t = 0.2."""
    O = oracle_lib
    prob = inputs.sod_100x10()
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = _oracle_state(O, mesh, prob["eos"], prob["d"], prob["v"], prob["p"])
    rc, t_o, n_o = O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], 1000,
                               first_order=True, stop_at_final=True)
    assert rc == 0 and t_o == 0.2
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], n_o + 7, first_order=True, stop_at_final=True)
        t, _, steps = ctx.time()
        assert t == 0.2 and steps == n_o
        _cmp_state(ctx, s, FINAL_TOL, "sod")
        # density stays in the range of the initial data (first-order HLLE is positive)
        d = ctx.get(fa.F_DENSITY)
        assert d.min() >= 0.125 - 1e-12 and d.max() <= 1.0 + 1e-12


def test_task_calls_after_a_run(oracle_lib):
    """A run that used up its step budget leaves the problem's step flag
    cleared; the task-level entry points called after it still act on the
    run's final state (time step, then one explicit half step by tasks)."""
    O = oracle_lib
    prob = inputs.quadrants_2d(48, steps=6)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = _oracle_state(O, mesh, prob["eos"], prob["d"], prob["v"], prob["p"])
    rc, t_o, n_o = O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], 6, flux_kind=prob["flux"])
    assert rc == 0 and n_o == 6
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], 6, flux_kind=prob["flux"])
        _cmp_state(ctx, s, 0.0, "run")
        dt = ctx.evaluate_time_step(prob["cfl"], prob["final_time"])
        assert dt == O.euler_time_step(mesh, s, prob["cfl"], prob["final_time"], t_o)
        ctx.evaluate_gradients()
        ctx.evaluate_fluxes(prob["flux"])
        ctx.apply_update(0.5)
        g = O.euler_gradients(mesh, s)
        F = O.euler_fluxes(mesh, s, g, prob["flux"])
        rc, _ = O.euler_apply_update(mesh, s, *prob["eos"], F, dt, 0.5)
        assert rc == 0
        _cmp_state(ctx, s, 0.0, "tasks after the run")


def test_graph_and_eager_agree_bitwise():
    prob = inputs.quadrants_2d(64, steps=23)
    out = []
    for use_graph in (False, True):
        with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
            ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
            ctx.advance(prob["cfl"], prob["final_time"], 23, flux_kind=prob["flux"], use_graph=use_graph)
            out.append((ctx.state(), ctx.time()))
    for k in out[0][0]:
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert out[0][1] == out[1][1]


def test_conservation_in_walled_box():
    """Mass and energy are conserved to round-off with wall boundaries (quadrant
    densities and pressures at rest, so no wall separation)."""
    prob = inputs.quadrants_2d(64, steps=20)
    prob["v"][:] = 0.0
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        vol = ctx.get(fa.F_CELL_VOLUME)

        def totals():
            st = ctx.state()
            E = st["d"] * (st["e"] + 0.5 * np.sum(st["v"] ** 2, axis=1))
            return np.sum(st["d"] * vol), np.sum(E * vol)

        m0, e0 = totals()
        ctx.advance(prob["cfl"], prob["final_time"], 20, flux_kind=fa.FLUX_HLLC | fa.BOUNDARY_WALL)
        m1, e1 = totals()
        assert abs(m1 - m0) <= 1e-13 * m0
        assert abs(e1 - e0) <= 1e-13 * e0


def test_negative_state_is_reported(oracle_lib):
    """A CFL of 50 drives the Sod state negative: the library reports the
    reference's 'Negative density or internal energy' error (apps/hydro/tasks.h:654)
    for the same cell the reference's serial loop stops at: the lowest failing
    cell of the first failing apply_update."""
    O = oracle_lib
    prob = inputs.sod_100x10()
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.euler_initial_state(mesh, prob["eos"][0], prob["eos"][1], prob["d"], prob["v"], prob["p"])
    bad, t = -1, 0.0
    for _ in range(20):
        dt = O.euler_time_step(mesh, s, 50.0, 0.2, t)
        for _half in range(2):
            F = O.euler_fluxes(mesh, s, O.euler_gradients(mesh, s, first_order=True), 0)
            rc, bad = O.euler_apply_update(mesh, s, prob["eos"][0], prob["eos"][1], F, dt, 0.5)
            if rc:
                break
            t += 0.5 * dt
        if bad >= 0:
            break
    assert bad >= 0
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        with pytest.raises(fa.FsaleError) as ei:
            ctx.advance(50.0, 0.2, 20, first_order=True)
            ctx.sync()
        assert ei.value.status == fa.ERR_NEG_STATE
        assert ctx.error_cell() == bad


def test_zero_wavespeed_reports_infinite_dt():
    coords, cells = inputs.box_quads(4, 4)
    with fa.EulerContext(coords, cells) as ctx:
        ctx.initial_conditions((1.4, 1.0), 1.0, np.zeros((16, 2)), 1.0)
        ctx.set(fa.F_SOUND_SPEED, 0.0)
        with pytest.raises(fa.FsaleError) as ei:
            ctx.evaluate_time_step(0.5, 1.0)
        assert ei.value.status == fa.ERR_DT
