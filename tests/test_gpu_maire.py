"""Maire Lagrangian (apps/maire_hydro) parity: HIP path vs CPU oracle, task by
task and over whole runs, plus the reference's sedov_2d golden.

The symmetry QR at boundary vertices amplifies one-ulp differences by ~1e5
over 20 steps (oracle experiments: a 1-ulp change of one wedge facet area
moves the sedov result by 1e-10), so this path is held bit-identical."""
import numpy as np
import pytest

import flecsale_amd as fa
from flecsale_amd import inputs
from conftest import load_golden, load_mesh, rel_linf

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-11
FINAL_TOL = 1e-9


def _cmp(ctx, s, mesh, tol, what):
    got = ctx.state()
    pairs = [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p), ("T", s.T), ("a", s.a), ("M", s.M), ("V", s.V)]
    for k, want in pairs:
        err = rel_linf(got[k], want)
        assert err <= tol, f"{what} {k}: {err:.3e}"
    err = rel_linf(ctx.get(fa.F_COORDS), mesh.geometry()["coords"])
    assert err <= tol, f"{what} coords: {err:.3e}"


def _setup(O, prob):
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
    ctx = fa.MaireContext(prob["coords"], prob["cells"])
    ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
    ctx.set_time(0.0, prob["dt0"])
    return mesh, s, ctx


def test_sedov_2d_matches_golden_and_oracle(oracle_lib):
    O = oracle_lib
    coords, cells = load_mesh("sedov_32x32")
    cen = O.Mesh(coords, cells).geometry()["cell_centroid"]
    prob = inputs.sedov_2d(coords, cells, cen)
    mesh, s, ctx = _setup(O, prob)
    with ctx:
        _cmp(ctx, s, mesh, 0.0, "ic")
        rc, t_o, n_o, dt_o = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"],
                                         prob["max_steps"], prob["dt0"])
        ctx.advance(prob["cfl"], prob["final_time"], prob["max_steps"])
        t, dt, steps = ctx.time()
        assert steps == 20 and t == t_o and dt == dt_o
        _cmp(ctx, s, mesh, FINAL_TOL, "sedov")
        g = load_golden("sedov_2d_1procs")
        st = ctx.state()
        assert abs(t - g["time"]) <= 1e-18
        for key, gk in [("d", "density"), ("e", "internal_energy"), ("p", "pressure"), ("v", "velocity")]:
            assert rel_linf(st[key], g[gk]) <= 1e-14, key
        assert rel_linf(ctx.get(fa.F_COORDS), g["vertex_coords"]) <= 1e-14


def test_tasks_match_oracle(oracle_lib):
    """Nodal solve, residual, time step and update on a jittered two-gas mesh."""
    O = oracle_lib
    prob = inputs.maire_bubble_2d(40, steps=1)
    mesh, s, ctx = _setup(O, prob)
    with ctx:
        ctx.evaluate_nodal_state()
        un, npc, Fpc = O.maire_nodal(mesh, s)
        assert rel_linf(ctx.get(fa.F_NODE_VELOCITY), un) <= STEP_TOL
        assert rel_linf(ctx.get(fa.F_CORNER_NORMAL), npc) <= STEP_TOL
        assert rel_linf(ctx.get(fa.F_CORNER_FORCE), Fpc) <= STEP_TOL
        ctx.evaluate_residual()
        dudt = O.maire_residual(mesh, un, npc, Fpc)
        assert rel_linf(ctx.get(fa.F_RESIDUAL), dudt) <= STEP_TOL
        dt = ctx.evaluate_time_step(prob["cfl"], prob["final_time"])
        dt_o = min(O.lib().om_maire_time_step(mesh.h, *prob["cfl"], prob["dt0"], s.a, dudt.reshape(-1)),
                   prob["final_time"])
        assert dt == dt_o


@pytest.mark.parametrize("n,steps", [(24, 40), (64, 12)])
def test_bubble_run_matches_oracle(oracle_lib, n, steps):
    """Jittered mesh, two gases (per-cell gamma), symmetry walls, shock driven from the left."""
    O = oracle_lib
    prob = inputs.maire_bubble_2d(n, steps=steps)
    mesh, s, ctx = _setup(O, prob)
    with ctx:
        rc, t_o, n_o, dt_o = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], steps,
                                         prob["dt0"])
        assert rc == 0
        ctx.advance(prob["cfl"], prob["final_time"], steps)
        t, dt, k = ctx.time()
        assert k == steps
        assert abs(t - t_o) <= 1e-15 * t_o
        _cmp(ctx, s, mesh, FINAL_TOL, "bubble")


def test_task_calls_after_a_run(oracle_lib):
    """A run that used up its step budget leaves the problem's step flag
    cleared; the task-level entry points called after it still act on the
    run's final state (the nodal solve, the residual and the time step)."""
    O = oracle_lib
    prob = inputs.maire_bubble_2d(16, steps=15)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
    rc, t_o, n_o, dt_o = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], 15, prob["dt0"])
    assert rc == 0 and n_o == 15
    un_o, npc_o, Fpc_o = O.maire_nodal(mesh, s)
    dudt_o = O.maire_residual(mesh, un_o, npc_o, Fpc_o)
    with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
        ctx.set_time(0.0, prob["dt0"])
        ctx.advance(prob["cfl"], prob["final_time"], 15)
        st = ctx.state()
        for k, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p)]:
            assert np.array_equal(st[k], want), k
        ctx.evaluate_nodal_state()
        assert np.array_equal(ctx.get(fa.F_NODE_VELOCITY), un_o)
        assert np.array_equal(ctx.get(fa.F_CORNER_NORMAL), npc_o)
        assert np.array_equal(ctx.get(fa.F_CORNER_FORCE), Fpc_o)
        ctx.evaluate_residual()
        assert np.array_equal(ctx.get(fa.F_RESIDUAL), dudt_o)
        dt = ctx.evaluate_time_step(prob["cfl"], prob["final_time"])
        assert dt == min(O.maire_time_step(mesh, prob["cfl"], dt_o, s.a, dudt_o), prob["final_time"] - t_o)


def test_run_to_final_time_inside_a_graph_chunk(oracle_lib):
    """The final time reached a few steps into a replayed chunk of steps:
    every field the run leaves (state, temperature, sound speed, volume,
    coordinates, centroids) equals the oracle's run to the same time."""
    O = oracle_lib
    prob = inputs.maire_bubble_2d(24, steps=200)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
    final_time = 2.5e-4
    rc, t_o, n_o, _ = O.maire_run(mesh, s, prob["cv"], prob["cfl"], final_time, 200, prob["dt0"])
    assert rc == 0 and 10 < n_o < 200 and n_o % 10 != 0
    geo = mesh.geometry()
    with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
        ctx.set_time(0.0, prob["dt0"])
        ctx.advance(prob["cfl"], final_time, 200)
        t, _, n = ctx.time()
        assert n == n_o and t == t_o
        st = ctx.state()
        for k, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p), ("T", s.T), ("a", s.a), ("V", s.V)]:
            assert np.array_equal(st[k], np.asarray(want).reshape(st[k].shape)), k
        assert np.array_equal(ctx.get(fa.F_COORDS), geo["coords"])
        assert np.array_equal(ctx.get(fa.F_CELL_CENTROID), geo["cell_centroid"])
        assert np.array_equal(ctx.get(fa.F_CELL_VOLUME), geo["cell_volume"])


def test_mass_and_total_energy_conserved():
    """Compatible discretisation: mass exactly, total energy to round-off (walls do no work)."""
    prob = inputs.maire_bubble_2d(32, steps=30)
    with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
        ctx.set_time(0.0, prob["dt0"])

        def totals():
            st = ctx.state()
            E = st["M"] * (st["e"] + 0.5 * np.sum(st["v"] ** 2, axis=1))
            return np.sum(st["d"] * st["V"]), np.sum(E)

        m0, e0 = totals()
        ctx.advance(prob["cfl"], prob["final_time"], 30)
        m1, e1 = totals()
        assert abs(m1 - m0) <= 1e-12 * m0
        assert abs(e1 - e0) <= 1e-12 * e0


def test_graph_and_eager_agree_bitwise():
    prob = inputs.maire_bubble_2d(32, steps=13)
    out = []
    for use_graph in (False, True):
        with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
            ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
            ctx.set_time(0.0, prob["dt0"])
            ctx.advance(prob["cfl"], prob["final_time"], 13, use_graph=use_graph)
            out.append((ctx.state(), ctx.get(fa.F_COORDS), ctx.time()))
    for k in out[0][0]:
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]
