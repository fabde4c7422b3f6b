"""3D Maire Lagrangian (apps/maire_hydro/3d) on hexes: HIP path vs the CPU
oracle, task by task and over whole runs, the reference's sedov_3d smoke case
(apps/maire_hydro/3d/CMakeLists.txt:24-32: sedov_20x20x20.g, 10 steps) and
the sedov_2d golden through its z-extrusion.

The kernels repeat the oracle's operation order, so results are compared
bit for bit (rel L-inf 0); the symmetry QR amplifies a one-ulp difference by
orders of magnitude over a run, which a tolerance would hide."""
import numpy as np
import pytest

import flecsale_amd as fa
from flecsale_amd import inputs
from conftest import load_golden, load_mesh, rel_linf

pytestmark = pytest.mark.gpu


def _setup(O, prob):
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
    ctx = fa.MaireContext(prob["coords"], prob["cells"])
    ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
    ctx.set_time(0.0, prob["dt0"])
    return mesh, s, ctx


def _bitwise(ctx, s, mesh, what):
    got = ctx.state()
    for k, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p), ("T", s.T), ("a", s.a), ("M", s.M),
                    ("V", s.V)]:
        assert np.array_equal(got[k], want), f"{what} {k}: {rel_linf(got[k], want):.3e}"
    x = ctx.get(fa.F_COORDS)
    want = mesh.geometry()["coords"]
    assert np.array_equal(x, want), f"{what} coords: {rel_linf(x, want):.3e}"


def _run_both(O, prob, steps):
    mesh, s, ctx = _setup(O, prob)
    rc, t_o, n_o, dt_o = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], steps, prob["dt0"])
    assert rc == 0 and n_o == steps
    ctx.advance(prob["cfl"], prob["final_time"], steps)
    t, dt, n = ctx.time()
    assert n == steps and t == t_o and dt == dt_o
    return mesh, s, ctx


def test_sedov_3d_smoke_case_matches_oracle(oracle_lib):
    O = oracle_lib
    coords, cells = load_mesh("sedov_20x20x20")
    geo = O.Mesh(coords, cells).geometry()
    prob = inputs.sedov_3d(coords, cells, geo["cell_centroid"], geo["cell_volume"])
    print("sedov_3d 20^3, 10 steps", flush=True)
    mesh, s, ctx = _run_both(O, prob, prob["max_steps"])
    with ctx:
        _bitwise(ctx, s, mesh, "sedov_3d")


def test_tasks_3d_match_oracle(oracle_lib):
    """Nodal solve, residual, time step and update on a jittered two-gas hex mesh."""
    O = oracle_lib
    prob = inputs.maire_bubble_3d(10, steps=1)
    print("tasks, bubble 10^3", flush=True)
    mesh, s, ctx = _setup(O, prob)
    with ctx:
        ctx.evaluate_nodal_state()
        un, npc, Fpc = O.maire_nodal(mesh, s)
        assert np.array_equal(ctx.get(fa.F_NODE_VELOCITY), un)
        assert np.array_equal(ctx.get(fa.F_CORNER_NORMAL), npc)
        assert np.array_equal(ctx.get(fa.F_CORNER_FORCE), Fpc)
        ctx.evaluate_residual()
        dudt = O.maire_residual(mesh, un, npc, Fpc)
        assert np.array_equal(ctx.get(fa.F_RESIDUAL), dudt)
        dt = ctx.evaluate_time_step(prob["cfl"], prob["final_time"])
        dt_o = min(O.lib().om_maire_time_step(mesh.h, *prob["cfl"], prob["dt0"], s.a, dudt.reshape(-1)),
                   prob["final_time"])
        assert dt == dt_o


@pytest.mark.parametrize("n,steps", [(9, 25), (16, 8)])
def test_bubble_3d_run_matches_oracle(oracle_lib, n, steps):
    """Jittered hexes, two gases (per-cell gamma), symmetry walls on all six sides."""
    print(f"bubble {n}^3, {steps} steps", flush=True)
    mesh, s, ctx = _run_both(oracle_lib, inputs.maire_bubble_3d(n, steps=steps), steps)
    with ctx:
        _bitwise(ctx, s, mesh, "bubble3d")


def test_extruded_sedov_2d_golden(oracle_lib):
    """Two z layers of the sedov_2d mesh: the reference's 2D golden, and the oracle bit for bit."""
    O = oracle_lib
    coords, cells = load_mesh("sedov_32x32")
    p2 = inputs.sedov_2d(coords, cells, O.Mesh(coords, cells).geometry()["cell_centroid"])
    x3, c3 = inputs.extrude_quads(coords, cells, 2, 0.5, (1, 2, 0))
    nc = len(c3)
    prob = dict(p2, coords=x3, cells=c3, gamma=np.full(nc, 1.4), d=np.tile(p2["d"], 2), v=np.zeros((nc, 3)),
                p=np.tile(p2["p"], 2))
    print("extruded sedov_2d, 2 layers", flush=True)
    mesh, s, ctx = _run_both(O, prob, 20)
    with ctx:
        _bitwise(ctx, s, mesh, "extruded")
        g = load_golden("sedov_2d_1procs")
        st = ctx.state()
        n2 = len(cells)
        assert rel_linf(st["d"][:n2], g["density"]) <= 1e-10
        assert rel_linf(st["v"][:n2][:, [2, 0]], g["velocity"]) <= 1e-10


def test_graph_and_eager_agree_bitwise_3d():
    prob = inputs.maire_bubble_3d(8, steps=7)
    out = []
    for use_graph in (False, True):
        with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
            ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
            ctx.set_time(0.0, prob["dt0"])
            ctx.advance(prob["cfl"], prob["final_time"], 7, use_graph=use_graph)
            out.append((ctx.state(), ctx.get(fa.F_COORDS), ctx.time()))
    for k in out[0][0]:
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_maire_3d_batch_alone_vs_mix(oracle_lib):
    """Three hex problems of uneven sizes in one batch: each bit-identical to
    its one-problem run, one of them also to the oracle."""
    O = oracle_lib
    mix = [(inputs.maire_bubble_3d(7, steps=5, seed=3), 5), (inputs.maire_bubble_3d(12, steps=9), 9),
           (inputs.maire_bubble_3d(5, steps=12, seed=9), 12)]
    alone = []
    for prob, n in mix:
        with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
            ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
            ctx.set_time(0.0, prob["dt0"])
            ctx.advance(prob["cfl"], prob["final_time"], n)
            st = ctx.state()
            st["x"] = ctx.get(fa.F_COORDS)
            alone.append((st, ctx.time()))
    nc = sum(len(p["cells"]) for p, _ in mix)
    nv = sum(len(p["coords"]) for p, _ in mix)
    print("3D batch of 3", flush=True)
    with fa.Batch(fa.SOLVER_MAIRE, 3, 8, 4, nc + 100, 4 * nc + 100, nv + 100, 27 * nc + 100) as b:
        ids = [b.add_maire(p["coords"], p["cells"], p["cv"], p["gamma"], p["d"], p["v"], p["p"], p["cfl"],
                           p["final_time"], n, p["dt0"]) for p, n in mix]
        b.step(14)
        st = b.status()
        for (p, n), pid, (want, (t, dt, steps)) in zip(mix, ids, alone):
            assert st[pid].error == 0 and st[pid].steps == n == steps and st[pid].time == t
            got = b.state(pid)
            for k in got:
                assert np.array_equal(got[k], want[k]), (pid, k)
        prob, n = mix[1]
        mesh = O.Mesh(prob["coords"], prob["cells"])
        s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
        O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], n, prob["dt0"])
        got = b.state(ids[1])
        for k, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p)]:
            assert np.array_equal(got[k], want), k


def test_task_sequence_3d_matches_oracle(oracle_lib):
    """The driver step as separate tasks (save, nodal, residual, dt, half
    update, nodal, residual, restore, full update), then a host write of the
    velocity before the next nodal solve: each equals the oracle."""
    O = oracle_lib
    prob = inputs.maire_bubble_3d(7, steps=2)
    print("task sequence, bubble 7^3", flush=True)
    mesh, s, ctx = _setup(O, prob)
    with ctx:
        rc, t_o, n_o, dt_o = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], 1, prob["dt0"])
        assert rc == 0
        ctx.save()
        ctx.evaluate_nodal_state()
        ctx.evaluate_residual()
        dt = ctx.evaluate_time_step(prob["cfl"], prob["final_time"])
        assert dt == dt_o
        ctx.apply_update(0.5)
        ctx.evaluate_nodal_state()
        ctx.evaluate_residual()
        ctx.restore()
        ctx.apply_update(1.0)
        _bitwise(ctx, s, mesh, "tasks")
        # a host write of the velocity reaches the next nodal solve
        v = s.v * 0.5 + 0.01
        ctx.set(fa.F_VELOCITY, v)
        s.v[:] = v
        ctx.evaluate_nodal_state()
        un, npc, Fpc = O.maire_nodal(mesh, s)
        assert np.array_equal(ctx.get(fa.F_NODE_VELOCITY), un)
        assert np.array_equal(ctx.get(fa.F_CORNER_FORCE), Fpc)


def test_autotune_keeps_3d_results(oracle_lib):
    """The autotuner times candidate shapes from a snapshot of a moved mesh
    and restores it, packed copies of state and coordinates included: steps
    after it equal the oracle's continuation bit for bit."""
    O = oracle_lib
    prob = inputs.maire_bubble_3d(6, steps=8, seed=5)
    mesh, s, ctx = _setup(O, prob)
    print("3D autotune regression 6^3", flush=True)
    with ctx:
        ctx.advance(prob["cfl"], prob["final_time"], 3)
        assert ctx.autotune(2) > 0
        ctx.advance(prob["cfl"], prob["final_time"], 5)
        rc, t_o, n_o, dt_o = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], 8, prob["dt0"])
        assert rc == 0 and n_o == 8
        _bitwise(ctx, s, mesh, "after autotune")
