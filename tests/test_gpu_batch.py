"""Ragged batches (fsale_batch_*): many problems of uneven sizes resident in
one context and advanced by the same persistent kernels.

The contract: a problem's result does not depend on what else is resident.
Every check compares a problem stepped inside a mix with the same problem in
a plain one-problem context, and requires bit-identical fields (the kernels
read only the problem's own cells and scalars).  One problem per solver is
also held against the CPU oracle at the run tolerance (1e-9 relative L-inf
at final time)."""
import numpy as np
import pytest

import flecsale_amd as fa
from flecsale_amd import inputs
from conftest import load_mesh, rel_linf

pytestmark = pytest.mark.gpu

FINAL_TOL = 1e-9


def _euler_alone(prob, nsteps):
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], nsteps, prob["flux"])
        t, dt, steps = ctx.time()
        return ctx.state(), (t, dt, steps)


def _maire_alone(prob, nsteps):
    with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
        ctx.set_time(0.0, prob["dt0"])
        ctx.advance(prob["cfl"], prob["final_time"], nsteps)
        st = ctx.state()
        st["x"] = ctx.get(fa.F_COORDS)
        return st, ctx.time()


def _euler_batch(problems, dim=2, vpc=4, max_problems=8):
    nc = sum(len(p["cells"]) for p in problems)
    nv = sum(len(p["coords"]) for p in problems)
    return fa.Batch(fa.SOLVER_EULER, dim, vpc, max_problems, nc + 100, 2 * nc + 4 * (nv + 100), nv + 100,
                    9 * nc + 100)


def _add_euler(b, prob, max_steps):
    return b.add_euler(prob["coords"], prob["cells"], prob["eos"], prob["d"], prob["v"], prob["p"], prob["cfl"],
                       prob["final_time"], max_steps, prob["flux"])


def _add_maire(b, prob, max_steps):
    return b.add_maire(prob["coords"], prob["cells"], prob["cv"], prob["gamma"], prob["d"], prob["v"],
                       prob["p"], prob["cfl"], prob["final_time"], max_steps, prob["dt0"])


def _same(got, want, what):
    assert set(got) <= set(want)
    for k in got:
        assert np.array_equal(got[k], want[k]), f"{what} {k}: {rel_linf(got[k], want[k]):.3e}"


def _euler_mix():
    sb_coords, sb_cells = inputs.box_quads(30, 20, -0.5, 0.5, -0.5, 0.5, jitter=0.1, seed=3)
    return [
        (inputs.quadrants_2d(n=40, steps=9), 9),
        (inputs.shock_box_2d(sb_coords, sb_cells, inputs.quad_centroids(sb_coords, sb_cells)), 5),
        (inputs.riemann_2d(33, "euler", index=4), 12),
        (inputs.riemann_2d(17, "euler", index=7), 3),
    ]


@pytest.mark.parametrize("use_graph", [True, False])
def test_euler_alone_vs_mix(use_graph):
    mix = _euler_mix()
    alone = [_euler_alone(p, n) for p, n in mix]
    with _euler_batch([p for p, _ in mix]) as b:
        ids = [_add_euler(b, p, n) for p, n in mix]
        b.step(max(n for _, n in mix) + 3, use_graph=use_graph)
        st = b.status()
        for (p, n), pid, (want, (t, dt, steps)) in zip(mix, ids, alone):
            s = st[pid]
            assert s.error == 0 and s.done == 1 and s.steps == n == steps
            assert s.time == t and s.time_step == dt
            _same(b.state(pid), want, f"problem {pid}")


def test_euler_admission_between_steps():
    """Retire a problem mid-run and admit a new one into the freed ranges;
    the survivor and the newcomer match their plain runs."""
    mix = _euler_mix()
    a, b_, c = mix[0][0], mix[2][0], mix[1][0]
    want_b, tb = _euler_alone(b_, 10)
    want_c, tc = _euler_alone(c, 6)
    with _euler_batch([a, b_]) as b:
        ia = _add_euler(b, a, 10)
        ib = _add_euler(b, b_, 10)
        b.step(4)
        b.remove(ia)
        ic = _add_euler(b, c, 6)
        b.step(3, use_graph=False)
        b.step(10)
        st = b.status()
        assert st[ib].steps == 10 and st[ic].steps == 6
        assert st[ib].time == tb[0] and st[ic].time == tc[0]
        _same(b.state(ib), want_b, "survivor")
        _same(b.state(ic), want_c, "newcomer")


def test_euler_batch_matches_oracle(oracle_lib):
    O = oracle_lib
    prob = inputs.riemann_2d(33, "euler", index=4)
    other = inputs.quadrants_2d(n=24, steps=12)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.euler_initial_state(mesh, prob["eos"][0], prob["eos"][1], prob["d"], prob["v"], prob["p"])
    rc, t_o, n_o = O.euler_run(mesh, s, prob["eos"][0], prob["eos"][1], prob["cfl"], prob["final_time"], 12,
                               prob["flux"])
    assert rc == 0 and n_o == 12
    with _euler_batch([prob, other]) as b:
        _add_euler(b, other, 12)
        pid = _add_euler(b, prob, 12)
        b.step(12)
        got = b.state(pid)
        assert b.status()[pid].time == t_o
        for k, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p)]:
            assert rel_linf(got[k], want) <= FINAL_TOL, k


def _maire_mix():
    coords, cells = load_mesh("sedov_32x32")
    cen = inputs.quad_centroids(coords, cells)
    return [
        (inputs.maire_bubble_2d(24, 6), 6),
        (inputs.sedov_2d(coords, cells, cen), 20),
        (inputs.riemann_2d(21, "maire", index=5), 9),
        (inputs.riemann_2d(40, "maire", index=8), 4),
    ]


def _maire_batch(problems, max_problems=8):
    nc = sum(len(p["cells"]) for p in problems)
    nv = sum(len(p["coords"]) for p in problems)
    return fa.Batch(fa.SOLVER_MAIRE, 2, 4, max_problems, nc + 100, 2 * nc + 4 * (nv + 100), nv + 100,
                    9 * nc + 100)


@pytest.mark.parametrize("use_graph", [True, False])
def test_maire_alone_vs_mix(use_graph):
    mix = _maire_mix()
    alone = [_maire_alone(p, n) for p, n in mix]
    with _maire_batch([p for p, _ in mix]) as b:
        ids = [_add_maire(b, p, n) for p, n in mix]
        b.step(max(n for _, n in mix) + 2, use_graph=use_graph)
        st = b.status()
        for (p, n), pid, (want, (t, dt, steps)) in zip(mix, ids, alone):
            s = st[pid]
            assert s.error == 0 and s.done == 1 and s.steps == n == steps
            assert s.time == t and s.time_step == dt
            _same(b.state(pid), want, f"problem {pid}")


def test_maire_batch_matches_oracle(oracle_lib):
    O = oracle_lib
    prob = inputs.riemann_2d(21, "maire", index=5)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
    rc, t_o, n_o, dt_o = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], 9, prob["dt0"])
    assert rc == 0 and n_o == 9
    with _maire_batch([prob, inputs.maire_bubble_2d(24, 6)]) as b:
        _add_maire(b, inputs.maire_bubble_2d(24, 6), 6)
        pid = _add_maire(b, prob, 9)
        b.step(9)
        got = b.state(pid)
        assert b.status()[pid].time == t_o
        for k, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p)]:
            assert rel_linf(got[k], want) <= FINAL_TOL, k
        assert rel_linf(got["x"], mesh.geometry()["coords"]) <= FINAL_TOL


def test_capacity_exhaustion_and_reuse():
    small = inputs.riemann_2d(16, "euler", index=1)
    with fa.Batch(fa.SOLVER_EULER, 2, 4, 3, 600, 1300, 700, 6000) as b:
        cap0 = b.capacity()
        assert cap0["cells"] == 600 and cap0["problems"] == 3
        i0 = _add_euler(b, small, 4)
        i1 = _add_euler(b, small, 4)
        with pytest.raises(fa.FsaleError) as e:
            _add_euler(b, small, 4)          # 3 x 256 cells > 600
        assert e.value.status == 4
        assert b.capacity()["cells"] == 600 - 512
        b.step(4)
        want, _ = _euler_alone(small, 4)
        _same(b.state(i0), want, "first")
        _same(b.state(i1), want, "second")
        b.remove(i0)
        i2 = _add_euler(b, small, 2)
        b.step(2)
        want2, _ = _euler_alone(small, 2)
        _same(b.state(i2), want2, "reused ranges")
        _same(b.state(i1), want, "untouched")


def test_failure_isolation():
    """A problem that fails (CFL 50 drives the state negative) is latched
    done with its error; the other problems of the batch are unaffected."""
    good = inputs.riemann_2d(24, "euler", index=2)
    bad = dict(inputs.sod_100x10(), cfl=50.0, flux=0)
    want, _ = _euler_alone(good, 6)
    with _euler_batch([good, bad]) as b:
        ig = _add_euler(b, good, 6)
        ib = _add_euler(b, bad, 6)
        b.step(6)
        st = b.status()
        assert st[ib].error == fa.ERR_NEG_STATE and st[ib].done == 1
        assert st[ig].error == 0 and st[ig].steps == 6
        _same(b.state(ig), want, "healthy problem")


@pytest.mark.parametrize("solver", ["euler", "maire"])
def test_launch_shape_never_changes_results(solver):
    """Persistent grids of any size and one-block-per-tile launches give
    bit-identical results; the autotuner restores the state it timed."""
    mix = _euler_mix() if solver == "euler" else _maire_mix()
    make, add = (_euler_batch, _add_euler) if solver == "euler" else (_maire_batch, _add_maire)
    results = []
    for shape in [0, 1, 3, 16, "tuned"]:
        with make([p for p, _ in mix]) as b:
            ids = [add(b, p, n) for p, n in mix]
            if shape == "tuned":
                before = {pid: b.state(pid) for pid in ids}
                ms = b.autotune(2)
                assert ms > 0
                for pid in ids:
                    _same(b.state(pid), before[pid], "state after autotune")
                assert all(s.steps == 0 for s in b.status().values())
            else:
                b.launch_shape(shape)
            b.step(7)
            b.step(6, use_graph=False)
            results.append({pid: b.state(pid) for pid in ids})
    for r in results[1:]:
        for pid in r:
            _same(r[pid], results[0][pid], f"problem {pid}")
