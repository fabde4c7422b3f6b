"""Parity at the benchmark sizes: the four-quadrant MUSCL problem on 1024^2
quads, the jittered-hex Kelvin-Helmholtz layer on 128^3 hexes and the Maire
bubble on a jittered 512^2 mesh, each a few steps through the C ABI against
the CPU oracle on identical inputs (the oracle runs on the host cores; a step
of these sizes takes it well under a second or two).

Tolerance (FP64): relative L-inf per field <= 1e-11 after every step; the
kernels keep the reference's operation order, so the fields are also asserted
bit-identical at the end.  Beside the oracle, size-independent invariants of
the same sizes: mass and total energy of a walled box (Eulerian: fluxes of
interior faces cancel pairwise; Lagrangian: compatible discretisation) are
conserved to round-off."""
import numpy as np
import pytest

import flecsale_amd as fa
from flecsale_amd import inputs
from conftest import rel_linf

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-11


def _euler_state(O, mesh, prob):
    return O.euler_initial_state(mesh, prob["eos"][0], prob["eos"][1], prob["d"], prob["v"], prob["p"])


def _cmp_euler(ctx, s, what, exact=False):
    got = ctx.state()
    for k, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p), ("T", s.T), ("a", s.a)]:
        err = rel_linf(got[k], want)
        assert err <= STEP_TOL, f"{what} {k}: {err:.3e}"
        if exact:
            assert np.array_equal(got[k], np.asarray(want).reshape(got[k].shape)), f"{what} {k} not bitwise"


def _euler_steps(O, prob, steps, what):
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = _euler_state(O, mesh, prob)
    flux = prob["flux"]
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        for step in range(steps):
            O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], 1, flux_kind=flux,
                        soln_time=ctx.time()[0])
            ctx.advance(prob["cfl"], prob["final_time"], 1, flux_kind=flux)
            _cmp_euler(ctx, s, f"{what} step {step}", exact=(step == steps - 1))


def test_muscl2d_1024_matches_oracle(oracle_lib):
    prob = inputs.quadrants_2d(1024, steps=3, flux=fa.FLUX_HLLC | fa.BOUNDARY_TRANSMISSIVE)
    _euler_steps(oracle_lib, prob, 3, "muscl2d 1024^2")


def test_muscl3d_128_matches_oracle(oracle_lib):
    prob = inputs.kelvin_helmholtz_3d(128, steps=2)
    _euler_steps(oracle_lib, prob, 2, "muscl3d 128^3")


def test_maire_512_matches_oracle(oracle_lib):
    O = oracle_lib
    steps = 3
    prob = inputs.maire_bubble_2d(512, steps=steps)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
    rc, t_o, n_o, dt_o = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], steps, prob["dt0"])
    assert rc == 0 and n_o == steps
    with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
        ctx.set_time(0.0, prob["dt0"])
        ctx.advance(prob["cfl"], prob["final_time"], steps)
        t, dt, k = ctx.time()
        assert k == steps and t == t_o and dt == dt_o
        got = ctx.state()
        for key in ("d", "v", "e", "p", "a", "V"):
            want = np.asarray(getattr(s, key)).reshape(got[key].shape)
            assert rel_linf(got[key], want) <= STEP_TOL, key
            assert np.array_equal(got[key], want), f"maire 512^2 {key} not bitwise"
        x = ctx.get(fa.F_COORDS)
        want = np.asarray(mesh.geometry()["coords"]).reshape(x.shape)
        assert np.array_equal(x, want), "maire 512^2 coordinates not bitwise"


# ------------------------------------------------------------ full step counts
# Each configuration to its stated step count through the device-resident
# driver loop (fsale_*_advance: graph-replayed chunks of steps, time-step
# partials and their finalize carried from step to step on the device),
# against the oracle run for the same steps.  Bound: relative L-inf per field
# <= 1e-11 at every chunk end and <= 1e-9 at the final step; the kernels keep
# the reference's operation order, so every chunk end is also bit-identical.

FINAL_TOL = 1e-9


def _euler_full(O, prob, steps, chunk, what):
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = _euler_state(O, mesh, prob)
    flux = prob["flux"]
    t_o = 0.0
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        done = 0
        while done < steps:
            n = min(chunk, steps - done)
            rc, t_o, k = O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], n,
                                     flux_kind=flux, soln_time=t_o)
            assert rc == 0 and k == n
            ctx.advance(prob["cfl"], prob["final_time"], n, flux_kind=flux)
            done += n
            t, _, k = ctx.time()
            assert k == done and t == t_o, (what, done)
            _cmp_euler(ctx, s, f"{what} step {done}", exact=True)
        got = ctx.state()
        for key, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p)]:
            assert rel_linf(got[key], want) <= FINAL_TOL


@pytest.mark.timeout(600)
def test_muscl2d_1024_200_steps_match_oracle(oracle_lib):
    """The four-quadrant problem to its 200 steps (chunks of 20)."""
    prob = inputs.quadrants_2d(1024, steps=200, flux=fa.FLUX_HLLC | fa.BOUNDARY_TRANSMISSIVE)
    _euler_full(oracle_lib, prob, 200, 20, "muscl2d 1024^2")


@pytest.mark.timeout(900)
def test_muscl3d_128_100_steps_match_oracle(oracle_lib):
    """The jittered-hex Kelvin-Helmholtz layer to its 100 steps (chunks of 10)."""
    prob = inputs.kelvin_helmholtz_3d(128, steps=100)
    _euler_full(oracle_lib, prob, 100, 10, "muscl3d 128^3")


@pytest.mark.timeout(600)
def test_maire_512_500_steps_match_oracle(oracle_lib):
    """The two-gas bubble to its 500 steps (chunks of 50).  The time step grows
    by 1.01 a step from 1e-5 until the acoustic limit C_acc min_length / a
    binds (past step ~250 here): the last steps' dt is below the growth
    limit's, so that branch is exercised and compared too."""
    O = oracle_lib
    steps, chunk = 500, 50
    prob = inputs.maire_bubble_2d(512, steps=steps)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.maire_initial_state(mesh, prob["gamma"], prob["cv"], prob["d"], prob["v"], prob["p"])
    t_o, dt_o = 0.0, prob["dt0"]
    with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
        ctx.set_time(0.0, prob["dt0"])
        done = 0
        while done < steps:
            rc, t_o, k, dt_o = O.maire_run(mesh, s, prob["cv"], prob["cfl"], prob["final_time"], chunk, dt_o,
                                           soln_time=t_o)
            assert rc == 0 and k == chunk
            ctx.advance(prob["cfl"], prob["final_time"], chunk)
            done += chunk
            t, dt, k = ctx.time()
            assert k == done and t == t_o and dt == dt_o, done
            got = ctx.state()
            for key in ("d", "v", "e", "p", "a", "V"):
                want = np.asarray(getattr(s, key)).reshape(got[key].shape)
                assert rel_linf(got[key], want) <= (FINAL_TOL if done == steps else STEP_TOL), (key, done)
                assert np.array_equal(got[key], want), f"maire 512^2 {key} not bitwise at step {done}"
        x = ctx.get(fa.F_COORDS)
        assert np.array_equal(x, np.asarray(mesh.geometry()["coords"]).reshape(x.shape))
    assert dt_o < 0.5 * prob["dt0"] * 1.01 ** steps


def test_muscl2d_1024_walled_box_conserves():
    """The 1024^2 mesh with reflecting walls and the quadrant densities and
    pressures at rest (no wall separation): total mass and energy after 10
    steps equal the initial ones to round-off (sums over 1M cells, relative
    1e-12)."""
    prob = inputs.quadrants_2d(1024, steps=10)
    prob["v"][:] = 0.0
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        vol = ctx.get(fa.F_CELL_VOLUME)

        def totals():
            st = ctx.state()
            E = st["d"] * (st["e"] + 0.5 * np.sum(st["v"] ** 2, axis=1))
            return np.sum(st["d"] * vol), np.sum(E * vol)

        m0, e0 = totals()
        ctx.advance(prob["cfl"], prob["final_time"], 10, flux_kind=fa.FLUX_HLLC | fa.BOUNDARY_WALL)
        assert ctx.time()[2] == 10
        m1, e1 = totals()
    assert abs(m1 - m0) <= 1e-12 * m0
    assert abs(e1 - e0) <= 1e-12 * e0


def test_maire_512_conserves():
    prob = inputs.maire_bubble_2d(512, steps=10)
    with fa.MaireContext(prob["coords"], prob["cells"]) as ctx:
        ctx.initial_conditions(prob["cv"], prob["gamma"], prob["d"], prob["v"], prob["p"])
        ctx.set_time(0.0, prob["dt0"])

        def totals():
            st = ctx.state()
            E = st["M"] * (st["e"] + 0.5 * np.sum(st["v"] ** 2, axis=1))
            return np.sum(st["M"]), np.sum(E)

        m0, e0 = totals()
        ctx.advance(prob["cfl"], prob["final_time"], 10)
        m1, e1 = totals()
    assert m1 == m0
    assert abs(e1 - e0) <= 1e-12 * e0
