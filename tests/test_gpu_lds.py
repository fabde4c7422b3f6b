"""LDS staging of the Eulerian driver kernels (2D quads): each tile of 256
cells is staged with the cells around it (its halo) when those are at most
80, and its gradients and fluxes read from LDS; other tiles gather from
memory.  Results do not depend on which: the tests hold staged runs, mixed
runs and memory-only runs to the oracle and to each other bit for bit."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import flecsale_amd as fa
from flecsale_amd import inputs
from conftest import rel_linf

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


STAGED = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import flecsale_amd as fa
from flecsale_amd import inputs
prob = inputs.quadrants_2d(64, steps=6)
with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
    lds = ctx.lds_staging()
    ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
    ctx.advance(prob["cfl"], prob["final_time"], 6, prob["flux"])
    print(json.dumps({"lds": lds, "time": ctx.time()[0], "state": {k: v.tolist() for k, v in ctx.state().items()}}))
"""


def test_blocked_order_stages_every_tile_and_matches_oracle(oracle_lib):
    """Spatial blocks of ~256 cells (the default order above 512^2 cells, set
    here for a 64 x 64 mesh) cut a structured mesh into 16 x 16 tiles with 68
    cells around each: every tile is staged, and six MUSCL steps equal the
    oracle bit for bit."""
    O = oracle_lib
    prob = inputs.quadrants_2d(64, steps=6)
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.euler_initial_state(mesh, *prob["eos"], prob["d"], prob["v"], prob["p"])
    rc, t_o, n_o = O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], 6, flux_kind=prob["flux"])
    assert rc == 0 and n_o == 6
    print("lds 64^2, 6 steps, blocked order", flush=True)
    env = dict(os.environ, FSALE_STORAGE_ORDER="blocks")
    r = subprocess.run([sys.executable, "-c", STAGED, ROOT], env=env, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["lds"] == 2 and out["time"] == t_o
    for k, want in [("d", s.d), ("v", s.v), ("e", s.e), ("p", s.p), ("T", s.T), ("a", s.a)]:
        got = np.asarray(out["state"][k])
        assert np.array_equal(got, want), f"{k}: {rel_linf(got, want):.3e}"


SCRIPT = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import flecsale_amd as fa
from flecsale_amd import inputs
sod = dict(inputs.sod_100x10())
sod["flux"] = fa.FLUX_HLLC
small = inputs.riemann_2d(20, "euler", index=2, steps=9)
out = {}
def alone(p, n):
    with fa.EulerContext(p["coords"], p["cells"]) as ctx:
        ctx.initial_conditions(p["eos"], p["d"], p["v"], p["p"])
        ctx.advance(p["cfl"], p["final_time"], n, p["flux"])
        return ctx.lds_staging(), {k: v.tolist() for k, v in ctx.state().items()}
out["sod_lds"], out["sod"] = alone(sod, 9)
out["small_lds"], out["small"] = alone(small, 9)
nc = len(sod["cells"]) + len(small["cells"])
nv = len(sod["coords"]) + len(small["coords"])
with fa.Batch(fa.SOLVER_EULER, 2, 4, 4, nc + 100, 3 * nc + 100, nv + 100, 9 * nc + 100) as b:
    ids = [b.add_euler(p["coords"], p["cells"], p["eos"], p["d"], p["v"], p["p"], p["cfl"], p["final_time"], 9,
                       p["flux"]) for p in (sod, small)]
    out["mix_lds"] = b.lds_staging()
    b.step(9)
    out["mix"] = [{k: v.tolist() for k, v in b.state(i).items()} for i in ids]
print(json.dumps(out))
"""


def _run(order):
    env = dict(os.environ, FSALE_STORAGE_ORDER=order)
    r = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], env=env, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_staged_and_memory_tiles_mix():
    """In the mesh's own order a 100 x 10 tube's tiles (2.6 rows of 100) have
    ~200 cells around them and gather from memory, while a 20 x 20 mesh's
    tiles are staged; in one batch both kinds of tile run in the same
    launches.  Each problem equals itself alone, and the identity-order runs
    equal the blocked order's (every tile staged)."""
    print("lds mix, identity order", flush=True)
    ident = _run("identity")
    print("lds mix, blocked order", flush=True)
    blk = _run("blocks")
    assert ident["sod_lds"] == 1 and ident["small_lds"] == 2 and ident["mix_lds"] == 1
    for out in (ident, blk):
        for got, want in zip(out["mix"], (out["sod"], out["small"])):
            for k in want:
                assert np.array_equal(np.asarray(got[k]), np.asarray(want[k])), k
    for key in ("sod", "small"):
        for k in ident[key]:
            assert np.array_equal(np.asarray(ident[key][k]), np.asarray(blk[key][k])), (key, k)


# ------------------------------------------------------------ 3D plane staging

def _kh_box(nx, ny, nz, steps=5, perm=None):
    """A Kelvin-Helmholtz-like shear state on a jittered nx x ny x nz hex box
    (cells numbered x fastest, or renumbered by perm)."""
    coords, cells = inputs.box_hexes(nx, ny, nz, jitter=0.1, seed=3)
    if perm is not None:
        cells = cells[perm]
    x = coords[cells].mean(axis=1)
    nc = len(cells)
    rng = np.random.default_rng(nx * 10000 + ny * 100 + nz)
    inner = np.abs(x[:, 1] - 0.5) < 0.25
    v = np.zeros((nc, 3))
    v[:, 0] = np.where(inner, 0.5, -0.5)
    v[:, 1] = 0.05 * rng.uniform(-1, 1, nc)
    v[:, 2] = 0.05 * rng.uniform(-1, 1, nc)
    return dict(coords=coords, cells=cells, eos=(1.4, 1.0), d=np.where(inner, 2.0, 1.0) + 0.1 * rng.random(nc),
                v=v, p=2.5 + 0.1 * rng.random(nc), cfl=1.0 / 3.0, final_time=1.0, max_steps=steps, flux=0)


def _oracle_run(O, prob):
    mesh = O.Mesh(prob["coords"], prob["cells"])
    s = O.euler_initial_state(mesh, *prob["eos"], prob["d"], prob["v"], prob["p"])
    rc, t, n = O.euler_run(mesh, s, *prob["eos"], prob["cfl"], prob["final_time"], prob["max_steps"], prob["flux"])
    assert rc == 0 and n == prob["max_steps"]
    return s, t


@pytest.mark.parametrize("shape,staged", [
    ((13, 11, 7), 2),     # rows of 13: tiles straddle rows and planes
    ((9, 8, 7), 2),
    ((2, 2, 2), 2),       # the smallest grid: one tile, every plane clipped
    ((64, 4, 3), 2),
    ((191, 2, 2), 2),     # the widest row the staging holds
    ((192, 2, 2), 0),     # one cell wider: the neighbour planes gather from memory
])
def test_structured_3d_grids_match_oracle(oracle_lib, shape, staged):
    """The gradient kernel stages the neighbour planes of a structured grid's
    tiles in LDS (fsale_lds_staging 2) when its rows hold at most 191 cells;
    staged or not, a run equals the oracle's bit for bit."""
    print("3D plane staging", shape, flush=True)
    prob = _kh_box(*shape)
    s, t = _oracle_run(oracle_lib, prob)
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        assert ctx.lds_staging() == staged
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], prob["max_steps"], flux_kind=prob["flux"])
        got = ctx.state()
        assert ctx.time()[0] == t
    for k in ("d", "v", "e", "p", "T", "a"):
        assert np.array_equal(got[k], np.asarray(getattr(s, k)).reshape(got[k].shape)), k


def test_structured_and_renumbered_3d_grids_mix(oracle_lib):
    """A grid numbered x fastest is staged; the same grid with its cells
    renumbered is not (its neighbour lists are not the grid's in that
    numbering).  In one batch both run in the same launches (the staged
    kernel with its memory path), each equal to the oracle bit for bit."""
    print("3D plane staging: mixed batch", flush=True)
    grid = _kh_box(11, 9, 8, steps=4)
    perm = np.random.default_rng(5).permutation(11 * 9 * 8)
    shuffled = _kh_box(11, 9, 8, steps=4, perm=perm)
    want = [_oracle_run(oracle_lib, p) for p in (grid, shuffled)]
    with fa.EulerContext(shuffled["coords"], shuffled["cells"]) as ctx:
        assert ctx.lds_staging() == 0
    nc = 2 * len(grid["cells"])
    nv = 2 * len(grid["coords"])
    with fa.Batch(fa.SOLVER_EULER, 3, 8, 4, nc + 512, 4 * nc + 512, nv + 512, 27 * nc + 512) as b:
        ids = [b.add_euler(p["coords"], p["cells"], p["eos"], p["d"], p["v"], p["p"], p["cfl"],
                           p["final_time"], p["max_steps"], flux_kind=p["flux"]) for p in (grid, shuffled)]
        assert b.lds_staging() == 1
        b.step(4)
        st = b.status()
        for pid, (s, t) in zip(ids, want):
            assert st[pid].steps == 4 and st[pid].time == t
            got = b.state(pid)
            for k in ("d", "v", "e", "p"):
                assert np.array_equal(got[k], np.asarray(getattr(s, k)).reshape(got[k].shape)), (pid, k)


# ------------------------------------------------ 3D fused step over bricks

@pytest.mark.parametrize("shape,flux", [
    ((8, 8, 4), 0),          # one brick: every side of it on the boundary (HLLE, walls)
    ((16, 8, 8), 2),         # 2 x 1 x 2 bricks (HLLC)
    ((24, 16, 12), 1),       # 3 x 2 x 3 bricks, a brick with neighbours on all six sides (Rusanov)
    ((16, 16, 8), 0x12),     # HLLC with transmissive boundaries
])
def test_brick_grids_match_oracle(oracle_lib, shape, flux):
    """Structured grids with nx, ny multiples of 8 and nz of 4 run the fused
    flux + update kernel over 8 x 8 x 4 bricks (fsale_lds_staging 2 + 4):
    each cell's face states meet its neighbours' in LDS, and the run equals
    the oracle's bit for bit."""
    print("3D bricks", shape, hex(flux), flush=True)
    prob = _kh_box(*shape, steps=5)
    prob["flux"] = flux
    s, t = _oracle_run(oracle_lib, prob)
    with fa.EulerContext(prob["coords"], prob["cells"]) as ctx:
        assert ctx.lds_staging() == 6
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], prob["max_steps"], flux_kind=prob["flux"])
        got = ctx.state()
        assert ctx.time()[0] == t
    for k in ("d", "v", "e", "p", "T", "a"):
        assert np.array_equal(got[k], np.asarray(getattr(s, k)).reshape(got[k].shape)), k


def test_brick_and_gathering_grids_mix(oracle_lib):
    """A batch holding a brick-shaped grid and one that is not runs the
    gathering fused kernel for both (fsale_lds_staging 2, no bricks); alone,
    the brick-shaped grid runs over bricks.  Every run equals the oracle."""
    print("3D bricks: mixed batch", flush=True)
    probs = [_kh_box(16, 8, 8, steps=4), _kh_box(13, 11, 7, steps=4)]
    want = [_oracle_run(oracle_lib, p) for p in probs]
    nc = sum(len(p["cells"]) for p in probs)
    nv = sum(len(p["coords"]) for p in probs)
    for members in ([0], [0, 1]):
        with fa.Batch(fa.SOLVER_EULER, 3, 8, 4, nc + 512, 4 * nc + 512, nv + 512, 27 * nc + 512) as b:
            ids = [b.add_euler(probs[i]["coords"], probs[i]["cells"], probs[i]["eos"], probs[i]["d"],
                               probs[i]["v"], probs[i]["p"], probs[i]["cfl"], probs[i]["final_time"],
                               probs[i]["max_steps"], flux_kind=probs[i]["flux"]) for i in members]
            assert b.lds_staging() == (6 if len(members) == 1 else 2)
            b.step(4)
            st = b.status()
            for pid, i in zip(ids, members):
                s, t = want[i]
                assert st[pid].steps == 4 and st[pid].time == t
                got = b.state(pid)
                for k in ("d", "v", "e", "p"):
                    assert np.array_equal(got[k], np.asarray(getattr(s, k)).reshape(got[k].shape)), (i, k)


def test_two_brick_grids_in_one_batch_graph_and_eager(oracle_lib):
    """Two brick-shaped grids of different extents share the brick launches
    (each brick finds its problem's strides); a graph-replayed run and an
    eager one both equal the oracle bit for bit."""
    print("3D bricks: two grids, graph and eager", flush=True)
    probs = [_kh_box(16, 8, 8, steps=6), _kh_box(8, 16, 12, steps=6)]
    probs[1]["flux"] = 2
    want = [_oracle_run(oracle_lib, p) for p in probs]
    nc = sum(len(p["cells"]) for p in probs)
    nv = sum(len(p["coords"]) for p in probs)
    for graph in (True, False):
        with fa.Batch(fa.SOLVER_EULER, 3, 8, 4, nc + 512, 4 * nc + 512, nv + 512, 27 * nc + 512) as b:
            ids = [b.add_euler(p["coords"], p["cells"], p["eos"], p["d"], p["v"], p["p"], p["cfl"],
                               p["final_time"], p["max_steps"], flux_kind=p["flux"]) for p in probs]
            assert b.lds_staging() == 6
            b.step(6, use_graph=graph)
            st = b.status()
            for pid, (s, t) in zip(ids, want):
                assert st[pid].steps == 6 and st[pid].time == t, (graph, pid)
                got = b.state(pid)
                for k in ("d", "v", "e", "p"):
                    assert np.array_equal(got[k], np.asarray(getattr(s, k)).reshape(got[k].shape)), (graph, pid, k)
