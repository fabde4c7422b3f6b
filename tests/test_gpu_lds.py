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
