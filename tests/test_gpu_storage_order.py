"""The device storage order of cells, vertices and faces (FSALE_STORAGE_ORDER:
identity, or a Z-order curve of the entities or of blocks of ~256 of them)
changes memory locality only: every ordered sum in the kernels follows the
original ids.  Each order runs in a process of its own (the mode is read once)
and must reproduce the identity order's results bit for bit, for the Eulerian
and Lagrangian steps and for the reported failing cell."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import flecsale_amd as fa
from flecsale_amd import inputs
out = {}
p = inputs.quadrants_2d(n=37, steps=6)
with fa.EulerContext(p["coords"], p["cells"]) as ctx:
    ctx.initial_conditions(p["eos"], p["d"], p["v"], p["p"])
    ctx.advance(p["cfl"], p["final_time"], 6, p["flux"])
    out["eul2d"] = {k: v.tolist() for k, v in ctx.state().items()}
p = inputs.kelvin_helmholtz_3d(n=9, steps=4)
with fa.EulerContext(p["coords"], p["cells"]) as ctx:
    ctx.initial_conditions(p["eos"], p["d"], p["v"], p["p"])
    ctx.advance(p["cfl"], p["final_time"], 4, p["flux"])
    out["eul3d"] = {k: v.tolist() for k, v in ctx.state().items()}
p = inputs.maire_bubble_2d(21, 7)
with fa.MaireContext(p["coords"], p["cells"]) as ctx:
    ctx.initial_conditions(p["cv"], p["gamma"], p["d"], p["v"], p["p"])
    ctx.set_time(0.0, p["dt0"])
    ctx.advance(p["cfl"], p["final_time"], 7)
    st = ctx.state()
    st["x"] = ctx.get(fa.F_COORDS)
    st["Fpc"] = ctx.get(fa.F_CORNER_FORCE)
    out["maire"] = {k: v.tolist() for k, v in st.items()}
p = inputs.sod_100x10()
with fa.EulerContext(p["coords"], p["cells"]) as ctx:
    ctx.initial_conditions(p["eos"], p["d"], p["v"], p["p"])
    try:
        ctx.advance(50.0, 0.2, 20, first_order=True)
        ctx.sync()
    except fa.FsaleError:
        pass
    out["error_cell"] = ctx.error_cell()
print(json.dumps(out))
"""


def _run(order):
    env = dict(os.environ, FSALE_STORAGE_ORDER=order)
    r = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], env=env, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_storage_orders_give_identical_results():
    ref = _run("identity")
    assert ref["error_cell"] >= 0
    for order in ("morton", "blocks"):
        got = _run(order)
        assert got["error_cell"] == ref["error_cell"], order
        for case in ("eul2d", "eul3d", "maire"):
            for k, v in ref[case].items():
                assert np.array_equal(np.array(got[case][k]), np.array(v)), (order, case, k)
