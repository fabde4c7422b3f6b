"""Partitioned runs on the GPU (flecsale_amd.partition): one context per rank,
the owned/ghost split of fsale_eul_set_owned_cells and device-side ghost
refreshes (fsale_halo_exchange).  The reference's two-rank regression tests
(apps/hydro/2d/CMakeLists.txt:50-62, apps/hydro/3d/CMakeLists.txt:50-62)
compare each rank's dump with its .std golden by numdiff --check-text
--absolute 1e-14: here with the partition of those runs read back from
their dumps; each rank's state is also bit-identical to the oracle's
partitioned run (tests/test_partition.py)."""
import gzip
import os
import shutil

import numpy as np
import pytest

import flecsale_amd as fa
from conftest import load_golden, load_mesh
from flecsale_amd import inputs, numdiff, partition
from test_partition import CASES, oracle_partitioned

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _problem(coords, cells, dim):
    import oracle as O
    cen = O.Mesh(coords, cells).geometry()["cell_centroid"]
    left = np.all(cen < 0, axis=1)
    return np.where(left, 0.125, 1.0), np.zeros((len(cells), dim)), np.where(left, 0.1, 1.0)


@pytest.mark.parametrize("dim", [2, 3])
def test_two_rank_regression_dumps(oracle_lib, dim, tmp_path):
    mesh_name, name, cfl, steps = CASES[dim]
    coords, cells = load_mesh(mesh_name)
    goldens = [load_golden(f"{name}_2procs_rank{r}") for r in (0, 1)]
    ranks = partition.from_rank_dumps(coords, cells, goldens)
    d, v, p = _problem(coords, cells, dim)
    meshes, states, t_o = oracle_partitioned(oracle_lib, coords, cells, ranks, d, v, p, cfl, 0.2, steps)
    with partition.PartitionedEuler(coords, cells, ranks) as run:
        run.initial_conditions((1.4, 1.0), d, v, p)
        run.run(cfl, 0.2, steps)
        assert run.time()[0] == t_o
        for r in (0, 1):
            st = run.state(r)
            for k in ("d", "v", "e", "p", "T", "a"):
                want = np.asarray(getattr(states[r], k)).reshape(st[k].shape)
                assert np.array_equal(st[k], want), (r, k)
            dump = run.dump(r, str(tmp_path / f"{name}-solution_rank{r:06d}.txt"))
            std = tmp_path / f"rank{r}.std"
            with gzip.open(os.path.join(HERE, "golden", "std", f"{name}_2procs_rank{r}.std.gz"), "rb") as src, \
                    open(std, "wb") as dst:
                shutil.copyfileobj(src, dst)
            failures = numdiff.compare(std, dump, abs_tol=1e-14, check_text=True)
            assert failures == [], failures[:5]


def test_strips_partition_matches_oracle(oracle_lib):
    """A partition of our own: three strips of the 64^2 four-quadrant problem
    (HLLC, transmissive boundaries), 12 steps, bitwise against the oracle's
    partitioned run, and within 1e-12 of the one-rank GPU run."""
    prob = inputs.quadrants_2d(64, steps=12, flux=fa.FLUX_HLLC | fa.BOUNDARY_TRANSMISSIVE)
    coords, cells = prob["coords"], prob["cells"]
    ranks = partition.strips(coords, cells, 3)
    meshes, states, t_o = oracle_partitioned(oracle_lib, coords, cells, ranks, prob["d"], prob["v"], prob["p"],
                                             prob["cfl"], prob["final_time"], 12, flux_kind=prob["flux"])
    with partition.PartitionedEuler(coords, cells, ranks) as run:
        run.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        run.run(prob["cfl"], prob["final_time"], 12, flux_kind=prob["flux"])
        assert run.time()[0] == t_o
        got = {k: np.zeros((len(cells),) + np.shape(states[0].v)[1:]) if k == "v" else np.zeros(len(cells))
               for k in ("d", "v", "e", "p")}
        for r, R in enumerate(ranks):
            st = run.state(r)
            for k in ("d", "v", "e", "p"):
                want = np.asarray(getattr(states[r], k)).reshape(st[k].shape)
                assert np.array_equal(st[k], want), (r, k)
                got[k][R.cells[:R.n_owned]] = st[k][:R.n_owned]
    with fa.EulerContext(coords, cells) as ctx:
        ctx.initial_conditions(prob["eos"], prob["d"], prob["v"], prob["p"])
        ctx.advance(prob["cfl"], prob["final_time"], 12, flux_kind=prob["flux"])
        one = ctx.state()
    for k in ("d", "v", "e", "p"):
        assert np.max(np.abs(got[k] - one[k])) <= 1e-12 * np.max(np.abs(one[k])), k


def test_partitioned_context_refuses_the_driver_loop():
    coords, cells = inputs.box_quads(16, 16)
    ranks = partition.strips(coords, cells, 2)
    with partition.PartitionedEuler(coords, cells, ranks) as run:
        run.initial_conditions((1.4, 1.0), np.ones(len(cells)), np.zeros((len(cells), 2)), np.ones(len(cells)))
        with pytest.raises(fa.FsaleError):
            run.ctx[0].advance(0.5, 1.0, 2)
        with pytest.raises(fa.FsaleError):
            fa._check(fa.lib().fsale_eul_set_owned_cells(run.ctx[0].h, len(ranks[0].cells) + 1))
