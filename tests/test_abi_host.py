"""CPU-side checks of the product library: it loads, exports every symbol
include/flecsale_amd.h declares, and its host connectivity builder matches
the oracle's mesh layer (face numbering, face->cell owner, neighbours)."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import flecsale_amd as fa
from flecsale_amd import inputs
from conftest import ROOT, load_mesh


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "flecsale_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fsale_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = fa.lib()
    declared = _header_symbols()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(fa.EXPORTS) == declared


def test_status_strings():
    L = fa.lib()
    assert L.fsale_status_string(0) == b"ok"
    assert b"negative" in L.fsale_status_string(fa.ERR_NEG_STATE)


def test_context_without_device_fails_loudly():
    if os.environ.get("HIP_VISIBLE_DEVICES", None) is None and os.path.exists("/dev/kfd"):
        pytest.skip("a device may be present")
    coords, cells = inputs.box_quads(4, 4)
    with pytest.raises(fa.FsaleError) as ei:
        fa.EulerContext(coords, cells)
    assert ei.value.status == fa.ERR_NO_DEVICE


@pytest.mark.parametrize("name", ["square_32x32", "sedov_32x32", "sedov_20x20x20"])
def test_connectivity_matches_oracle_reference_meshes(oracle_lib, name):
    coords, cells = load_mesh(name)
    got = fa.connectivity(coords, cells)
    want = oracle_lib.Mesh(coords, cells).topology()
    for k in ("face_cells", "cell_faces", "nbr_off", "nbr"):
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.parametrize("gen", [
    lambda: inputs.box_quads(7, 5, jitter=0.2),
    lambda: inputs.box_quads(1, 9),
    lambda: inputs.box_hexes(3, 4, 2, jitter=0.1),
    lambda: inputs.box_hexes(1, 1, 1),
])
def test_connectivity_matches_oracle_generated(oracle_lib, gen):
    coords, cells = gen()
    got = fa.connectivity(coords, cells)
    want = oracle_lib.Mesh(coords, cells).topology()
    for k in ("face_cells", "cell_faces", "nbr_off", "nbr"):
        assert np.array_equal(got[k], want[k]), k


def test_connectivity_invariants():
    coords, cells = inputs.box_quads(16, 9)
    t = fa.connectivity(coords, cells)
    nx, ny = 16, 9
    assert len(t["face_cells"]) == nx * (ny + 1) + ny * (nx + 1)
    fc = t["face_cells"]
    assert np.all(fc[:, 0] >= 0)
    inner = fc[:, 1] >= 0
    assert np.all(fc[inner, 0] < fc[inner, 1])     # cells[0] is the lower id
    assert np.sum(~inner) == 2 * (nx + ny)
    # each cell lists each of its faces once
    assert all(len(set(r)) == 4 for r in t["cell_faces"])
    # vertex neighbours of an interior quad: 8, corner: 3
    cnt = np.diff(t["nbr_off"])
    assert cnt.max() == 8 and cnt.min() == 3


def test_bad_mesh_rejected():
    coords, cells = inputs.box_quads(2, 2)
    bad = cells.copy()
    bad[0, 0] = 999
    with pytest.raises(fa.FsaleError) as ei:
        fa.connectivity(coords, bad)
    assert ei.value.status == 2


def test_qr_solve_matches_oracle_bitwise(oracle_lib):
    """The library's register-resident QR (the nodal kernel's symmetry solve,
    physically swapped pivot columns) against the oracle's restatement of
    flecsale::linalg::qr, bit for bit: random systems, the Maire augmented
    systems [M N^T; N 0] with one and two symmetry rows, and degenerate ones
    (zero columns, which exercise the pivot search over processed columns and
    the persistent Householder vector)."""
    import flecsale_amd as fa
    rng = np.random.default_rng(11)
    cases = []
    for n in (1, 2, 3, 4):
        for _ in range(200):
            cases.append((rng.standard_normal((n, n)) * rng.choice([1e-3, 1.0, 1e3]), rng.standard_normal(n)))
    for _ in range(300):
        M = rng.standard_normal((2, 2))
        M = M @ M.T + 0.1 * np.eye(2)
        s1, s2 = rng.standard_normal(2), rng.standard_normal(2)
        b = np.concatenate([rng.standard_normal(2), [0.0, 0.0]])
        cases.append((np.block([[M, s1[:, None]], [s1[None, :], np.zeros((1, 1))]]), b[:3]))
        cases.append((np.block([[M, np.stack([s1, s2], 1)], [np.stack([s1, s2], 0), np.zeros((2, 2))]]), b))
        axis = np.array([[1.0, 0.0], [0.0, 1.0]])[rng.integers(2)] * rng.uniform(0.01, 0.1)
        cases.append((np.block([[M, axis[:, None]], [axis[None, :], np.zeros((1, 1))]]), b[:3]))
    for n in (3, 4):
        for _ in range(50):
            A = rng.standard_normal((n, n))
            A[:, rng.integers(n)] = 0.0
            cases.append((A, rng.standard_normal(n)))
        cases.append((np.zeros((n, n)), rng.standard_normal(n)))
    for A, b in cases:
        want = oracle_lib.qr_solve(A, b)
        got = fa.qr_solve(A, b)
        assert np.array_equal(got, want, equal_nan=True), (A, b, got, want)
