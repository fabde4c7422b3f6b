"""Write the committed fixtures under tests/golden/ from the reference's own test data.

Run in the development container only (it reads /root/reference, which the GPU
machines do not have):  python tests/golden/make_fixtures.py

Outputs (all plain numpy .npz, loaded with allow_pickle=False):
  meshes/<name>.npz   coords [nv, dim] float64, cells [nc, vpc] int32 (0-based,
                      Exodus node order) read from data/meshes/<name>.g
  <case>.npz          the reference's regression golden (.std text dump) as
                      arrays: time, iterations, cell_centroid, density,
                      internal_energy, pressure, velocity, vertex_coords.
  exodus/<file>       the Exodus-II mesh files themselves (data/meshes), for
                      the library's own reader and the regression drivers.
  std/<case>.std.gz   the .std goldens as text (gzip), for the numdiff
                      comparison of the drivers' dumps (one rank, and each
                      rank of the two-rank runs).

The Lua input decks under lua/ are written for this repository, not taken
from the reference.

Mesh sources: data/meshes/sedov_32x32.g and sedov_20x20x20.g are
netCDF-classic (64-bit offset) read with scipy.io.netcdf_file;
data/meshes/square_32x32.g is netCDF-4/HDF5, read through the h5dump tool's
raw binary export so no digits are lost.
Golden sources: apps/hydro/2d/shock_box_2d-solution_rank000000.txt.1procs.std,
apps/hydro/3d/shock_box_3d-solution_rank000000.txt.1procs.std,
apps/maire_hydro/2d/sedov_2d-solution_rank000000.txt.1procs.std and the
2-rank dumps beside them.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
from scipy.io import netcdf_file

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
H5DUMP = "/opt/conda/bin/h5dump"


def read_netcdf(path):
    f = netcdf_file(path, "r", mmap=False)
    dim = f.dimensions["num_dim"]
    names = ["coordx", "coordy", "coordz"][:dim]
    coords = np.stack([np.asarray(f.variables[n][:], dtype=np.float64) for n in names], axis=1)
    cells = np.asarray(f.variables["connect1"][:], dtype=np.int64) - 1
    f.close()
    return coords, cells.astype(np.int32)


def read_hdf5(path):
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, dt in [("coordx", "<f8"), ("coordy", "<f8"), ("connect1", "<i4")]:
            fn = os.path.join(td, name + ".bin")
            subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", fn, path], check=True,
                           stdout=subprocess.DEVNULL)
            out[name] = np.fromfile(fn, dtype=dt)
    coords = np.stack([out["coordx"], out["coordy"]], axis=1)
    cells = out["connect1"].reshape(-1, 4).astype(np.int64) - 1
    return coords, cells.astype(np.int32)


def read_golden(path, dim):
    lines = open(path).read().splitlines()
    time = float(lines[0].split(":")[1])
    iters = int(lines[1].split(":")[1])
    i = lines.index("# BEGIN CELLS")
    ncell = int(lines[i + 1].split(":")[1])
    rows = np.array([[float(t) for t in l.split()] for l in lines[i + 3:i + 3 + ncell]])
    j = lines.index("# BEGIN VERTICES")
    nvert = int(lines[j + 1].split(":")[1])
    vrows = np.array([[float(t) for t in l.split()] for l in lines[j + 3:j + 3 + nvert]])
    return dict(
        time=np.float64(time), iterations=np.int64(iters),
        cell_lid=rows[:, 0].astype(np.int64), cell_gid=rows[:, 1].astype(np.int64),
        cell_centroid=rows[:, 2:2 + dim], density=rows[:, 2 + dim],
        internal_energy=rows[:, 3 + dim], pressure=rows[:, 4 + dim],
        velocity=rows[:, 5 + dim:5 + 2 * dim],
        vertex_lid=vrows[:, 0].astype(np.int64), vertex_gid=vrows[:, 1].astype(np.int64),
        vertex_coords=vrows[:, 2:2 + dim])


def main():
    os.makedirs(os.path.join(HERE, "meshes"), exist_ok=True)
    meshes = {
        "square_32x32": read_hdf5(f"{REF}/data/meshes/square_32x32.g"),
        "sedov_32x32": read_netcdf(f"{REF}/data/meshes/sedov_32x32.g"),
        "sedov_20x20x20": read_netcdf(f"{REF}/data/meshes/sedov_20x20x20.g"),
    }
    for name, (coords, cells) in meshes.items():
        np.savez_compressed(os.path.join(HERE, "meshes", name + ".npz"), coords=coords, cells=cells)
        print(name, coords.shape, cells.shape)
    goldens = {
        "shock_box_2d_1procs": (f"{REF}/apps/hydro/2d/shock_box_2d-solution_rank000000.txt.1procs.std", 2),
        "shock_box_2d_2procs_rank0": (f"{REF}/apps/hydro/2d/shock_box_2d-solution_rank000000.txt.2procs.std", 2),
        "shock_box_2d_2procs_rank1": (f"{REF}/apps/hydro/2d/shock_box_2d-solution_rank000001.txt.2procs.std", 2),
        "shock_box_3d_1procs": (f"{REF}/apps/hydro/3d/shock_box_3d-solution_rank000000.txt.1procs.std", 3),
        "shock_box_3d_2procs_rank0": (f"{REF}/apps/hydro/3d/shock_box_3d-solution_rank000000.txt.2procs.std", 3),
        "shock_box_3d_2procs_rank1": (f"{REF}/apps/hydro/3d/shock_box_3d-solution_rank000001.txt.2procs.std", 3),
        "sedov_2d_1procs": (f"{REF}/apps/maire_hydro/2d/sedov_2d-solution_rank000000.txt.1procs.std", 2),
    }
    for name, (path, dim) in goldens.items():
        g = read_golden(path, dim)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **g)
        print(name, g["time"], g["iterations"], g["density"].shape)


    # the data files themselves: Exodus meshes and text goldens
    import gzip
    import shutil
    os.makedirs(os.path.join(HERE, "exodus"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "std"), exist_ok=True)
    for name in ["square_32x32.g", "sedov_32x32.g", "sedov_20x20x20.g", "cube_3k_tet.g", "uniform_32.exo"]:
        shutil.copyfile(f"{REF}/data/meshes/{name}", os.path.join(HERE, "exodus", name))
    for name in ["shock_box_2d_1procs", "shock_box_3d_1procs", "sedov_2d_1procs", "shock_box_2d_2procs_rank0",
                 "shock_box_2d_2procs_rank1", "shock_box_3d_2procs_rank0", "shock_box_3d_2procs_rank1"]:
        with open(goldens[name][0], "rb") as src, gzip.GzipFile(os.path.join(HERE, "std", name + ".std.gz"), "wb",
                                                                mtime=0) as dst:
            shutil.copyfileobj(src, dst)
        print("copied", name)


if __name__ == "__main__":
    sys.exit(main())
