"""Partitioned meshes: the reference's multi-rank apps/hydro runs on one device.

The reference colours the mesh (ParMETIS through flecsi-sp) and runs one MPI
rank per colour.  Each rank holds its owned cells -- exclusive ones, then
shared ones (ghosts of another rank) -- followed by its ghost cells, one
layer of cells sharing a vertex with an owned cell; FleCSI copies owner
values into ghosts after a task writes a field that a later task reads on
the ghosts (apps/hydro/driver.h:317-413): the state after apply_update, the
gradients before evaluate_fluxes.  evaluate_fluxes recomputes the faces of
shared cells on every rank instead of exchanging fluxes
(apps/hydro/tasks.h:466-534); evaluate_time_step and apply_update work on the
owned cells (tasks.h:350,614) and the time step is a min-reduction over the
ranks (driver.h:329-336).

Here each rank's cells are a context of their own on the one GPU, built
from the rank's local mesh (its cells in local order, its vertices in local
order), with the owned/ghost split set by fsale_eul_set_owned_cells and the
ghost refreshes done by fsale_halo_exchange, device to device.  The gradient
neighbour lists and face orders follow each rank's local numbering, as the
reference's per-rank meshes do, so a rank's results are the reference's
rank's (the two-rank shock_box goldens, digit for digit), not the one-rank
run's bits.

    ranks = partition.from_rank_dumps(coords, cells, [dump0, dump1])   # the reference's layout
    ranks = partition.strips(coords, cells, 4)                          # or a partition of our own
    run = partition.PartitionedEuler(coords, cells, ranks)
    run.initial_conditions((1.4, 1.0), d, v, p)   # per global cell
    run.run(cfl=0.5, final_time=0.2, max_steps=20)
    run.dump(0, "shock_box_2d-solution_rank000000.txt")
"""
import ctypes as ct

import numpy as np

import flecsale_amd as fa


class RankMesh:
    """One rank's cells (global ids, local order: owned first), the number
    owned, its vertices (global ids, local order), the ids its dump prints,
    and for every ghost the owner (rank, local id)."""

    def __init__(self, cells, n_owned, vertices, cell_gid=None, vertex_gid=None):
        self.cells = np.ascontiguousarray(cells, dtype=np.int64)
        self.n_owned = int(n_owned)
        self.vertices = np.ascontiguousarray(vertices, dtype=np.int64)
        self.cell_gid = self.cells if cell_gid is None else np.asarray(cell_gid, dtype=np.int64)
        self.vertex_gid = self.vertices if vertex_gid is None else np.asarray(vertex_gid, dtype=np.int64)
        self.ghost_owner = None

    def local_mesh(self, coords, cells):
        """Coordinates and cell -> vertex table of the rank's sub-mesh, in local ids."""
        local = np.full(len(coords), -1, dtype=np.int64)
        local[self.vertices] = np.arange(len(self.vertices))
        lc = local[np.asarray(cells)[self.cells]]
        if np.any(lc < 0):
            raise ValueError("a rank's cell uses a vertex the rank does not list")
        return np.ascontiguousarray(np.asarray(coords)[self.vertices]), lc.astype(np.int32)


def _link_ghosts(ranks, ncells):
    """Owner of every global cell and, per rank, (owner rank, owner local id) of its ghosts."""
    owner = np.full(ncells, -1, dtype=np.int64)
    owner_local = np.full(ncells, -1, dtype=np.int64)
    for r, R in enumerate(ranks):
        own = R.cells[:R.n_owned]
        if np.any(owner[own] >= 0):
            raise ValueError("a cell is owned by two ranks")
        owner[own] = r
        owner_local[own] = np.arange(R.n_owned)
    if np.any(owner < 0):
        raise ValueError("a cell is owned by no rank")
    for R in ranks:
        g = R.cells[R.n_owned:]
        R.ghost_owner = np.stack([owner[g], owner_local[g]], axis=1) if len(g) else np.zeros((0, 2), np.int64)
    return ranks


def _vertex_neighbours(cells, ncells, nverts):
    """CSR of the cells around each vertex."""
    cells = np.asarray(cells)
    flat = cells.reshape(-1)
    order = np.argsort(flat, kind="stable")
    off = np.zeros(nverts + 1, dtype=np.int64)
    np.add.at(off, flat + 1, 1)
    return np.cumsum(off), (order // cells.shape[1])


def strips(coords, cells, nranks, axis=None):
    """A partition of our own: cells cut into `nranks` slabs of equal count
    along `axis` (default: the longest extent) by centroid; each rank's
    ghosts are the cells sharing a vertex with its owned cells.  Local order
    as FleCSI lays a rank out: exclusive cells, shared cells, ghosts, each by
    global id; local vertices in order of first use by the local cells."""
    coords = np.asarray(coords)
    cells = np.asarray(cells)
    nc, nv = len(cells), len(coords)
    cen = coords[cells].mean(axis=1)
    if axis is None:
        axis = int(np.argmax(coords.max(axis=0) - coords.min(axis=0)))
    order = np.lexsort((np.arange(nc), cen[:, axis]))
    part = np.empty(nc, dtype=np.int64)
    part[order] = np.arange(nc) * nranks // nc
    voff, vcell = _vertex_neighbours(cells, nc, nv)
    ranks = []
    ghost_sets = []
    for r in range(nranks):
        own = np.nonzero(part == r)[0]
        verts = np.unique(cells[own].reshape(-1))
        near = np.unique(np.concatenate([vcell[voff[v]:voff[v + 1]] for v in verts]))
        ghost_sets.append(near[part[near] != r])
    shared = np.zeros(nc, dtype=bool)
    for g in ghost_sets:
        shared[g] = True
    for r in range(nranks):
        own = np.nonzero(part == r)[0]
        local = np.concatenate([own[~shared[own]], own[shared[own]], ghost_sets[r]])
        _, first = np.unique(cells[local].reshape(-1), return_index=True)
        verts = cells[local].reshape(-1)[np.sort(first)]
        ranks.append(RankMesh(local, len(own), verts))
    return _link_ghosts(ranks, nc)


def from_rank_dumps(coords, cells, dumps):
    """The partition of the reference's own multi-rank run, read back from its
    per-rank dumps (apps/hydro/tasks.h:771-849: each rank prints its cells and
    vertices in local order with their global ids).  A dump's cells are
    matched to the mesh's by centroid; its vertex ids are the mesh's.  FleCSI
    numbers each rank's owned cells contiguously, so the owner of a global id
    is the rank whose exclusive cells (those no other rank lists) start that
    range.  dumps: dicts with cell_gid, cell_centroid, vertex_gid."""
    from scipy.spatial import cKDTree
    coords = np.asarray(coords)
    cells = np.asarray(cells)
    nc = len(cells)
    cen = coords[cells].mean(axis=1)
    tree = cKDTree(cen)
    spacing = np.min(np.linalg.norm(coords[cells[:, 1]] - coords[cells[:, 0]], axis=1))
    mapped = []
    for r, dmp in enumerate(dumps):
        dist, ids = tree.query(np.asarray(dmp["cell_centroid"]))
        if np.any(dist > 0.25 * spacing) or len(np.unique(ids)) != len(ids):
            raise ValueError(f"rank {r}: dump centroids do not match the mesh's cells")
        mapped.append(ids.astype(np.int64))
    count = np.zeros(nc, dtype=np.int64)
    for ids in mapped:
        np.add.at(count, ids, 1)
    gids = [np.asarray(d["cell_gid"], dtype=np.int64) for d in dumps]
    # each dump lists its exclusive cells (listed by no other rank), then its
    # shared and its ghost cells; its owned cells -- the exclusive ones and a
    # leading part of the rest -- carry one contiguous range of global ids,
    # the ranges following each other in rank order
    excl = [int(np.argmax(count[m] > 1)) if np.any(count[m] > 1) else len(m) for m in mapped]
    order = sorted(range(len(dumps)), key=lambda r: int(gids[r][:excl[r]].min()))
    n_owned = [0] * len(dumps)

    def search(k, offset, owner):
        if k == len(order):
            return offset == nc and np.all(owner >= 0)
        r = order[k]
        for n in range(excl[r], len(mapped[r]) + 1):
            g = gids[r][:n]
            if g.min() != offset or g.max() != offset + n - 1 or len(np.unique(g)) != n:
                continue
            cells_r = mapped[r][:n]
            if np.any(owner[cells_r] >= 0):
                continue
            owner[cells_r] = r
            n_owned[r] = n
            if search(k + 1, offset + n, owner):
                return True
            owner[cells_r] = -1
        return False

    if not search(0, 0, np.full(nc, -1, dtype=np.int64)):
        raise ValueError("the dumps' cells do not split into contiguous owned ranges")
    ranks = [RankMesh(mapped[r], n_owned[r], np.asarray(d["vertex_gid"], dtype=np.int64), cell_gid=gids[r],
                      vertex_gid=np.asarray(d["vertex_gid"]))
             for r, d in enumerate(dumps)]
    return _link_ghosts(ranks, nc)


class PartitionedEuler:
    """The apps/hydro step loop over the ranks of a partition, one context
    per rank, ghosts refreshed device to device."""

    def __init__(self, coords, cells, ranks):
        self.ranks = ranks
        self.coords = np.asarray(coords)
        self.cells = np.asarray(cells)
        self.ctx = []
        self.halos = []
        self.eos = None
        for R in ranks:
            lx, lc = R.local_mesh(coords, cells)
            ctx = fa.EulerContext(lx, lc)
            self.ctx.append(ctx)
        # one plan per (ghost rank, owner rank) pair
        for r, R in enumerate(ranks):
            for o in range(len(ranks)):
                sel = np.nonzero(R.ghost_owner[:, 0] == o)[0] if len(R.ghost_owner) else np.zeros(0, np.int64)
                if len(sel) == 0:
                    continue
                dst = np.ascontiguousarray(R.n_owned + sel, dtype=np.int32)
                src = np.ascontiguousarray(R.ghost_owner[sel, 1], dtype=np.int32)
                h = ct.c_void_p()
                fa._check(fa.lib().fsale_halo_create(self.ctx[r].h, self.ctx[o].h, len(dst), dst.ctypes.data,
                                                     src.ctypes.data, ct.byref(h)))
                self.halos.append(h)
        self.steps = 0

    def close(self):
        for h in self.halos:
            fa.lib().fsale_halo_destroy(h)
        self.halos = []
        for c in self.ctx:
            c.close()
        self.ctx = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _exchange(self, what):
        for h in self.halos:
            fa._check(fa.lib().fsale_halo_exchange(h, what))

    def initial_conditions(self, eos, d, v, p):
        """initial_conditions (apps/hydro/tasks.h:46-68) on every rank's cells
        (ghosts included), from per-global-cell d, v, p."""
        self.eos = eos
        d, v, p = np.asarray(d), np.asarray(v), np.asarray(p)
        for R, ctx in zip(self.ranks, self.ctx):
            ctx.initial_conditions(eos, d[R.cells], v[R.cells], p[R.cells])
            fa._check(fa.lib().fsale_eul_set_owned_cells(ctx.h, R.n_owned))
        self.steps = 0

    def time(self):
        return self.ctx[0].time()

    def run(self, cfl, final_time, max_steps, flux_kind=fa.FLUX_HLLE):
        """apps/hydro/driver.h:317-413: per step the min over ranks of each
        rank's time step, then twice gradients -> ghost gradients -> fluxes
        -> apply_update(1/2) -> ghost state."""
        for _ in range(max_steps):
            t = self.ctx[0].time()[0]
            dt = min(ctx.evaluate_time_step(cfl, final_time) for ctx in self.ctx)
            for ctx in self.ctx:
                ctx.set_time(t, dt)
            for _half in range(2):
                for ctx in self.ctx:
                    ctx.evaluate_gradients()
                self._exchange(fa.HALO_GRADIENTS)
                for ctx in self.ctx:
                    ctx.evaluate_fluxes(flux_kind)
                for ctx in self.ctx:
                    ctx.apply_update(0.5)
                self._exchange(fa.HALO_STATE)
            self.steps += 1
        for ctx in self.ctx:
            ctx.sync()

    def state(self, rank):
        return self.ctx[rank].state()

    def dump(self, rank, path):
        """The dump task's file for one rank (apps/hydro/tasks.h:771-849): its
        cells and vertices in local order with their global ids."""
        R, ctx = self.ranks[rank], self.ctx[rank]
        st = ctx.state()
        cen = ctx.get(fa.F_CELL_CENTROID)
        x = ctx.get(fa.F_COORDS)
        t = ctx.time()[0]
        N = cen.shape[1]
        with open(path, "w") as f:
            f.write("# Solution time: %.14e\n# Number iterations: %d\n" % (t, self.steps))
            f.write("# BEGIN CELLS\n# Total number: %d\n# local_id global_id " % len(R.cells))
            f.write("".join("centroid(%d) " % k for k in range(N)))
            f.write("density internal_energy pressure ")
            f.write("".join("velocity(%d) " % k for k in range(N)) + "\n")
            for c in range(len(R.cells)):
                f.write("%d %d " % (c, R.cell_gid[c]) + "".join("%.14e " % q for q in cen[c]) +
                        "%.14e %.14e %.14e " % (st["d"][c], st["e"][c], st["p"][c]) +
                        "".join("%.14e " % q for q in st["v"][c]) + "\n")
            f.write("# END CELLS\n# BEGIN VERTICES\n# Total number: %d\n# local_id global_id " % len(R.vertices))
            f.write("".join("coordinate(%d) " % k for k in range(N)) + "\n")
            for i in range(len(R.vertices)):
                f.write("%d %d " % (i, R.vertex_gid[i]) + "".join("%.14e " % q for q in x[i]) + "\n")
            f.write("# END VERTICES\n")
        return path
