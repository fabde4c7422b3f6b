"""Problem inputs: meshes and initial conditions (the role of the reference's
apps/hydro/{2d,3d}/inputs.cc and apps/maire_hydro/2d/inputs*.cc).

Everything is generated on the host, deterministically, so the device path
and the CPU restatement consume bit-identical inputs.  Random numbers come
from splitmix64 over (seed, stream, counter) -- a counter-based generator, so
any array can be regenerated independently of the others.

Structured meshes are numbered row-major with x fastest; quads are
counter-clockwise from their lower-left vertex, hexes follow the Exodus hex8
order (bottom face counter-clockwise seen from below the cell's top).
"""
import math

import numpy as np

MASTER_SEED = 20261203
_M64 = (1 << 64) - 1


def splitmix64(x):
    """splitmix64 finaliser over a uint64 array (Steele, Lea, Flood 2014)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, stream, n, lo=0.0, hi=1.0):
    """n doubles in [lo, hi) from counters 0..n-1 of (seed, stream)."""
    with np.errstate(over="ignore"):
        base = splitmix64(np.uint64(seed) ^ (np.uint64(stream) * np.uint64(0xD1B54A32D192ED03)))
        z = splitmix64(base + np.arange(n, dtype=np.uint64))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return lo + (hi - lo) * u


def box_quads(nx, ny, x0=0.0, x1=1.0, y0=0.0, y1=1.0, jitter=0.0, seed=MASTER_SEED, stream=1):
    """Quad mesh of [x0,x1]x[y0,y1]; interior vertices moved by U[-jitter, jitter]*spacing per axis."""
    xs = x0 + (x1 - x0) * np.arange(nx + 1) / nx
    ys = y0 + (y1 - y0) * np.arange(ny + 1) / ny
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    coords = np.stack([X.ravel(), Y.ravel()], axis=1)
    if jitter:
        nv = coords.shape[0]
        i = np.arange(nv) % (nx + 1)
        j = np.arange(nv) // (nx + 1)
        interior = (i > 0) & (i < nx) & (j > 0) & (j < ny)
        dx, dy = (x1 - x0) / nx, (y1 - y0) / ny
        jx = uniform(seed, stream, nv, -jitter, jitter) * dx
        jy = uniform(seed, stream + 1, nv, -jitter, jitter) * dy
        coords[interior, 0] += jx[interior]
        coords[interior, 1] += jy[interior]
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    i, j = i.ravel(), j.ravel()
    v00 = j * (nx + 1) + i
    cells = np.stack([v00, v00 + 1, v00 + nx + 2, v00 + nx + 1], axis=1).astype(np.int32)
    return coords, cells


def box_hexes(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), jitter=0.0, seed=MASTER_SEED, stream=11):
    """Hex mesh of a box; interior vertices moved by U[-jitter, jitter]*spacing per axis."""
    ax = [lo[d] + (hi[d] - lo[d]) * np.arange(n + 1) / n for d, n in enumerate((nx, ny, nz))]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    coords = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    if jitter:
        nv = coords.shape[0]
        i = np.arange(nv) % (nx + 1)
        j = (np.arange(nv) // (nx + 1)) % (ny + 1)
        k = np.arange(nv) // ((nx + 1) * (ny + 1))
        interior = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
        for d, n in enumerate((nx, ny, nz)):
            h = (hi[d] - lo[d]) / n
            jd = uniform(seed, stream + d, nv, -jitter, jitter) * h
            coords[interior, d] += jd[interior]
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    v0 = k * sz + j * sy + i
    cells = np.stack([v0, v0 + sx, v0 + sx + sy, v0 + sy,
                      v0 + sz, v0 + sz + sx, v0 + sz + sx + sy, v0 + sz + sy], axis=1).astype(np.int32)
    return coords, cells


def quad_centroids(coords, cells):
    """Area-weighted centroids (shoelace about the origin, as the library's geometry)."""
    p = coords[cells]
    x, y = p[..., 0], p[..., 1]
    xn, yn = np.roll(x, -1, axis=1), np.roll(y, -1, axis=1)
    cr = x * yn - xn * y
    a2 = np.zeros(len(cells))
    sx = np.zeros(len(cells))
    sy = np.zeros(len(cells))
    for i in range(cells.shape[1]):   # sequential sums, same order as the device
        a2 = a2 + cr[:, i]
        sx = sx + (x[:, i] + xn[:, i]) * cr[:, i]
        sy = sy + (y[:, i] + yn[:, i]) * cr[:, i]
    area = 0.5 * a2
    return np.stack([sx / (6.0 * area), sy / (6.0 * area)], axis=1)


# ---------------------------------------------------------------- problems
# Each returns a dict: coords, cells and solver inputs.

def shock_box_2d(coords, cells, centroids):
    """apps/hydro/2d/inputs.cc:27-60 (CFL 1/2, t_f 0.2, 20 steps, gamma 1.4)."""
    x = centroids
    left = (x[:, 0] < 0) & (x[:, 1] < 0)
    d = np.where(left, 0.125, 1.0)
    p = np.where(left, 0.1, 1.0)
    return dict(coords=coords, cells=cells, eos=(1.4, 1.0), d=d, v=np.zeros((len(d), 2)), p=p,
                cfl=0.5, final_time=0.2, max_steps=20, flux=0)


def shock_box_3d(coords, cells, centroids):
    """apps/hydro/3d/inputs.cc:38-75 (CFL 1/3, t_f 0.2, 12 steps)."""
    x = centroids
    left = (x[:, 0] < 0) & (x[:, 1] < 0) & (x[:, 2] < 0)
    d = np.where(left, 0.125, 1.0)
    p = np.where(left, 0.1, 1.0)
    return dict(coords=coords, cells=cells, eos=(1.4, 1.0), d=d, v=np.zeros((len(d), 3)), p=p,
                cfl=1.0 / 3.0, final_time=0.2, max_steps=12, flux=0)


def sedov_2d(coords, cells, centroids):
    """apps/maire_hydro/2d/inputs_regression.cc:36-100 (32x32 on [0,1.2]^2)."""
    dx = 1.2 / 32
    dy = 1.2 / 32
    vol = dx * dy
    delta_r = np.finfo(np.float64).eps + math.sqrt((dx / 2) ** 2 + (dy / 2) ** 2)
    x = centroids
    r = np.sqrt(x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1])
    p = np.where(r < delta_r, (1.4 - 1) * 1.0 * 0.244816 / vol, 1.e-6)
    nc = len(p)
    return dict(coords=coords, cells=cells, cv=1.0, gamma=np.full(nc, 1.4), d=np.ones(nc),
                v=np.zeros((nc, 2)), p=p, cfl=(0.25, 0.1, 1.01), final_time=1.0, max_steps=20, dt0=1.e-5)


def sod_100x10():
    """Sod tube on [0,1]x[0,0.1], 100x10 quads, first order, HLLE, CFL 0.5, to t=0.2."""
    coords, cells = box_quads(100, 10, 0.0, 1.0, 0.0, 0.1)
    x = quad_centroids(coords, cells)
    left = x[:, 0] < 0.5
    d = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    return dict(coords=coords, cells=cells, eos=(1.4, 1.0), d=d, v=np.zeros((len(d), 2)), p=p,
                cfl=0.5, final_time=0.2, max_steps=100000, flux=0, first_order=True, stop_at_final=True)


def quadrants_2d(n=1024, steps=200, flux=2 | 0x10):
    """Four-quadrant Riemann problem on [0,1]^2, n x n quads, split at (0.5, 0.5).

    Transmissive boundaries by default (flux |= 0x10): with the reference's
    reflecting walls the supersonic SW state pulls away from the corner walls
    and the reference arithmetic reaches negative internal energy at step 6."""
    coords, cells = box_quads(n, n)
    x = quad_centroids(coords, cells)
    east, north = x[:, 0] >= 0.5, x[:, 1] >= 0.5
    nc = len(cells)
    d = np.empty(nc)
    u = np.zeros((nc, 2))
    p = np.empty(nc)
    ne, nw = east & north, ~east & north
    sw, se = ~east & ~north, east & ~north
    d[ne], p[ne] = 1.5, 1.5
    d[nw], u[nw, 0], p[nw] = 0.5323, 1.206, 0.3
    d[sw], u[sw, 0], u[sw, 1], p[sw] = 0.138, 1.206, 1.206, 0.029
    d[se], u[se, 1], p[se] = 0.5323, 1.206, 0.3
    return dict(coords=coords, cells=cells, eos=(1.4, 1.0), d=d, v=u, p=p, cfl=0.5, final_time=1.0,
                max_steps=steps, flux=flux)


def kelvin_helmholtz_3d(n=128, steps=100, seed=MASTER_SEED):
    """KH shear layer on a jittered n^3 hex mesh of [0,1]^3."""
    coords, cells = box_hexes(n, n, n, jitter=0.1, seed=seed)
    x = coords[cells].mean(axis=1)   # layer membership only; the solver uses its own centroids
    inner = np.abs(x[:, 1] - 0.5) < 0.25
    nc = len(cells)
    d = np.where(inner, 2.0, 1.0)
    v = np.zeros((nc, 3))
    v[:, 0] = np.where(inner, 0.5, -0.5)
    xi = uniform(seed, 21, nc, -1.0, 1.0)
    v[:, 1] = 0.01 * np.sin(4 * math.pi * x[:, 0]) * xi
    p = np.full(nc, 2.5)
    return dict(coords=coords, cells=cells, eos=(1.4, 1.0), d=d, v=v, p=p, cfl=1.0 / 3.0, final_time=1.0,
                max_steps=steps, flux=0)


def maire_bubble_2d(n=512, steps=500, seed=MASTER_SEED):
    """Two-gas shock-bubble problem on a jittered n x n quad mesh of [0,1]^2 (Maire solver)."""
    coords, cells = box_quads(n, n, jitter=0.15, seed=seed, stream=31)
    x = quad_centroids(coords, cells)
    nc = len(cells)
    g = np.full(nc, 1.4)
    d = np.ones(nc)
    v = np.zeros((nc, 2))
    p = np.ones(nc)
    r = np.hypot(x[:, 0] - 0.5, x[:, 1] - 0.5)
    bub = r < 0.15
    g[bub], d[bub] = 1.67, 0.138
    shock = x[:, 0] < 0.1
    d[shock], v[shock, 0], p[shock] = 1.376, 0.394, 1.5698
    return dict(coords=coords, cells=cells, cv=1.0, gamma=g, d=d, v=v, p=p, cfl=(0.25, 0.1, 1.01),
                final_time=1.0, max_steps=steps, dt0=1.e-5)


def sedov_3d(coords, cells, centroids, volumes):
    """apps/maire_hydro/3d/inputs.cc:30-101 (the reference's sedov_20x20x20 smoke
    case): energy 0.106384 in the cells within sqrt(3) cbrt(V) / 2 (+ eps) of the
    origin, CFL (0.25, 0.1, 1.01), dt0 1e-5, 10 steps; x / y / z symmetry walls."""
    x = centroids
    delta_r = np.finfo(np.float64).eps + math.sqrt(3) * np.cbrt(volumes) / 2
    r = np.sqrt(x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1] + x[:, 2] * x[:, 2])
    p = np.where(r < delta_r, (1.4 - 1) * 1.0 * 0.106384 / volumes, 1.e-6)
    nc = len(p)
    return dict(coords=coords, cells=cells, cv=1.0, gamma=np.full(nc, 1.4), d=np.ones(nc),
                v=np.zeros((nc, 3)), p=p, cfl=(0.25, 0.1, 1.01), final_time=1.0, max_steps=10, dt0=1.e-5)


def maire_bubble_3d(n=64, steps=100, seed=MASTER_SEED):
    """The two-gas shock-bubble problem on a jittered n^3 hex mesh of [0,1]^3: a
    sphere r < 0.15 about the centre (gamma 1.67, rho 0.138), the shocked state
    at x < 0.1 driven to +x, symmetry walls."""
    coords, cells = box_hexes(n, n, n, jitter=0.15, seed=seed, stream=41)
    x = coords[cells].mean(axis=1)
    nc = len(cells)
    g = np.full(nc, 1.4)
    d = np.ones(nc)
    v = np.zeros((nc, 3))
    p = np.ones(nc)
    r = np.sqrt(((x - 0.5) ** 2).sum(axis=1))
    bub = r < 0.15
    g[bub], d[bub] = 1.67, 0.138
    shock = x[:, 0] < 0.1
    d[shock], v[shock, 0], p[shock] = 1.376, 0.394, 1.5698
    return dict(coords=coords, cells=cells, cv=1.0, gamma=g, d=d, v=v, p=p, cfl=(0.25, 0.1, 1.01),
                final_time=1.0, max_steps=steps, dt0=1.e-5)


def extrude_quads(coords, cells, nz, h, perm=(0, 1, 2)):
    """The hex mesh of a 2D quad mesh extruded through nz layers of height h
    (bottom quad, then the quad above it: Exodus hex8 order), its axes then
    permuted by `perm` (a cyclic permutation keeps the hexes positive)."""
    nv2 = len(coords)
    xs = np.zeros((nv2 * (nz + 1), 3))
    for j in range(nz + 1):
        xs[j * nv2:(j + 1) * nv2, :2] = coords
        xs[j * nv2:(j + 1) * nv2, 2] = j * h
    layers = [np.concatenate([cells + j * nv2, cells + (j + 1) * nv2], axis=1) for j in range(nz)]
    return np.ascontiguousarray(xs[:, list(perm)]), np.concatenate(layers).astype(np.int32)


_SQUARES = {}


def _unit_square(side):
    """side x side quad mesh of [0,1]^2 and its centroids, built once per side
    (a request stream reuses a few hundred sizes; the arrays are read-only)."""
    m = _SQUARES.get(side)
    if m is None:
        coords, cells = box_quads(side, side)
        x = quad_centroids(coords, cells)
        for a in (coords, cells, x):
            a.setflags(write=False)
        m = _SQUARES[side] = (coords, cells, x)
    return m


def riemann_2d(side, solver, seed=MASTER_SEED, index=0, steps=100):
    """Two-state Riemann problem on a side x side quad mesh of [0,1]^2.

    The split line passes through the box centre at angle theta ~ U[0, pi);
    each side draws (rho, p) ~ U[0.1, 1] and a velocity u ~ U[-1, 1] along the
    split normal; gamma is 1.4 or 1.67.  Draws come from substream
    1000 + index, so problem `index` of a stream is reproducible alone.
    solver: "euler" (MUSCL, HLLE, walls, CFL 1/2) or "maire" (CFL
    {0.25, 0.1, 1.01}, dt0 1e-5, symmetry boundaries)."""
    r = uniform(seed, 1000 + index, 8)
    theta = math.pi * r[0]
    n = np.array([math.cos(theta), math.sin(theta)])
    rho = 0.1 + 0.9 * r[1:3]
    pre = 0.1 + 0.9 * r[3:5]
    un = -1.0 + 2.0 * r[5:7]
    gamma = 1.4 if r[7] < 0.5 else 1.67
    coords, cells, x = _unit_square(side)
    right = ((x[:, 0] - 0.5) * n[0] + (x[:, 1] - 0.5) * n[1]) >= 0
    k = right.astype(np.int64)
    d = rho[k]
    p = pre[k]
    v = un[k][:, None] * n[None, :]
    nc = len(cells)
    if solver == "euler":
        return dict(coords=coords, cells=cells, eos=(gamma, 1.0), d=d, v=v, p=p, cfl=0.5, final_time=10.0,
                    max_steps=steps, flux=0)
    return dict(coords=coords, cells=cells, cv=1.0, gamma=np.full(nc, gamma), d=d, v=v, p=p,
                cfl=(0.25, 0.1, 1.01), final_time=10.0, max_steps=steps, dt0=1.e-5)


def stream_sides(count=4096, seed=MASTER_SEED, lo=32, hi=512):
    """Mesh sides of the mixed-size request stream: round(exp(U[ln lo, ln hi]))."""
    u = uniform(seed, 999, count)
    return np.rint(np.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * u)).astype(np.int64)


def stream_arrivals(count, rate, seed=MASTER_SEED):
    """Poisson arrival times (s) at `rate` requests/s: exponential gaps."""
    u = uniform(seed, 998, count)
    return np.cumsum(-np.log1p(-u) / rate)


# ------------------------------------------------------------ byte models
def euler_bytes_per_step(nc, nf, nv, dim):
    """Algorithmic (task-dataflow compulsory) HBM bytes of one Eulerian step."""
    if dim == 2:
        return 1036 * nc + 32 * nv + 264 * nf
    return 1628 * nc + 48 * nv + 336 * nf


def maire_bytes_per_step(nc, nv, ncorner):
    return 708 * nc + 312 * nv + 144 * ncorner
