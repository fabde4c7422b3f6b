"""CPU oracle for FleCSALE's hydro paths -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker (or the timed CPU baseline).
The product (flecsale_amd) never loads it.

ctypes wrapper over oracle/liboracle.so (built from flecsale_oracle.c by
oracle/Makefile), which restates apps/hydro/tasks.h, apps/maire_hydro/tasks.h,
flecsale/eqns/*.h, flecsale/eos/ideal_gas.h and flecsale/linalg/qr.h loop by
loop.  Arrays are numpy, C-contiguous; vectors are interleaved [n, dim].
"""
import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
OM_ERR_CORNERS = 5   # flecsale_oracle.c: a vertex with more than 16 corners

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = ct.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)


def load(path):
    """Use the restatement built at `path` (e.g. liboracle_native.so, built for
    the host it runs on) instead of liboracle.so; before any other call."""
    global LIB
    if _lib is not None and path != LIB:
        raise RuntimeError("the oracle library is already loaded from " + LIB)
    LIB = path


def set_threads(n):
    """OpenMP threads of the restatement's parallel loops."""
    lib().om_set_threads(int(n))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ct.CDLL(LIB)
        L.om_set_threads.argtypes = [ct.c_int]
        L.om_mesh_create.restype = _vp
        L.om_mesh_create.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, _dp, _ip]
        L.om_mesh_destroy.argtypes = [_vp]
        L.om_mesh_counts.argtypes = [_vp, _ip]
        L.om_mesh_topology.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        L.om_mesh_geometry.argtypes = [_vp] * 8
        L.om_mesh_set_coords.argtypes = [_vp, _dp]
        L.oe_state_from_pressure.argtypes = [ct.c_int, ct.c_double, ct.c_double, _dp, _dp, _dp, _dp, _dp]
        L.oe_time_step.restype = ct.c_int
        L.oe_time_step.argtypes = [_vp, _dp, _dp, ct.c_double, ct.c_double, ct.c_double,
                                   ct.POINTER(ct.c_double)]
        L.oe_gradients.argtypes = [_vp] + [_dp] * 10 + [ct.c_int]
        L.oe_fluxes.argtypes = [_vp] + [_dp] * 11 + [ct.c_int, _dp]
        L.oe_apply_update.restype = ct.c_int
        L.oe_apply_update.argtypes = [_vp, ct.c_double, ct.c_double, _dp, ct.c_double, ct.c_double] + \
            [_dp] * 6 + [ct.POINTER(ct.c_int)]
        L.oe_run.restype = ct.c_int
        L.oe_run.argtypes = [_vp, ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_int,
                             ct.c_int, ct.c_int, ct.c_int] + [_dp] * 6 + \
            [ct.POINTER(ct.c_double), ct.POINTER(ct.c_int)]
        L.om_qr_solve.argtypes = [ct.c_int, _dp, _dp]
        L.om_solve3.argtypes = [_dp, _dp, _dp]
        L.om_maire_set_bc.argtypes = [ct.POINTER(ct.c_int), _dp]
        L.om_set_variant.argtypes = [ct.c_int]
        L.oe_face_flux.argtypes = [ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp]
        L.om_mesh_set_orders.argtypes = [_vp, _vp, _vp]
        L.om_eos_eval.argtypes = [ct.c_double, ct.c_double, ct.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.om_maire_init.argtypes = [_vp, _dp, ct.c_double, _dp, _dp] + [_dp] * 6
        L.om_maire_nodal.restype = ct.c_int
        L.om_maire_nodal.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.om_maire_residual.argtypes = [_vp, _dp, _dp, _dp, _dp]
        L.om_maire_time_step.restype = ct.c_double
        L.om_maire_time_step.argtypes = [_vp, ct.c_double, ct.c_double, ct.c_double, ct.c_double, _dp, _dp]
        L.om_maire_apply_update.restype = ct.c_int
        L.om_maire_apply_update.argtypes = [_vp, ct.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                            ct.POINTER(ct.c_int)]
        L.om_maire_from_energy.argtypes = [ct.c_int, _dp, ct.c_double, _dp, _dp, _dp, _dp, _dp]
        L.om_maire_run.restype = ct.c_int
        L.om_maire_run.argtypes = [_vp, _dp, ct.c_double, ct.c_double, ct.c_double, ct.c_double,
                                   ct.c_double, ct.c_int, ct.POINTER(ct.c_double)] + [_dp] * 8 + \
            [ct.POINTER(ct.c_double), ct.POINTER(ct.c_int)]
        _lib = L
    return _lib


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


class Mesh:
    """Connectivity + geometry of a quad (2D) or hex (3D) mesh (burton replacement)."""

    def __init__(self, coords, cells):
        coords = _c(coords)
        cells = _c(cells, np.int32)
        self.dim = coords.shape[1]
        self.nv = coords.shape[0]
        self.nc, self.vpc = cells.shape
        self.cells = cells
        self.h = lib().om_mesh_create(self.dim, self.nc, self.nv, self.vpc, coords.ravel(), cells.ravel())
        if not self.h:
            raise ValueError("bad mesh")
        cnt = np.zeros(8, np.int32)
        lib().om_mesh_counts(self.h, cnt)
        self.nf, self.nnbr, _, _, _, _, self.fpc, self.vpf = [int(x) for x in cnt]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.om_mesh_destroy(self.h)
            self.h = None

    def topology(self):
        cf = np.zeros((self.nc, self.fpc), np.int32)
        fc = np.zeros((self.nf, 2), np.int32)
        fv = np.zeros((self.nf, self.vpf), np.int32)
        off = np.zeros(self.nc + 1, np.int32)
        nbr = np.zeros(self.nnbr, np.int32)
        lib().om_mesh_topology(self.h, _ptr(cf), _ptr(fc), _ptr(fv), _ptr(off), _ptr(nbr))
        return dict(cell_faces=cf, face_cells=fc, face_verts=fv, nbr_off=off, nbr=nbr)

    def geometry(self):
        d = self.dim
        g = dict(cell_centroid=np.zeros((self.nc, d)), cell_volume=np.zeros(self.nc),
                 cell_min_length=np.zeros(self.nc), face_centroid=np.zeros((self.nf, d)),
                 face_normal=np.zeros((self.nf, d)), face_area=np.zeros(self.nf),
                 coords=np.zeros((self.nv, d)))
        lib().om_mesh_geometry(self.h, *[_ptr(g[k]) for k in
                                         ["cell_centroid", "cell_volume", "cell_min_length",
                                          "face_centroid", "face_normal", "face_area", "coords"]])
        return g

    def set_coords(self, coords):
        lib().om_mesh_set_coords(self.h, _c(coords).ravel())


# ---------------------------------------------------------------- Eulerian

class EulerState:
    """Cell state of apps/hydro (d, v, e, p, T, a), FP64."""

    def __init__(self, nc, dim):
        self.d = np.zeros(nc)
        self.v = np.zeros((nc, dim))
        self.e = np.zeros(nc)
        self.p = np.zeros(nc)
        self.T = np.zeros(nc)
        self.a = np.zeros(nc)

    def copy(self):
        s = EulerState(0, 1)
        for k in "d v e p T a".split():
            setattr(s, k, getattr(self, k).copy())
        return s

    def arrays(self):
        return [self.d, self.v, self.e, self.p, self.T, self.a]


def euler_initial_state(mesh, gamma, cv, d, v, p):
    """initial_conditions (apps/hydro/tasks.h:46-68)."""
    s = EulerState(mesh.nc, mesh.dim)
    s.d[:] = d
    s.v[:] = v
    s.p[:] = p
    lib().oe_state_from_pressure(mesh.nc, gamma, cv, s.d, s.p, s.e, s.T, s.a)
    return s


def euler_time_step(mesh, s, cfl, final_time, soln_time):
    dt = ct.c_double(0)
    rc = lib().oe_time_step(mesh.h, s.v.reshape(-1), s.a, cfl, final_time, soln_time, ct.byref(dt))
    if rc:
        raise RuntimeError("infinite delta t")
    return dt.value


def euler_gradients(mesh, s, first_order=False):
    N, nc = mesh.dim, mesh.nc
    g = dict(d=np.zeros((nc, N)), v=np.zeros((nc, N, N)), e=np.zeros((nc, N)),
             p=np.zeros((nc, N)), a=np.zeros((nc, N)))
    lib().oe_gradients(mesh.h, s.d, s.v.reshape(-1), s.e, s.p, s.a, g["d"].reshape(-1),
                       g["v"].reshape(-1), g["e"].reshape(-1), g["p"].reshape(-1),
                       g["a"].reshape(-1), int(first_order))
    return g


def euler_fluxes(mesh, s, g, flux_kind=0):
    F = np.zeros((mesh.nf, mesh.dim + 2))
    lib().oe_fluxes(mesh.h, s.d, s.v.reshape(-1), s.e, s.p, s.T, s.a, g["d"].reshape(-1),
                    g["v"].reshape(-1), g["e"].reshape(-1), g["p"].reshape(-1),
                    g["a"].reshape(-1), int(flux_kind), F.reshape(-1))
    return F


def euler_apply_update(mesh, s, gamma, cv, F, dt, factor):
    bad = ct.c_int(-1)
    rc = lib().oe_apply_update(mesh.h, gamma, cv, _c(F).reshape(-1), dt, factor, s.d, s.v.reshape(-1),
                               s.e, s.p, s.T, s.a, ct.byref(bad))
    return rc, bad.value


def euler_run(mesh, s, gamma, cv, cfl, final_time, max_steps, flux_kind=0, first_order=False,
              stop_at_final=False, soln_time=0.0):
    """apps/hydro/driver.h:317-413; updates s in place, returns (rc, time, steps)."""
    t = ct.c_double(soln_time)
    n = ct.c_int(0)
    rc = lib().oe_run(mesh.h, gamma, cv, cfl, final_time, max_steps, int(flux_kind), int(first_order),
                      int(stop_at_final), s.d, s.v.reshape(-1), s.e, s.p, s.T, s.a, ct.byref(t),
                      ct.byref(n))
    return rc, t.value, n.value


def face_flux(kind, wl, wr, n):
    """The oracle's flux functions on face states: wl, wr [nf, N+4] as
    (d, v[N], p, e, a); n [nf, N] unit normals; kind 0 HLLE, 1 Rusanov,
    2 HLLC, 3 physical flux of wl, 4 wall flux of wl.  Returns [nf, N+2]."""
    wl = _c(np.atleast_2d(wl)); wr = _c(np.atleast_2d(wr)); n = _c(np.atleast_2d(n))
    nf, N = n.shape
    F = np.zeros((nf, N + 2))
    lib().oe_face_flux(int(kind), N, nf, wl.reshape(-1), wr.reshape(-1), n.reshape(-1), F.reshape(-1))
    return F


def set_variant(v):
    """Plant a known mistake in the oracle (tests only); 0 restores the reference arithmetic."""
    lib().om_set_variant(int(v))


def eos_eval(gamma, cv, d, p):
    """ideal_gas_t: e(d,p), p(d,e), a(e), T(e) (flecsale/eos/ideal_gas.h:178-226)."""
    d = _c(d); p = _c(p); n = len(d)
    e, pp, a, T = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    lib().om_eos_eval(gamma, cv, n, d, p, e, pp, a, T)
    return e, pp, a, T


# ---------------------------------------------------------------- QR

def qr_solve(A, b):
    """flecsale::linalg::qr on a square system (row-major copy); returns x."""
    A = _c(A).copy()
    b = _c(b).copy()
    lib().om_qr_solve(A.shape[0], A.reshape(-1), b)
    return b


def solve3(M, r):
    """The 3x3 nodal solve (adjugate over determinant)."""
    u = np.zeros(3)
    lib().om_solve3(_c(M).reshape(-1).copy(), _c(r).copy(), u)
    return u


# ---------------------------------------------------------------- Maire
BC_SYMMETRY, BC_PRESSURE, BC_VELOCITY, BC_NONE = 0, 1, 2, 3


def maire_set_boundary(kinds=(0, 0, 0), values=None):
    """Boundary condition per symmetry tag (x, y, z faces): kinds BC_*, values
    [3][3] (pressure in [t][0], velocity in [t][:dim]); global to the oracle."""
    k = (ct.c_int * 3)(*[int(x) for x in kinds])
    v = np.zeros((3, 3)) if values is None else np.ascontiguousarray(values, dtype=np.float64).reshape(3, 3)
    lib().om_maire_set_bc(k, v.reshape(-1).copy())


class MaireState:
    """Cell state of apps/maire_hydro (V, M, v, p, d, e, T, a) + per-cell gamma."""

    def __init__(self, nc, dim=2):
        self.V = np.zeros(nc)
        self.M = np.zeros(nc)
        self.v = np.zeros((nc, dim))
        self.p = np.zeros(nc)
        self.d = np.zeros(nc)
        self.e = np.zeros(nc)
        self.T = np.zeros(nc)
        self.a = np.zeros(nc)
        self.g = np.zeros(nc)


def maire_initial_state(mesh, gamma, cv, den, v, p):
    """initial_conditions (apps/maire_hydro/tasks.h:71-101)."""
    s = MaireState(mesh.nc, mesh.dim)
    s.g[:] = gamma
    s.v[:] = v
    s.p[:] = p
    den = _c(np.broadcast_to(den, (mesh.nc,)))
    lib().om_maire_init(mesh.h, s.g, cv, den, s.p, s.V, s.M, s.d, s.e, s.T, s.a)
    return s


def maire_run(mesh, s, cv, cfl, final_time, max_steps, dt0, soln_time=0.0):
    """apps/maire_hydro/driver.h:330-538; cfl = (accoustic, volume, growth)."""
    t = ct.c_double(soln_time)
    dt = ct.c_double(dt0)
    n = ct.c_int(0)
    rc = lib().om_maire_run(mesh.h, s.g, cv, cfl[0], cfl[1], cfl[2], final_time, max_steps, ct.byref(dt),
                            s.V, s.M, s.v.reshape(-1), s.p, s.d, s.e, s.T, s.a, ct.byref(t), ct.byref(n))
    if rc == OM_ERR_CORNERS:
        raise ValueError("oracle maire_run: a vertex has more than 16 corners")
    return rc, t.value, n.value, dt.value


def maire_nodal(mesh, s):
    """evaluate_nodal_state (apps/maire_hydro/tasks.h:221-453): nodal
    velocities [nv, dim] and corner normals / forces [nc*vpc, dim]."""
    nv, ncn, N = mesh.nv, mesh.nc * mesh.vpc, mesh.dim
    un = np.zeros((nv, N)); npc = np.zeros((ncn, N)); Fpc = np.zeros((ncn, N))
    rc = lib().om_maire_nodal(mesh.h, s.v.reshape(-1), s.p, s.d, s.a, un.reshape(-1), npc.reshape(-1),
                              Fpc.reshape(-1))
    if rc == OM_ERR_CORNERS:
        raise ValueError("oracle maire_nodal: a vertex has more than 16 corners")
    return un, npc, Fpc


def maire_time_step(mesh, cfl, prev_dt, a, dudt):
    """evaluate_time_step (apps/maire_hydro/tasks.h:142-190): the minimum of
    the acoustic, volume and growth limits; cfl = (accoustic, volume, growth)."""
    return lib().om_maire_time_step(mesh.h, cfl[0], cfl[1], cfl[2], prev_dt, _c(a), _c(dudt).reshape(-1))


def maire_residual(mesh, un, npc, Fpc):
    dudt = np.zeros((mesh.nc, mesh.dim + 2))
    lib().om_maire_residual(mesh.h, _c(un).reshape(-1), _c(npc).reshape(-1), _c(Fpc).reshape(-1),
                            dudt.reshape(-1))
    return dudt
