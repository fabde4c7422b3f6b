"""flecsale_amd -- MI355X implementation of FleCSALE's hydro time-step paths.

Python binding (ctypes) over the C ABI of libflecsale_amd.so
(include/flecsale_amd.h).  The classes mirror the reference's task interface:

  EulerContext   apps/hydro/tasks.h       initial_conditions, evaluate_time_step,
                                          evaluate_gradients, evaluate_fluxes,
                                          apply_update (+ the driver loop: advance)
  MaireContext   apps/maire_hydro/tasks.h initial_conditions, evaluate_nodal_state,
                                          evaluate_residual, evaluate_time_step,
                                          apply_update, save/restore (+ advance)

Errors raise FsaleError with the library's status (the reference throws
THROW_RUNTIME_ERROR at the same points).  The library must be built
(python -m flecsale_amd.build); there is no CPU fallback.
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FSALE_LIB: another build of the library to load (e.g. one compiled with other options)
LIB_PATH = os.environ.get("FSALE_LIB") or os.path.join(HERE, "libflecsale_amd.so")

FLUX_HLLE, FLUX_RUSANOV, FLUX_HLLC = 0, 1, 2
BOUNDARY_WALL, BOUNDARY_TRANSMISSIVE = 0x00, 0x10   # OR-ed into flux_kind
SOLVER_EULER, SOLVER_MAIRE = 0, 1
HALO_STATE, HALO_GRADIENTS = 0, 1   # fsale_halo_exchange field sets
# Maire boundary condition kinds (fsale_maire_set_boundary)
BC_SYMMETRY, BC_PRESSURE, BC_VELOCITY, BC_NONE = 0, 1, 2, 3

(F_DENSITY, F_VELOCITY, F_INTERNAL_ENERGY, F_PRESSURE, F_TEMPERATURE, F_SOUND_SPEED,
 F_GRAD_DENSITY, F_GRAD_VELOCITY, F_GRAD_ENERGY, F_GRAD_PRESSURE, F_GRAD_SOUND_SPEED,
 F_FLUX, F_CELL_CENTROID, F_CELL_VOLUME, F_CELL_MIN_LENGTH, F_FACE_CENTROID, F_FACE_NORMAL,
 F_FACE_AREA, F_COORDS, F_MASS, F_VOLUME, F_NODE_VELOCITY, F_CORNER_NORMAL, F_CORNER_FORCE,
 F_RESIDUAL, F_GAMMA) = range(26)

K_EUL_DT, K_EUL_GRAD, K_EUL_FLUX, K_EUL_UPDATE = 0, 1, 2, 3
K_MAIRE_NODAL, K_MAIRE_RESIDUAL, K_MAIRE_DT, K_MAIRE_MOVE, K_MAIRE_UPDATE, K_MAIRE_COPY = 4, 5, 6, 7, 8, 9
KERNEL_NAMES = ["eul_dt", "eul_grad", "eul_flux", "eul_update", "maire_nodal", "maire_residual",
                "maire_dt", "maire_move", "maire_update", "maire_copy", "eul_fluxupd", "maire_nodal_bnd"]
K_EUL_FLUXUPD = 10
K_MAIRE_NODAL_BND = 11

ERR_ARG = 1
ERR_MESH = 2
ERR_HIP = 3
ERR_OOM = 4
ERR_NEG_STATE = 5
ERR_DT = 6
ERR_STATE = 7
ERR_NO_DEVICE = 8

# every symbol include/flecsale_amd.h declares
EXPORTS = [
    "fsale_ctx_create", "fsale_ctx_destroy", "fsale_sync", "fsale_mesh_counts", "fsale_get_topology",
    "fsale_mesh_connectivity", "fsale_get_field", "fsale_set_field", "fsale_get_time", "fsale_set_time",
    "fsale_status_string", "fsale_last_message", "fsale_qr_solve", "fsale_error_cell", "fsale_set_profiling",
    "fsale_get_profile", "fsale_reset_profile",
    "fsale_eul_initial_conditions", "fsale_eul_evaluate_time_step", "fsale_eul_evaluate_gradients",
    "fsale_eul_evaluate_fluxes", "fsale_eul_apply_update", "fsale_eul_advance", "fsale_eul_time_step_noalloc",
    "fsale_eul_set_owned_cells", "fsale_halo_create", "fsale_halo_exchange", "fsale_halo_destroy",
    "fsale_maire_initial_conditions", "fsale_maire_evaluate_nodal_state", "fsale_maire_evaluate_residual",
    "fsale_maire_set_boundary", "fsale_batch_set_maire_boundary",
    "fsale_maire_evaluate_time_step", "fsale_maire_apply_update", "fsale_maire_save", "fsale_maire_restore",
    "fsale_maire_advance",
    "fsale_batch_create", "fsale_batch_add_euler", "fsale_batch_add_maire", "fsale_batch_remove",
    "fsale_batch_step", "fsale_batch_status", "fsale_batch_get_field", "fsale_batch_get_fields",
    "fsale_batch_capacity",
    "fsale_launch_shape", "fsale_get_launch_shape", "fsale_lds_staging", "fsale_autotune",
    "fsale_server_create", "fsale_server_destroy", "fsale_server_clock", "fsale_server_submit",
    "fsale_server_poll", "fsale_server_times", "fsale_server_result", "fsale_server_stats",
    "fsale_exodus_info", "fsale_exodus_read", "fsale_dump", "fsale_output", "fsale_exodus_write",
]


class FsaleError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{message} (status {status})")
        self.status = status


class MeshDesc(ct.Structure):
    _fields_ = [("dim", ct.c_int32), ("num_cells", ct.c_int32), ("num_vertices", ct.c_int32),
                ("verts_per_cell", ct.c_int32), ("coords", ct.c_void_p), ("cells", ct.c_void_p)]


class Eos(ct.Structure):
    _fields_ = [("gamma", ct.c_double), ("cv", ct.c_double)]


class EulerParams(ct.Structure):
    _fields_ = [("cfl", ct.c_double), ("final_time", ct.c_double), ("flux_kind", ct.c_int32),
                ("first_order", ct.c_int32), ("stop_at_final", ct.c_int32), ("max_steps", ct.c_int64)]


class MaireParams(ct.Structure):
    _fields_ = [("cfl_accoustic", ct.c_double), ("cfl_volume", ct.c_double), ("cfl_growth", ct.c_double),
                ("final_time", ct.c_double), ("initial_time_step", ct.c_double), ("max_steps", ct.c_int64)]


class ProblemStatus(ct.Structure):
    _fields_ = [("problem", ct.c_int32), ("error", ct.c_int32), ("steps", ct.c_int64), ("time", ct.c_double),
                ("time_step", ct.c_double), ("done", ct.c_int32)]


class ServerConfig(ct.Structure):
    _fields_ = [("mode", ct.c_int32), ("dim", ct.c_int32), ("verts_per_cell", ct.c_int32),
                ("max_problems", ct.c_int32),
                ("euler_cells", ct.c_int64), ("euler_faces", ct.c_int64), ("euler_vertices", ct.c_int64),
                ("euler_neighbours", ct.c_int64),
                ("maire_cells", ct.c_int64), ("maire_faces", ct.c_int64), ("maire_vertices", ct.c_int64),
                ("maire_neighbours", ct.c_int64),
                ("chunk_steps", ct.c_int32), ("lookahead", ct.c_int32), ("eager", ct.c_int32),
                ("reserved", ct.c_int32)]


class Request(ct.Structure):
    _fields_ = [("solver", ct.c_int32), ("mesh", MeshDesc), ("eos", Eos), ("euler", EulerParams),
                ("cv", ct.c_double), ("gamma", ct.c_void_p), ("maire", MaireParams),
                ("d", ct.c_void_p), ("v", ct.c_void_p), ("p", ct.c_void_p)]


_lib = None


def lib():
    """Load the HIP library; raises if it was not built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run python -m flecsale_amd.build")
        L = ct.CDLL(LIB_PATH)
        vp, i32, dp = ct.c_void_p, ct.c_int32, ct.c_double
        L.fsale_status_string.restype = ct.c_char_p
        L.fsale_status_string.argtypes = [ct.c_int]
        L.fsale_last_message.restype = ct.c_char_p
        L.fsale_qr_solve.argtypes = [i32, vp, vp, vp]
        L.fsale_exodus_info.argtypes = [ct.c_char_p, vp, vp, vp, vp]
        L.fsale_exodus_read.argtypes = [ct.c_char_p, vp, vp]
        L.fsale_dump.argtypes = [vp, ct.c_char_p]
        L.fsale_output.argtypes = [vp, ct.c_char_p, ct.c_char_p, ct.c_int64, ct.c_int32]
        L.fsale_exodus_write.argtypes = [ct.c_char_p, ct.c_char_p, ct.c_int32, ct.c_int32, vp, ct.c_int32,
                                         ct.c_int32, vp, ct.c_double, ct.c_int32, vp, vp]
        L.fsale_ctx_create.argtypes = [ct.POINTER(MeshDesc), ct.c_int, ct.POINTER(vp)]
        L.fsale_ctx_destroy.argtypes = [vp]
        L.fsale_sync.argtypes = [vp]
        L.fsale_mesh_counts.argtypes = [vp, vp]
        L.fsale_get_topology.argtypes = [vp, vp, vp, vp, vp]
        L.fsale_mesh_connectivity.argtypes = [ct.POINTER(MeshDesc), vp, vp, vp, vp, vp]
        L.fsale_get_field.argtypes = [vp, ct.c_int, vp]
        L.fsale_set_field.argtypes = [vp, ct.c_int, vp]
        L.fsale_get_time.argtypes = [vp, ct.POINTER(dp), ct.POINTER(dp), ct.POINTER(ct.c_int64)]
        L.fsale_set_time.argtypes = [vp, dp, dp]
        L.fsale_error_cell.argtypes = [vp, ct.POINTER(ct.c_int64)]
        L.fsale_set_profiling.argtypes = [vp, ct.c_int]
        L.fsale_get_profile.argtypes = [vp, ct.c_int, ct.POINTER(dp), ct.POINTER(ct.c_int64)]
        L.fsale_reset_profile.argtypes = [vp]
        L.fsale_eul_initial_conditions.argtypes = [vp, Eos, vp, vp, vp]
        L.fsale_eul_evaluate_time_step.argtypes = [vp, dp, dp, ct.POINTER(dp)]
        L.fsale_eul_evaluate_gradients.argtypes = [vp, ct.c_int]
        L.fsale_eul_evaluate_fluxes.argtypes = [vp, ct.c_int]
        L.fsale_eul_apply_update.argtypes = [vp, Eos, dp]
        L.fsale_eul_advance.argtypes = [vp, Eos, dp, dp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int]
        L.fsale_eul_set_owned_cells.argtypes = [vp, i32]
        L.fsale_halo_create.argtypes = [vp, vp, i32, vp, vp, ct.POINTER(vp)]
        L.fsale_halo_exchange.argtypes = [vp, ct.c_int]
        L.fsale_halo_destroy.argtypes = [vp]
        L.fsale_eul_time_step_noalloc.argtypes = [i32, i32, i32] + [vp] * 6 + [dp, dp, dp, vp, vp, vp]
        L.fsale_maire_initial_conditions.argtypes = [vp, dp, vp, vp, vp, vp]
        L.fsale_maire_evaluate_nodal_state.argtypes = [vp]
        L.fsale_maire_set_boundary.argtypes = [vp, ct.c_int, ct.c_int, vp]
        L.fsale_batch_set_maire_boundary.argtypes = [vp, i32, ct.c_int, ct.c_int, vp]
        L.fsale_maire_evaluate_residual.argtypes = [vp]
        L.fsale_maire_evaluate_time_step.argtypes = [vp, dp, dp, dp, dp, ct.POINTER(dp)]
        L.fsale_maire_apply_update.argtypes = [vp, dp]
        L.fsale_maire_save.argtypes = [vp]
        L.fsale_maire_restore.argtypes = [vp]
        L.fsale_maire_advance.argtypes = [vp, dp, dp, dp, dp, ct.c_int, ct.c_int]
        i64 = ct.c_int64
        L.fsale_batch_create.argtypes = [ct.c_int, ct.c_int, ct.c_int, i32, i64, i64, i64, i64, ct.POINTER(vp)]
        L.fsale_batch_add_euler.argtypes = [vp, ct.POINTER(MeshDesc), Eos, ct.POINTER(EulerParams), vp, vp, vp,
                                            ct.POINTER(i32)]
        L.fsale_batch_add_maire.argtypes = [vp, ct.POINTER(MeshDesc), dp, vp, ct.POINTER(MaireParams), vp, vp, vp,
                                            ct.POINTER(i32)]
        L.fsale_batch_remove.argtypes = [vp, i32]
        L.fsale_batch_step.argtypes = [vp, ct.c_int, ct.c_int]
        L.fsale_batch_status.argtypes = [vp, i32, ct.POINTER(ProblemStatus), ct.POINTER(i32)]
        L.fsale_batch_get_field.argtypes = [vp, i32, ct.c_int, vp]
        L.fsale_batch_get_fields.argtypes = [vp, i32, i32, vp, vp]
        L.fsale_batch_capacity.argtypes = [vp, ct.POINTER(i64), ct.POINTER(i64), ct.POINTER(i64),
                                           ct.POINTER(i64), ct.POINTER(i32)]
        L.fsale_launch_shape.argtypes = [vp, i32, i32]
        L.fsale_get_launch_shape.argtypes = [vp, vp]
        L.fsale_lds_staging.argtypes = [vp, ct.POINTER(i32)]
        L.fsale_autotune.argtypes = [vp, i32, ct.POINTER(dp)]
        L.fsale_server_create.argtypes = [ct.POINTER(ServerConfig), dp, ct.POINTER(vp)]
        L.fsale_server_destroy.argtypes = [vp]
        L.fsale_server_clock.argtypes = []
        L.fsale_server_clock.restype = dp
        L.fsale_server_submit.argtypes = [vp, ct.POINTER(Request), dp, ct.POINTER(i64)]
        L.fsale_server_poll.argtypes = [vp, ct.POINTER(i32), ct.POINTER(i32)]
        L.fsale_server_times.argtypes = [vp, i64, ct.POINTER(dp), ct.POINTER(dp), ct.POINTER(dp)]
        L.fsale_server_stats.argtypes = [vp, vp]
        L.fsale_server_result.argtypes = [vp, i64, ct.c_int, ct.POINTER(ProblemStatus), vp, vp, vp, vp, vp]
        for name in EXPORTS:
            f = getattr(L, name)
            if f.restype is ct.c_int or f.restype is None:
                f.restype = ct.c_int
        L.fsale_server_clock.restype = dp
        L.fsale_status_string.restype = ct.c_char_p
        L.fsale_last_message.restype = ct.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        L = lib()
        raise FsaleError(rc, (L.fsale_last_message() or L.fsale_status_string(rc)).decode())


def _desc(coords, cells):
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    d = MeshDesc(coords.shape[1], cells.shape[0], coords.shape[0], cells.shape[1],
                 coords.ctypes.data, cells.ctypes.data)
    return d, (coords, cells)


def qr_solve(A, b):
    """flecsale::linalg::qr solve of A x = b (n <= 4) with the nodal kernel's arithmetic, on the host."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(b))
    _check(lib().fsale_qr_solve(len(b), A.ctypes.data, b.ctypes.data, x.ctypes.data))
    return x


def connectivity(coords, cells):
    """Host-only connectivity (fsale_mesh_connectivity): face_cells, cell_faces, nbr_off, nbr."""
    d, keep = _desc(coords, cells)
    counts = np.zeros(5, np.int32)
    _check(lib().fsale_mesh_connectivity(ct.byref(d), counts.ctypes.data, None, None, None, None))
    nc, nv, nf, nn, dim = [int(x) for x in counts]
    fpc = cells.shape[1] if dim == 2 else 6
    fc = np.zeros((nf, 2), np.int32)
    cf = np.zeros((nc, fpc), np.int32)
    off = np.zeros(nc + 1, np.int32)
    nbr = np.zeros(max(nn, 1), np.int32)
    _check(lib().fsale_mesh_connectivity(ct.byref(d), counts.ctypes.data, fc.ctypes.data, cf.ctypes.data,
                                         off.ctypes.data, nbr.ctypes.data))
    return dict(face_cells=fc, cell_faces=cf, nbr_off=off, nbr=nbr[:nn])


def write_exodus(path, coords, cells, time=0.0, variables=None, title=""):
    """One time step of an Exodus-II file (fsale_exodus_write): the mesh
    (coords [nv, dim], cells [nc, vpc] 0-based) and the cell variables of
    `variables` ({name: values [nc]}, in order)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    items = list((variables or {}).items())
    vals = [np.ascontiguousarray(np.broadcast_to(v, (len(cells),)), dtype=np.float64) for _, v in items]
    names = (ct.c_char_p * max(1, len(items)))(*[k.encode() for k, _ in items])
    ptrs = (ct.c_void_p * max(1, len(items)))(*[v.ctypes.data for v in vals])
    _check(lib().fsale_exodus_write(os.fsencode(path), title.encode(), coords.shape[1], len(coords),
                                    coords.ctypes.data, len(cells), cells.shape[1], cells.ctypes.data,
                                    float(time), len(items), ct.cast(names, ct.c_void_p),
                                    ct.cast(ptrs, ct.c_void_p)))


def read_exodus(path):
    """Exodus-II mesh file (fsale_exodus_read): (coords [nv, dim] float64,
    cells [nc, vpc] int32, 0-based, file order)."""
    b = os.fsencode(path)
    dim, nc, nv, vpc = ct.c_int32(), ct.c_int32(), ct.c_int32(), ct.c_int32()
    _check(lib().fsale_exodus_info(b, ct.byref(dim), ct.byref(nc), ct.byref(nv), ct.byref(vpc)))
    coords = np.zeros((nv.value, dim.value))
    cells = np.zeros((nc.value, vpc.value), np.int32)
    _check(lib().fsale_exodus_read(b, coords.ctypes.data, cells.ctypes.data))
    return coords, cells


class _Context:
    solver = None

    def __init__(self, coords, cells):
        d, keep = _desc(coords, cells)
        h = ct.c_void_p()
        _check(lib().fsale_ctx_create(ct.byref(d), self.solver, ct.byref(h)))
        self.h = h
        self.dim = d.dim
        counts = np.zeros(5, np.int32)
        _check(lib().fsale_mesh_counts(self.h, counts.ctypes.data))
        self.nc, self.nv, self.nf, self.nnbr, _ = [int(x) for x in counts]
        self.vpc = d.verts_per_cell

    def close(self):
        if getattr(self, "h", None):
            lib().fsale_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def dump(self, path):
        """The dump task's regression text file (fsale_dump)."""
        _check(lib().fsale_dump(self.h, os.fsencode(path)))

    def output(self, prefix, iteration, postfix="exo", rank=0):
        """The output task's Exodus-II file (fsale_output); returns its path."""
        _check(lib().fsale_output(self.h, os.fsencode(prefix), os.fsencode(postfix), iteration, rank))
        sep = "." if self.solver == SOLVER_EULER else "_"
        return f"{prefix}_rank{rank:06d}{sep}{iteration:06d}.{postfix}"

    _shapes = None

    def _shape(self, field):
        N, nc, nf, nv = self.dim, self.nc, self.nf, self.nv
        table = {
            F_DENSITY: (nc,), F_VELOCITY: (nc, N), F_INTERNAL_ENERGY: (nc,), F_PRESSURE: (nc,),
            F_TEMPERATURE: (nc,), F_SOUND_SPEED: (nc,), F_GRAD_DENSITY: (nc, N),
            F_GRAD_VELOCITY: (nc, N, N), F_GRAD_ENERGY: (nc, N), F_GRAD_PRESSURE: (nc, N),
            F_GRAD_SOUND_SPEED: (nc, N), F_FLUX: (nf, N + 2), F_CELL_CENTROID: (nc, N),
            F_CELL_VOLUME: (nc,), F_CELL_MIN_LENGTH: (nc,), F_FACE_CENTROID: (nf, N),
            F_FACE_NORMAL: (nf, N), F_FACE_AREA: (nf,), F_COORDS: (nv, N), F_MASS: (nc,),
            F_VOLUME: (nc,), F_NODE_VELOCITY: (nv, N), F_CORNER_NORMAL: (nc * self.vpc, N),
            F_CORNER_FORCE: (nc * self.vpc, N), F_RESIDUAL: (nc, N + 2), F_GAMMA: (nc,),
        }
        return table[field]

    def get(self, field):
        out = np.zeros(self._shape(field))
        _check(lib().fsale_get_field(self.h, field, out.ctypes.data))
        return out

    def set(self, field, values):
        a = np.ascontiguousarray(np.broadcast_to(values, self._shape(field)), dtype=np.float64)
        _check(lib().fsale_set_field(self.h, field, a.ctypes.data))

    def sync(self):
        _check(lib().fsale_sync(self.h))

    def time(self):
        t, dt, n = ct.c_double(), ct.c_double(), ct.c_int64()
        _check(lib().fsale_get_time(self.h, ct.byref(t), ct.byref(dt), ct.byref(n)))
        return t.value, dt.value, n.value

    def set_time(self, t, dt):
        _check(lib().fsale_set_time(self.h, t, dt))

    def error_cell(self):
        c = ct.c_int64()
        _check(lib().fsale_error_cell(self.h, ct.byref(c)))
        return c.value

    def set_profiling(self, enable):
        _check(lib().fsale_set_profiling(self.h, int(enable)))

    def profile(self, kernel):
        """(total milliseconds, launches) of one kernel id since the last reset."""
        t, n = ct.c_double(), ct.c_int64()
        _check(lib().fsale_get_profile(self.h, int(kernel), ct.byref(t), ct.byref(n)))
        return t.value, n.value

    def reset_profile(self):
        _check(lib().fsale_reset_profile(self.h))

    def launch_shape(self, blocks_per_cu, kernel=-1):
        """Blocks per CU of one kernel id (-1: all); 0 = one block per tile."""
        _check(lib().fsale_launch_shape(self.h, int(kernel), int(blocks_per_cu)))

    def get_launch_shape(self):
        out = np.zeros(len(KERNEL_NAMES), np.int32)
        _check(lib().fsale_get_launch_shape(self.h, out.ctypes.data))
        return {KERNEL_NAMES[k]: int(out[k]) for k in range(len(KERNEL_NAMES))}

    def lds_staging(self):
        """Eulerian driver kernels' LDS staging: 2 every tile, 1 some tiles, 0 none
        (3D: plus 4 when the fused step runs over bricks)."""
        on = ct.c_int32()
        _check(lib().fsale_lds_staging(self.h, ct.byref(on)))
        return int(on.value)

    def autotune(self, nsteps=3):
        """Pick each kernel's launch shape by timing real steps; returns the tuned step's kernel ms."""
        ms = ct.c_double()
        _check(lib().fsale_autotune(self.h, int(nsteps), ct.byref(ms)))
        return ms.value

    def topology(self):
        fpc = self.vpc if self.dim == 2 else 6
        fc = np.zeros((self.nf, 2), np.int32)
        cf = np.zeros((self.nc, fpc), np.int32)
        off = np.zeros(self.nc + 1, np.int32)
        nbr = np.zeros(max(self.nnbr, 1), np.int32)
        _check(lib().fsale_get_topology(self.h, fc.ctypes.data, cf.ctypes.data, off.ctypes.data,
                                        nbr.ctypes.data))
        return dict(face_cells=fc, cell_faces=cf, nbr_off=off, nbr=nbr[:self.nnbr])

    def state(self):
        return dict(d=self.get(F_DENSITY), v=self.get(F_VELOCITY), e=self.get(F_INTERNAL_ENERGY),
                    p=self.get(F_PRESSURE), T=self.get(F_TEMPERATURE), a=self.get(F_SOUND_SPEED))


class EulerContext(_Context):
    """Device-resident apps/hydro state over one mesh."""
    solver = SOLVER_EULER

    def initial_conditions(self, eos, d, v, p):
        self.eos = Eos(*eos)
        d = np.ascontiguousarray(np.broadcast_to(d, (self.nc,)), dtype=np.float64)
        v = np.ascontiguousarray(np.broadcast_to(v, (self.nc, self.dim)), dtype=np.float64)
        p = np.ascontiguousarray(np.broadcast_to(p, (self.nc,)), dtype=np.float64)
        _check(lib().fsale_eul_initial_conditions(self.h, self.eos, d.ctypes.data, v.ctypes.data,
                                                  p.ctypes.data))

    def evaluate_time_step(self, cfl, final_time, sync=True):
        dt = ct.c_double()
        _check(lib().fsale_eul_evaluate_time_step(self.h, cfl, final_time, ct.byref(dt) if sync else None))
        return dt.value if sync else None

    def evaluate_gradients(self, first_order=False):
        _check(lib().fsale_eul_evaluate_gradients(self.h, int(first_order)))

    def evaluate_fluxes(self, flux_kind=FLUX_HLLE):
        _check(lib().fsale_eul_evaluate_fluxes(self.h, int(flux_kind)))

    def apply_update(self, factor):
        _check(lib().fsale_eul_apply_update(self.h, self.eos, factor))

    def advance(self, cfl, final_time, nsteps, flux_kind=FLUX_HLLE, first_order=False,
                stop_at_final=False, use_graph=True):
        _check(lib().fsale_eul_advance(self.h, self.eos, cfl, final_time, int(nsteps), int(flux_kind),
                                       int(first_order), int(stop_at_final), int(use_graph)))


class MaireContext(_Context):
    """Device-resident apps/maire_hydro state over one 2D quad or 3D hex mesh."""
    solver = SOLVER_MAIRE

    def initial_conditions(self, cv, gamma, d, v, p):
        g = np.ascontiguousarray(np.broadcast_to(gamma, (self.nc,)), dtype=np.float64)
        d = np.ascontiguousarray(np.broadcast_to(d, (self.nc,)), dtype=np.float64)
        v = np.ascontiguousarray(np.broadcast_to(v, (self.nc, self.dim)), dtype=np.float64)
        p = np.ascontiguousarray(np.broadcast_to(p, (self.nc,)), dtype=np.float64)
        _check(lib().fsale_maire_initial_conditions(self.h, cv, g.ctypes.data, d.ctypes.data,
                                                    v.ctypes.data, p.ctypes.data))

    def set_boundary(self, tag, kind, value=None):
        """Boundary condition of the faces of symmetry tag `tag` (0 x, 1 y, 2 z):
        BC_SYMMETRY, BC_PRESSURE (value = p), BC_VELOCITY (value = u), BC_NONE."""
        v = np.zeros(3)
        if value is not None:
            a = np.atleast_1d(np.asarray(value, dtype=np.float64))
            v[:len(a)] = a
        _check(lib().fsale_maire_set_boundary(self.h, int(tag), int(kind), v.ctypes.data))

    def evaluate_nodal_state(self):
        _check(lib().fsale_maire_evaluate_nodal_state(self.h))

    def evaluate_residual(self):
        _check(lib().fsale_maire_evaluate_residual(self.h))

    def evaluate_time_step(self, cfl, final_time, sync=True):
        dt = ct.c_double()
        _check(lib().fsale_maire_evaluate_time_step(self.h, cfl[0], cfl[1], cfl[2], final_time,
                                                    ct.byref(dt) if sync else None))
        return dt.value if sync else None

    def apply_update(self, factor):
        _check(lib().fsale_maire_apply_update(self.h, factor))

    def save(self):
        _check(lib().fsale_maire_save(self.h))

    def restore(self):
        _check(lib().fsale_maire_restore(self.h))

    def advance(self, cfl, final_time, nsteps, use_graph=True):
        _check(lib().fsale_maire_advance(self.h, cfl[0], cfl[1], cfl[2], final_time, int(nsteps),
                                         int(use_graph)))

    def state(self):
        s = super().state()
        s.update(M=self.get(F_MASS), V=self.get(F_VOLUME))
        return s


class Batch:
    """A ragged batch of independent problems of one solver (fsale_batch_*).

    Each resident problem runs the reference's own step loop with its own
    mesh, EOS, CFL and clock; problems are admitted (add_euler / add_maire)
    and retired (remove) between steps.
    """

    def __init__(self, solver, dim, verts_per_cell, max_problems, cells, faces, vertices, neighbours):
        h = ct.c_void_p()
        _check(lib().fsale_batch_create(int(solver), int(dim), int(verts_per_cell), int(max_problems),
                                        int(cells), int(faces), int(vertices), int(neighbours), ct.byref(h)))
        self.h = h
        self.solver = solver
        self.dim = dim
        self.vpc = verts_per_cell
        self.meshes = {}

    def close(self):
        if getattr(self, "h", None):
            lib().fsale_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_euler(self, coords, cells, eos, d, v, p, cfl, final_time, max_steps, flux_kind=FLUX_HLLE,
                  first_order=False, stop_at_final=False):
        desc, keep = _desc(coords, cells)
        nc = desc.num_cells
        d = np.ascontiguousarray(np.broadcast_to(d, (nc,)), dtype=np.float64)
        v = np.ascontiguousarray(np.broadcast_to(v, (nc, self.dim)), dtype=np.float64)
        p = np.ascontiguousarray(np.broadcast_to(p, (nc,)), dtype=np.float64)
        prm = EulerParams(cfl, final_time, int(flux_kind), int(first_order), int(stop_at_final), int(max_steps))
        out = ct.c_int32()
        _check(lib().fsale_batch_add_euler(self.h, ct.byref(desc), Eos(*eos), ct.byref(prm), d.ctypes.data,
                                           v.ctypes.data, p.ctypes.data, ct.byref(out)))
        self.meshes[out.value] = keep
        return out.value

    def add_maire(self, coords, cells, cv, gamma, d, v, p, cfl, final_time, max_steps, initial_time_step):
        desc, keep = _desc(coords, cells)
        nc = desc.num_cells
        g = np.ascontiguousarray(np.broadcast_to(gamma, (nc,)), dtype=np.float64)
        d = np.ascontiguousarray(np.broadcast_to(d, (nc,)), dtype=np.float64)
        v = np.ascontiguousarray(np.broadcast_to(v, (nc, self.dim)), dtype=np.float64)
        p = np.ascontiguousarray(np.broadcast_to(p, (nc,)), dtype=np.float64)
        prm = MaireParams(cfl[0], cfl[1], cfl[2], final_time, initial_time_step, int(max_steps))
        out = ct.c_int32()
        _check(lib().fsale_batch_add_maire(self.h, ct.byref(desc), cv, g.ctypes.data, ct.byref(prm),
                                           d.ctypes.data, v.ctypes.data, p.ctypes.data, ct.byref(out)))
        self.meshes[out.value] = keep
        return out.value

    def remove(self, problem):
        _check(lib().fsale_batch_remove(self.h, int(problem)))
        self.meshes.pop(problem, None)

    def set_maire_boundary(self, problem, tag, kind, value=None):
        """MaireContext.set_boundary for one resident problem."""
        v = np.zeros(3)
        if value is not None:
            a = np.atleast_1d(np.asarray(value, dtype=np.float64))
            v[:len(a)] = a
        _check(lib().fsale_batch_set_maire_boundary(self.h, int(problem), int(tag), int(kind), v.ctypes.data))

    def step(self, nsteps, use_graph=True):
        _check(lib().fsale_batch_step(self.h, int(nsteps), int(use_graph)))

    def status(self):
        """{problem: ProblemStatus} of every resident problem."""
        n = max(len(self.meshes), 1)
        arr = (ProblemStatus * n)()
        cnt = ct.c_int32()
        _check(lib().fsale_batch_status(self.h, n, arr, ct.byref(cnt)))
        return {arr[i].problem: arr[i] for i in range(cnt.value)}

    def get(self, problem, field):
        coords, cells = self.meshes[problem]
        nc, nv, N = cells.shape[0], coords.shape[0], self.dim
        shapes = {F_DENSITY: (nc,), F_VELOCITY: (nc, N), F_INTERNAL_ENERGY: (nc,), F_PRESSURE: (nc,),
                  F_TEMPERATURE: (nc,), F_SOUND_SPEED: (nc,), F_CELL_CENTROID: (nc, N), F_CELL_VOLUME: (nc,),
                  F_COORDS: (nv, N), F_MASS: (nc,), F_VOLUME: (nc,), F_NODE_VELOCITY: (nv, N),
                  F_GAMMA: (nc,)}
        out = np.zeros(shapes[field])
        _check(lib().fsale_batch_get_field(self.h, int(problem), int(field), out.ctypes.data))
        return out

    def state(self, problem):
        s = dict(d=self.get(problem, F_DENSITY), v=self.get(problem, F_VELOCITY),
                 e=self.get(problem, F_INTERNAL_ENERGY), p=self.get(problem, F_PRESSURE),
                 T=self.get(problem, F_TEMPERATURE), a=self.get(problem, F_SOUND_SPEED))
        if self.solver == SOLVER_MAIRE:
            s["x"] = self.get(problem, F_COORDS)
        return s

    def capacity(self):
        c, f, v, n = ct.c_int64(), ct.c_int64(), ct.c_int64(), ct.c_int64()
        p = ct.c_int32()
        _check(lib().fsale_batch_capacity(self.h, ct.byref(c), ct.byref(f), ct.byref(v), ct.byref(n),
                                          ct.byref(p)))
        return dict(cells=c.value, faces=f.value, vertices=v.value, neighbours=n.value, problems=p.value)


    def launch_shape(self, blocks_per_cu, kernel=-1):
        """Blocks per CU of one kernel id (-1: all); 0 = one block per tile."""
        _check(lib().fsale_launch_shape(self.h, int(kernel), int(blocks_per_cu)))

    def get_launch_shape(self):
        out = np.zeros(len(KERNEL_NAMES), np.int32)
        _check(lib().fsale_get_launch_shape(self.h, out.ctypes.data))
        return {KERNEL_NAMES[k]: int(out[k]) for k in range(len(KERNEL_NAMES))}

    def lds_staging(self):
        """Eulerian driver kernels' LDS staging: 2 every tile, 1 some tiles, 0 none
        (3D: plus 4 when the fused step runs over bricks)."""
        on = ct.c_int32()
        _check(lib().fsale_lds_staging(self.h, ct.byref(on)))
        return int(on.value)

    def autotune(self, nsteps=3):
        """Pick each kernel's launch shape by timing real steps; returns the tuned step's kernel ms."""
        ms = ct.c_double()
        _check(lib().fsale_autotune(self.h, int(nsteps), ct.byref(ms)))
        return ms.value


class Server:
    """Serving loop (fsale_server_*): requests are admitted into resident
    batches while capacity allows and advance together (mode 0), or run one
    at a time in plain contexts (mode 1)."""

    def __init__(self, mode=0, dim=2, verts_per_cell=4, max_problems=1024, euler=(0, 0, 0, 0), maire=(0, 0, 0, 0),
                 chunk_steps=10, lookahead=64, frac_hbm=0.5, graphs=True):
        cfg = ServerConfig(int(mode), int(dim), int(verts_per_cell), int(max_problems), *[int(x) for x in euler],
                           *[int(x) for x in maire], int(chunk_steps), int(lookahead), int(not graphs), 0)
        h = ct.c_void_p()
        _check(lib().fsale_server_create(ct.byref(cfg), float(frac_hbm), ct.byref(h)))
        self.h = h
        self.dim = dim
        self.shapes = {}

    def close(self):
        if getattr(self, "h", None):
            lib().fsale_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def clock():
        return lib().fsale_server_clock()

    def submit(self, prob, solver, arrival=None):
        """Queue one problem dict (inputs.riemann_2d and the like); returns its id."""
        desc, keep = _desc(prob["coords"], prob["cells"])
        nc = desc.num_cells
        d = np.ascontiguousarray(np.broadcast_to(prob["d"], (nc,)), dtype=np.float64)
        v = np.ascontiguousarray(np.broadcast_to(prob["v"], (nc, self.dim)), dtype=np.float64)
        p = np.ascontiguousarray(np.broadcast_to(prob["p"], (nc,)), dtype=np.float64)
        q = Request()
        q.solver = int(solver)
        q.mesh = desc
        q.d, q.v, q.p = d.ctypes.data, v.ctypes.data, p.ctypes.data
        if solver == SOLVER_EULER:
            q.eos = Eos(*prob["eos"])
            q.euler = EulerParams(prob["cfl"], prob["final_time"], int(prob["flux"]),
                                  int(prob.get("first_order", False)), int(prob.get("stop_at_final", False)),
                                  int(prob["max_steps"]))
        else:
            g = np.ascontiguousarray(np.broadcast_to(prob["gamma"], (nc,)), dtype=np.float64)
            q.cv = prob["cv"]
            q.gamma = g.ctypes.data
            c = prob["cfl"]
            q.maire = MaireParams(c[0], c[1], c[2], prob["final_time"], prob["dt0"], int(prob["max_steps"]))
        out = ct.c_int64()
        t = self.clock() if arrival is None else float(arrival)
        _check(lib().fsale_server_submit(self.h, ct.byref(q), t, ct.byref(out)))
        self.shapes[out.value] = (nc, desc.num_vertices, solver)
        return out.value

    def poll(self):
        """One scheduling round: (requests completed in it, requests queued or running)."""
        c, p = ct.c_int32(), ct.c_int32()
        _check(lib().fsale_server_poll(self.h, ct.byref(c), ct.byref(p)))
        return c.value, p.value

    def stats(self):
        """Host seconds per phase of the batched loop, rounds, admissions and the
        GPU seconds of the Eulerian and Lagrangian batches' rounds."""
        out = np.zeros(8)
        _check(lib().fsale_server_stats(self.h, out.ctypes.data))
        return dict(admit_s=out[0], wait_s=out[1], retire_s=out[2], launch_s=out[3], rounds=int(out[4]),
                    admissions=int(out[5]), gpu_euler_s=out[6], gpu_maire_s=out[7])

    def times(self, rid):
        a, b, c = ct.c_double(), ct.c_double(), ct.c_double()
        _check(lib().fsale_server_times(self.h, int(rid), ct.byref(a), ct.byref(b), ct.byref(c)))
        return a.value, b.value, c.value

    def result(self, rid, keep=False, fields=True):
        """(ProblemStatus, {d, v, e, p[, x]}) of a finished request."""
        nc, nv, solver = self.shapes[rid]
        st = ProblemStatus()
        out = {}
        ptrs = [None] * 5
        if fields:
            out = dict(d=np.zeros(nc), v=np.zeros((nc, self.dim)), e=np.zeros(nc), p=np.zeros(nc))
            if solver == SOLVER_MAIRE:
                out["x"] = np.zeros((nv, self.dim))
            ptrs = [out[k].ctypes.data if k in out else None for k in ("d", "v", "e", "p", "x")]
        _check(lib().fsale_server_result(self.h, int(rid), int(keep), ct.byref(st), *ptrs))
        if not keep:
            self.shapes.pop(rid, None)
        return st, out
