"""fsale_eul_time_step_noalloc (no device allocation: its partials live in a
file-scope device array) against fsale_eul_evaluate_time_step, the context
entry point it varies, on the same meshes and states: two calls in a row, the
second on a smaller mesh, then a 3D pair.  The time step is a maximum and a
few exactly rounded operations, so the bar is bit equality."""
import ctypes as ct

import numpy as np
import pytest

import flecsale_amd as fa
from flecsale_amd import inputs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Dev:
    """Device buffers through the HIP runtime the library itself uses."""

    def __init__(self):
        fa.lib()
        self.hip = ct.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        self.hip.hipFree.argtypes = [ct.c_void_p]
        self.hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ct.c_void_p()
        assert self.hip.hipMalloc(ct.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)


def _entry():
    return fa.lib().fsale_eul_time_step_noalloc


def _problem(dim, n, seed):
    if dim == 2:
        coords, cells = inputs.box_quads(n, (3 * n) // 4, jitter=0.2, seed=seed)
    else:
        coords, cells = inputs.box_hexes(n, n - 2, n - 4, jitter=0.1, seed=seed)
    nc = len(cells)
    rng = np.random.default_rng(seed)
    return coords, cells, rng.uniform(0.2, 2.0, nc), rng.uniform(-1, 1, (nc, dim)), rng.uniform(0.1, 3.0, nc)


def _compare(dev, f, dim, n, seed, cfl=0.45, final_time=10.0):
    coords, cells, d, v, p = _problem(dim, n, seed)
    with fa.EulerContext(coords, cells) as ctx:
        ctx.initial_conditions((1.4, 1.0), d, v, p)
        want = ctx.evaluate_time_step(cfl, final_time)
        t0 = ctx.time()[0]
        topo = ctx.topology()
        args = [ctx.get(fa.F_FACE_NORMAL), ctx.get(fa.F_FACE_AREA), ctx.get(fa.F_CELL_VOLUME),
                ctx.get(fa.F_VELOCITY), ctx.get(fa.F_SOUND_SPEED)]
    nc, fpc = topo["cell_faces"].shape
    ptrs = [dev.put(topo["cell_faces"].astype(np.int32))] + [dev.put(a) for a in args]
    dt_dev = dev.put(np.array([-1.0]))
    st_dev = dev.put(np.array([-7], np.int32))
    rc = f(dim, nc, fpc, *ptrs, cfl, final_time, t0, dt_dev, st_dev, None)
    assert rc == 0, fa.lib().fsale_last_message()
    got = dev.get(dt_dev, np.zeros(1))[0]
    st = dev.get(st_dev, np.zeros(1, np.int32))[0]
    assert st == 0
    assert got == want, (dim, n, got, want)
    return nc


def test_noalloc_time_step_equals_context_entry_point():
    dev = _Dev()
    f = _entry()
    try:
        big = _compare(dev, f, 2, 400, 3)
        small = _compare(dev, f, 2, 48, 4)
        assert small < big
        big3 = _compare(dev, f, 3, 40, 5)
        small3 = _compare(dev, f, 3, 12, 6)
        assert small3 < big3
    finally:
        dev.close()


def test_noalloc_time_step_errors():
    """A gas with no sound speed and at rest has no time step: the device
    status word says so (the reference throws "infinite delta t"); sizes
    beyond the scratch array are refused before any launch."""
    dev = _Dev()
    f = _entry()
    try:
        coords, cells, d, v, p = _problem(2, 16, 7)
        with fa.EulerContext(coords, cells) as ctx:
            topo = ctx.topology()
            geo = [ctx.get(fa.F_FACE_NORMAL), ctx.get(fa.F_FACE_AREA), ctx.get(fa.F_CELL_VOLUME)]
        nc, fpc = topo["cell_faces"].shape
        ptrs = [dev.put(topo["cell_faces"].astype(np.int32))] + [dev.put(a) for a in geo]
        ptrs += [dev.put(np.zeros((nc, 2))), dev.put(np.zeros(nc))]
        dt_dev = dev.put(np.array([-1.0]))
        st_dev = dev.put(np.array([-7], np.int32))
        assert f(2, nc, fpc, *ptrs, 0.5, 1.0, 0.0, dt_dev, st_dev, None) == 0
        assert dev.get(st_dev, np.zeros(1, np.int32))[0] == fa.ERR_DT
        assert dev.get(dt_dev, np.zeros(1))[0] == -1.0
        assert f(2, (1 << 24) + 1, fpc, *ptrs, 0.5, 1.0, 0.0, dt_dev, st_dev, None) == fa.ERR_ARG
        assert b"scratch limit" in fa.lib().fsale_last_message()
    finally:
        dev.close()
