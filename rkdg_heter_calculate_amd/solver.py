"""Python host mirror of the reference's solver surface, over librkdg's C ABI.

``Grid`` stands for CUnstructuredGrid's input side (reference mesh files, or the
split-quad generator), ``Solver`` for CCUDARkdgSolver + RkdgAdvance
(multi_gpuRkdg/src/cudarkdgsolver.cu, src/rkdgadvance.cpp): create from a mesh and the
main.conf parameters, set the initial condition, step, read the state back, write the
.plt output.  Every numerical call goes through librkdg.so on the GPU.
"""
import ctypes
import math

import numpy as np

from . import _lib as L


def _np_copy(ptr, n, dtype):
    if not ptr or n == 0:
        return None
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Grid:
    """Owning triangle mesh (rkdg_grid)."""

    def __init__(self, g):
        self._g = g
        lib = L.load()
        self._lib = lib
        E, NV = g.n_elements, g.n_vertices
        self.n_elements, self.n_vertices = E, NV
        self.vx = _np_copy(g.vx, NV, np.float64)
        self.vy = _np_copy(g.vy, NV, np.float64)
        self.tri = _np_copy(g.tri, 3 * E, np.int32).reshape(E, 3)
        self.neighbour = _np_copy(g.neighbour, 3 * E, np.int32).reshape(E, 3)
        self.shared_edge = _np_copy(g.shared_edge, 3 * E, np.int32).reshape(E, 3)
        self.boundary = _np_copy(g.boundary, 3 * E, np.int32).reshape(E, 3)
        part = _np_copy(g.partition, E, np.int32)
        self.partition = part

    @classmethod
    def splitquad(cls, nx, ny, x0=0.0, y0=0.0, lx=1.0, ly=1.0, periodic=True, diagonal=0, jitter=0.0,
                  seed=0, bc=L.BC_OUTFLOW):
        lib = L.load()
        g = L.Grid()
        rc = lib.rkdg_grid_splitquad(nx, ny, x0, y0, lx, ly, int(periodic), int(diagonal), float(jitter),
                                     ctypes.c_uint64(seed), bc, ctypes.byref(g))
        if rc != L.OK:
            raise L.RkdgError(rc, lib.rkdg_grid_last_error().decode())
        return cls(g)

    @classmethod
    def read(cls, meshconf, partition=None, box=15.0):
        """Reference mesh format (src/unstructuredgrid.cpp:14-60,193-305)."""
        lib = L.load()
        g = L.Grid()
        rc = lib.rkdg_grid_read(meshconf.encode(), partition.encode() if partition else None, box, ctypes.byref(g))
        if rc != L.OK:
            raise L.RkdgError(rc, lib.rkdg_grid_last_error().decode())
        return cls(g)

    def __del__(self):
        g = getattr(self, "_g", None)
        if g is not None and getattr(self, "_lib", None) is not None:
            self._lib.rkdg_grid_free(ctypes.byref(g))
            self._g = None

    def write(self, directory):
        """Reference mesh format (mesh.conf + *.dat) into an existing directory; the arrays
        written are this object's (edits to them are written)."""
        g = self._sync()
        rc = self._lib.rkdg_grid_write(str(directory).encode(), ctypes.byref(g))
        if rc != L.OK:
            raise L.RkdgError(rc, self._lib.rkdg_grid_last_error().decode())

    def _sync(self):
        """An rkdg_grid view of the numpy arrays held here."""
        g = L.Grid()
        self._wkeep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._wkeep.append(a)
            return a

        g.n_vertices, g.n_elements = self.n_vertices, self.n_elements
        g.vx = L.dptr(arr(self.vx, np.float64))
        g.vy = L.dptr(arr(self.vy, np.float64))
        g.tri = L.iptr(arr(self.tri, np.int32))
        g.neighbour = L.iptr(arr(self.neighbour, np.int32))
        g.shared_edge = L.iptr(arr(self.shared_edge, np.int32))
        g.boundary = L.iptr(arr(self.boundary, np.int32))
        g.partition = L.iptr(arr(self.partition, np.int32)) if self.partition is not None else None
        return g

    def element_coords(self):
        return self.vx[self.tri], self.vy[self.tri]

    def locality_order(self):
        order = np.empty(self.n_elements, np.int32)
        L.check(self._lib.rkdg_grid_locality_order(ctypes.byref(self._g), L.iptr(order)), None, "locality order")
        return order

    def write_plt(self, path, U, Np, flags=None, gamma=1.4):
        U = np.ascontiguousarray(U, dtype=np.float64)
        f = None if flags is None else np.ascontiguousarray(flags, dtype=np.int32)
        rc = self._lib.rkdg_write_plt(path.encode(), ctypes.byref(self._g), Np, L.dptr(U),
                                      L.iptr(f) if f is not None else None, gamma)
        L.check(rc, None, "write_plt")


def freestream_state(gamma=1.4, alpha_deg=5.0, mach=0.4, rhoref=1.0, pref=1.0):
    """Far-field conserved state of src/rkdgadvance.cpp:31-34 (alpha in degrees, as main.conf)."""
    alpha = alpha_deg * math.atan(1.0) * 4 / 180
    ut = math.sqrt(gamma * pref / rhoref) * mach
    return [rhoref, rhoref * ut * math.cos(alpha), rhoref * ut * math.sin(alpha),
            0.5 * rhoref * ut * ut + pref / (gamma - 1)]


def vortex_params(gamma=1.4, beta=5.0, uinf=1.0, vinf=1.0, xc=5.0, yc=5.0, lx=10.0, ly=10.0, t=0.0):
    return [gamma, beta, uinf, vinf, xc, yc, lx, ly, t]


def riemann4_params(gamma=1.4, xs=0.5, ys=0.5, states=None):
    if states is None:  # the four-quadrant configuration of the 2.1M-triangle workload
        states = [(1.5, 0.0, 0.0, 1.5), (0.5323, 1.206, 0.0, 0.3), (0.138, 1.206, 1.206, 0.029),
                  (0.5323, 0.0, 1.206, 0.3)]
    p = [gamma, xs, ys]
    for s in states:
        p += list(s)
    return p


_FLUX = {"lf": L.FLUX_LF, "roe": L.FLUX_ROE}
_DT = {"reference": L.DT_REFERENCE, "max": L.DT_MAX, "fixed": L.DT_FIXED}
_PREC = {"fp64": L.PREC_FP64, "fp32": L.PREC_FP32_MFMA, "bf16": L.PREC_BF16_MFMA}


def make_params(degree=2, flux="lf", limiter=False, tvb_M=50.0, gamma=1.4, cfl=0.18, dt_policy="reference",
                reduction_threads=512, dt_fixed=0.0, precision="fp64", freestream=None, block=0, device=0):
    lib = L.load()
    p = L.Params()
    lib.rkdg_default_params(ctypes.byref(p))
    p.degree = degree
    p.flux = _FLUX[flux]
    p.limiter = L.LIMITER_TVB if limiter else L.LIMITER_NONE
    p.tvb_M = tvb_M
    p.gamma = gamma
    p.cfl = cfl
    p.dt_policy = _DT[dt_policy]
    p.reduction_threads = reduction_threads
    p.dt_fixed = dt_fixed
    p.precision = _PREC[precision]
    p.block = block
    p.device = device
    fs = freestream if freestream is not None else freestream_state(gamma)
    for i in range(4):
        p.freestream[i] = fs[i]
    return p


def mesh_view(grid, order=None, dt_id=None):
    """Non-owning rkdg_mesh of a Grid; keeps the arrays alive on the returned object."""
    m = L.Mesh()
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    vx, vy = arr(grid.vx, np.float64), arr(grid.vy, np.float64)
    tri, nb = arr(grid.tri, np.int32), arr(grid.neighbour, np.int32)
    se, bd = arr(grid.shared_edge, np.int32), arr(grid.boundary, np.int32)
    m.n_vertices, m.n_elements = grid.n_vertices, grid.n_elements
    m.vx, m.vy, m.tri, m.neighbour = L.dptr(vx), L.dptr(vy), L.iptr(tri), L.iptr(nb)
    m.shared_edge, m.boundary = L.iptr(se), L.iptr(bd)
    if order is not None:
        m.order = L.iptr(arr(order, np.int32))
    if dt_id is not None:
        m.dt_id = L.iptr(arr(dt_id, np.int32))
    m._keep = keep
    return m


class Solver:
    """One mesh on one GPU (CCUDARkdgSolver + RkdgAdvance)."""

    def __init__(self, grid, order=None, dt_id=None, **kw):
        self.lib = L.load()
        self.grid = grid
        self.params = make_params(**kw)
        self.degree = self.params.degree
        self.Np = (self.degree + 1) * (self.degree + 2) // 2
        self.E = grid.n_elements
        self._mesh = mesh_view(grid, order, dt_id)
        h = ctypes.c_void_p()
        rc = self.lib.rkdg_create(ctypes.byref(self._mesh), ctypes.byref(self.params), ctypes.byref(h))
        if rc != L.OK:
            raise L.RkdgError(rc, self.lib.rkdg_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.rkdg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        return L.check(rc, self.h, what)

    def state_shape(self):
        return (4, self.Np, self.E)

    def project_ic(self, kind, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._chk(self.lib.rkdg_project_ic(self.h, kind, L.dptr(p), len(p)), "project_ic")

    def set_state(self, U, basis=L.BASIS_ELEMENT):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.state_shape())
        self._chk(self.lib.rkdg_set_state(self.h, basis, L.dptr(U)), "set_state")

    def get_state(self, basis=L.BASIS_ELEMENT):
        U = np.empty(self.state_shape(), np.float64)
        self._chk(self.lib.rkdg_get_state(self.h, basis, L.dptr(U)), "get_state")
        return U

    def compute_dt(self):
        dt = ctypes.c_double()
        self._chk(self.lib.rkdg_compute_dt(self.h, ctypes.byref(dt)), "compute_dt")
        return dt.value

    def stage(self, s, dt):
        self._chk(self.lib.rkdg_stage(self.h, s, dt), "stage")

    def step(self):
        dt = ctypes.c_double()
        self._chk(self.lib.rkdg_step(self.h, ctypes.byref(dt)), "step")
        return dt.value

    def advance(self, nsteps, t_end=0.0):
        n = ctypes.c_int32()
        t = ctypes.c_double()
        self._chk(self.lib.rkdg_advance(self.h, nsteps, t_end, ctypes.byref(n), ctypes.byref(t)), "advance")
        return n.value, t.value

    def residual(self):
        r = np.zeros(2)
        self._chk(self.lib.rkdg_residual(self.h, L.dptr(r)), "residual")
        return r

    def time(self):
        t = ctypes.c_double()
        s = ctypes.c_int64()
        self._chk(self.lib.rkdg_time(self.h, ctypes.byref(t), ctypes.byref(s)), "time")
        return t.value, s.value

    def flags(self):
        f = np.empty(self.E, np.int32)
        self._chk(self.lib.rkdg_get_flags(self.h, L.iptr(f)), "flags")
        return f

    def sync(self):
        self._chk(self.lib.rkdg_sync(self.h), "sync")

    def time_stage(self, s=1, reps=20):
        ms = ctypes.c_double()
        self._chk(self.lib.rkdg_time_stage(self.h, s, reps, ctypes.byref(ms)), "time_stage")
        return ms.value

    def time_advance(self, nsteps):
        out = np.zeros(4)
        self._chk(self.lib.rkdg_time_advance(self.h, nsteps, L.dptr(out)), "time_advance")
        return {"total_ms": out[0], "stage_ms": out[1], "launches": int(out[2]), "limiter_ms": out[3]}

    def time_steps(self, nsteps):
        """Event-timed ms of `nsteps` steps of advance() (hipGraph-replayed, device dt)."""
        ms = ctypes.c_double()
        self._chk(self.lib.rkdg_time_steps(self.h, nsteps, ctypes.byref(ms)), "time_steps")
        return ms.value

    def error_vs_exact(self, kind, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        out = np.zeros(8)
        self._chk(self.lib.rkdg_error_vs_exact(self.h, kind, L.dptr(p), len(p), L.dptr(out)), "error_vs_exact")
        return {"l2": out[:4].tolist(), "linf": out[4:].tolist()}

    def info(self):
        out = np.zeros(5, np.int64)
        self._chk(self.lib.rkdg_info(self.h, out.ctypes.data_as(L.P_int64)), "info")
        return {"Np": int(out[0]), "Nqv": int(out[1]), "Nqe": int(out[2]), "E": int(out[3]), "bytes": int(out[4])}


class Batch:
    """Many independent meshes advanced together, one launch per stage (rkdg_batch_*).

    params: one keyword dict per instance (make_params' keywords); instances share every
    field but the mesh, the initial state and cfl (rkdg_batch_create rejects the rest)."""

    def __init__(self, grids, params):
        self.lib = L.load()
        n = len(grids)
        if n != len(params):
            raise ValueError("one parameter set per mesh")
        self.grids = grids
        self._views = [mesh_view(g) for g in grids]
        meshes = (L.Mesh * n)(*self._views)
        prm = (L.Params * n)(*[make_params(**p) for p in params])
        self.degree = prm[0].degree
        self.Np = (self.degree + 1) * (self.degree + 2) // 2
        h = ctypes.c_void_p()
        rc = self.lib.rkdg_batch_create(n, meshes, prm, ctypes.byref(h))
        if rc != L.OK:
            raise L.RkdgError(rc, self.lib.rkdg_batch_last_error(None).decode())
        self.h = h
        self.n = n

    def close(self):
        if getattr(self, "h", None):
            self.lib.rkdg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        return L.bcheck(rc, self.h, what)

    def project_ic(self, i, kind, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._chk(self.lib.rkdg_batch_project_ic(self.h, i, kind, L.dptr(p), len(p)), "batch_project_ic")

    def set_state(self, i, U, basis=L.BASIS_ELEMENT):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(4, self.Np, self.grids[i].n_elements)
        self._chk(self.lib.rkdg_batch_set_state(self.h, i, basis, L.dptr(U)), "batch_set_state")

    def get_state(self, i, basis=L.BASIS_ELEMENT):
        U = np.empty((4, self.Np, self.grids[i].n_elements), np.float64)
        self._chk(self.lib.rkdg_batch_get_state(self.h, i, basis, L.dptr(U)), "batch_get_state")
        return U

    def advance(self, nsteps, use_graph=True):
        self._chk(self.lib.rkdg_batch_advance(self.h, nsteps, int(use_graph)), "batch_advance")

    def sync(self):
        self._chk(self.lib.rkdg_batch_sync(self.h), "batch_sync")

    def time(self, i):
        t = ctypes.c_double()
        self._chk(self.lib.rkdg_batch_time(self.h, i, ctypes.byref(t)), "batch_time")
        return t.value

    def autotune(self):
        best = ctypes.c_int32()
        self._chk(self.lib.rkdg_batch_autotune(self.h, ctypes.byref(best)), "batch_autotune")
        return best.value

    def set_tiles(self, staged, block, per_cu):
        self._chk(self.lib.rkdg_batch_set_tiles(self.h, int(staged), int(block), int(per_cu)), "batch_set_tiles")

    def tiles(self):
        out = np.zeros(3, np.int32)
        self._chk(self.lib.rkdg_batch_get_tiles(self.h, L.iptr(out)), "batch_get_tiles")
        return {"staged": int(out[0]), "block": int(out[1]), "per_cu": int(out[2])}

    def info(self):
        out = np.zeros(4, np.int64)
        self._chk(self.lib.rkdg_batch_info(self.h, out.ctypes.data_as(L.P_int64)), "batch_info")
        return {"instances": int(out[0]), "elements": int(out[1]), "block": int(out[2]), "bytes": int(out[3])}
