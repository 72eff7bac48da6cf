"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py,
always as the checker or the reported CPU baseline, never as the measured product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the oracle (the sanitizer build, liboracle_san.so)
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libref.so")
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build with __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        lib.orc_info_size.restype = ctypes.c_int
        lib.orc_triangle_info.argtypes = [ctypes.c_int, D, D, D]
        lib.orc_basis_value.restype = ctypes.c_double
        lib.orc_basis_value.argtypes = [ctypes.c_int, D, D, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        lib.orc_create.restype = ctypes.c_void_p
        lib.orc_create.argtypes = [ctypes.c_int, D, D, I, I, I, I, I, D]
        lib.orc_destroy.argtypes = [ctypes.c_void_p]
        for n in ("orc_set_state", "orc_get_state", "orc_get_mass", "orc_residual", "orc_residual_change"):
            getattr(lib, n).argtypes = [ctypes.c_void_p, D]
        lib.orc_get_flags.argtypes = [ctypes.c_void_p, I]
        lib.orc_compute_dt.restype = ctypes.c_double
        lib.orc_compute_dt.argtypes = [ctypes.c_void_p]
        lib.orc_stage.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        lib.orc_advance.restype = ctypes.c_int
        lib.orc_advance.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, D, D]
        lib.orc_project.argtypes = [ctypes.c_void_p, D, D, ctypes.c_int, D]
        lib.orc_vortex_exact.argtypes = [ctypes.c_int, D, D, D, D]
        lib.orc_roe_flux.argtypes = [ctypes.c_double, D, D, ctypes.c_double, ctypes.c_double, D]
        lib.orc_set_threads.argtypes = [ctypes.c_int]
        lib.orc_get_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(D)


def _i(a):
    return a.ctypes.data_as(I)


def sizes(P):
    Np = (P + 1) * (P + 2) // 2
    return Np, (7 if P <= 2 else 12), (3 if P <= 2 else 4)


def triangle_info(P, x, y):
    """Per-element precompute, dict of arrays (layout of orc_triangle_info)."""
    lib = load()
    Np, Nqv, Nqe = sizes(P)
    out = np.zeros(lib.orc_info_size(P))
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    lib.orc_triangle_info(P, _d(x), _d(y), _d(out))
    return split_info(out, Np, Nqv, Nqe)


def split_info(out, Np, Nqv, Nqe):
    o = 0
    r = {}
    for name, n, shp in (("area", 1, None), ("perimeter", 1, None), ("normal", 6, (6,)), ("bary", 2, (2,)),
                         ("coeff", Np * Np, (Np, Np)), ("mass", Np, (Np,)), ("vol_w", Nqv, (Nqv,)),
                         ("edge_w", 3 * Nqe, (3 * Nqe,)), ("vol_bf", Np * Nqv, (Np, Nqv)),
                         ("vol_bdf", 2 * Np * Nqv, (2 * Np, Nqv)), ("edge_bf", Np * 3 * Nqe, (Np, 3 * Nqe))):
        v = out[o:o + n]
        r[name] = v[0] if shp is None else v.reshape(shp)
        o += n
    return r


class Oracle:
    """FP64 CPU restatement of the reference solver on one mesh."""

    def __init__(self, ex, ey, neighbour, shared_edge, boundary, P=2, flux=0, limiter=0, dt_policy=0,
                 reduction_threads=512, gamma=1.4, cfl=0.18, tvb_M=50.0, dt_fixed=0.0, freestream=(1, 0, 0, 2.5),
                 dt_id=None):
        self.lib = load()
        E = ex.shape[0]
        self.E, self.P = E, P
        self.Np = sizes(P)[0]
        self.ex = np.ascontiguousarray(ex, dtype=np.float64)
        self.ey = np.ascontiguousarray(ey, dtype=np.float64)
        nb = np.ascontiguousarray(np.asarray(neighbour, dtype=np.int32).T)  # [3][E]
        se = np.ascontiguousarray(np.asarray(shared_edge, dtype=np.int32).T)
        bc = np.ascontiguousarray(np.asarray(boundary, dtype=np.int32).T)
        ip = np.array([P, flux, limiter, dt_policy, reduction_threads], dtype=np.int32)
        dp = np.array([gamma, cfl, tvb_M, dt_fixed] + list(freestream), dtype=np.float64)
        did = None if dt_id is None else np.ascontiguousarray(dt_id, dtype=np.int32)
        self.h = self.lib.orc_create(E, _d(self.ex), _d(self.ey), _i(nb), _i(se), _i(bc),
                                     _i(did) if did is not None else None, _i(ip), _d(dp))

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def shape(self):
        return (4, self.Np, self.E)

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        self.lib.orc_set_state(self.h, _d(U))

    def get_state(self):
        U = np.empty(self.shape())
        self.lib.orc_get_state(self.h, _d(U))
        return U

    def mass(self):
        M = np.empty((self.Np, self.E))
        self.lib.orc_get_mass(self.h, _d(M))
        return M

    def residual(self):
        R = np.empty(self.shape())
        self.lib.orc_residual(self.h, _d(R))
        return R

    def compute_dt(self):
        return self.lib.orc_compute_dt(self.h)

    def stage(self, s, dt):
        self.lib.orc_stage(self.h, s, dt)

    def advance(self, nsteps, t_end=0.0, t0=0.0):
        t = np.array([t0])
        log = np.zeros(max(nsteps, 1))
        n = self.lib.orc_advance(self.h, nsteps, t_end, _d(t), _d(log))
        return n, float(t[0]), log[:n]

    def residual_change(self):
        r = np.zeros(2)
        self.lib.orc_residual_change(self.h, _d(r))
        return r

    def flags(self):
        f = np.zeros(self.E, np.int32)
        self.lib.orc_get_flags(self.h, _i(f))
        return f

    def project(self, kind, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self.lib.orc_project(self.h, _d(self.ex), _d(self.ey), kind, _d(p))


def set_threads(n):
    load().orc_set_threads(int(n))


def get_threads():
    return load().orc_get_threads()


def vortex_exact(px, py, params):
    lib = load()
    px = np.ascontiguousarray(px, dtype=np.float64).ravel()
    py = np.ascontiguousarray(py, dtype=np.float64).ravel()
    p = np.ascontiguousarray(params, dtype=np.float64)
    U = np.zeros((px.size, 4))
    lib.orc_vortex_exact(px.size, _d(px), _d(py), _d(p), _d(U))
    return U


def roe_flux(gamma, UL, UR, nx, ny):
    lib = load()
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    F = np.zeros(4)
    lib.orc_roe_flux(gamma, _d(UL), _d(UR), nx, ny, _d(F))
    return F


def l2_norms(U, mass):
    """Per-variable L2 norm of a modal field, sqrt(sum_e sum_i M_i c_i^2) (orthogonal basis)."""
    return np.sqrt(np.einsum("vie,ie->v", U * U, mass))


def load_ref():
    """The reference's own host precompute (oracle/_ref/libref.so), or None when not built."""
    if not os.path.exists(REF_PATH):
        return None
    lib = ctypes.CDLL(REF_PATH)
    lib.ref_triangle_info.argtypes = [ctypes.c_int, D, D, D]
    lib.ref_grid.argtypes = [ctypes.c_char_p, ctypes.c_char_p, I, I, I, I, D]
    lib.ref_config.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_char_p,
                               ctypes.c_int]
    return lib
