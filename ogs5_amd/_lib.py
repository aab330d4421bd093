"""ctypes binding of the C ABI in include/ogsx.h (libogsx.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("OGSX_LIB") or os.path.join(_HERE, "libogsx.so")

OK = 0
LINE2, QUAD4, HEX8, TET4 = 1, 2, 3, 5
LIQUID_FLOW, GROUNDWATER_FLOW, HEAT_TRANSPORT = 1, 2, 3
BICGSTAB, CG = 2, 5
PRECOND_NONE, PRECOND_JACOBI = 0, 1
FP64, FP32_IR = 0, 1
DIRICHLET_ROW, DIRICHLET_SYMMETRIC = 0, 1
VEC_X, VEC_B, VEC_UOLD, VEC_IC = 0, 1, 2, 3
PROF_SPMV64, PROF_SPMV32, PROF_ASSEMBLY, PROF_VELOCITY = 0, 1, 2, 3


class OgsxError(RuntimeError):
    pass


class Elements(C.Structure):
    _fields_ = [
        ("process", C.c_int), ("dim", C.c_int), ("n_gauss", C.c_int), ("cond_stride", C.c_int),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
        ("mass_coef", C.c_void_p), ("cond", C.c_void_p),
        ("tuf", C.c_double), ("mu", C.c_double), ("geo_area", C.c_double),
        ("adv_coef", C.c_double), ("vel", C.c_void_p), ("rho_g", C.c_double),
    ]


IR_RESTART, IR_RELIABLE, IR_RELIABLE_SR = 0, 1, 2


class SolverConfig(C.Structure):
    _fields_ = [("method", C.c_int), ("precond", C.c_int), ("max_iter", C.c_int),
                ("tol", C.c_double), ("inner_tol", C.c_double), ("max_outer", C.c_int),
                ("ir_mode", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer", C.c_int), ("converged", C.c_int),
                ("error", C.c_double), ("bnorm", C.c_double), ("ret", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every exported symbol and its signature (mirrors include/ogsx.h).
_P = C.c_void_p
_I64 = C.c_int64
SIGNATURES = {
    "ogsx_last_error": (C.c_char_p, []),
    "ogsx_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ogsx_version": (C.c_char_p, []),
    "ogsx_pattern_create": (C.c_int, [_I64, _I64, C.c_int, _P, C.POINTER(_P)]),
    "ogsx_pattern_info": (C.c_int, [_P, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(C.c_int)]),
    "ogsx_pattern_export": (C.c_int, [_P, _P, _P, _P]),
    "ogsx_pattern_colors": (C.c_int, [_P, _P]),
    "ogsx_pattern_destroy": (C.c_int, [_P]),
    "ogsx_system_create": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "ogsx_system_destroy": (C.c_int, [_P]),
    "ogsx_system_set_elements": (C.c_int, [_P, C.POINTER(Elements)]),
    "ogsx_system_initialize": (C.c_int, [_P]),
    "ogsx_assemble": (C.c_int, [_P, C.c_double, C.c_double, C.c_int]),
    "ogsx_gauss_velocity": (C.c_int, [_P, _P, _P]),
    "ogsx_add_rhs": (C.c_int, [_P, _I64, _P, _P]),
    "ogsx_set_dirichlet": (C.c_int, [_P, _I64, _P, _P, C.c_int]),
    "ogsx_solve": (C.c_int, [_P, C.POINTER(SolverConfig), C.POINTER(SolveInfo)]),
    "ogsx_set_vector": (C.c_int, [_P, C.c_int, _P]),
    "ogsx_get_vector": (C.c_int, [_P, C.c_int, _P]),
    "ogsx_copy_vector": (C.c_int, [_P, C.c_int, C.c_int]),
    "ogsx_get_matrix": (C.c_int, [_P, _P]),
    "ogsx_spmv": (C.c_int, [_P, _P, _P]),
    "ogsx_write_bin": (C.c_int, [_P, C.c_char_p]),
    "ogsx_profile_enable": (C.c_int, [_P, C.c_int]),
    "ogsx_profile_read": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(_I64),
                                    C.POINTER(C.c_double)]),
    "ogsx_set_spmv_variant": (C.c_int, [_P, C.c_int]),
    "ogsx_get_spmv_variant": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "ogsx_get_spmv_formats": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ogsx_spmv_tuning": (C.c_int, [_P, C.c_int, _P, _P, C.POINTER(C.c_int)]),
    "ogsx_bench_spmv": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]),
    "ogsx_stream_copy": (C.c_int, [_I64, C.c_int, C.POINTER(C.c_double)]),
    "ogsx_synchronize": (C.c_int, [_P]),
    "ogsx_set_persistent": (C.c_int, [C.c_int]),
    "ogsx_get_persistent": (C.c_int, [_P, _P]),
    "ogsx_persistent_giveup_probe": (C.c_int, [_P]),
    "ogsx_solve_crs_noalloc": (C.c_int, [_I64, _P, _P, _P, _P, _P, _P,
                                         C.POINTER(SolverConfig), _P, C.POINTER(SolveInfo)]),
    "ogsx_pattern_footprint": (C.c_int, [_P, C.c_int, C.POINTER(_I64)]),
    "ogsx_system_bytes": (C.c_int, [_P, C.POINTER(_I64)]),
    "ogsx_group_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_P)]),
    "ogsx_group_destroy": (C.c_int, [_P]),
    "ogsx_group_add": (C.c_int, [_P, _P]),
    "ogsx_group_remove": (C.c_int, [_P, _P]),
    "ogsx_group_start": (C.c_int, [_P, C.POINTER(_P), C.c_int, C.POINTER(SolverConfig)]),
    "ogsx_group_step": (C.c_int, [_P, C.POINTER(_P), C.c_int, _P, C.POINTER(SolverConfig)]),
    "ogsx_group_advance": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    "ogsx_solve_status": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(SolveInfo)]),
    "ogsx_group_solve": (C.c_int, [_P, C.POINTER(SolverConfig), C.POINTER(SolveInfo)]),
}

_lib = None
_lib_lock = threading.Lock()


def lib():
    """The loaded libogsx.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise OgsxError(f"{LIB_PATH} is missing: run __graft_entry__.build() first "
                            "(the HIP library has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != OK:
        msg = lib().ogsx_last_error().decode()
        raise OgsxError(f"{what} failed ({rc}): {msg}")


def device_count() -> int:
    c = C.c_int(0)
    check(lib().ogsx_device_count(C.byref(c)), "ogsx_device_count")
    return c.value


def stream_copy_gbs(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Achieved (read + write) GB/s of a plain device copy of nbytes."""
    g = C.c_double()
    check(lib().ogsx_stream_copy(int(nbytes), int(reps), C.byref(g)), "ogsx_stream_copy")
    return g.value


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def solver_config(method=CG, precond=PRECOND_JACOBI, tol=1e-10, max_iter=10000,
                  inner_tol=None, max_outer=50, ir_mode=None) -> SolverConfig:
    """ogsx_solver_config; for IR systems CG defaults to reliable updates
    (inner_tol = delta 0.1), BiCGStab to restarted refinement (inner
    tolerance 1e-3; 1e-4 for restarted CG)."""
    if ir_mode is None:
        ir_mode = IR_RELIABLE if method == CG else IR_RESTART
    if inner_tol is None:
        inner_tol = (0.1 if ir_mode in (IR_RELIABLE, IR_RELIABLE_SR)
                     else 1e-3 if method == BICGSTAB else 1e-4)
    return SolverConfig(method, precond, max_iter, tol, inner_tol, max_outer, ir_mode)


class Pattern:
    """CRS sparsity pattern + element colouring (ogsx_pattern_create)."""

    def __init__(self, n_nodes: int, etype: int, conn: np.ndarray):
        conn = _i64(conn)
        self._keep = conn
        h = C.c_void_p()
        n_elem = conn.shape[0] if conn.ndim == 2 else conn.size // {1: 2, 2: 4, 3: 8, 5: 4}[etype]
        check(lib().ogsx_pattern_create(int(n_nodes), int(n_elem), int(etype), _ptr(conn),
                                        C.byref(h)), "ogsx_pattern_create")
        self.h = h
        self.etype = etype
        self.n_elem = int(n_elem)
        n, nnz, nc = C.c_int64(), C.c_int64(), C.c_int()
        check(lib().ogsx_pattern_info(h, C.byref(n), C.byref(nnz), C.byref(nc)), "info")
        self.n, self.nnz, self.n_colors = n.value, nnz.value, nc.value

    def export(self):
        ptr = np.empty(self.n + 1, np.int64)
        col = np.empty(self.nnz, np.int64)
        diag = np.empty(self.n, np.int64)
        check(lib().ogsx_pattern_export(self.h, _ptr(ptr), _ptr(col), _ptr(diag)), "export")
        return ptr, col, diag

    def colors(self):
        c = np.empty(self.n_elem, np.int32)
        check(lib().ogsx_pattern_colors(self.h, _ptr(c)), "colors")
        return c

    def footprint(self, precision: int = FP64) -> int:
        """Device bytes of a system on this pattern (ogsx_pattern_footprint)."""
        b = C.c_int64()
        check(lib().ogsx_pattern_footprint(self.h, int(precision), C.byref(b)),
              "ogsx_pattern_footprint")
        return b.value

    def close(self):
        """Frees the pattern; refused (OgsxError) while systems on it exist."""
        if self.h:
            check(lib().ogsx_pattern_destroy(self.h), "ogsx_pattern_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class System:
    """One device-resident linear system (ogsx_system_*)."""

    def __init__(self, pattern: Pattern, precision: int = FP64):
        self.pattern = pattern
        h = C.c_void_p()
        check(lib().ogsx_system_create(pattern.h, int(precision), C.byref(h)), "ogsx_system_create")
        self.h = h
        self.n = pattern.n
        self.nnz = pattern.nnz
        self.precision = precision
        self._keep = []
        self.n_gauss = 3

    def set_elements(self, *, process, dim, x, y=None, z=None, mass_coef, cond, cond_stride=1,
                     n_gauss=3, tuf=1.0, mu=1.0, geo_area=1.0, adv_coef=0.0, vel=None,
                     rho_g=0.0):
        arrs = [_f64(x), _f64(y) if y is not None else None, _f64(z) if z is not None else None,
                _f64(mass_coef), _f64(cond), _f64(vel) if vel is not None else None]
        self._keep = arrs
        e = Elements(process, dim, n_gauss, cond_stride,
                     *(None if a is None else _ptr(a).value for a in arrs[:5]),
                     tuf, mu, geo_area, adv_coef,
                     None if arrs[5] is None else _ptr(arrs[5]).value, rho_g)
        check(lib().ogsx_system_set_elements(self.h, C.byref(e)), "ogsx_system_set_elements")
        self.n_gauss = int(n_gauss)

    def initialize(self):
        check(lib().ogsx_system_initialize(self.h), "ogsx_system_initialize")

    def assemble(self, dt, theta, steady=False):
        check(lib().ogsx_assemble(self.h, float(dt), float(theta), int(bool(steady))),
              "ogsx_assemble")

    def gauss_velocity(self, heat: "System" = None, host: bool = True):
        """Darcy velocity at the Gauss points from this LIQUID_FLOW system's
        solution (ogsx_gauss_velocity).  heat: a HEAT_TRANSPORT system on the
        same mesh that takes it as its advection velocity on the device.
        Returns the [n_elem, 3, ngp] array when host is true."""
        out = None
        if host:
            et, ng = self.pattern.etype, self.n_gauss
            ngp = {LINE2: ng, QUAD4: ng * ng, HEX8: ng ** 3, TET4: 5}[et]
            out = np.empty((self.pattern.n_elem, 3, ngp))
        check(lib().ogsx_gauss_velocity(self.h, heat.h if heat is not None else None,
                                        _ptr(out) if out is not None else None),
              "ogsx_gauss_velocity")
        return out

    def add_rhs(self, idx, val):
        idx, val = _i64(idx), _f64(val)
        check(lib().ogsx_add_rhs(self.h, idx.size, _ptr(idx), _ptr(val)), "ogsx_add_rhs")

    def set_dirichlet(self, idx, val, mode=DIRICHLET_ROW):
        idx, val = _i64(idx), _f64(val)
        check(lib().ogsx_set_dirichlet(self.h, idx.size, _ptr(idx), _ptr(val), int(mode)),
              "ogsx_set_dirichlet")

    def solve(self, method=CG, precond=PRECOND_JACOBI, tol=1e-10, max_iter=10000,
              inner_tol=None, max_outer=50, ir_mode=None) -> dict:
        """IR systems: CG defaults to reliable updates (inner_tol = delta 0.1),
        BiCGStab to restarted refinement (inner_tol 1e-4)."""
        cfg = solver_config(method, precond, tol, max_iter, inner_tol, max_outer, ir_mode)
        info = SolveInfo()
        check(lib().ogsx_solve(self.h, C.byref(cfg), C.byref(info)), "ogsx_solve")
        return info.as_dict()

    def set_vector(self, which, v):
        v = _f64(v)
        assert v.size == self.n
        check(lib().ogsx_set_vector(self.h, which, _ptr(v)), "ogsx_set_vector")

    def get_vector(self, which) -> np.ndarray:
        v = np.empty(self.n, np.float64)
        check(lib().ogsx_get_vector(self.h, which, _ptr(v)), "ogsx_get_vector")
        return v

    def copy_vector(self, dst, src):
        check(lib().ogsx_copy_vector(self.h, dst, src), "ogsx_copy_vector")

    def get_matrix(self) -> np.ndarray:
        a = np.empty(self.nnz, np.float64)
        check(lib().ogsx_get_matrix(self.h, _ptr(a)), "ogsx_get_matrix")
        return a

    def spmv(self, x) -> np.ndarray:
        x = _f64(x)
        y = np.empty(self.n, np.float64)
        check(lib().ogsx_spmv(self.h, _ptr(x), _ptr(y)), "ogsx_spmv")
        return y

    def write_bin(self, path):
        check(lib().ogsx_write_bin(self.h, str(path).encode()), "ogsx_write_bin")

    def profile(self, on=True):
        check(lib().ogsx_profile_enable(self.h, int(on)), "ogsx_profile_enable")

    def profile_read(self, kind):
        ms, n, b = C.c_double(), C.c_int64(), C.c_double()
        check(lib().ogsx_profile_read(self.h, kind, C.byref(ms), C.byref(n), C.byref(b)),
              "ogsx_profile_read")
        return ms.value, n.value, b.value

    def set_spmv_variant(self, lanes_per_row):
        check(lib().ogsx_set_spmv_variant(self.h, int(lanes_per_row)), "ogsx_set_spmv_variant")

    def spmv_variant(self) -> int:
        v = C.c_int()
        check(lib().ogsx_get_spmv_variant(self.h, C.byref(v)), "ogsx_get_spmv_variant")
        return v.value

    def spmv_format(self, fp32=False) -> int:
        """SpMV format in use for the FP64 (or, for IR systems, FP32) matrix:
        the tuner's fastest, a forced variant, one shared in a group, or 25
        (the symmetric diagonal copy of a symmetric system on a banded
        pattern, chosen by structure at the solve)."""
        f64, f32 = C.c_int(), C.c_int()
        check(lib().ogsx_get_spmv_formats(self.h, C.byref(f64), C.byref(f32)),
              "ogsx_get_spmv_formats")
        return f32.value if fp32 else f64.value

    def bench_spmv(self, fp32=False, reps=20):
        ms, b = C.c_double(), C.c_double()
        check(lib().ogsx_bench_spmv(self.h, int(bool(fp32)), int(reps), C.byref(ms), C.byref(b)),
              "ogsx_bench_spmv")
        return ms.value, b.value

    def spmv_tuning(self):
        f = np.zeros(32, np.int32)
        ms = np.zeros(32, np.float64)
        cnt = C.c_int()
        check(lib().ogsx_spmv_tuning(self.h, 32, _ptr(f), _ptr(ms), C.byref(cnt)),
              "ogsx_spmv_tuning")
        return [(int(a), float(b)) for a, b in zip(f[:cnt.value], ms[:cnt.value])]

    def synchronize(self):
        check(lib().ogsx_synchronize(self.h), "ogsx_synchronize")

    def device_bytes(self) -> int:
        b = C.c_int64()
        check(lib().ogsx_system_bytes(self.h, C.byref(b)), "ogsx_system_bytes")
        return b.value

    def close(self):
        if getattr(self, "h", None):
            lib().ogsx_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StepParams(C.Structure):
    """ogsx_step_params: one member's time step in ogsx_group_step."""
    _fields_ = [("dt", C.c_double), ("theta", C.c_double), ("steady", C.c_int),
                ("dirichlet", C.c_int), ("sources", C.c_int), ("shift", C.c_int)]


class Group:
    """Systems solved together (ogsx_group_*): one launch per Krylov phase
    over all running members; each member's results are those of
    System.solve on it alone."""

    def __init__(self, precision: int = FP64, use_graphs: bool = True):
        h = C.c_void_p()
        check(lib().ogsx_group_create(int(precision), int(bool(use_graphs)), C.byref(h)),
              "ogsx_group_create")
        self.h = h
        self.precision = precision
        self.members = []

    def add(self, system: "System"):
        check(lib().ogsx_group_add(self.h, system.h), "ogsx_group_add")
        self.members.append(system)

    def remove(self, system: "System"):
        check(lib().ogsx_group_remove(self.h, system.h), "ogsx_group_remove")
        self.members.remove(system)

    def start(self, systems, **cfg):
        arr = (C.c_void_p * len(systems))(*[s.h.value for s in systems])
        c = solver_config(**cfg)
        check(lib().ogsx_group_start(self.h, arr, len(systems), C.byref(c)), "ogsx_group_start")

    def step(self, systems, params, **cfg):
        """ogsx_group_step: one time step of each listed system in batched
        launches (params: one dict per system with dt, theta, steady,
        dirichlet (-1 for none), sources, shift), then their solves."""
        if not systems:
            return
        arr = (C.c_void_p * len(systems))(*[s.h.value for s in systems])
        ps = (StepParams * len(systems))()
        for k, p in enumerate(params):
            ps[k] = StepParams(float(p["dt"]), float(p["theta"]), int(bool(p.get("steady", 0))),
                               int(p.get("dirichlet", -1)), int(bool(p.get("sources", 0))),
                               int(bool(p.get("shift", 0))))
        c = solver_config(**cfg)
        check(lib().ogsx_group_step(self.h, arr, len(systems), C.cast(ps, C.c_void_p),
                                    C.byref(c)), "ogsx_group_step")

    def advance(self, max_iterations: int = 1 << 30) -> int:
        n = C.c_int()
        check(lib().ogsx_group_advance(self.h, int(max_iterations), C.byref(n)),
              "ogsx_group_advance")
        return n.value

    def solve(self, **cfg):
        infos = (SolveInfo * max(1, len(self.members)))()
        c = solver_config(**cfg)
        check(lib().ogsx_group_solve(self.h, C.byref(c), infos), "ogsx_group_solve")
        return [infos[i].as_dict() for i in range(len(self.members))]

    def close(self):
        if getattr(self, "h", None):
            lib().ogsx_group_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_status(system: "System"):
    """(active, info) of the system's last solve."""
    a = C.c_int()
    info = SolveInfo()
    check(lib().ogsx_solve_status(system.h, C.byref(a), C.byref(info)), "ogsx_solve_status")
    return bool(a.value), info.as_dict()


# Rows the zero-allocation solve takes at most (its fixed device array).
NOALLOC_MAX_ROWS = 1 << 19


def solve_crs_noalloc(n: int, row_ptr_dev: int, col_dev: int, val_dev: int, b_dev: int,
                      x_dev: int, *, bnorm: float = 1.0, stream: int = 0, **cfg_kw):
    """ogsx_solve_crs_noalloc on device addresses (ints): the FP64 solve of
    ogsx_solve on a caller-held CRS matrix, with no allocation.  Returns
    (info, ||b||); pass ||b|| back as `bnorm` to carry CheckNormRHS's
    reference between solves."""
    cfg = solver_config(**cfg_kw)
    info = SolveInfo()
    bn = C.c_double(bnorm)
    check(lib().ogsx_solve_crs_noalloc(int(n), C.c_void_p(row_ptr_dev), C.c_void_p(col_dev),
                                       C.c_void_p(val_dev), C.c_void_p(b_dev),
                                       C.c_void_p(x_dev), C.byref(bn), C.byref(cfg),
                                       C.c_void_p(stream or None), C.byref(info)),
          "ogsx_solve_crs_noalloc")
    return info.as_dict(), bn.value


def set_persistent(mode: int):
    """Persistent Krylov launches: 0 off, 1 auto (small systems), 2 always."""
    check(lib().ogsx_set_persistent(int(mode)), "ogsx_set_persistent")


def get_persistent():
    """(mode, device supports cooperative launches)."""
    m, sup = C.c_int(), C.c_int()
    check(lib().ogsx_get_persistent(C.byref(m), C.byref(sup)), "ogsx_get_persistent")
    return m.value, bool(sup.value)


def persistent_giveup_probe() -> int:
    """Runs the barrier give-up diagnostic; 1 when the blocks gave up."""
    a = C.c_int(0)
    check(lib().ogsx_persistent_giveup_probe(C.byref(a)), "ogsx_persistent_giveup_probe")
    return a.value
