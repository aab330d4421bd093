"""ctypes wrapper of oracle/liboracle.so (the CPU restatement in oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as a product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# OGSX_ORACLE_LIB: another build of oracle.c (the sanitizer runs use one)
LIB_PATH = os.environ.get("OGSX_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_DUMP = os.path.join(HERE, "_ref", "ogs5_ref_dump")

LINE2, QUAD4, HEX8, TET4 = 1, 2, 3, 5
LIQUID_FLOW, HEAT_TRANSPORT = 1, 3


class Problem(C.Structure):
    _fields_ = [
        ("etype", C.c_int), ("process", C.c_int), ("dim", C.c_int), ("n_gauss", C.c_int),
        ("n_nodes", C.c_int64), ("n_elem", C.c_int64),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("conn", C.c_void_p),
        ("mass_coef", C.c_void_p), ("cond", C.c_void_p), ("cond_stride", C.c_int),
        ("tuf", C.c_double), ("mu", C.c_double), ("geo_area", C.c_double),
        ("adv_coef", C.c_double), ("vel", C.c_void_p), ("rho_g", C.c_double),
        ("dt", C.c_double), ("theta", C.c_double), ("steady", C.c_int), ("u_old", C.c_void_p),
    ]


class Eqs(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("nnz", C.c_int64), ("ptr", C.c_void_p), ("col", C.c_void_p),
        ("diag", C.c_void_p), ("a", C.c_void_p), ("b", C.c_void_p), ("x", C.c_void_p),
        ("bnorm", C.c_double), ("error", C.c_double), ("iter", C.c_int), ("max_iter", C.c_int),
        ("tol", C.c_double), ("method", C.c_int), ("precond", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            "orc_gauss_pkt": (C.c_double, [C.c_long, C.c_long]),
            "orc_gauss_fkt": (C.c_double, [C.c_long, C.c_long]),
            "orc_sample_tet5": (None, [C.c_int, P]),
            "orc_ngp": (C.c_int, [C.c_int, C.c_int]),
            "orc_shape_pool": (None, [C.c_int, C.c_int, P, P]),
            "orc_pattern_ptr": (C.c_int64, [C.c_int64, C.c_int64, C.c_int, P, P]),
            "orc_pattern_fill": (None, [C.c_int64, C.c_int64, C.c_int, P, P, P, P]),
            "orc_element_matrices": (None, [C.POINTER(Problem), C.c_int64, P, P, P, P]),
            "orc_element_system": (None, [C.POINTER(Problem), C.c_int64, P, P]),
            "orc_gauss_velocity": (None, [C.POINTER(Problem), P, P]),
            "orc_jds_build": (C.c_int64, [C.c_int64, P, P, P, P, P, P]),
            "orc_jds_spmv": (None, [C.c_int64, C.c_int64, P, P, P, P, P, P]),
            "orc_eqs_init": (None, [C.POINTER(Eqs), C.c_int64, C.c_int64, P, P, P, P, P, P]),
            "orc_eqs_config": (None, [C.POINTER(Eqs), C.c_int, C.c_int, C.c_double, C.c_int]),
            "orc_eqs_initialize": (None, [C.POINTER(Eqs)]),
            "orc_assemble": (None, [C.POINTER(Eqs), C.POINTER(Problem)]),
            "orc_add_rhs": (None, [C.POINTER(Eqs), C.c_int64, P, P]),
            "orc_set_known_x": (None, [C.POINTER(Eqs), C.c_int64, C.c_double]),
            "orc_set_known_x_sym": (None, [C.POINTER(Eqs), C.c_int64, C.c_double]),
            "orc_spmv": (None, [C.POINTER(Eqs), P, P]),
            "orc_precond_jacobi": (None, [C.POINTER(Eqs), P, P]),
            "orc_dot": (C.c_double, [C.c_int64, P, P]),
            "orc_cg": (C.c_int, [C.POINTER(Eqs)]),
            "orc_bicgstab": (C.c_int, [C.POINTER(Eqs)]),
            "orc_solver": (C.c_int, [C.POINTER(Eqs)]),
            "orc_write_bin": (C.c_int, [C.POINTER(Eqs), C.c_char_p]),
            "orc_dense_multi": (None, [P, C.c_int, C.c_int, P, P, C.c_double]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p).value


NPE = {LINE2: 2, QUAD4: 4, HEX8: 8, TET4: 4}


def pattern(n_nodes, conn):
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    n_elem, npe = conn.shape
    ptr = np.empty(n_nodes + 1, np.int64)
    nnz = lib().orc_pattern_ptr(n_nodes, n_elem, npe, _p(conn), _p(ptr))
    col = np.empty(nnz, np.int64)
    diag = np.empty(n_nodes, np.int64)
    lib().orc_pattern_fill(n_nodes, n_elem, npe, _p(conn), _p(ptr), _p(col), _p(diag))
    return ptr, col, diag


class OracleProblem:
    """Holds numpy arrays alive and exposes the orc_problem struct."""

    def __init__(self, mesh, *, process, mass_coef, cond, cond_stride=1, n_gauss=3, tuf=1.0,
                 mu=1.0, geo_area=1.0, adv_coef=0.0, vel=None, rho_g=0.0, dt=1.0, theta=1.0,
                 steady=False, u_old=None):
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        self.arrays = dict(x=f(mesh.x), y=f(mesh.y), z=f(mesh.z),
                           conn=np.ascontiguousarray(mesh.conn, dtype=np.int64),
                           mass_coef=f(mass_coef), cond=f(cond), vel=f(vel),
                           u_old=f(u_old if u_old is not None else np.zeros(mesh.n_nodes)))
        a = self.arrays
        self.s = Problem(mesh.etype, process, mesh.dim, n_gauss, mesh.n_nodes, mesh.n_elem,
                         _p(a["x"]), _p(a["y"]), _p(a["z"]), _p(a["conn"]), _p(a["mass_coef"]),
                         _p(a["cond"]), cond_stride, tuf, mu, geo_area, adv_coef, _p(a["vel"]),
                         rho_g, dt, theta, int(bool(steady)), _p(a["u_old"]))
        self.npe = NPE[mesh.etype]

    def set(self, **kw):
        for k, v in kw.items():
            if k in ("u_old", "vel"):
                self.arrays[k] = np.ascontiguousarray(v, dtype=np.float64)
                setattr(self.s, k, _p(self.arrays[k]))
            elif k == "steady":
                self.s.steady = int(bool(v))
            else:
                setattr(self.s, k, v)

    def element_matrices(self, e):
        n = self.npe
        M, L, A, G = (np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n)), np.zeros(n))
        lib().orc_element_matrices(C.byref(self.s), e, _p(M), _p(L), _p(A), _p(G))
        return M, L, A, G

    def gauss_velocity(self, u):
        """[n_elem][3][ngp] Gauss-point Darcy velocity from nodal values u."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        ngp = lib().orc_ngp(self.s.etype, self.s.n_gauss)
        vel = np.empty((self.s.n_elem, 3, ngp))
        lib().orc_gauss_velocity(C.byref(self.s), _p(u), _p(vel))
        return vel

    def element_system(self, e):
        n = self.npe
        lhs, rhs = np.zeros((n, n)), np.zeros(n)
        lib().orc_element_system(C.byref(self.s), e, _p(lhs), _p(rhs))
        return lhs, rhs


class OracleEqs:
    """Linear_EQS restatement with the reference's bNorm persistence."""

    def __init__(self, ptr, col, diag):
        self.ptr, self.col, self.diag = ptr, col, diag
        n = ptr.size - 1
        self.n, self.nnz = n, int(ptr[-1])
        self.a = np.zeros(self.nnz)
        self.b = np.zeros(n)
        self.x = np.zeros(n)
        self.s = Eqs()
        lib().orc_eqs_init(C.byref(self.s), n, self.nnz, _p(ptr), _p(col), _p(diag), _p(self.a),
                           _p(self.b), _p(self.x))

    def config(self, method, precond=1, tol=1e-10, max_iter=10000):
        lib().orc_eqs_config(C.byref(self.s), method, precond, tol, max_iter)

    def initialize(self):
        lib().orc_eqs_initialize(C.byref(self.s))

    def assemble(self, prob: OracleProblem):
        lib().orc_assemble(C.byref(self.s), C.byref(prob.s))

    def add_rhs(self, idx, val):
        idx = np.ascontiguousarray(idx, np.int64)
        val = np.ascontiguousarray(val, np.float64)
        lib().orc_add_rhs(C.byref(self.s), idx.size, _p(idx), _p(val))

    def set_known_x(self, idx, val, symmetric=False):
        f = lib().orc_set_known_x_sym if symmetric else lib().orc_set_known_x
        for i, g in zip(np.asarray(idx).ravel(), np.asarray(val).ravel()):
            f(C.byref(self.s), int(i), float(g))

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n)
        lib().orc_spmv(C.byref(self.s), _p(x), _p(y))
        return y

    def solve(self):
        ret = lib().orc_solver(C.byref(self.s))
        return dict(ret=ret, iterations=self.s.iter, error=self.s.error, bnorm=self.s.bnorm)

    def write_bin(self, path):
        return lib().orc_write_bin(C.byref(self.s), str(path).encode())

    def dense(self):
        A = np.zeros((self.n, self.n))
        for i in range(self.n):
            A[i, self.col[self.ptr[i]:self.ptr[i + 1]]] = self.a[self.ptr[i]:self.ptr[i + 1]]
        return A


def read_bin(path):
    """Reads a Write_BIN file (FileIO/MathIO/CRSIO.h:70 long variant)."""
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw, np.int64, 1, 0)[0])
    off = 8
    ptr = np.frombuffer(raw, np.int64, n + 1, off).copy()
    off += 8 * (n + 1)
    nnz = int(ptr[-1])
    col = np.frombuffer(raw, np.int64, nnz, off).copy()
    off += 8 * nnz
    a = np.frombuffer(raw, np.float64, nnz, off).copy()
    off += 8 * nnz
    b = np.frombuffer(raw, np.float64, n, off).copy()
    return n, ptr, col, a, b


def ref_dump_lines():
    """Lines printed by the reference-built oracle/_ref/ogs5_ref_dump, or None."""
    import subprocess
    if not os.path.exists(REF_DUMP):
        return None
    return subprocess.run([REF_DUMP], check=True, capture_output=True, text=True).stdout.splitlines()


def jds_spmv(ptr, col, a, x):
    """JDS layout of the CRS pattern and y = A x in JDS storage order."""
    n, nnz = ptr.size - 1, int(ptr[-1])
    n2o = np.empty(max(n, 1), np.int64)
    ncol = np.empty(max(n, 1), np.int64)
    jcol = np.empty(max(nnz, 1), np.int64)
    src = np.empty(max(nnz, 1), np.int64)
    ptr = np.ascontiguousarray(ptr, np.int64)
    col = np.ascontiguousarray(col, np.int64)
    maxc = lib().orc_jds_build(n, _p(ptr), _p(col), _p(n2o), _p(ncol), _p(jcol), _p(src))
    aj = np.ascontiguousarray(np.asarray(a, np.float64)[src[:nnz]])
    y = np.empty(n)
    lib().orc_jds_spmv(n, maxc, _p(ncol), _p(n2o), _p(jcol), _p(aj),
                       _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y, dict(n2o=n2o[:n], ncol=ncol[:maxc], jcol=jcol[:nnz], src=src[:nnz])
