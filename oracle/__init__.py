"""ctypes front-end of the FP64 CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` as the checker.  The product path never
imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OMesh(C.Structure):
    _fields_ = [("np", C.c_int), ("v", _dp), ("pptr", _ip), ("pv", _ip), ("p2p", _ip),
                ("nbc", C.c_int), ("bc", _ip), ("tptr", _ip), ("txy", _dp), ("bb", _dp)]


class OEq(C.Structure):
    _fields_ = [("kind", C.c_int), ("beta", C.c_double * 6), ("gamma", C.c_double),
                ("bc_kind", C.c_int), ("bc_par", C.c_double * 8), ("periodic_trace", C.c_int),
                ("jacobian", C.c_int)]


class OSolver(C.Structure):
    _fields_ = [("gmres_mode", C.c_int), ("restart", C.c_int), ("rtol", C.c_double), ("maxit", C.c_int),
                ("pc_kind", C.c_int), ("newton_tol", C.c_double), ("newton_maxit", C.c_int),
                ("dirk_tableau", C.c_int)]


class OStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int), ("gmres_iters", C.c_int), ("linear_solves", C.c_int),
                ("last_resid", C.c_double)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Oracle:
    """Oracle bound to one mesh; arrays are kept alive by this object."""

    def __init__(self, mesh):
        self.mesh = mesh
        self._keep = dict(v=np.ascontiguousarray(mesh.v, np.float64),
                          pptr=np.ascontiguousarray(mesh.pptr, np.int32),
                          pv=np.ascontiguousarray(mesh.pv, np.int32),
                          p2p=np.ascontiguousarray(mesh.p2p, np.int32),
                          bc=np.ascontiguousarray(mesh.bc, np.int32),
                          tptr=np.ascontiguousarray(mesh.tptr, np.int32),
                          txy=np.ascontiguousarray(mesh.txy, np.float64),
                          bb=np.ascontiguousarray(mesh.bb, np.float64))
        k = self._keep
        self.m = OMesh(mesh.n_cells, _d(k["v"]), _i(k["pptr"]), _i(k["pv"]), _i(k["p2p"]), len(k["bc"]),
                       _i(k["bc"]), _i(k["tptr"]), _d(k["txy"]), _d(k["bb"]))
        self.n = mesh.n_cells
        rp = np.zeros(self.n + 1, np.int32)
        nnzb = lib().o_jacobian_pattern(C.byref(self.m), _i(rp), None)
        ci = np.zeros(nnzb, np.int32)
        lib().o_jacobian_pattern(C.byref(self.m), _i(rp), _i(ci))
        self.rowptr, self.colind = rp, ci

    @staticmethod
    def eq(spec) -> OEq:
        e = OEq()
        e.kind = spec.kind
        for j in range(6):
            e.beta[j] = spec.beta[j]
        e.gamma = spec.gamma
        e.bc_kind = spec.bc_kind
        for j in range(8):
            e.bc_par[j] = spec.bc_par[j] if j < len(spec.bc_par) else 0.0
        e.periodic_trace = spec.periodic_trace
        e.jacobian = spec.jacobian
        return e

    def mass(self, deg, nc):
        bl = nc * (deg + 1) * (deg + 2) // 2
        out = np.zeros((self.n, bl, bl))
        lib().o_mass(C.byref(self.m), deg, nc, _d(out))
        return out

    def assemble(self, spec, deg, u, t=0.0):
        u = np.ascontiguousarray(u, np.float64)
        b = np.zeros_like(u)
        e = self.eq(spec)
        lib().o_assemble(C.byref(self.m), C.byref(e), deg, _d(u), C.c_double(t), _d(b))
        return b

    def jacobian(self, spec, deg, u, t=0.0):
        u = np.ascontiguousarray(u, np.float64)
        bl = spec.nc * (deg + 1) * (deg + 2) // 2
        blocks = np.zeros((len(self.colind), bl, bl))
        e = self.eq(spec)
        lib().o_jacobian(C.byref(self.m), C.byref(e), deg, _d(u), C.c_double(t), _i(self.rowptr), _d(blocks))
        return blocks

    def l2norm(self, deg, nc, u):
        out = np.zeros(nc)
        u = np.ascontiguousarray(u, np.float64)
        lib().o_l2norm(C.byref(self.m), deg, nc, _d(u), _d(out))
        return out

    def interp_points(self, deg):
        """MeshFn::interp's fitting points, (np, (deg+1)^2, 2)."""
        out = np.zeros((self.n, (deg + 1) ** 2, 2))
        lib().o_interp_points(C.byref(self.m), deg, _d(out))
        return out

    def interp_values(self, deg, nc, vals):
        """Least-squares fit (MeshFn::interp) to values (np, (deg+1)^2, nc) at interp_points."""
        vals = np.ascontiguousarray(vals, np.float64)
        out = np.zeros(self.n * nc * (deg + 1) * (deg + 2) // 2)
        assert lib().o_interp_values(C.byref(self.m), deg, nc, _d(vals), _d(out)) == 0
        return out

    def quad_points(self, deg2):
        """The degree-deg2 polygon rule points of all cells: (count, 3) of x, y, weight x area."""
        cnt = lib().o_quad_count(C.byref(self.m), deg2)
        out = np.zeros((cnt, 3))
        lib().o_quad_points(C.byref(self.m), deg2, _d(out))
        return out

    def l2_error_values(self, deg, nc, u, f):
        """MeshFn::L2Error of u against values f (count, nc) at quad_points(2 deg)."""
        out = np.zeros(nc)
        u = np.ascontiguousarray(u, np.float64)
        f = np.ascontiguousarray(f, np.float64)
        lib().o_l2_error_values(C.byref(self.m), deg, nc, _d(u), _d(f), _d(out))
        return out

    def dirk_step(self, spec, deg, nstages, u, dt, t, solver):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros_like(u)
        st = OStats()
        lib().o_dirk_step(C.byref(self.m), C.byref(self.eq(spec)), deg, nstages, _d(u), C.c_double(dt),
                          C.c_double(t), C.byref(osolver(solver)), _d(out), C.byref(st))
        return out, st

    def irk_step(self, spec, deg, nstages, u, dt, t, solver):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros_like(u)
        st = OStats()
        lib().o_irk_step(C.byref(self.m), C.byref(self.eq(spec)), deg, nstages, _d(u), C.c_double(dt),
                         C.c_double(t), C.byref(osolver(solver)), _d(out), C.byref(st))
        return out, st

    def forward_euler_step(self, spec, deg, u, dt, t):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros_like(u)
        lib().o_forward_euler_step(C.byref(self.m), C.byref(self.eq(spec)), deg, _d(u), C.c_double(dt),
                                   C.c_double(t), _d(out))
        return out

    def rk2_step(self, spec, deg, u, dt, t):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros_like(u)
        lib().o_rk2_step(C.byref(self.m), C.byref(self.eq(spec)), deg, _d(u), C.c_double(dt), C.c_double(t),
                         _d(out))
        return out

    def rk4_step(self, spec, deg, u, dt, t):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros_like(u)
        lib().o_rk4_step(C.byref(self.m), C.byref(self.eq(spec)), deg, _d(u), C.c_double(dt), C.c_double(t),
                         _d(out))
        return out


def osolver(s) -> OSolver:
    return OSolver(s.gmres_mode, s.restart, s.rtol, s.maxit, s.pc_kind, s.newton_tol, s.newton_maxit,
                   getattr(s, "dirk_tableau", 0))


# ---- block-sparse algebra ---------------------------------------------------------

def bsr_matvec(rowptr, colind, A, x):
    n = len(rowptr) - 1
    bl = A.shape[1]
    A = np.ascontiguousarray(A, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(n * bl)
    lib().o_bsr_matvec(n, bl, _i(np.ascontiguousarray(rowptr, np.int32)), _i(np.ascontiguousarray(colind, np.int32)),
                       _d(A), _d(x), _d(y))
    return y


def dilu_factor(rowptr, colind, A):
    n = len(rowptr) - 1
    bl = A.shape[1]
    A = np.ascontiguousarray(A, np.float64)
    F = np.zeros_like(A)
    info = lib().o_dilu_factor(n, bl, _i(np.ascontiguousarray(rowptr, np.int32)),
                               _i(np.ascontiguousarray(colind, np.int32)), _d(A), _d(F))
    return F, info


def dilu_apply(rowptr, colind, F, x):
    n = len(rowptr) - 1
    bl = F.shape[1]
    x = np.array(x, np.float64, copy=True)
    lib().o_dilu_apply(n, bl, _i(np.ascontiguousarray(rowptr, np.int32)), _i(np.ascontiguousarray(colind, np.int32)),
                       _d(np.ascontiguousarray(F)), _d(x))
    return x


def bjacobi_factor(rowptr, colind, A):
    n = len(rowptr) - 1
    bl = A.shape[1]
    LU = np.zeros((n, bl, bl))
    piv = np.zeros(n * bl, np.int32)
    info = lib().o_bjacobi_factor(n, bl, _i(np.ascontiguousarray(rowptr, np.int32)),
                                  _i(np.ascontiguousarray(colind, np.int32)), _d(np.ascontiguousarray(A)), _d(LU),
                                  _i(piv))
    return LU, piv, info


def bjacobi_apply(LU, piv, x):
    n, bl, _ = LU.shape
    x = np.array(x, np.float64, copy=True)
    lib().o_bjacobi_apply(n, bl, _d(LU), _i(piv), _d(x))
    return x


def inv(A):
    bl = A.shape[0]
    out = np.zeros_like(A)
    info = lib().o_inv(bl, _d(np.ascontiguousarray(A)), _d(out))
    return out, info


def gmres(rowptr, colind, A, b, x0=None, m=20, tol=1e-14, maxit=90000, pc_kind=2, mode=0):
    """Returns (x, tol_out, iterations)."""
    n = len(rowptr) - 1
    bl = A.shape[1]
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colind = np.ascontiguousarray(colind, np.int32)
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n * bl) if x0 is None else np.array(x0, np.float64, copy=True)
    piv = None
    if pc_kind == 2:
        pcd, _ = dilu_factor(rowptr, colind, A)
    elif pc_kind == 1:
        pcd, piv, _ = bjacobi_factor(rowptr, colind, A)
    else:
        pcd = np.zeros(1)
    t = C.c_double(tol)
    it = C.c_int(maxit)
    lib().o_gmres(n, bl, _i(rowptr), _i(colind), _d(A), pc_kind, _d(pcd), _i(piv) if piv is not None else None,
                  _d(b), _d(x), m, C.byref(t), C.byref(it), mode)
    return x, t.value, it.value


def jacobi_solve(rowptr, colind, A, b, x0=None, tol=1e-12, maxit=1000):
    n = len(rowptr) - 1
    bl = A.shape[1]
    LU, piv, _ = bjacobi_factor(rowptr, colind, A)
    x = np.zeros(n * bl) if x0 is None else np.array(x0, np.float64, copy=True)
    t = C.c_double(tol)
    it = C.c_int(maxit)
    lib().o_jacobi_solve(n, bl, _i(np.ascontiguousarray(rowptr, np.int32)), _i(np.ascontiguousarray(colind, np.int32)),
                         _d(np.ascontiguousarray(A)), _d(LU), _i(piv), _d(np.ascontiguousarray(b)), _d(x),
                         C.byref(t), C.byref(it))
    return x, t.value, it.value


def legendre(p, x):
    P = np.zeros(p + 1)
    dP = np.zeros(p + 1)
    lib().o_legendre(p, C.c_double(x), _d(P), _d(dP))
    return P, dP


def basis(p, xi, eta):
    bs = (p + 1) * (p + 2) // 2
    phi, dx, dy = np.zeros(bs), np.zeros(bs), np.zeros(bs)
    lib().o_basis(p, C.c_double(xi), C.c_double(eta), _d(phi), _d(dx), _d(dy))
    return phi, dx, dy


def rule_tables(deg2):
    nt, nl = C.c_int(), C.c_int()
    lib().o_rule_sizes(deg2, C.byref(nt), C.byref(nl))
    tri = np.zeros((nt.value, 3))
    lin = np.zeros((nl.value, 2))
    lib().o_rule_tables(deg2, _d(tri), _d(lin))
    return tri, lin


def set_threads(n: int):
    """Host threads of the oracle's loops (1: the serial order the tests pin; more: for timing)."""
    lib().o_set_threads(int(n))


def eval_bc(spec, x, y, t=0.0):
    """The Dirichlet state of ``spec`` (bc_kind, bc_par) at (x, y, t): ``o_eval_bc``."""
    e = Oracle.eq(spec)
    out = np.zeros(spec.nc)
    lib().o_eval_bc(C.byref(e), C.c_double(x), C.c_double(y), C.c_double(t), _d(out))
    return out


def dirk_tableau(ns, variant=0):
    """variant (3 stages): 0 the reference's tableau, 1 Alexander's L-stable SDIRK3."""
    A, b, c = np.zeros(25), np.zeros(5), np.zeros(5)
    lib().o_dirk_tableau_v(ns, variant, _d(A), _d(b), _d(c))
    return A[:ns * ns].reshape(ns, ns), b[:ns], c[:ns]


def irk_tableau(ns):
    A, b, c = np.zeros(9), np.zeros(3), np.zeros(3)
    lib().o_irk_tableau(ns, _d(A), _d(b), _d(c))
    return A[:ns * ns].reshape(ns, ns), b[:ns], c[:ns]


def _setup_signatures():
    L = lib()
    L.o_jacobian_pattern.restype = C.c_int
    L.o_dilu_factor.restype = C.c_int
    L.o_bjacobi_factor.restype = C.c_int
    L.o_inv.restype = C.c_int
    L.o_polygon_integral_monomial.restype = C.c_double
    L.o_polygon_integral_monomial.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]


def polygon_monomial(orc: Oracle, i, deg2, px, py):
    _setup_signatures()
    return lib().o_polygon_integral_monomial(C.byref(orc.m), i, deg2, px, py)
