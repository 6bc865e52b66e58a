"""ctypes binding of ``libcpolyg_hip.so`` (the C ABI in ``include/cpolyg.h``).

Device vectors are passed as torch tensors on ``cuda`` (torch is only the
allocator and stream provider here) or as raw device addresses.  Every call
checks the returned status and raises :class:`CpgError` with the library's
message.  There is no CPU fallback: loading fails loudly when the library has
not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import spec as _spec

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpolyg_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "cpolyg.h")

_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class CpgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"cpolyg status {status}: {msg}")
        self.status = status


class EqParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("beta", C.c_double * 6), ("gamma", C.c_double), ("bc_kind", C.c_int),
                ("bc_par", C.c_double * 8), ("periodic_trace", C.c_int), ("jacobian", C.c_int)]


class SolverParams(C.Structure):
    _fields_ = [("gmres_mode", C.c_int), ("restart", C.c_int), ("rtol", C.c_double), ("maxit", C.c_int),
                ("pc_kind", C.c_int), ("precision", C.c_int), ("newton_tol", C.c_double),
                ("newton_maxit", C.c_int), ("inner_iters", C.c_int), ("newton_guess", C.c_int),
                ("newton_backtrack", C.c_int), ("newton_operator", C.c_int), ("pc_refresh", C.c_int),
                ("dirk_tableau", C.c_int), ("pc_level_launches", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int), ("gmres_iters", C.c_int), ("linear_solves", C.c_int),
                ("status", C.c_int), ("last_resid", C.c_double), ("ms_residual", C.c_double),
                ("ms_jacobian", C.c_double), ("ms_factor", C.c_double), ("ms_gmres", C.c_double),
                ("ms_total", C.c_double), ("backtracks", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    """Load the library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _lib.cpg_last_error.restype = C.c_char_p
        _lib.cpg_status_string.restype = C.c_char_p
    return _lib


def _chk(st: int):
    if st != 0:
        raise CpgError(st, lib().cpg_last_error().decode())


def declared_symbols() -> list[str]:
    """Function names declared in include/cpolyg.h."""
    import re

    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cpg_[a-z0-9_]+)\s*\(", txt)))


def ptr(t) -> int:
    """Device address of a torch tensor (or an int address)."""
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _ia(a):
    a = np.ascontiguousarray(a, np.int32)
    return a, a.ctypes.data_as(_ip)


def _da(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(_dp)


class Bsr:
    """Device BSR matrix (``BlockMatrix``)."""

    def __init__(self, rowptr, colind, bl, handle=None):
        self.rowptr = np.ascontiguousarray(rowptr, np.int32)
        self.colind = np.ascontiguousarray(colind, np.int32)
        self.n = len(self.rowptr) - 1
        self.bl = int(bl)
        self.nnzb = len(self.colind)
        if handle is None:
            h = _vp()
            _chk(lib().cpg_bsr_create(self.n, self.bl, self.nnzb, self.rowptr.ctypes.data_as(_ip),
                                      self.colind.ctypes.data_as(_ip), C.byref(h)))
            self.h = h
        else:
            self.h = handle

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpg_bsr_destroy(self.h)
            self.h = None

    def upload(self, blocks, stream=None):
        b, bp = _da(blocks)
        assert b.size == self.nnzb * self.bl * self.bl
        _chk(lib().cpg_bsr_upload(self.h, bp, _stream(stream)))
        _chk(lib().cpg_stream_sync(_stream(stream)))

    def download(self):
        out = np.zeros((self.nnzb, self.bl, self.bl))
        _chk(lib().cpg_bsr_download(self.h, out.ctypes.data_as(_dp), None))
        return out

    def blocks_colmajor(self):
        """Blocks as (nnzb, bl, bl) with [s, r, c] = entry (r, c)."""
        return self.download().transpose(0, 2, 1)

    def scale(self, s, stream=None):
        _chk(lib().cpg_bsr_scale(self.h, C.c_double(s), _stream(stream)))

    def matvec(self, x, y, stream=None):
        _chk(lib().cpg_bsr_matvec(self.h, _vp(ptr(x)), _vp(ptr(y)), _stream(stream)))

    def autotune(self, precision=_spec.PREC_FP64, reps=5, stream=None):
        """Time this matrix's row stream stored in `precision` for 1, 2 and 4 wavefronts per
        block row and keep the fastest shape for (precision, bl); returns (best, [ms1, ms2, ms4])."""
        best = C.c_int()
        ms = (C.c_float * 3)()
        _chk(lib().cpg_stream_autotune(self.h, precision, reps, C.byref(best), ms, _stream(stream)))
        return best.value, list(ms)


def set_stream_wpr(precision, bl, wpr):
    """Wavefronts per block row of the row-stream kernels for (storage precision, block size)."""
    _chk(lib().cpg_stream_set_wpr(precision, bl, wpr))


def get_stream_wpr(precision, bl):
    v = C.c_int()
    _chk(lib().cpg_stream_get_wpr(precision, bl, C.byref(v)))
    return v.value


class Prec:
    """Device preconditioner (``BlockJacobi`` / ``BlockILU0``)."""

    def __init__(self, A: Bsr, kind=_spec.PC_DILU, precision=_spec.PREC_FP64):
        h = _vp()
        _chk(lib().cpg_prec_create(A.h, kind, precision, C.byref(h)))
        self.h = h
        self.A = A
        self.kind = kind
        self.precision = precision

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpg_prec_destroy(self.h)
            self.h = None

    def set_inner(self, iterations: int):
        _chk(lib().cpg_prec_set_inner(self.h, int(iterations)))

    def factor(self, A: Bsr | None = None, stream=None):
        _chk(lib().cpg_prec_factor(self.h, (A or self.A).h, _stream(stream)))

    def apply(self, x, stream=None):
        _chk(lib().cpg_prec_apply(self.h, _vp(ptr(x)), _stream(stream)))

    def factor_coupled(self, E, stage_jacobians, coef, stream=None):
        """Block DILU of the Radau IIA stage-coupled Newton matrix (``cpg_prec_factor_coupled``):
        block (s, t) of slot q is coef[s, t] J_s[q] (+ E's mass on the diagonal when s == t)."""
        J0, J1, J2 = stage_jacobians
        c = np.ascontiguousarray(np.asarray(coef, np.float64).reshape(9))
        _chk(lib().cpg_prec_factor_coupled(self.h, E.h, J0.h, J1.h, J2.h, c.ctypes.data_as(_dp), _stream(stream)))

    def download(self):
        nblk = self.A.nnzb if self.kind == _spec.PC_DILU else self.A.n
        out = np.zeros((nblk, self.A.bl, self.A.bl))
        _chk(lib().cpg_prec_download(self.h, out.ctypes.data_as(_dp)))
        return out

    def levels(self):
        n = C.c_int()
        _chk(lib().cpg_prec_levels(self.h, C.byref(n)))
        return n.value

    def apply_to(self, x, out, stream=None):
        """out = M^-1 x (distinct vectors)."""
        _chk(lib().cpg_prec_apply_to(self.h, _vp(ptr(x)), _vp(ptr(out)), _stream(stream)))

    def set_flow(self, enable=-1, spin_limit=-1):
        """Block DILU apply in one launch (1) or one launch per level (0); spin_limit bounds each
        row's wait for the rows it reads (0: give up at once).  -1 leaves a setting unchanged."""
        _chk(lib().cpg_prec_set_flow(self.h, int(enable), int(spin_limit)))

    def flow_in_use(self):
        v = C.c_int()
        _chk(lib().cpg_prec_flow_state(self.h, C.byref(v)))
        return bool(v.value)


class Krylov:
    def __init__(self, n, bl, restart, mode):
        h = _vp()
        _chk(lib().cpg_krylov_create(n, bl, restart, mode, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpg_krylov_destroy(self.h)
            self.h = None

    def gmres(self, A: Bsr, P: Prec | None, b, x, tol, maxit, stream=None):
        t = C.c_double(tol)
        it = C.c_int(maxit)
        _chk(lib().cpg_gmres(self.h, A.h, P.h if P is not None else None, _vp(ptr(b)), _vp(ptr(x)), C.byref(t),
                             C.byref(it), _stream(stream)))
        return t.value, it.value

    def jacobi(self, A: Bsr, P: Prec | None, b, x, tol, maxit, stream=None):
        """BlockMatrix::jacobi (src/BlockMatrix.cpp:51-77); returns (||A x - b||_2, iterations)."""
        t = C.c_double(tol)
        it = C.c_int(maxit)
        _chk(lib().cpg_jacobi(self.h, A.h, P.h if P is not None else None, _vp(ptr(b)), _vp(ptr(x)), C.byref(t),
                              C.byref(it), _stream(stream)))
        return t.value, it.value


def eq_params(s: _spec.EquationSpec) -> EqParams:
    e = EqParams()
    e.kind = s.kind
    for j in range(6):
        e.beta[j] = s.beta[j]
    e.gamma = s.gamma
    e.bc_kind = s.bc_kind
    for j in range(8):
        e.bc_par[j] = s.bc_par[j] if j < len(s.bc_par) else 0.0
    e.periodic_trace = s.periodic_trace
    e.jacobian = s.jacobian
    return e


def solver_params(s: _spec.SolverSpec) -> SolverParams:
    return SolverParams(s.gmres_mode, s.restart, s.rtol, s.maxit, s.pc_kind, s.precision, s.newton_tol,
                        s.newton_maxit, s.inner_iters, s.newton_guess, s.newton_backtrack, s.newton_operator,
                        s.pc_refresh, s.dirk_tableau, s.pc_level_launches)


class Mesh:
    def __init__(self, m):
        self.mesh = m
        self._keep = [np.ascontiguousarray(a, dt) for a, dt in (
            (m.v, np.float64), (m.pptr, np.int32), (m.pv, np.int32), (m.p2p, np.int32), (m.bc, np.int32),
            (m.tptr, np.int32), (m.txy, np.float64), (m.bb, np.float64))]
        v, pptr, pv, p2p, bc, tptr, txy, bb = self._keep
        h = _vp()
        _chk(lib().cpg_mesh_create(m.n_cells, len(v), v.ctypes.data_as(_dp), pptr.ctypes.data_as(_ip),
                                   pv.ctypes.data_as(_ip), p2p.ctypes.data_as(_ip), len(bc), bc.ctypes.data_as(_ip),
                                   tptr.ctypes.data_as(_ip), txy.ctypes.data_as(_dp), bb.ctypes.data_as(_dp),
                                   C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpg_mesh_destroy(self.h)
            self.h = None


class Equation:
    """Device ``Equation`` (residual, Jacobian, mass matrix) for one mesh and degree."""

    def __init__(self, mesh: Mesh, s: _spec.EquationSpec, deg: int):
        self.mesh = mesh
        self.spec = s
        self.deg = deg
        self.params = eq_params(s)
        h = _vp()
        _chk(lib().cpg_equation_create(mesh.h, C.byref(self.params), deg, C.byref(h)))
        self.h = h
        n, bl, nnzb = C.c_int(), C.c_int(), C.c_int()
        _chk(lib().cpg_equation_info(self.h, C.byref(n), C.byref(bl), C.byref(nnzb)))
        self.n, self.bl, self.nnzb = n.value, bl.value, nnzb.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpg_equation_destroy(self.h)
            self.h = None

    def pattern(self):
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(self.nnzb, np.int32)
        _chk(lib().cpg_jacobian_pattern(self.h, rp.ctypes.data_as(_ip), ci.ctypes.data_as(_ip)))
        return rp, ci

    def jacobian_matrix(self) -> Bsr:
        h = _vp()
        _chk(lib().cpg_jacobian_create(self.h, C.byref(h)))
        rp, ci = self.pattern()
        return Bsr(rp, ci, self.bl, handle=h)

    def residual(self, u, t, b, stream=None):
        _chk(lib().cpg_assemble_residual(self.h, _vp(ptr(u)), C.c_double(t), _vp(ptr(b)), _stream(stream)))

    def jacobian(self, u, t, J: Bsr, scale=1.0, add_mass=False, stream=None):
        _chk(lib().cpg_assemble_jacobian(self.h, _vp(ptr(u)), C.c_double(t), C.c_double(scale), int(add_mass), J.h,
                                         _stream(stream)))

    def newton_apply(self, u, t, scale, v, y, stream=None):
        """y = M v + scale J(u, t) v without assembling J (``cpg_newton_apply``)."""
        _chk(lib().cpg_newton_apply(self.h, _vp(ptr(u)), C.c_double(t), C.c_double(scale), _vp(ptr(v)),
                                    _vp(ptr(y)), _stream(stream)))

    def newton_prepare(self, u, t, stream=None):
        """The (u, t)-dependent point data of the Newton operator (``cpg_newton_prepare``)."""
        _chk(lib().cpg_newton_prepare(self.h, _vp(ptr(u)), C.c_double(t), _stream(stream)))

    def newton_apply_prepared(self, scale, v, y, stream=None):
        """y = M v + scale J(u, t) v for the (u, t) of the last ``newton_prepare``."""
        _chk(lib().cpg_newton_apply_prepared(self.h, C.c_double(scale), _vp(ptr(v)), _vp(ptr(y)), _stream(stream)))

    def mass_matvec(self, x, y, stream=None):
        _chk(lib().cpg_mass_matvec(self.h, _vp(ptr(x)), _vp(ptr(y)), _stream(stream)))

    def mass(self):
        out = np.zeros((self.n, self.bl, self.bl))
        _chk(lib().cpg_mass_download(self.h, out.ctypes.data_as(_dp)))
        return out

    def l2norm(self, u, stream=None):
        out = np.zeros(self.spec.nc)
        _chk(lib().cpg_l2norm(self.h, _vp(ptr(u)), out.ctypes.data_as(_dp), _stream(stream)))
        return out

    def l2norm_noalloc(self, u, stream=None):
        """``MeshFn::L2Norm`` through ``cpg_l2norm_noalloc`` (no device allocation)."""
        out = np.zeros(self.spec.nc)
        _chk(lib().cpg_l2norm_noalloc(self.h, _vp(ptr(u)), out.ctypes.data_as(_dp), _stream(stream)))
        return out

    @staticmethod
    def _field_par(par):
        p = np.zeros(8)
        p[:len(par)] = par
        return p

    def interp_field(self, field, par, t, u, stream=None):
        """u = MeshFn::interp of an analytic field (``cpg_interp_field``; spec.FIELD_*)."""
        p = self._field_par(par)
        _chk(lib().cpg_interp_field(self.h, int(field), p.ctypes.data_as(_dp), C.c_double(t), _vp(ptr(u)),
                                    _stream(stream)))

    def interp_points(self, xy, stream=None):
        """xy (device, n (p+1)^2 x 2) = the fitting points of MeshFn::interp."""
        _chk(lib().cpg_interp_points(self.h, _vp(ptr(xy)), _stream(stream)))

    def interp_values(self, vals, u, stream=None):
        """u = the least-squares fit to vals (device, n (p+1)^2 x nc) at the fitting points."""
        _chk(lib().cpg_interp_values(self.h, _vp(ptr(vals)), _vp(ptr(u)), _stream(stream)))

    def l2_error_field(self, u, field, par, t, stream=None):
        """MeshFn::L2Error of u against an analytic field at time t, per component."""
        p = self._field_par(par)
        out = np.zeros(self.spec.nc)
        _chk(lib().cpg_l2_error_field(self.h, _vp(ptr(u)), int(field), p.ctypes.data_as(_dp), C.c_double(t),
                                      out.ctypes.data_as(_dp), _stream(stream)))
        return out


class Stepper:
    """Device ``TimeStepper`` (DIRK / backward Euler / IRK / RK4 / RK2 / forward Euler)."""

    def __init__(self, eq: Equation, method: int, nstages: int, solver: _spec.SolverSpec):
        self.eq = eq
        self.sp = solver_params(solver)
        h = _vp()
        _chk(lib().cpg_stepper_create(eq.h, method, nstages, C.byref(self.sp), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpg_stepper_destroy(self.h)
            self.h = None

    def advance(self, u, dt, t, unp1, stream=None) -> StepStats:
        st = StepStats()
        _chk(lib().cpg_stepper_advance(self.h, _vp(ptr(u)), C.c_double(dt), C.c_double(t), _vp(ptr(unp1)),
                                       C.byref(st), _stream(stream)))
        return st

    def enable_kernel_timing(self, on=True):
        _chk(lib().cpg_stepper_enable_kernel_timing(self.h, int(on)))

    def kernel_timing(self):
        nm, na = C.c_int(), C.c_int()
        mm, ma = C.c_double(), C.c_double()
        _chk(lib().cpg_stepper_kernel_timing(self.h, C.byref(nm), C.byref(mm), C.byref(na), C.byref(ma)))
        return nm.value, mm.value, na.value, ma.value


class Batch:
    """Batched implicit DIRK steps over many independent instances (``cpg_batch``).

    ``classes`` is a list of ``(Equation, cell_ptr)``: each equation is built on a union mesh
    (:func:`cpolyg_amd.mesh.union_mesh`) whose parts are the instances of that class."""

    EAGER, GRAPH, PERSISTENT = 0, 1, 2

    def __init__(self, classes, nstages: int, solver: _spec.SolverSpec, use_graph: bool = True, mode=None):
        """``mode``: ``Batch.EAGER``, ``Batch.GRAPH`` (restart cycles replayed from hipGraphs) or
        ``Batch.PERSISTENT`` (one cooperative launch per batched solve); by default ``use_graph``
        picks between the first two."""
        if mode is None:
            mode = self.GRAPH if use_graph else self.EAGER
        self.mode = int(mode)
        self.classes = classes
        self.sp = solver_params(solver)
        n = len(classes)
        eqs = (C.c_void_p * n)(*[e.h.value if isinstance(e.h, C.c_void_p) else e.h for e, _ in classes])
        self._cptr = [np.ascontiguousarray(cp, np.int32) for _, cp in classes]
        ninst = (C.c_int * n)(*[len(cp) - 1 for cp in self._cptr])
        cps = (_ip * n)(*[cp.ctypes.data_as(_ip) for cp in self._cptr])
        h = _vp()
        _chk(lib().cpg_batch_create(n, eqs, ninst, cps, int(nstages), C.byref(self.sp), self.mode, C.byref(h)))
        self.h = h
        self.ninst = sum(len(cp) - 1 for cp in self._cptr)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cpg_batch_destroy(self.h)
            self.h = None

    def info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _chk(lib().cpg_batch_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"instances": a.value, "chunks": b.value, "graph_captured": bool(c.value)}

    def set_teams(self, wg_per_cu: int = 0, max_team: int = 0, spin_limit: int = 0):
        """Persistent mode: workgroups per CU (0: autotuned over the first solves), the most
        workgroups per instance (0: as many as its rows can use) and the team barriers' wait
        limit in sleep rounds (0: the default)."""
        _chk(lib().cpg_batch_set_teams(self.h, int(wg_per_cu), int(max_team), C.c_uint(int(spin_limit))))

    def teams(self):
        """Persistent mode, after a solve: {"grid", "wg_per_cu", "team_wg", "team_ms" (per instance)}."""
        g, w = C.c_int(), C.c_int()
        tw = np.zeros(self.ninst, np.int32)
        tm = np.zeros(self.ninst, np.float64)
        _chk(lib().cpg_batch_teams(self.h, C.byref(g), C.byref(w), tw.ctypes.data_as(_ip), tm.ctypes.data_as(_dp)))
        return {"grid": g.value, "wg_per_cu": w.value, "team_wg": tw, "team_ms": tm}

    def advance(self, us, dts, t, outs, stream=None):
        """One step of every instance; ``us``/``outs`` are per-class device vectors, ``dts``
        one dt per instance.  Returns (StepStats, per-instance Arnoldi steps)."""
        n = len(self.classes)
        uarr = (C.c_void_p * n)(*[ptr(u) for u in us])
        oarr = (C.c_void_p * n)(*[ptr(o) for o in outs])
        dt = np.ascontiguousarray(dts, np.float64)
        assert len(dt) == self.ninst
        its = np.zeros(self.ninst, np.int32)
        st = StepStats()
        _chk(lib().cpg_batch_advance(self.h, uarr, dt.ctypes.data_as(_dp), C.c_double(t), oarr, C.byref(st),
                                     its.ctypes.data_as(_ip), _stream(stream)))
        return st, its
