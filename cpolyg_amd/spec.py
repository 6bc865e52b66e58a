"""Equation and solver descriptors shared by the HIP library and its callers.

These mirror the knobs the reference hard-codes (``eqns/Advection.h``,
``eqns/Euler.h``, ``src/TimeIntegration.cpp:4-5``, ``src/Jacobian.cpp:75-81``)
as plain values that cross the C ABI.
"""
from __future__ import annotations

import dataclasses
import math

ADVECTION = 0
EULER = 1

BC_ZERO = 0
BC_CONSTANT = 1
BC_ISENTROPIC_VORTEX = 2
BC_REF_VORTEX = 3

PC_NONE = 0
PC_BJACOBI = 1
PC_DILU = 2
PC_INNER_GMRES = 3   # FP32 GMRES(k) with DILU inside flexible FP64 GMRES

GMRES_LEFT_REF = 0
GMRES_FLEXIBLE = 1

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 2
PREC_BF16_D32 = 3    # block DILU: L and DU blocks in BF16, the diagonal inverses in FP32

DIRK3_REFERENCE = 0  # the reference's 3-stage tableau (src/TimeIntegration.cpp:71-77; not A-stable)
DIRK3_L_STABLE = 1   # Alexander's L-stable SDIRK3 (not in the reference)

OP_ASSEMBLED = 0     # GMRES operator: BSR matvec of the assembled Newton matrix (reference)
OP_MATRIX_FREE = 1   # the same matrix applied at the quadrature points (cpg_newton_apply)
# analytic fields of cpg_interp_field / cpg_l2_error_field (parameters in include/cpolyg.h)
FIELD_GAUSSIAN = 0        # cx, cy, a, vx, vy, Lx, Ly
FIELD_VORTEX = 1          # beta, x0, y0, uinf, vinf, gamma, Lx, Ly
FIELD_SINE_DENSITY = 2    # ux, uy, p, gamma, amp


@dataclasses.dataclass
class EquationSpec:
    """``kind``: ADVECTION or EULER.

    Advection velocity is affine, ``bx = b0 + b1 x + b2 y``, ``by = b3 + b4 x + b5 y``:
    ``ConstantAdvection`` has only ``b0, b3``; the reference ``Advection`` default
    (``eqns/Advection.cpp:8-16``) is ``(-1 + 2y, 1 - 2x)``.

    ``periodic_trace``: 0 reproduces the reference's periodic trace
    (``src/BoundaryConditions.cpp:18``), 1 evaluates the neighbour's full series.
    ``jacobian``: 0 reproduces the reference's Euler Jacobian (off-diagonal blocks
    built from the row cell's state, ``src/Equation.cpp:84,109``, and its x flux
    Jacobian entry ``H - (gamma-1)|u|^2``, ``eqns/Euler.cpp:157``); 1 is the exact
    derivative of the residual.  Both give the same converged Newton solution.
    """
    kind: int = ADVECTION
    beta: tuple = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
    gamma: float = 1.4
    bc_kind: int = BC_ZERO
    bc_par: tuple = (0.0,) * 8
    periodic_trace: int = 0
    jacobian: int = 0

    @property
    def nc(self) -> int:
        return 4 if self.kind == EULER else 1

    @staticmethod
    def constant_advection(bx: float, by: float, **kw) -> "EquationSpec":
        return EquationSpec(ADVECTION, (bx, 0.0, 0.0, by, 0.0, 0.0), **kw)

    @staticmethod
    def reference_advection(**kw) -> "EquationSpec":
        return EquationSpec(ADVECTION, (-1.0, 0.0, 2.0, 1.0, -2.0, 0.0), **kw)

    @staticmethod
    def rotating_advection(**kw) -> "EquationSpec":
        """``(-(y - 1/2), x - 1/2)``."""
        return EquationSpec(ADVECTION, (0.5, 0.0, -1.0, -0.5, 1.0, 0.0), **kw)

    @staticmethod
    def euler(gamma: float = 1.4, **kw) -> "EquationSpec":
        return EquationSpec(EULER, (0.0,) * 6, gamma, **kw)


@dataclasses.dataclass
class SolverSpec:
    """Newton-Krylov settings.  Reference values: GMRES(20), tol 1e-14, 90000
    iterations (``src/Jacobian.cpp:75-81``); Newton 5e-13 absolute, 100 iterations
    (``src/TimeIntegration.cpp:4-5``)."""
    gmres_mode: int = GMRES_LEFT_REF
    restart: int = 20
    rtol: float = 1e-14
    maxit: int = 90000
    pc_kind: int = PC_DILU
    newton_tol: float = 5e-13
    newton_maxit: int = 100
    precision: int = PREC_FP64
    inner_iters: int = 20
    newton_guess: int = 0        # 0: k_s = 0 (reference); 1: previous stage's k
    newton_backtrack: int = 0    # 1: halve updates that lead to a non-finite residual
    newton_operator: int = OP_ASSEMBLED
    pc_refresh: int = 0          # 1: with OP_MATRIX_FREE, refresh the preconditioner once per stage
    dirk_tableau: int = DIRK3_REFERENCE   # 3-stage DIRK: reference tableau or Alexander's L-stable SDIRK3
    pc_level_launches: int = 0   # 1: block DILU applies one launch per colour level (else one launch per apply)


def bs_of(deg: int) -> int:
    return (deg + 1) * (deg + 2) // 2


def isentropic_vortex(x, y, t=0.0, beta=5.0, x0=5.0, y0=5.0, uinf=1.0, vinf=1.0, gamma=1.4):
    """Isentropic vortex (Yee, Sandham, Djomehri 1999) advected by the free stream;
    conserved variables ``(rho, rho u, rho v, rho E)``.  Accepts numpy arrays."""
    import numpy as np

    dx = x - x0 - uinf * t
    dy = y - y0 - vinf * t
    r2 = dx * dx + dy * dy
    ex = np.exp(0.5 * (1 - r2))
    u = uinf - beta / (2 * math.pi) * dy * ex
    v = vinf + beta / (2 * math.pi) * dx * ex
    T = 1 - (gamma - 1) * beta ** 2 / (8 * gamma * math.pi ** 2) * ex * ex
    rho = T ** (1 / (gamma - 1))
    P = rho * T
    return np.stack([rho, rho * u, rho * v, P / (gamma - 1) + 0.5 * rho * (u * u + v * v)], axis=-1)
