"""GPU DIRK / backward-Euler / IRK / explicit (RK4, RK2, forward Euler) steps against the oracle's steppers."""
import numpy as np
import pytest

import oracle as O
from cpolyg_amd import hip, setups
from cpolyg_amd.spec import (DIRK3_L_STABLE, EquationSpec, SolverSpec, GMRES_FLEXIBLE, GMRES_LEFT_REF, OP_ASSEMBLED,
                             OP_MATRIX_FREE, PC_DILU, PC_BJACOBI, PREC_BF16, PREC_BF16_D32, PREC_FP32, PREC_FP64, bs_of)
from problems import rel, unit_periodic_mesh, vortex_mesh, vortex_state

pytestmark = pytest.mark.gpu


def _dev(cuda, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(cuda)


@pytest.mark.parametrize("nstages", [3, 1])
@pytest.mark.parametrize("jacobian", [0, 1])
def test_dirk_euler_matches_oracle(cuda, nstages, jacobian):
    m = vortex_mesh(8)
    deg = 2
    spec = EquationSpec.euler(jacobian=jacobian)
    u = vortex_state(m, deg)
    solver = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=20, rtol=1e-10, maxit=4000, pc_kind=PC_DILU,
                        newton_tol=5e-12, newton_maxit=40)
    dt = 0.5
    uo, st = O.Oracle(m).dirk_step(spec, deg, nstages, u, dt, 0.0, solver)
    E = hip.Equation(hip.Mesh(m), spec, deg)
    S = hip.Stepper(E, 0, nstages, solver)
    ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    gs = S.advance(ud, dt, 0.0, outd)
    assert gs.newton_iters == st.newton_iters
    assert abs(gs.gmres_iters - st.gmres_iters) <= gs.linear_solves
    # converged Newton solutions agree to the solver tolerances
    assert rel(outd.cpu().numpy(), uo) < 1e-9


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32, PREC_BF16])
@pytest.mark.parametrize("op", [OP_ASSEMBLED, OP_MATRIX_FREE])
def test_dirk_flexible_mixed_precision(cuda, prec, op):
    m = vortex_mesh(8)
    deg = 3
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    solver = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=30, rtol=1e-10, maxit=4000, pc_kind=PC_DILU,
                        newton_tol=1e-11, newton_maxit=40, precision=prec, newton_operator=op)
    dt = 0.5
    ref = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=30, rtol=1e-10, maxit=4000, pc_kind=PC_DILU,
                     newton_tol=1e-11, newton_maxit=40)
    uo, st = O.Oracle(m).dirk_step(spec, deg, 3, u, dt, 0.0, ref)
    E = hip.Equation(hip.Mesh(m), spec, deg)
    S = hip.Stepper(E, 0, 3, solver)
    ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    gs = S.advance(ud, dt, 0.0, outd)
    assert gs.last_resid < 1e-11
    assert rel(outd.cpu().numpy(), uo) < 1e-9


def test_dirk3_l_stable_25_steps_mixed_bf16(cuda):
    """25 steps at CFL 20 with the benchmark's solver (Alexander's L-stable SDIRK3, BF16 DILU factors
    in FP64 flexible GMRES(50) to 1e-10, matrix-free Newton operator, preconditioner refreshed once
    per step, each stage's Newton started from the previous stage's increment, backtracking on)
    on a 64-cell vortex at p = 3, every step against the oracle's trajectory (FP64 DILU, Newton
    from zero, the same tableau) started from the same state."""
    m = vortex_mesh(8)
    deg = 3
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    dt = setups.dg_time_step(20.0, 10.0 / 8, deg, setups.euler_max_speed(u, deg))
    ref = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=50, rtol=1e-10, maxit=4000, pc_kind=PC_DILU,
                     newton_tol=5e-13, newton_maxit=40, dirk_tableau=DIRK3_L_STABLE)
    solver = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=50, rtol=1e-10, maxit=20000, pc_kind=PC_DILU,
                        newton_tol=5e-13, newton_maxit=100, precision=PREC_BF16, newton_guess=1, newton_backtrack=1,
                        newton_operator=OP_MATRIX_FREE, pc_refresh=2, dirk_tableau=DIRK3_L_STABLE)
    orc = O.Oracle(m)
    S = hip.Stepper(hip.Equation(hip.Mesh(m), spec, deg), 0, 3, solver)
    uo, ud, t = u.copy(), _dev(cuda, u), 0.0
    worst = 0.0
    for k in range(25):
        uo, _ = orc.dirk_step(spec, deg, 3, uo, dt, t, ref)
        out = _dev(cuda, np.zeros_like(u))
        gs = S.advance(ud, dt, t, out)
        assert gs.last_resid < 5e-13
        ud, t = out, t + dt
        err = rel(ud.cpu().numpy(), uo)
        worst = max(worst, err)
        assert err < 1e-9, (k, err)
    print(f"25 L-stable SDIRK3 steps at CFL 20 (dt {dt:.4f}): max rel. difference from the oracle {worst:.2e}")


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32, PREC_BF16, PREC_BF16_D32])
def test_dirk3_euler_p4_matches_oracle(cuda, prec):
    """Euler at p = 4 (60 x 60 blocks): one DIRK3 step (L-stable tableau) with FP64, FP32 and BF16
    DILU factors in flexible GMRES, matrix-free Newton operator, against the oracle's step."""
    m = vortex_mesh(6)
    deg = 4
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    dt = setups.dg_time_step(20.0, 10.0 / 6, deg, setups.euler_max_speed(u, deg))
    kw = dict(gmres_mode=GMRES_FLEXIBLE, restart=50, rtol=1e-10, maxit=4000, pc_kind=PC_DILU, newton_tol=1e-12,
              newton_maxit=40, dirk_tableau=DIRK3_L_STABLE)
    uo, st = O.Oracle(m).dirk_step(spec, deg, 3, u, dt, 0.0, SolverSpec(**kw))
    S = hip.Stepper(hip.Equation(hip.Mesh(m), spec, deg), 0, 3,
                    SolverSpec(**kw, precision=prec, newton_operator=OP_MATRIX_FREE, pc_refresh=2))
    ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    gs = S.advance(ud, dt, 0.0, outd)
    assert gs.last_resid < 1e-12
    assert rel(outd.cpu().numpy(), uo) < 1e-9


@pytest.mark.parametrize("variant", [0, DIRK3_L_STABLE])
def test_dirk5_and_dirk3_variants_match_oracle(cuda, variant):
    """The 5-stage DIRK (Cooper and Sayfy, src/TimeIntegration.cpp:78-89) and the selected 3-stage
    tableau, reference GMRES mode, against the oracle's steps."""
    m = vortex_mesh(6)
    deg = 2
    spec = EquationSpec.euler(jacobian=1)
    u = vortex_state(m, deg)
    solver = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=20, rtol=1e-11, maxit=4000, pc_kind=PC_DILU,
                        newton_tol=1e-12, newton_maxit=40, dirk_tableau=variant)
    orc = O.Oracle(m)
    E = hip.Equation(hip.Mesh(m), spec, deg)
    for ns in (5, 3):
        uo, st = orc.dirk_step(spec, deg, ns, u, 0.4, 0.0, solver)
        S = hip.Stepper(E, 0, ns, solver)
        ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
        gs = S.advance(ud, 0.4, 0.0, outd)
        assert gs.newton_iters == st.newton_iters
        assert rel(outd.cpu().numpy(), uo) < 1e-9


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_BF16])
def test_dirk_stage_lagged_preconditioner(cuda, prec):
    """Matrix-free operator with the preconditioner refreshed once per stage: the Newton
    iteration is unchanged, so the step matches the oracle's (which refreshes every iteration)."""
    m = vortex_mesh(8)
    deg = 3
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    kw = dict(gmres_mode=GMRES_FLEXIBLE, restart=30, rtol=1e-10, maxit=4000, pc_kind=PC_DILU, newton_tol=1e-11,
              newton_maxit=40)
    uo, st = O.Oracle(m).dirk_step(spec, deg, 3, u, 0.5, 0.0, SolverSpec(**kw))
    S = hip.Stepper(hip.Equation(hip.Mesh(m), spec, deg), 0, 3,
                    SolverSpec(**kw, precision=prec, newton_operator=OP_MATRIX_FREE, pc_refresh=1))
    ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    gs = S.advance(ud, 0.5, 0.0, outd)
    assert gs.newton_iters == st.newton_iters
    assert gs.last_resid < 1e-11
    assert rel(outd.cpu().numpy(), uo) < 1e-9


def test_backward_euler_advection_bjacobi(cuda):
    m = unit_periodic_mesh(8)
    deg = 2
    spec = EquationSpec.constant_advection(1.0, 1.0)
    u = setups.interp(m, deg, setups.gaussian(a=50.0), 1)
    solver = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=20, rtol=1e-12, maxit=4000, pc_kind=PC_BJACOBI,
                        newton_tol=1e-13, newton_maxit=10)
    uo, st = O.Oracle(m).dirk_step(spec, deg, 1, u, 0.05, 0.0, solver)
    E = hip.Equation(hip.Mesh(m), spec, deg)
    S = hip.Stepper(E, 0, 1, solver)
    ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    gs = S.advance(ud, 0.05, 0.0, outd)
    assert gs.newton_iters == st.newton_iters
    assert rel(outd.cpu().numpy(), uo) < 1e-10


@pytest.mark.parametrize("method,name", [(2, "rk4_step"), (4, "rk2_step"), (5, "forward_euler_step")])
def test_rk4_advection(cuda, method, name):
    """The explicit steps (RK4, the reference's RK2, forward Euler) against the oracle's."""
    m = unit_periodic_mesh(8)
    deg = 2
    spec = EquationSpec.constant_advection(1.0, 1.0)
    u = setups.interp(m, deg, setups.gaussian(), 1)
    dt = 0.1 * m.h_min()
    uo = getattr(O.Oracle(m), name)(spec, deg, u, dt, 0.0)
    E = hip.Equation(hip.Mesh(m), spec, deg)
    S = hip.Stepper(E, method, 4, SolverSpec())
    ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    S.advance(ud, dt, 0.0, outd)
    assert rel(outd.cpu().numpy(), uo) < 1e-12


@pytest.mark.parametrize("ns", [2, 3])
def test_irk_radau_matches_oracle(cuda, ns):
    """IRK::advance: stage-major coupled system, DILU, GMRES warm-started from the previous
    Newton correction (src/TimeIntegration.cpp:188-281)."""
    m = vortex_mesh(6)
    deg = 1
    spec = EquationSpec.euler(jacobian=1)
    u = vortex_state(m, deg)
    solver = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=20, rtol=1e-11, maxit=4000, pc_kind=PC_DILU,
                        newton_tol=1e-12, newton_maxit=30)
    dt = 0.4
    uo, st = O.Oracle(m).irk_step(spec, deg, ns, u, dt, 0.0, solver)
    E = hip.Equation(hip.Mesh(m), spec, deg)
    S = hip.Stepper(E, 1, ns, solver)
    ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    gs = S.advance(ud, dt, 0.0, outd)
    assert abs(gs.newton_iters - st.newton_iters) <= 1
    assert rel(outd.cpu().numpy(), uo) < 1e-9


@pytest.mark.parametrize("ns,prec,refresh,pc", [(3, PREC_FP64, 0, PC_BJACOBI), (3, PREC_BF16, 2, PC_BJACOBI),
                                                (2, PREC_FP32, 0, PC_BJACOBI), (3, PREC_FP64, 0, PC_DILU),
                                                (3, PREC_BF16, 2, PC_DILU), (3, PREC_FP32, 1, PC_DILU)])
def test_irk_matrix_free_coupled(cuda, ns, prec, refresh, pc):
    """Radau IIA with the coupled Newton system applied matrix-free and preconditioned by block
    Jacobi on the cell-interleaved (ns bl) x (ns bl) diagonal blocks (120 x 120 for 3 stages at
    p = 3) or by block DILU of the cell-interleaved coupled matrix: the converged step equals the
    oracle's (stage-major assembled system, DILU, left GMRES)."""
    m = vortex_mesh(6)
    deg = 3
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    ref = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=20, rtol=1e-11, maxit=4000, pc_kind=PC_DILU,
                     newton_tol=1e-12, newton_maxit=30)
    dt = 0.4
    uo, st = O.Oracle(m).irk_step(spec, deg, ns, u, dt, 0.0, ref)
    solver = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=40, rtol=1e-11, maxit=20000, pc_kind=pc,
                        newton_tol=1e-12, newton_maxit=30, precision=prec, newton_operator=OP_MATRIX_FREE,
                        pc_refresh=refresh)
    S = hip.Stepper(hip.Equation(hip.Mesh(m), spec, deg), 1, ns, solver)
    ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    gs = S.advance(ud, dt, 0.0, outd)
    assert gs.last_resid < 1e-12
    assert rel(outd.cpu().numpy(), uo) < 1e-9


STEP_IRK_W = 3


@pytest.mark.parametrize("ns,prec,pc", [(3, PREC_FP64, PC_DILU), (3, PREC_BF16, PC_DILU), (2, PREC_FP64, PC_BJACOBI),
                                        (3, PREC_FP32, PC_BJACOBI)])
def test_irk_w_form_matches_oracle(cuda, ns, prec, pc):
    """Radau IIA in the form of IRK::newAdvance (unknowns w = A k, stage-block-diagonal
    preconditioner of A^-1[s][s] M - dt J_s, coupled matrix applied matrix-free): the converged
    step solves the same stage equations as IRK::advance, so it equals the oracle's step."""
    m = vortex_mesh(6)
    deg = 3
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    ref = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=20, rtol=1e-11, maxit=4000, pc_kind=PC_DILU,
                     newton_tol=1e-12, newton_maxit=30)
    dt = 0.4
    uo, st = O.Oracle(m).irk_step(spec, deg, ns, u, dt, 0.0, ref)
    solver = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=40, rtol=1e-11, maxit=20000, pc_kind=pc,
                        newton_tol=1e-12, newton_maxit=30, precision=prec, newton_operator=OP_MATRIX_FREE,
                        pc_refresh=2)
    S = hip.Stepper(hip.Equation(hip.Mesh(m), spec, deg), STEP_IRK_W, ns, solver)
    ud, outd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    gs = S.advance(ud, dt, 0.0, outd)
    assert gs.last_resid < 1e-12
    assert rel(outd.cpu().numpy(), uo) < 1e-9


# ---- failure paths: each ends with a status, every kernel exits normally -------------------
@pytest.mark.parametrize("method,ns,pc", [(0, 3, PC_DILU), (STEP_IRK_W, 3, PC_DILU), (1, 3, PC_BJACOBI)])
def test_nonfinite_state_reports_status(cuda, method, ns, pc):
    """A NaN in the state makes the first Newton residual non-finite: the step returns
    CPG_ERR_NONFINITE before any linear solve."""
    m = vortex_mesh(4)
    deg = 1
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    u[7] = np.nan
    solver = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=20, rtol=1e-10, maxit=200, pc_kind=pc,
                        newton_tol=1e-12, newton_maxit=5, newton_operator=OP_MATRIX_FREE)
    S = hip.Stepper(hip.Equation(hip.Mesh(m), spec, deg), method, ns, solver)
    with pytest.raises(hip.CpgError) as e:
        S.advance(_dev(cuda, u), 0.1, 0.0, _dev(cuda, np.zeros_like(u)))
    assert e.value.status == 5


def test_nonphysical_result_reports_its_own_step(cuda):
    """A forward Euler step with a huge dt from a finite state: the state it produces has
    negative pressures, and the step itself (not the next one) returns CPG_ERR_NONFINITE."""
    m = vortex_mesh(4)
    deg = 1
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    assert np.isfinite(u).all()
    S = hip.Stepper(hip.Equation(hip.Mesh(m), spec, deg), 5, 1, SolverSpec())
    with pytest.raises(hip.CpgError) as e:
        S.advance(_dev(cuda, u), 1e6, 0.0, _dev(cuda, np.zeros_like(u)))
    assert e.value.status == 5
    assert "non-physical state" in str(e.value)


def test_newton_iteration_limit_reports_status(cuda):
    """One Newton iteration against a 1e-30 tolerance: CPG_ERR_NOT_CONVERGED, with the
    iterations done and the last residual in the message."""
    m = vortex_mesh(4)
    deg = 2
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    solver = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=20, rtol=1e-10, maxit=400, pc_kind=PC_DILU,
                        newton_tol=1e-30, newton_maxit=1, newton_operator=OP_MATRIX_FREE)
    S = hip.Stepper(hip.Equation(hip.Mesh(m), spec, deg), 0, 3, solver)
    with pytest.raises(hip.CpgError) as e:
        S.advance(_dev(cuda, u), 0.2, 0.0, _dev(cuda, np.zeros_like(u)))
    assert e.value.status == 6
    assert "Newton" in str(e.value)


@pytest.mark.parametrize("method,ns,pc,prec", [(0, 3, PC_DILU, PREC_BF16), (1, 3, PC_BJACOBI, PREC_BF16),
                                               (1, 3, PC_DILU, PREC_FP32), (STEP_IRK_W, 2, PC_DILU, PREC_FP64)])
def test_factor_kept_across_steps(cuda, method, ns, pc, prec):
    """pc_refresh 3 keeps the preconditioner's factor across steps (renewed when GMRES slows
    down or dt changes): three steps, the third with another dt, match three oracle steps."""
    m = vortex_mesh(6)
    deg = 3
    spec = EquationSpec.euler(periodic_trace=1, jacobian=1)
    u = vortex_state(m, deg)
    ref = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=20, rtol=1e-11, maxit=4000, pc_kind=PC_DILU,
                     newton_tol=1e-12, newton_maxit=30)
    orc = O.Oracle(m)
    solver = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=40, rtol=1e-11, maxit=20000, pc_kind=pc,
                        newton_tol=1e-12, newton_maxit=30, precision=prec, newton_operator=OP_MATRIX_FREE,
                        pc_refresh=3, newton_guess=1 if method == 0 else 0)
    S = hip.Stepper(hip.Equation(hip.Mesh(m), spec, deg), method, ns, solver)
    uo, ud, t = u.copy(), _dev(cuda, u), 0.0
    for dt in (0.3, 0.3, 0.2):
        if method == 0:
            uo, _ = orc.dirk_step(spec, deg, ns, uo, dt, t, ref)
        else:
            uo, _ = orc.irk_step(spec, deg, ns, uo, dt, t, ref)
        out = _dev(cuda, np.zeros_like(u))
        gs = S.advance(ud, dt, t, out)
        assert gs.last_resid < 1e-12
        ud, t = out, t + dt
        assert rel(ud.cpu().numpy(), uo) < 1e-9
