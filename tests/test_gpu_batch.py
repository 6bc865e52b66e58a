"""Batched implicit steps (cpg_batch): every instance of a batch against the oracle's step
of that instance alone, graph replay against eager launches, and uneven convergence."""
import numpy as np
import pytest

import oracle as O
from cpolyg_amd import hip, setups
from cpolyg_amd.spec import EquationSpec, GMRES_FLEXIBLE, GMRES_LEFT_REF, PC_DILU, PREC_FP32, PREC_FP64, SolverSpec
from problems import rel

pytestmark = pytest.mark.gpu


def _solver(prec=PREC_FP32, restart=30):
    return SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=restart, rtol=1e-10, maxit=3000, pc_kind=PC_DILU,
                      precision=prec, newton_tol=1e-12, newton_maxit=10)


def _small_batch(count=6, seed=7):
    return setups.batch_advection(count=count, seed=seed, nmin=40, nmax=220, cfl=5.0)


def _run(cuda, classes, meta, nstages, solver, use_graph, steps=1, mode=None, teams=None):
    import torch
    spec = EquationSpec.rotating_advection()
    eqs = [(hip.Equation(hip.Mesh(um), spec, p), cp) for p, um, cp, _, _, _ in classes]
    B = hip.Batch(eqs, nstages, solver, use_graph=use_graph, mode=mode)
    if teams is not None:
        B.set_teams(*teams)
    us = [torch.from_numpy(u0).to(cuda) for _, _, _, u0, _, _ in classes]
    outs = [torch.zeros_like(u) for u in us]
    dts = [d for _, _, d in meta]
    its_all = []
    for _ in range(steps):
        st, its = B.advance(us, dts, 0.0, outs)
        its_all.append(its)
        us, outs = outs, us
    return [u.cpu().numpy() for u in us], st, its_all, B


@pytest.mark.parametrize("nstages,mode", [(1, hip.Batch.GRAPH), (2, hip.Batch.GRAPH), (1, hip.Batch.PERSISTENT),
                                          (2, hip.Batch.PERSISTENT)])
def test_batch_matches_oracle_per_instance(cuda, nstages, mode):
    classes, meta = _small_batch()
    got, st, its, B = _run(cuda, classes, meta, nstages, _solver(), True, mode=mode)
    assert B.info()["graph_captured"] == (mode == hip.Batch.GRAPH)
    spec = EquationSpec.rotating_advection()
    ref = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=40, rtol=1e-13, maxit=4000, pc_kind=PC_DILU,
                     newton_tol=1e-13, newton_maxit=10)
    q = 0
    for c, (p, um, cp, u0, ids, meshes) in enumerate(classes):
        bs = (p + 1) * (p + 2) // 2
        for k in range(len(cp) - 1):
            lo, hi = cp[k] * bs, cp[k + 1] * bs
            m = meshes[k]   # the instance's own mesh
            uo, _ = O.Oracle(m).dirk_step(spec, p, nstages, u0[lo:hi], meta[q][2], 0.0, ref)
            assert rel(got[c][lo:hi], uo) < 1e-8, (c, k, rel(got[c][lo:hi], uo))
            assert its[0][q] > 0
            q += 1


def test_batch_graph_replay_matches_eager(cuda):
    classes, meta = _small_batch(count=5, seed=11)
    a, sa, ia, Ba = _run(cuda, classes, meta, 2, _solver(), True, steps=3)
    b, sb, ib, Bb = _run(cuda, classes, meta, 2, _solver(), False, steps=3)
    assert Ba.info()["graph_captured"] and not Bb.info()["graph_captured"]
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)      # deterministic kernels: every replay identical
    for x, y in zip(ia, ib):
        np.testing.assert_array_equal(x, y)


def test_batch_uneven_convergence_fp64(cuda):
    """Instances of very different size and degree converge after different Arnoldi counts in
    the same cycles; each must still meet its own tolerance (FP64 factors, short restart)."""
    classes, meta = setups.batch_advection(count=8, seed=3, nmin=30, nmax=400, cfl=8.0)
    got, st, its, B = _run(cuda, classes, meta, 1, _solver(PREC_FP64, restart=8), True)
    assert len(set(its[0].tolist())) > 1
    spec = EquationSpec.rotating_advection()
    ref = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=40, rtol=1e-13, maxit=4000, pc_kind=PC_DILU,
                     newton_tol=1e-13, newton_maxit=10)
    q = 0
    for c, (p, um, cp, u0, ids, meshes) in enumerate(classes):
        bs = (p + 1) * (p + 2) // 2
        for k in range(len(cp) - 1):
            lo, hi = cp[k] * bs, cp[k + 1] * bs
            m = meshes[k]
            uo, _ = O.Oracle(m).dirk_step(spec, p, 1, u0[lo:hi], meta[q][2], 0.0, ref)
            assert rel(got[c][lo:hi], uo) < 1e-8
            q += 1


def _persistent(cuda, classes, meta, solver, nstages=2, steps=2, teams=(1, 0, 0)):
    return _run(cuda, classes, meta, nstages, solver, False, steps=steps, mode=hip.Batch.PERSISTENT, teams=teams)


def test_persistent_batch_matches_graph_batch(cuda):
    """The cooperative solver against the launch-per-phase one on the same batch: the same
    steps to rounding, the same Arnoldi counts within one per instance and solve."""
    classes, meta = _small_batch(count=7, seed=5)
    a, sa, ia, _ = _run(cuda, classes, meta, 2, _solver(), True, steps=2)
    b, sb, ib, Bp = _persistent(cuda, classes, meta, _solver())
    for x, y in zip(a, b):
        assert rel(y, x) < 1e-10
    for x, y in zip(ia, ib):
        assert np.abs(np.asarray(x) - np.asarray(y)).max() <= 2
    t = Bp.teams()
    assert t["grid"] == int(t["team_wg"].sum()) and (t["team_wg"] >= 1).all()


def test_persistent_batch_independent_of_team_sizes(cuda):
    """Reductions go over fixed chunks in chunk order and every row has one summation order,
    so the result is bitwise the same for every team layout (one workgroup per instance, and
    teams as large as the rows allow)."""
    classes, meta = _small_batch(count=6, seed=9)
    a, _, ia, Ba = _persistent(cuda, classes, meta, _solver(), teams=(4, 1, 0))
    b, _, ib, Bb = _persistent(cuda, classes, meta, _solver(), teams=(4, 0, 0))
    assert (Ba.teams()["team_wg"] != Bb.teams()["team_wg"]).any()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(ia, ib):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("pc", ["bjacobi", "none"])
def test_persistent_batch_other_preconditioners(cuda, pc):
    from cpolyg_amd.spec import PC_BJACOBI, PC_NONE
    # without a preconditioner (the mass matrix's scaling left in) GMRES needs small p-1 systems
    if pc == "bjacobi":
        classes, meta = setups.batch_advection(count=5, seed=13, nmin=30, nmax=300, cfl=2.0)
    else:
        classes, meta = setups.batch_advection(count=4, seed=13, nmin=20, nmax=60, degs=(1,), cfl=0.5)
    solver = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=20 if pc == "bjacobi" else 60, rtol=1e-10, maxit=20000,
                        pc_kind=PC_BJACOBI if pc == "bjacobi" else PC_NONE, precision=PREC_FP64,
                        newton_tol=1e-12, newton_maxit=10)
    got, st, its, B = _persistent(cuda, classes, meta, solver, nstages=1, steps=1, teams=(2, 0, 0))
    spec = EquationSpec.rotating_advection()
    ref = SolverSpec(gmres_mode=GMRES_LEFT_REF, restart=40, rtol=1e-13, maxit=4000, pc_kind=PC_DILU,
                     newton_tol=1e-13, newton_maxit=10)
    q = 0
    for c, (p, um, cp, u0, ids, meshes) in enumerate(classes):
        bs = (p + 1) * (p + 2) // 2
        for k in range(len(cp) - 1):
            lo, hi = cp[k] * bs, cp[k + 1] * bs
            uo, _ = O.Oracle(meshes[k]).dirk_step(spec, p, 1, u0[lo:hi], meta[q][2], 0.0, ref)
            assert rel(got[c][lo:hi], uo) < 1e-8
            q += 1


def test_persistent_batch_barrier_limit_reports_and_recovers(cuda):
    """A team barrier that waits past its limit (forced here from the host with a one-round
    limit) ends the launch with an error in every workgroup; the next solve, with the default
    limit, starts from clean barrier counters and is correct."""
    classes, meta = _small_batch(count=6, seed=7)
    ref, _, _, _ = _persistent(cuda, classes, meta, _solver(), steps=1)
    import torch
    spec = EquationSpec.rotating_advection()
    eqs = [(hip.Equation(hip.Mesh(um), spec, p), cp) for p, um, cp, _, _, _ in classes]
    B = hip.Batch(eqs, 2, _solver(), mode=hip.Batch.PERSISTENT)
    B.set_teams(4, 0, 1)
    us = [torch.from_numpy(u0).to(cuda) for _, _, _, u0, _, _ in classes]
    outs = [torch.zeros_like(u) for u in us]
    dts = [d for _, _, d in meta]
    with pytest.raises(hip.CpgError) as e:
        B.advance(us, dts, 0.0, outs)
    assert (B.teams()["team_wg"] > 1).any()
    assert "team barrier" in str(e.value)
    B.set_teams(1, 0, 0)
    B.advance(us, dts, 0.0, outs)
    for x, y in zip(ref, outs):
        np.testing.assert_array_equal(x, y.cpu().numpy())
