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


def _run(cuda, classes, meta, nstages, solver, use_graph, steps=1):
    import torch
    spec = EquationSpec.rotating_advection()
    eqs = [(hip.Equation(hip.Mesh(um), spec, p), cp) for p, um, cp, _, _, _ in classes]
    B = hip.Batch(eqs, nstages, solver, use_graph=use_graph)
    us = [torch.from_numpy(u0).to(cuda) for _, _, _, u0, _, _ in classes]
    outs = [torch.zeros_like(u) for u in us]
    dts = [d for _, _, d in meta]
    its_all = []
    for _ in range(steps):
        st, its = B.advance(us, dts, 0.0, outs)
        its_all.append(its)
        us, outs = outs, us
    return [u.cpu().numpy() for u in us], st, its_all, B


@pytest.mark.parametrize("nstages", [1, 2])
def test_batch_matches_oracle_per_instance(cuda, nstages):
    classes, meta = _small_batch()
    got, st, its, B = _run(cuda, classes, meta, nstages, _solver(), True)
    assert B.info()["graph_captured"]
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
