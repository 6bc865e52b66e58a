"""GPU parity of the block-sparse algebra: BSR matvec, block DILU and block Jacobi
(FP64 / FP32 / BF16 factors) and GMRES, through the C ABI, against the oracle."""
import numpy as np
import pytest

import oracle as O
from cpolyg_amd import hip
from cpolyg_amd.spec import (EquationSpec, PC_BJACOBI, PC_DILU, PREC_BF16, PREC_BF16_D32, PREC_FP32, PREC_FP64,
                             GMRES_FLEXIBLE, GMRES_LEFT_REF)
from problems import dense, newton_matrix, rel, vortex_mesh, vortex_state, unit_periodic_mesh

pytestmark = pytest.mark.gpu


def _euler_system(deg=3, nx=8):
    m = vortex_mesh(nx)
    orc = O.Oracle(m)
    u = vortex_state(m, deg, perturb=0.02)
    B = newton_matrix(orc, EquationSpec.euler(), deg, u)
    return m, orc, B


def _adv_system(deg, nx=8):
    m = unit_periodic_mesh(nx)
    orc = O.Oracle(m)
    rng = np.random.default_rng(5)
    from cpolyg_amd.spec import bs_of
    u = rng.standard_normal(m.n_cells * bs_of(deg))
    B = newton_matrix(orc, EquationSpec.reference_advection(), deg, u, dt=0.05)
    return m, orc, B


def _dev(torch_cuda, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(torch_cuda)


@pytest.mark.parametrize("kind,deg", [("euler", 3), ("euler", 4), ("euler", 1), ("adv", 1), ("adv", 2), ("adv", 3)])
def test_bsr_matvec(cuda, kind, deg):
    m, orc, B = _euler_system(deg) if kind == "euler" else _adv_system(deg)
    bl = B.shape[1]
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    np.testing.assert_array_equal(A.download(), B)
    x = np.random.default_rng(1).standard_normal(m.n_cells * bl)
    xd = _dev(cuda, x)
    yd = _dev(cuda, np.zeros_like(x))
    A.matvec(xd, yd)
    y = yd.cpu().numpy()
    yo = O.bsr_matvec(orc.rowptr, orc.colind, B, x)
    assert rel(y, yo) < 1e-13


@pytest.mark.parametrize("deg", [3, 4, 1])
def test_dilu_fp64_factor_and_apply(cuda, deg):
    m, orc, B = _euler_system(deg)
    bl = B.shape[1]
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    P = hip.Prec(A, PC_DILU, PREC_FP64)
    assert 2 <= P.levels() <= m.n_colours   # contiguous levels: at most one per colour
    P.factor()
    F = P.download()
    Fo, info = O.dilu_factor(orc.rowptr, orc.colind, B)
    assert info == 0
    assert rel(F, Fo) < 1e-11
    x = np.random.default_rng(2).standard_normal(m.n_cells * bl)
    xd = _dev(cuda, x)
    P.apply(xd)
    assert rel(xd.cpu().numpy(), O.dilu_apply(orc.rowptr, orc.colind, Fo, x)) < 1e-11


def _permuted(orc, B, seed):
    """The same matrix with rows and columns renumbered by a random permutation (each row's diagonal
    block first, then the others by column): DILU levels that are not colour segments."""
    rp, ci = np.asarray(orc.rowptr), np.asarray(orc.colind)
    n = len(rp) - 1
    perm = np.random.default_rng(seed).permutation(n)   # new row r is old row perm[r]
    inv = np.argsort(perm)
    rp2, ci2, blk = [0], [], []
    for r in range(n):
        i = perm[r]
        ent = sorted(range(rp[i], rp[i + 1]), key=lambda q: (inv[ci[q]] != r, inv[ci[q]]))
        ci2 += [int(inv[ci[q]]) for q in ent]
        blk += [B[q] for q in ent]
        rp2.append(len(ci2))
    return np.array(rp2, np.int32), np.array(ci2, np.int32), np.array(blk)


@pytest.mark.parametrize("case,prec", [(c, p) for c in ("colour-ordered", "permuted")
                                       for p in (PREC_FP64, PREC_FP32, PREC_BF16)]
                         + [("colour-ordered", PREC_BF16_D32), ("colour-ordered-p4", PREC_BF16_D32)])
def test_dilu_single_launch_matches_level_launches(cuda, case, prec):
    """The one-launch DILU apply (rows waiting on completion counters; with FP32 diagonal inverses one
    launch per sweep) against one launch per level: bitwise equal over alternating applies (the
    counters advance by one per apply), on a colour-ordered matrix (contiguous levels, the last
    forward level finishing its rows) and on a randomly renumbered one (levels of scattered rows);
    FP64 also against the oracle."""
    if case == "colour-ordered-p4":
        m, orc, B = _euler_system(4, nx=6)
        rp, ci = orc.rowptr, orc.colind
    elif case == "colour-ordered":
        m, orc, B = _euler_system(3)
        rp, ci = orc.rowptr, orc.colind
    else:
        m, orc, B = _euler_system(1, nx=6)
        rp, ci, B = _permuted(orc, B, 1)
    bl = B.shape[1]
    A = hip.Bsr(rp, ci, bl)
    A.upload(B)
    P = hip.Prec(A, PC_DILU, prec)
    P.factor()
    assert P.flow_in_use()
    if case == "permuted":
        assert P.levels() > m.n_colours   # not the colour segments
    x = np.random.default_rng(7).standard_normal(m.n_cells * bl)
    outs = []
    for mode in (1, 0, 1, 1, 0, 1):
        P.set_flow(mode)
        xd = _dev(cuda, x)
        P.apply(xd)
        outs.append(xd.cpu().numpy())
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])
    if prec == PREC_FP64:
        Fo, info = O.dilu_factor(rp, ci, B)
        assert info == 0
        assert rel(outs[0], O.dilu_apply(rp, ci, Fo, x)) < 1e-11


@pytest.mark.parametrize("kind,prec", [(PC_DILU, PREC_BF16), (PC_DILU, PREC_FP64), (PC_DILU, PREC_BF16_D32),
                                       (PC_BJACOBI, PREC_FP32)])
def test_prec_apply_out_of_place_matches_in_place(cuda, kind, prec):
    """cpg_prec_apply_to (z = M^-1 x, the flexible-GMRES form) leaves x alone and gives the in-place
    result bitwise."""
    import torch
    m, orc, B = _euler_system(3)
    A = hip.Bsr(orc.rowptr, orc.colind, B.shape[1])
    A.upload(B)
    P = hip.Prec(A, kind, prec)
    P.factor()
    x = np.random.default_rng(9).standard_normal(m.n_cells * B.shape[1])
    xd = _dev(cuda, x)
    z = torch.full_like(xd, np.nan)
    P.apply_to(xd, z)
    np.testing.assert_array_equal(xd.cpu().numpy(), x)
    P.apply(xd)
    np.testing.assert_array_equal(z.cpu().numpy(), xd.cpu().numpy())
    with pytest.raises(hip.CpgError):
        P.apply_to(xd, xd)


def test_dilu_single_launch_wait_limit_reports_and_recovers(cuda):
    """A wait limit of 0 polls makes every row that reads another row give up at once: the kernel
    still ends (the counters stay consistent), the next apply reports the failed one, and the
    preconditioner continues with one launch per level until the single launch is enabled again."""
    import torch
    m, orc, B = _euler_system(3)
    A = hip.Bsr(orc.rowptr, orc.colind, B.shape[1])
    A.upload(B)
    P = hip.Prec(A, PC_DILU, PREC_FP64)
    P.factor()
    x = np.random.default_rng(8).standard_normal(m.n_cells * B.shape[1])
    P.set_flow(0)
    xd = _dev(cuda, x)
    P.apply(xd)
    ref = xd.cpu().numpy()
    P.set_flow(1, 0)
    P.apply(_dev(cuda, x))
    torch.cuda.synchronize()
    with pytest.raises(hip.CpgError, match="waited more than 0 polls"):
        P.apply(_dev(cuda, x))
    assert not P.flow_in_use()
    xd = _dev(cuda, x)
    P.apply(xd)
    np.testing.assert_array_equal(xd.cpu().numpy(), ref)
    P.set_flow(1, 1 << 22)
    assert P.flow_in_use()
    for _ in range(2):
        xd = _dev(cuda, x)
        P.apply(xd)
        np.testing.assert_array_equal(xd.cpu().numpy(), ref)


@pytest.mark.parametrize("deg", [3, 4])
@pytest.mark.parametrize("prec,tol", [(PREC_FP32, 5e-6), (PREC_BF16, 2e-2), (PREC_BF16_D32, 2e-2)])
def test_dilu_reduced_precision(cuda, prec, tol, deg):
    m, orc, B = _euler_system(deg, nx=8 if deg == 3 else 6)
    bl = B.shape[1]
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    P = hip.Prec(A, PC_DILU, prec)
    P.factor()
    F = P.download()
    Fo, _ = O.dilu_factor(orc.rowptr, orc.colind, B)
    # stored factors are the FP64 factors rounded once: every element within half an ulp of its
    # own magnitude (unit roundoff: FP32 2^-24, BF16 2^-8 with its 8-bit significand, the per-block
    # power-of-two scale changing no significand), so within that of its block's largest element
    half_ulp = np.full(len(F), 2.0 ** -24 if prec == PREC_FP32 else 2.0 ** -8)
    if prec == PREC_BF16_D32:   # the diagonal inverses in FP32, the L and DU blocks in BF16
        half_ulp[orc.rowptr[:-1]] = 2.0 ** -24
    err = np.abs(F - Fo).reshape(len(F), -1).max(1) / np.maximum(np.abs(Fo).reshape(len(F), -1).max(1), 1e-300)
    assert np.all(err <= half_ulp * (1 + 1e-6) + 1e-11)
    x = np.random.default_rng(3).standard_normal(m.n_cells * bl)
    xd = _dev(cuda, x)
    P.apply(xd)
    got = xd.cpu().numpy()
    # the apply of the stored factors: FP32 partial sums and FP32 gathered vectors, against the
    # oracle's FP64 apply of the same (rounded) factors
    e_apply = rel(got, O.dilu_apply(orc.rowptr, orc.colind, F, x))
    # end to end against the FP64 factors: the factor rounding amplified by the sweeps
    e_total = rel(got, O.dilu_apply(orc.rowptr, orc.colind, Fo, x))
    print(f"p={deg} prec={prec}: factor {err.max():.2e}, apply vs rounded factors {e_apply:.2e}, vs FP64 {e_total:.2e}")
    assert e_apply < (1e-6 if prec == PREC_FP32 else 1e-5)
    if prec == PREC_BF16_D32 and deg == 4:
        # FP32 diagonal inverses bring the sweep closer to the FP64 one: 9.3e-3 here, against
        # 2.9e-2 with every factor block in BF16
        assert e_total < 1.5e-2
    assert e_total < (tol if deg == 3 else 3 * tol)


def _rne_bf16(x):
    """x rounded to the nearest BF16 (8 significand bits), ties to even, exactly in FP64."""
    m, e = np.frexp(x)
    return np.ldexp(np.rint(m * 256.0) / 256.0, e)


@pytest.mark.parametrize("deg", [3, 4])
def test_bf16_factors_are_correctly_rounded(cuda, deg):
    """BF16 factors are the FP64 factors rounded once to nearest even relative to the block's
    power-of-two scale (max |block| / s in (0.5, 1]): the block-Jacobi inverses of the diagonal
    blocks, which the factor kernel computes in FP64 for every storage precision, against the
    FP64 factor rounded here.  Bitwise."""
    m, orc, B = _euler_system(deg, nx=8 if deg == 3 else 6)
    A = hip.Bsr(orc.rowptr, orc.colind, B.shape[1])
    A.upload(B)
    P64 = hip.Prec(A, PC_BJACOBI, PREC_FP64)
    P64.factor()
    F64 = P64.download()
    P16 = hip.Prec(A, PC_BJACOBI, PREC_BF16)
    P16.factor()
    F16 = P16.download()
    mx = np.abs(F64).reshape(len(F64), -1).max(1)
    s = np.ldexp(1.0, np.clip(np.frexp(mx)[1], -125, 127))[:, None, None]
    np.testing.assert_array_equal(F16, _rne_bf16(F64 / s) * s)


# Radau IIA, 3 stages (Hairer & Wanner, Solving ODEs II, table IV.5.6)
_R6 = np.sqrt(6.0)
RADAU3 = np.array([[(88 - 7 * _R6) / 360, (296 - 169 * _R6) / 1800, (-2 + 3 * _R6) / 225],
                   [(296 + 169 * _R6) / 1800, (88 + 7 * _R6) / 360, (-2 - 3 * _R6) / 225],
                   [(16 - _R6) / 36, (16 + _R6) / 36, 1 / 9]])


@pytest.mark.parametrize("prec,products,ftol,atol", [(PREC_FP64, "fp64", 1e-11, 1e-11),
                                                     (PREC_FP32, "fp32", 1e-5, 2e-5),
                                                     (PREC_FP32, "fp64", 1e-6, 5e-6),
                                                     (PREC_BF16, "fp32", 8e-3, 2e-2),
                                                     (PREC_BF16, "fp64", 8e-3, 2e-2)])
def test_coupled_dilu_matches_oracle(cuda, monkeypatch, prec, products, ftol, atol):
    """Block DILU of the Radau IIA stage-coupled Newton matrix in cell-interleaved order, its
    120 x 120 blocks formed inside the factor from three stage Jacobians and the mass matrix,
    against the oracle's DILU of the same matrix assembled explicitly (block (s, t) of slot q =
    -dt A[s][t] J_s[q], plus the mass block on the diagonal when s == t).  FP32 and BF16 factors
    form their Schur and DU products on FP32 MFMA unless CPOLYG_COUPLED_FP64_PRODUCTS=1: FP32
    rounding of the products puts the FP32 factor within 1e-5 of the oracle (2.5e-6 measured)
    against 1e-6 with FP64 products rounded once on store."""
    if products == "fp64" and prec != PREC_FP64:
        monkeypatch.setenv("CPOLYG_COUPLED_FP64_PRODUCTS", "1")
    else:
        monkeypatch.delenv("CPOLYG_COUPLED_FP64_PRODUCTS", raising=False)
    deg, bl = 3, 40
    m = vortex_mesh(6)
    orc = O.Oracle(m)
    spec = EquationSpec.euler()
    E = hip.Equation(hip.Mesh(m), spec, deg)
    coef = -0.3 * RADAU3
    Jo = [orc.jacobian(spec, deg, vortex_state(m, deg, perturb=0.02, seed=s)) for s in range(3)]
    M = orc.mass(deg, 4)
    nnzb = len(orc.colind)
    # blocks are column-major: B[q][c][r] is entry (r, c)
    Bc = np.zeros((nnzb, 3 * bl, 3 * bl))
    for s in range(3):
        for t in range(3):
            Bc[:, bl * t:bl * (t + 1), bl * s:bl * (s + 1)] = coef[s, t] * Jo[s]
    for i in range(m.n_cells):
        for s in range(3):
            Bc[orc.rowptr[i], bl * s:bl * (s + 1), bl * s:bl * (s + 1)] += M[i]
    Js = []
    for s in range(3):
        J = hip.Bsr(orc.rowptr, orc.colind, bl)
        J.upload(Jo[s])
        Js.append(J)
    P = hip.Prec(hip.Bsr(orc.rowptr, orc.colind, 3 * bl), PC_DILU, prec)
    P.factor_coupled(E, Js, coef)
    F = P.download()
    Fo, info = O.dilu_factor(orc.rowptr, orc.colind, Bc)
    assert info == 0
    err = np.abs(F - Fo).reshape(nnzb, -1).max(1) / np.abs(Fo).reshape(nnzb, -1).max(1)
    assert err.max() < ftol
    x = np.random.default_rng(4).standard_normal(m.n_cells * 3 * bl)
    xd = _dev(cuda, x)
    P.apply(xd)
    assert rel(xd.cpu().numpy(), O.dilu_apply(orc.rowptr, orc.colind, Fo, x)) < atol


@pytest.mark.parametrize("prec,tol", [(PREC_FP64, 1e-11), (PREC_FP32, 5e-6), (PREC_BF16, 2e-2)])
def test_block_jacobi(cuda, prec, tol):
    m, orc, B = _euler_system(3)
    bl = B.shape[1]
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    P = hip.Prec(A, PC_BJACOBI, prec)
    P.factor()
    LU, piv, _ = O.bjacobi_factor(orc.rowptr, orc.colind, B)
    x = np.random.default_rng(4).standard_normal(m.n_cells * bl)
    xd = _dev(cuda, x)
    P.apply(xd)
    assert rel(xd.cpu().numpy(), O.bjacobi_apply(LU, piv, x)) < tol


@pytest.mark.parametrize("pc", [PC_BJACOBI, None])
def test_jacobi_iteration_matches_oracle(cuda, pc):
    """cpg_jacobi (BlockMatrix::jacobi, omega = 1) against the oracle: a fixed number of steps
    (tol 0) gives the same iterate and residual norm; to convergence, the same iteration count."""
    m, orc, B = _adv_system(2)
    bl = B.shape[1]
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    P = None
    if pc is not None:
        P = hip.Prec(A, PC_BJACOBI, PREC_FP64)
        P.factor()
    K = hip.Krylov(m.n_cells, bl, 4, GMRES_LEFT_REF)
    b = np.random.default_rng(8).standard_normal(m.n_cells * bl)
    bd = _dev(cuda, b)
    if pc is None:
        # P = I: x <- x + (b - A x), a Richardson iteration; scale the system so that it converges
        B = B / np.abs(B).max() * 0.5
        A.upload(B)
        xd = _dev(cuda, np.zeros_like(b))
        tol, its = K.jacobi(A, None, bd, xd, 0.0, 5)
        x = np.zeros_like(b)
        Ad = dense(orc.rowptr, orc.colind, B)
        for _ in range(5):
            r = Ad @ x - b
            x = x - r
        assert its == 5 and rel(xd.cpu().numpy(), x) < 1e-12
        return
    xd = _dev(cuda, np.zeros_like(b))
    tol, its = K.jacobi(A, P, bd, xd, 0.0, 9)
    xo, tolo, itso = O.jacobi_solve(orc.rowptr, orc.colind, B, b, tol=0.0, maxit=9)
    assert its == itso == 9 and rel(xd.cpu().numpy(), xo) < 1e-12 and abs(tol - tolo) < 1e-10 * tolo
    xd = _dev(cuda, np.zeros_like(b))
    tol, its = K.jacobi(A, P, bd, xd, 1e-10, 1000)
    xo, tolo, itso = O.jacobi_solve(orc.rowptr, orc.colind, B, b, tol=1e-10, maxit=1000)
    assert its == itso < 1000 and tol < 1e-10 and rel(xd.cpu().numpy(), xo) < 1e-9
    # no iterations: the reference's empty residual
    xd = _dev(cuda, np.zeros_like(b))
    assert K.jacobi(A, P, bd, xd, 1e-10, 0) == (0.0, 0)


@pytest.mark.parametrize("mode", [GMRES_LEFT_REF, GMRES_FLEXIBLE])
@pytest.mark.parametrize("pc", [PC_DILU, PC_BJACOBI])
def test_gmres_matches_oracle(cuda, mode, pc):
    m, orc, B = _euler_system(3)
    bl = B.shape[1]
    n = m.n_cells
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    P = hip.Prec(A, pc, PREC_FP64)
    P.factor()
    b = np.random.default_rng(6).standard_normal(n * bl)
    K = hip.Krylov(n, bl, 20, mode)
    bd = _dev(cuda, b)
    xd = _dev(cuda, np.zeros_like(b))
    tol, it = K.gmres(A, P, bd, xd, 1e-10, 2000)
    xo, tolo, ito = O.gmres(orc.rowptr, orc.colind, B, b, m=20, tol=1e-10, maxit=2000, pc_kind=pc, mode=mode)
    x = xd.cpu().numpy()
    assert abs(it - ito) <= 1, (it, ito)
    assert tol < 1e-10 and tolo < 1e-10
    r = b - O.bsr_matvec(orc.rowptr, orc.colind, B, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8
    assert rel(x, xo) < 1e-7


def test_gmres_zero_rhs(cuda):
    m, orc, B = _euler_system(1)
    bl = B.shape[1]
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    P = hip.Prec(A, PC_DILU, PREC_FP64)
    P.factor()
    K = hip.Krylov(m.n_cells, bl, 10, GMRES_LEFT_REF)
    bd = _dev(cuda, np.zeros(m.n_cells * bl))
    xd = _dev(cuda, np.ones(m.n_cells * bl))
    tol, it = K.gmres(A, P, bd, xd, 1e-10, 100)
    assert it == 0 and tol == 0.0
    assert np.all(xd.cpu().numpy() == 0.0)


def test_singular_block_reports_status(cuda):
    m, orc, B = _euler_system(1)
    B = B.copy()
    B[orc.rowptr[3]] = 0.0      # zero diagonal block of row 3
    bl = B.shape[1]
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    P = hip.Prec(A, PC_BJACOBI, PREC_FP64)
    with pytest.raises(hip.CpgError) as e:
        P.factor()
    assert e.value.status == 4


def _blocks_120(n):
    """n 120 x 120 blocks that exercise the pivot search: Gaussian, row-permuted with a zero
    diagonal, graded over 16 orders of magnitude, random +-1 (ties), and diagonally dominant."""
    rng = np.random.default_rng(11)
    out = []
    for q in range(n):
        kind = q % 5
        if kind == 0:
            X = rng.standard_normal((120, 120))
        elif kind == 1:
            X = rng.standard_normal((120, 120))
            X = X[rng.permutation(120)]
            np.fill_diagonal(X, 0.0)
        elif kind == 2:
            d1 = 10.0 ** rng.uniform(-8, 8, 120)
            d2 = 10.0 ** rng.uniform(-8, 8, 120)
            X = d1[:, None] * rng.standard_normal((120, 120)) * d2[None, :]
        elif kind == 3:
            X = rng.choice([-1.0, 1.0], size=(120, 120))
        else:
            X = rng.standard_normal((120, 120)) + 120.0 * np.eye(120)
        out.append(X)
    return np.array(out)


def test_blocked_inverse_120(cuda):
    """The 120 x 120 inverses (blocked Gauss-Jordan on FP64 MFMA) against LAPACK's, per block
    relative to cond(block) * 2^-52; the FP32 and BF16 factors are the FP64 ones rounded once
    (bitwise); a block with a zero column reports the singular status."""
    m = vortex_mesh(4)
    orc = O.Oracle(m)
    n, nnzb = m.n_cells, len(orc.colind)
    B = np.zeros((nnzb, 120, 120))
    diag = np.array([orc.rowptr[i] + list(orc.colind[orc.rowptr[i]:orc.rowptr[i + 1]]).index(i)
                     for i in range(n)])
    B[diag] = _blocks_120(n)
    A = hip.Bsr(orc.rowptr, orc.colind, 120)
    A.upload(B)
    P64 = hip.Prec(A, PC_BJACOBI, PREC_FP64)
    P64.factor()
    F64 = P64.download()
    for i in range(n):
        ref = np.linalg.inv(B[diag[i]])
        err = np.abs(F64[i] - ref).max() / np.abs(ref).max()
        bound = 120 * np.linalg.cond(B[diag[i]]) * 2.0 ** -52
        assert err < bound, (i, i % 5, err, bound)
    P32 = hip.Prec(A, PC_BJACOBI, PREC_FP32)
    P32.factor()
    np.testing.assert_array_equal(P32.download(), F64.astype(np.float32).astype(np.float64))
    P16 = hip.Prec(A, PC_BJACOBI, PREC_BF16)
    P16.factor()
    mx = np.abs(F64).reshape(n, -1).max(1)
    s = np.ldexp(1.0, np.clip(np.frexp(mx)[1], -125, 127))[:, None, None]
    np.testing.assert_array_equal(P16.download(), _rne_bf16(F64 / s) * s)
    B[diag[5], 37, :] = 0.0     # download layout [c][r]: column 37 of block 5 is zero
    A.upload(B)
    P = hip.Prec(A, PC_BJACOBI, PREC_FP64)
    with pytest.raises(hip.CpgError) as e:
        P.factor()
    assert e.value.status == 4


@pytest.mark.parametrize("prec", [PREC_FP32, PREC_BF16])
def test_inner_gmres_preconditioner_in_flexible_gmres(cuda, prec):
    """FP64 flexible GMRES preconditioned by FP32 GMRES(k) + DILU reaches the FP64 tolerance;
    the solution matches the oracle's FP64 solve."""
    from cpolyg_amd.spec import PC_INNER_GMRES
    m, orc, B = _euler_system(3)
    bl = B.shape[1]
    n = m.n_cells
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    P = hip.Prec(A, PC_INNER_GMRES, prec)
    P.set_inner(8)
    P.factor()
    b = np.random.default_rng(7).standard_normal(n * bl)
    # one application approximates A^-1 b far better than the DILU sweep alone
    zd = _dev(cuda, b)
    P.apply(zd)
    r_inner = b - O.bsr_matvec(orc.rowptr, orc.colind, B, zd.cpu().numpy())
    Fo, _ = O.dilu_factor(orc.rowptr, orc.colind, B)
    r_dilu = b - O.bsr_matvec(orc.rowptr, orc.colind, B, O.dilu_apply(orc.rowptr, orc.colind, Fo, b))
    assert np.linalg.norm(r_inner) < 0.2 * np.linalg.norm(r_dilu)
    K = hip.Krylov(n, bl, 20, GMRES_FLEXIBLE)
    bd = _dev(cuda, b)
    xd = _dev(cuda, np.zeros_like(b))
    tol, it = K.gmres(A, P, bd, xd, 1e-10, 500)
    x = xd.cpu().numpy()
    r = b - O.bsr_matvec(orc.rowptr, orc.colind, B, x)
    assert tol < 1e-10 and np.linalg.norm(r) / np.linalg.norm(b) < 2e-10
    xo, _, _ = O.gmres(orc.rowptr, orc.colind, B, b, m=20, tol=1e-12, maxit=2000, pc_kind=PC_DILU, mode=GMRES_FLEXIBLE)
    assert rel(x, xo) < 1e-7


def test_flexible_graphs_follow_reallocated_handles(cuda):
    """Regression: the flexible solver replays its captured cycles only for the handles and
    buffers it captured.  Changing the inner iteration count between two solves (which
    reallocates the inner basis), and destroying and re-creating the preconditioner, give
    bitwise the result of a fresh Krylov handle."""
    import torch
    from cpolyg_amd.spec import PC_INNER_GMRES
    m, orc, B = _euler_system(3)
    bl = B.shape[1]
    n = m.n_cells
    A = hip.Bsr(orc.rowptr, orc.colind, bl)
    A.upload(B)
    b = _dev(cuda, np.random.default_rng(11).standard_normal(n * bl))

    def solve(K, P):
        x = torch.zeros_like(b)
        K.gmres(A, P, b, x, 1e-10, 500)
        return x.cpu().numpy()

    K = hip.Krylov(n, bl, 20, GMRES_FLEXIBLE)
    P = hip.Prec(A, PC_INNER_GMRES, PREC_FP32)
    P.set_inner(4)
    P.factor()
    solve(K, P)
    P.set_inner(9)
    P.factor()
    got = solve(K, P)
    np.testing.assert_array_equal(got, solve(hip.Krylov(n, bl, 20, GMRES_FLEXIBLE), P))
    del P
    P2 = hip.Prec(A, PC_INNER_GMRES, PREC_FP32)
    P2.set_inner(6)
    P2.factor()
    got = solve(K, P2)
    np.testing.assert_array_equal(got, solve(hip.Krylov(n, bl, 20, GMRES_FLEXIBLE), P2))


@pytest.mark.parametrize("kind", [PC_DILU, PC_BJACOBI])
def test_factor_is_deterministic(cuda, kind):
    """Regression: repeated factorizations of one matrix are bitwise identical (the pivot of
    the in-LDS Gauss-Jordan is broadcast behind a barrier)."""
    m = vortex_mesh(32)
    orc = O.Oracle(m)
    u = vortex_state(m, 3, perturb=0.02)
    B = newton_matrix(orc, EquationSpec.euler(jacobian=1), 3, u, dt=0.05)
    A = hip.Bsr(orc.rowptr, orc.colind, B.shape[1])
    A.upload(B)
    P = hip.Prec(A, kind, PREC_FP64)
    P.factor()
    F0 = P.download()
    for _ in range(4):
        P.factor()
        np.testing.assert_array_equal(P.download(), F0)
    if kind == PC_DILU:
        Fo, _ = O.dilu_factor(orc.rowptr, orc.colind, B)
        assert rel(F0, Fo) < 1e-10


def test_empty_matrix_operations_are_no_ops(cuda):
    """A BSR with no rows: matvec, DILU / block-Jacobi factor and apply return at once."""
    import torch
    A = hip.Bsr(np.zeros(1, np.int32), np.zeros(0, np.int32), 40)
    x = torch.zeros(0, dtype=torch.float64, device=cuda)
    y = torch.zeros(0, dtype=torch.float64, device=cuda)
    A.matvec(x, y)
    for kind in (PC_DILU, PC_BJACOBI):
        P = hip.Prec(A, kind, PREC_BF16)
        P.factor()
        P.apply(x)


def test_coupled_dilu_rejects_mismatched_inputs(cuda):
    """The coupled factor checks its handles before any launch: a 40-wide preconditioner, or a
    120-wide one over another pattern, is refused with a status."""
    deg, bl = 3, 40
    m = vortex_mesh(4)
    orc = O.Oracle(m)
    E = hip.Equation(hip.Mesh(m), EquationSpec.euler(), deg)
    Js = [hip.Bsr(orc.rowptr, orc.colind, bl) for _ in range(3)]
    P40 = hip.Prec(hip.Bsr(orc.rowptr, orc.colind, bl), PC_DILU, PREC_FP64)
    with pytest.raises(hip.CpgError) as e:
        P40.factor_coupled(E, Js, -0.1 * RADAU3)
    assert e.value.status == 7
    n = m.n_cells
    diag_only = hip.Bsr(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), 3 * bl)
    Pd = hip.Prec(diag_only, PC_DILU, PREC_FP64)
    with pytest.raises(hip.CpgError) as e:
        Pd.factor_coupled(E, Js, -0.1 * RADAU3)
    assert e.value.status == 1
