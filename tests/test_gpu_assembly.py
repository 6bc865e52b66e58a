"""GPU parity of the DG assembly (mass matrix, residual, Jacobian written into BSR
slots, Newton matrix M - dt a J, L2 norm) through the C ABI, against the oracle."""
import numpy as np
import pytest

import oracle as O
from cpolyg_amd import hip, setups
from cpolyg_amd.mesh import periodic_rectangle
from cpolyg_amd.spec import EquationSpec, BC_ISENTROPIC_VORTEX, BC_ZERO, bs_of
from problems import box_mesh, rel, unit_periodic_mesh, vortex_mesh, vortex_state

pytestmark = pytest.mark.gpu


def _dev(cuda, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(cuda)


def _cases():
    out = []
    for deg in (1, 2, 3, 4):
        for trace in (0, 1):
            out.append(("euler-periodic", deg, dict(periodic_trace=trace, jacobian=0)))
        out.append(("euler-periodic", deg, dict(periodic_trace=1, jacobian=1)))
        out.append(("adv-periodic", deg, dict(periodic_trace=0)))
        out.append(("adv-box", deg, {}))
    out.append(("euler-box", 3, dict(jacobian=1)))
    out.append(("adv-quad", 2, dict(periodic_trace=0)))
    return out


def _setup(kind, deg, kw):
    rng = np.random.default_rng(deg)
    if kind.startswith("euler"):
        if kind == "euler-box":
            m = box_mesh(60)
            # vortex centred in the unit box, Dirichlet data = moving exact vortex
            bc_par = (1.0, 0.5, 0.5, 1.0, 1.0, 0, 0, 0)
            spec = EquationSpec.euler(bc_kind=BC_ISENTROPIC_VORTEX, bc_par=bc_par, **kw)
            u = setups.interp(m, deg, setups.vortex(beta=1.0, x0=0.5, y0=0.5), 4)
        else:
            m = vortex_mesh(8)
            spec = EquationSpec.euler(**kw)
            u = vortex_state(m, deg, perturb=0.02, seed=deg)
    else:
        if kind == "adv-box":
            m = box_mesh(60)
            spec = EquationSpec.reference_advection(bc_kind=BC_ZERO)
        elif kind == "adv-quad":
            m = periodic_rectangle(4, 4, 1.0, 1.0, colour=True)
            spec = EquationSpec.constant_advection(1.0, 0.5, **kw)
        else:
            m = unit_periodic_mesh(8)
            spec = EquationSpec.constant_advection(1.0, 1.0, **kw)
        u = rng.standard_normal(m.n_cells * bs_of(deg))
    return m, spec, u


@pytest.mark.parametrize("kind,deg,kw", _cases())
def test_residual_and_jacobian(cuda, kind, deg, kw):
    m, spec, u = _setup(kind, deg, kw)
    orc = O.Oracle(m)
    mesh = hip.Mesh(m)
    E = hip.Equation(mesh, spec, deg)
    rp, ci = E.pattern()
    np.testing.assert_array_equal(rp, orc.rowptr)
    np.testing.assert_array_equal(ci, orc.colind)
    # mass
    Mo = orc.mass(deg, spec.nc)
    assert rel(E.mass(), Mo) < 1e-13
    # residual
    t = 0.37
    ud = _dev(cuda, u)
    bd = _dev(cuda, np.zeros_like(u))
    E.residual(ud, t, bd)
    bo = orc.assemble(spec, deg, u, t)
    assert rel(bd.cpu().numpy(), bo) < 1e-12
    # Jacobian, raw and as a Newton matrix
    J = E.jacobian_matrix()
    E.jacobian(ud, t, J, scale=1.0, add_mass=False)
    Jo = orc.jacobian(spec, deg, u, t)
    assert rel(J.download(), Jo) < 1e-12
    sc = -0.3 * 0.71
    E.jacobian(ud, t, J, scale=sc, add_mass=True)
    Bo = sc * Jo
    Bo[orc.rowptr[:-1]] += Mo
    assert rel(J.download(), Bo) < 1e-12
    # L2 norm
    nrm = E.l2norm(ud)
    assert rel(nrm, orc.l2norm(deg, spec.nc, u)) < 1e-12
    # the Newton matrix applied at the quadrature points against the oracle's assembled
    # matrix times v and against the BSR matvec of the GPU-assembled matrix
    v = np.random.default_rng(7 + deg).standard_normal(u.shape)
    vd = _dev(cuda, v)
    yd = _dev(cuda, np.zeros_like(u))
    E.newton_apply(ud, t, sc, vd, yd)
    bl = E.bl
    yo = np.einsum("scr,sc->sr", Bo, v.reshape(-1, bl)[orc.colind])
    yo = np.add.reduceat(yo, orc.rowptr[:-1], axis=0).ravel()
    assert rel(yd.cpu().numpy(), yo) < 1e-12
    zd = _dev(cuda, np.zeros_like(u))
    J.matvec(vd, zd)
    assert rel(yd.cpu().numpy(), zd.cpu().numpy()) < 1e-12
    # the same in two parts: prepare for (u, t), then products for two scales
    E.newton_prepare(ud, t)
    for s2 in (sc, 2.0 * sc):
        E.newton_apply_prepared(s2, vd, zd)
        yo2 = np.einsum("scr,sc->sr", s2 * Jo, v.reshape(-1, bl)[orc.colind])
        yo2 = np.add.reduceat(yo2, orc.rowptr[:-1], axis=0)
        yo2 += np.einsum("icr,ic->ir", Mo, v.reshape(-1, bl))
        assert rel(zd.cpu().numpy(), yo2.ravel()) < 1e-12


def _golden():
    import glob
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return sorted(glob.glob(os.path.join(here, "literal_*.npz")))


@pytest.mark.parametrize("path", _golden(), ids=lambda p: p.rsplit("/", 1)[-1])
def test_assembly_matches_literal_fixtures(cuda, path):
    """The GPU mass matrix, residual and Jacobian in reference mode (the reference's periodic
    trace and Jacobian form) against the golden vectors of the literal restatement of the
    reference's loops (p = 1, 2, 3; advection and Euler)."""
    from cpolyg_amd.mesh import PolyMesh
    z = np.load(path)
    m = PolyMesh(z["v"], z["pptr"], z["pv"], z["p2p"], z["bc"], z["tptr"], z["txy"], z["bb"])
    deg = int(z["deg"])
    spec = EquationSpec.euler() if int(z["kind"]) == 1 else EquationSpec(kind=0, beta=tuple(z["beta"]))
    E = hip.Equation(hip.Mesh(m), spec, deg)
    rp, ci = E.pattern()
    np.testing.assert_array_equal(rp, z["rowptr"])
    np.testing.assert_array_equal(ci, z["colind"])
    assert rel(E.mass(), z["mass"]) < 1e-13
    ud = _dev(cuda, z["u"])
    bd = _dev(cuda, np.zeros_like(z["u"]))
    E.residual(ud, 0.0, bd)
    assert rel(bd.cpu().numpy(), z["residual"]) < 1e-12
    J = E.jacobian_matrix()
    E.jacobian(ud, 0.0, J, scale=1.0, add_mass=False)
    assert rel(J.download(), z["jacobian"]) < 1e-12


def test_mass_matvec(cuda):
    m = vortex_mesh(8)
    deg = 3
    E = hip.Equation(hip.Mesh(m), EquationSpec.euler(), deg)
    x = np.random.default_rng(0).standard_normal(m.n_cells * 40)
    xd, yd = _dev(cuda, x), _dev(cuda, np.zeros_like(x))
    E.mass_matvec(xd, yd)
    Mo = O.Oracle(m).mass(deg, 4)
    y = np.einsum("icr,ic->ir", Mo, x.reshape(-1, 40)).ravel()
    assert rel(yd.cpu().numpy(), y) < 1e-13


def test_free_stream_preserved(cuda):
    m = vortex_mesh(8)
    deg = 3
    E = hip.Equation(hip.Mesh(m), EquationSpec.euler(), deg)
    U0 = np.array([1.0, 1.0, 0.5, 1.0 / 0.4 + 0.5 * 1.25])
    u = setups.interp(m, deg, lambda x, y: np.broadcast_to(U0, x.shape + (4,)), 4)
    ud, bd = _dev(cuda, u), _dev(cuda, np.zeros_like(u))
    E.residual(ud, 0.0, bd)
    assert np.abs(bd.cpu().numpy()).max() < 1e-11


# ---- analytic fields: MeshFn::interp and MeshFn::L2Error on the device ---------------------
def _np_field(kind, par, t, x, y):
    """The fields of cpg_interp_field, in numpy (nearest periodic image where Lx, Ly > 0)."""
    from cpolyg_amd.spec import FIELD_GAUSSIAN, FIELD_VORTEX, isentropic_vortex

    def wrap(d, L):
        return d - L * np.rint(d / L) if L > 0 else d

    if kind == FIELD_GAUSSIAN:
        cx, cy, a, vx, vy, Lx, Ly = par
        dx, dy = wrap(x - cx - vx * t, Lx), wrap(y - cy - vy * t, Ly)
        return np.exp(-a * (dx * dx + dy * dy))[..., None]
    if kind == FIELD_VORTEX:
        beta, x0, y0, uinf, vinf, gam, Lx, Ly = par
        # the centre's nearest image, then the unwrapped formula at t = 0
        dx, dy = wrap(x - x0 - uinf * t, Lx), wrap(y - y0 - vinf * t, Ly)
        return isentropic_vortex(dx, dy, 0.0, beta, 0.0, 0.0, uinf, vinf, gam)
    ux, uy, p, gam, amp = par
    rho = 1 + amp * np.sin(2 * np.pi * (x - ux * t)) * np.sin(2 * np.pi * (y - uy * t))
    return np.stack([rho, rho * ux, rho * uy, p / (gam - 1) + 0.5 * rho * (ux * ux + uy * uy) + 0 * rho], axis=-1)


@pytest.mark.parametrize("kind,deg", [(0, 2), (0, 4), (1, 3), (1, 1), (2, 4), (2, 2)])
def test_interp_and_l2_error_match_oracle(cuda, kind, deg):
    """Projection of an analytic field (and of caller values at the fitting points) and the L2
    error against the field at a later time equal the oracle's restatement (Householder least
    squares, its own sub-triangle rule) on the same mesh."""
    import torch
    from cpolyg_amd.spec import FIELD_GAUSSIAN, FIELD_VORTEX

    if kind == FIELD_GAUSSIAN:
        m = unit_periodic_mesh(8)
        spec = EquationSpec.reference_advection()
        par = (0.5, 0.4, 60.0, 0.9, -0.6, 1.0, 1.0)   # crosses the periodic boundary by t = 0.7
    elif kind == FIELD_VORTEX:
        m = vortex_mesh(8)
        spec = EquationSpec.euler()
        par = (5.0, 5.0, 5.0, 1.0, 1.0, 1.4, 10.0, 10.0)
    else:
        m = unit_periodic_mesh(8)
        spec = EquationSpec.euler()
        par = (1.0, 0.5, 1.0, 1.4, 0.2)
    E = hip.Equation(hip.Mesh(m), spec, deg)
    orc = O.Oracle(m)
    nc, bs, n = spec.nc, bs_of(deg), m.n_cells
    Q = (deg + 1) ** 2
    xy = torch.empty(n * Q * 2, dtype=torch.float64, device=cuda)
    E.interp_points(xy)
    pts = orc.interp_points(deg)
    assert np.abs(xy.cpu().numpy().reshape(n, Q, 2) - pts).max() < 1e-13
    t = 0.7
    u = torch.empty(n * nc * bs, dtype=torch.float64, device=cuda)
    E.interp_field(kind, par, t, u)
    vals = _np_field(kind, par, t, pts[..., 0], pts[..., 1])
    uo = orc.interp_values(deg, nc, vals)
    assert rel(u.cpu().numpy(), uo) < 1e-12
    u2 = torch.empty_like(u)
    E.interp_values(_dev(cuda, vals), u2)
    assert rel(u2.cpu().numpy(), uo) < 1e-12
    for t2 in (t, t + 0.05):
        q = orc.quad_points(2 * deg)
        eo = orc.l2_error_values(deg, nc, uo, _np_field(kind, par, t2, q[:, 0], q[:, 1]))
        eg = E.l2_error_field(u, kind, par, t2)
        assert np.all(np.abs(eg - eo) <= 1e-9 * eo + 1e-14), (t2, eg, eo)
