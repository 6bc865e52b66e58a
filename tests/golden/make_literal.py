"""Writes tests/golden/literal_*.npz: golden vectors from a literal pure-Python
restatement of the reference's loop structure on tiny meshes.

The C oracle factors the quadrature (one state evaluation per point, derivatives of
the basis evaluated directly).  This script instead follows the reference step by
step, so index, sign or transposition mistakes in the factored oracle show up:
  * Leg2D as a double sum over the coefficient triangle (src/Legendre.h:51-68),
  * derivative series through LegDerX / LegDerY (src/Legendre.cpp:34-124),
  * Equation::assemble: one polygon and one line quadrature per test function
    (src/Equation.cpp:152-191, integrands :20-57),
  * Equation::jacobian: one quadrature per (trial, test) pair, results scattered to
    rows componentIndices+k, columns componentIndices+j (src/Equation.cpp:193-279),
  * periodic traces through BoundaryConditions::boundaryValue with
    m = int(0.5*(sqrt(4*n_rows+1)-1)) (src/BoundaryConditions.cpp:7-37),
  * Euler fluxes and alphaDerivative (eqns/Euler.cpp), first index on ties,
  * MassMatrix (src/MassMatrix.cpp:30-94), BlockILU0 (src/Preconditioners.cpp:49-158)
    with dense inverses, and BlockMatrix::gmres (src/BlockMatrix.cpp:79-163).
Legendre polynomials come from the Bonnet recurrence (the reference calls
gsl_sf_legendre_Pl).  Run:  python tests/golden/make_literal.py  (p <= 2 cases), then
python tests/golden/make_literal.py p3  (the p = 3 cases, minutes of pure Python)
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cpolyg_amd import setups  # noqa: E402
from cpolyg_amd.mesh import box_voronoi, jittered_lattice_seeds, periodic_rectangle, periodic_voronoi, uniform_seeds  # noqa: E402

# ---- quadrature tables (src/Quadrature.cpp:103-171) ---------------------------------
LIN = {
    2: [(-0.5773502691896257, 1.0), (0.5773502691896257, 1.0)],
    4: [(-0.3399810435848563, 0.6521451548625461), (0.3399810435848563, 0.6521451548625461),
        (-0.8611363115940526, 0.3478548451374538), (0.8611363115940526, 0.3478548451374538)],
    6: [(0.6612093864662645, 0.3607615730481386), (-0.6612093864662645, 0.3607615730481386),
        (-0.2386191860831969, 0.4679139345726910), (0.2386191860831969, 0.4679139345726910),
        (-0.9324695142031521, 0.1713244923791704), (0.9324695142031521, 0.1713244923791704)],
    10: [(0.1488743389816312, 0.2955242247147529), (-0.1488743389816312, 0.2955242247147529),
         (-0.4333953941292472, 0.2692667193099963), (0.4333953941292472, 0.2692667193099963),
         (-0.6794095682990244, 0.2190863625159820), (0.6794095682990244, 0.2190863625159820),
         (-0.8650633666889845, 0.1494513491505806), (0.8650633666889845, 0.1494513491505806),
         (-0.9739065285171717, 0.0666713443086881), (0.9739065285171717, 0.0666713443086881)],
}
TRI = {
    2: [(0.16666666666667, 0.16666666666667, 0.33333333333333), (0.16666666666667, 0.66666666666667, 0.33333333333333),
        (0.66666666666667, 0.16666666666667, 0.33333333333333)],
    4: [(0.44594849091597, 0.44594849091597, 0.22338158967801), (0.44594849091597, 0.10810301816807, 0.22338158967801),
        (0.10810301816807, 0.44594849091597, 0.22338158967801), (0.09157621350977, 0.09157621350977, 0.10995174365532),
        (0.09157621350977, 0.81684757298046, 0.10995174365532), (0.81684757298046, 0.09157621350977, 0.10995174365532)],
    6: [(0.24928674517091, 0.24928674517091, 0.11678627572638), (0.24928674517091, 0.50142650965818, 0.11678627572638),
        (0.50142650965818, 0.24928674517091, 0.11678627572638), (0.06308901449150, 0.06308901449150, 0.05084490637021),
        (0.06308901449150, 0.87382197101700, 0.05084490637021), (0.87382197101700, 0.06308901449150, 0.05084490637021),
        (0.31035245103378, 0.63650249912140, 0.08285107561837), (0.63650249912140, 0.05314504984482, 0.08285107561837),
        (0.05314504984482, 0.31035245103378, 0.08285107561837), (0.63650249912140, 0.31035245103378, 0.08285107561837),
        (0.31035245103378, 0.05314504984482, 0.08285107561837), (0.05314504984482, 0.63650249912140, 0.08285107561837)],
}


def rules(deg2):
    """Quadrature::polygonIntegral / lineIntegral selection (src/Quadrature.cpp:20-27, 74-81):
    deg <= 2, 4, 6 take the 3-, 6-, 12-point triangle and 2-, 4-, 6-point line rules (the
    25-point / 10-point pair, for deg > 6, is not needed by the p <= 3 fixtures)."""
    k = 2 if deg2 <= 2 else 4 if deg2 <= 4 else 6 if deg2 <= 6 else None
    if k is None:
        raise ValueError(f"rule of degree {deg2} is not tabulated here")
    return TRI[k], LIN[k]


def Pl(l, x):
    p0, p1 = 1.0, x
    if l == 0:
        return 1.0
    for n in range(1, l):
        p0, p1 = p1, ((2 * n + 1) * x * p1 - n * p0) / (n + 1)
    return p1


def Leg2D(x, y, m, c):
    val = 0.0
    k = 0
    for i in range(m):
        for j in range(m - i):
            val += c[k] * Pl(i, x) * Pl(j, y)
            k += 1
    return val


def LegDerX(m, a_c):
    c = list(a_c)
    cp = [0.0] * (m * (m + 1) // 2)
    for i in range(m - 1, 2, -1):
        for j in range(m - i):
            cp[-(i - 1) * (i - 2 * (1 + m)) // 2 + j] = (2 * i - 1) * c[-i * (i - 1 - 2 * m) // 2 + j]
            c[-(i - 2) * (i - 3 - 2 * m) // 2 + j] += c[-i * (i - 1 - 2 * m) // 2 + j]
    if m > 2:
        for j in range(m - 2):
            cp[m + j] = 3 * c[2 * m - 1 + j]
    for j in range(m - 1):
        cp[j] = c[m + j]
    return cp


def LegDerY(m, a_c):
    c = list(a_c)
    cp = [0.0] * (m * (m + 1) // 2)
    idx = lambda i: -i * (i - 1 - 2 * m) // 2  # noqa: E731
    for j in range(m - 1, 2, -1):
        for i in range(m - j):
            cp[idx(i) + j - 1] = (2 * j - 1) * c[idx(i) + j]
            c[idx(i) + j - 2] += c[idx(i) + j]
    if m > 2:
        for i in range(m - 2):
            cp[idx(i) + 1] = 3 * c[idx(i) + 2]
    for i in range(m - 1):
        cp[idx(i)] = c[idx(i) + 1]
    return cp


class Msh:
    def __init__(self, m):
        self.m = m
        self.np = m.n_cells
        self.v = m.v
        self.p = [list(m.pv[m.pptr[i]:m.pptr[i + 1]]) for i in range(self.np)]
        self.p2p = [list(m.p2p[m.pptr[i]:m.pptr[i + 1]]) for i in range(self.np)]
        self.bc = {-1 - k: row for k, row in enumerate(m.bc)}
        self.bb = m.bb
        self.tri = [m.txy[m.tptr[i]:m.tptr[i + 1]] for i in range(self.np)]

    def local(self, i, x, y):
        b = self.bb[i]
        return 2.0 * (x - b[0]) / b[2] - 1, 2.0 * (y - b[1]) / b[3] - 1

    def normal(self, i, a, b):
        pts = self.v[self.p[i]]
        cx, cy = pts[:, 0].mean(), pts[:, 1].mean()
        dx, dy = self.v[a][0] - self.v[b][0], self.v[a][1] - self.v[b][1]
        ln = math.sqrt(dx * dx + dy * dy)
        sv = float(-dy * (self.v[a][0] - cx) + dx * (self.v[a][1] - cy))
        s = int(0 < sv) - int(sv < 0)
        return -s * dy / ln, s * dx / ln

    def periodic_local(self, b, x, y):
        row = self.bc[b]
        a1, a2, p2 = row[3], row[5], row[2]
        return self.local(p2, x + (self.v[a2][0] - self.v[a1][0]), y + (self.v[a2][1] - self.v[a1][1]))


def polygon_integral(msh, f, i, deg2):
    tri, _ = rules(deg2)
    integ = 0.0
    for T in msh.tri[i]:
        x1, y1, x2, y2, x3, y3 = T
        area = 0.5 * abs(x1 * y2 + x2 * y3 + x3 * y1 - x1 * y3 - x3 * y2 - x2 * y1)
        z = 0.0
        for r, s, w in tri:
            z = z + f(x1 * (1 - r - s) + x2 * r + x3 * s, y1 * (1 - r - s) + y2 * r + y3 * s) * w
        integ = integ + z * area
    return integ


def line_integral(msh, f, a, b, deg2):
    _, lin = rules(deg2)
    x1, y1 = msh.v[a]
    x2, y2 = msh.v[b]
    length = math.sqrt((x2 - x1) ** 2 + (y2 - y1) ** 2)
    integ = 0.0
    for xq, w in lin:
        integ = integ + w * f(x1 + 0.5 * (xq + 1) * (x2 - x1), y1 + 0.5 * (xq + 1) * (y2 - y1))
    return integ * 0.5 * length


# ---- equations ----------------------------------------------------------------------
class Euler:
    nc = 4

    def __init__(self, g=1.4):
        self.g = g

    def flux(self, U):
        g = self.g
        rho, ru, rv, rE = U
        u, v = ru / rho, rv / rho
        P = (g - 1) * (rE - 0.5 * (ru * ru + rv * rv) / rho)
        return (np.array([ru, ru * u + P, ru * v, (rE + P) * u]), np.array([rv, ru * v, rv * v + P, (rE + P) * v]),
                math.sqrt(g * P / rho), u, v)

    def fluxFunction(self, U, x, y):
        fx, fy, *_ = self.flux(U)
        return np.stack([fx, fy], 1)

    def numericalFlux(self, Um, Up, x, y, nx, ny):
        fxm, fym, cm, um, vm = self.flux(Um)
        fxp, fyp, cp, up, vp = self.flux(Up)
        vnm, vnp = um * nx + vm * ny, up * nx + vp * ny
        alpha = max(abs(vnm - cm), abs(vnm), abs(vnm + cm), abs(vnp - cp), abs(vnp), abs(vnp + cp))
        return 0.5 * ((fxp + fxm) * nx + (fyp + fym) * ny + alpha * (Um - Up))

    def fluxJacobian(self, U, x, y):
        g = self.g
        rho = U[0]
        u, v, E = U[1] / rho, U[2] / rho, U[3] / rho
        usq = u * u + v * v
        P = (g - 1) * rho * (E - 0.5 * usq)
        H = E + P / rho
        A = np.array([[0, 1, 0, 0], [-u * u + 0.5 * (g - 1) * usq, (3 - g) * u, -(g - 1) * v, g - 1],
                      [-u * v, v, u, 0], [u * (0.5 * (g - 1) * usq - H), H - (g - 1) * usq, -(g - 1) * u * v, g * u]])
        B = np.array([[0, 0, 1, 0], [-u * v, v, u, 0], [-v * v + 0.5 * (g - 1) * usq, -(g - 1) * u, (3 - g) * v, g - 1],
                      [v * (0.5 * (g - 1) * usq - H), -(g - 1) * u * v, H - (g - 1) * v * v, g * v]])
        return A, B

    def alphaDerivative(self, U1, U2, nx, ny):
        g = self.g
        u1, v1, u2, v2 = U1[1] / U1[0], U1[2] / U1[0], U2[1] / U2[0], U2[2] / U2[0]
        P1 = (g - 1) * (U1[3] - 0.5 * (U1[1] ** 2 + U1[2] ** 2) / U1[0])
        c1 = math.sqrt(g * P1 / U1[0])
        P2 = (g - 1) * (U2[3] - 0.5 * (U2[1] ** 2 + U2[2] ** 2) / U2[0])
        c2 = math.sqrt(g * P2 / U2[0])
        vn1, vn2 = u1 * nx + v1 * ny, u2 * nx + v2 * ny
        eigs = [vn1 - c1, vn1, vn1 + c1, vn2 - c2, vn2, vn2 + c2]
        a = [abs(e) for e in eigs]
        idx = a.index(max(a))
        d = np.zeros(4)
        if idx <= 2:
            rho = U1[0]
            E = U1[3] / rho
            d[:] = [-u1 / rho * nx - v1 / rho * ny, 1 / rho * nx, 1 / rho * ny, 0]
            usq = u1 * u1 + v1 * v1
            aa = math.sqrt(g * g - g)
            b = rho * math.sqrt(4 * E - 2 * usq)
            corr = np.array([aa * (usq - E) / b, -aa * u1 / b, -aa * v1 / b, aa / b])
            if idx == 0:
                d -= corr
            elif idx == 2:
                d += corr
            if eigs[idx] < 0:
                d *= -1
        return max(a), d

    def numericalFluxJacobian(self, U, U2, x, y, nx, ny, sgn):
        A, B = self.fluxJacobian(U, x, y)
        alpha, ap = self.alphaDerivative(U, U2, nx, ny)
        return 0.5 * (A * nx + B * ny + sgn * alpha * np.eye(4) + np.outer(U - U2, ap))


class Advection:
    nc = 1

    def __init__(self, beta):
        self.beta = beta

    def b(self, x, y):
        b = self.beta
        return b[0] + b[1] * x + b[2] * y, b[3] + b[4] * x + b[5] * y

    def fluxFunction(self, U, x, y):
        bx, by = self.b(x, y)
        return np.array([[bx * U[0], by * U[0]]])

    def numericalFlux(self, Um, Up, x, y, nx, ny):
        bx, by = self.b(x, y)
        bn = bx * nx + by * ny
        return np.array([Um[0] * bn if bn > 0 else Up[0] * bn])

    def fluxJacobian(self, U, x, y):
        bx, by = self.b(x, y)
        return np.array([[bx]]), np.array([[by]])

    def numericalFluxJacobian(self, U, U2, x, y, nx, ny, sgn):
        bx, by = self.b(x, y)
        bn = nx * bx + ny * by
        return np.array([[bn if sgn * bn > 0 else 0.0]])


def boundary_value(msh, x, y, U, b, dirichlet):
    """BoundaryConditions::boundaryValue; U is (n_rows, nc)."""
    row = msh.bc[b]
    if not row[0]:
        return np.asarray(dirichlet(x, y), float).reshape(-1)
    xl, yl = msh.periodic_local(b, x, y)
    m = int(0.5 * (math.sqrt(4 * U.shape[0] + 1) - 1))
    return np.array([Leg2D(xl, yl, m, U[:, c]) for c in range(U.shape[1])])


def assemble(msh, eqn, u, deg, dirichlet):
    bs = (deg + 1) * (deg + 2) // 2
    nc = eqn.nc
    m = deg + 1
    b = np.zeros_like(u)
    for i in range(msh.np):
        w, h = msh.bb[i][2], msh.bb[i][3]
        UMinus = u[i]
        for j in range(bs):
            psi = [0.0] * bs
            psi[j] = 1.0
            psi_x = [c * 2.0 / w for c in LegDerX(m, psi)]
            psi_y = [c * 2.0 / h for c in LegDerY(m, psi)]

            def vol(x, y):
                xx, yy = msh.local(i, x, y)
                vars_ = np.array([Leg2D(xx, yy, m, UMinus[:, c]) for c in range(nc)])
                F = eqn.fluxFunction(vars_, x, y)
                return F[:, 0] * Leg2D(xx, yy, m, psi_x) + F[:, 1] * Leg2D(xx, yy, m, psi_y)

            row = polygon_integral(msh, vol, i, 2 * deg)
            nv1 = len(msh.p[i])
            for e in range(nv1):
                a1, b1 = msh.p[i][e], msh.p[i][(e + 1) % nv1]
                nx, ny = msh.normal(i, a1, b1)
                iPlus = msh.p2p[i][e]
                UPlus = u[iPlus] if iPlus >= 0 else (u[msh.bc[iPlus][2]] if msh.bc[iPlus][0] else None)

                def bnd(x, y):
                    xm, ym = msh.local(i, x, y)
                    psiVal = Leg2D(xm, ym, m, psi)
                    vm = np.array([Leg2D(xm, ym, m, UMinus[:, c]) for c in range(nc)])
                    if iPlus < 0:
                        vp = boundary_value(msh, x, y, UPlus, iPlus, dirichlet)
                    else:
                        xp, yp = msh.local(iPlus, x, y)
                        vp = np.array([Leg2D(xp, yp, m, UPlus[:, c]) for c in range(nc)])
                    return psiVal * eqn.numericalFlux(vm, vp, x, y, nx, ny)

                row = row - line_integral(msh, bnd, a1, b1, 2 * deg)
            b[i][j, :] = row
    return b


def jacobian(msh, eqn, f, deg, dirichlet, rowptr):
    bs = (deg + 1) * (deg + 2) // 2
    nc = eqn.nc
    bl = nc * bs
    m = deg + 1
    comp = np.arange(nc) * bs
    blocks = np.zeros((rowptr[-1], bl, bl))   # [slot, row, col]
    for i in range(msh.np):
        w, h = msh.bb[i][2], msh.bb[i][3]
        d = rowptr[i]
        U = f[i]
        for j in range(bs):
            phi = [0.0] * bs
            phi[j] = 1.0
            for k in range(bs):
                psi = [0.0] * bs
                psi[k] = 1.0
                psi_x = [c * 2.0 / w for c in LegDerX(m, psi)]
                psi_y = [c * 2.0 / h for c in LegDerY(m, psi)]

                def volj(x, y):
                    xx, yy = msh.local(i, x, y)
                    vars_ = np.array([Leg2D(xx, yy, m, U[:, c]) for c in range(nc)])
                    A, B = eqn.fluxJacobian(vars_, x, y)
                    return Leg2D(xx, yy, m, phi) * (A * Leg2D(xx, yy, m, psi_x) + B * Leg2D(xx, yy, m, psi_y))

                blocks[d][np.ix_(comp + k, comp + j)] += polygon_integral(msh, volj, i, 2 * deg)
                nv1 = len(msh.p[i])
                blk = d + 1
                for e in range(nv1):
                    a1, b1 = msh.p[i][e], msh.p[i][(e + 1) % nv1]
                    nx, ny = msh.normal(i, a1, b1)
                    nbr = msh.p2p[i][e]
                    UN = f[nbr] if nbr >= 0 else (f[msh.bc[nbr][2]] if msh.bc[nbr][0] else None)

                    def bndj(x, y, iPhi):
                        xs, ys = msh.local(i, x, y)
                        psiVal = Leg2D(xs, ys, m, psi)
                        vars_ = np.array([Leg2D(xs, ys, m, U[:, c]) for c in range(nc)])
                        if iPhi >= 0:
                            xp, yp = msh.local(iPhi, x, y)
                            phiVal = Leg2D(xp, yp, m, phi)
                        else:
                            phiVal = boundary_value(msh, x, y, np.array(phi)[:, None], iPhi, dirichlet)[0]
                        if nbr >= 0:
                            xn, yn = msh.local(nbr, x, y)
                            vars2 = np.array([Leg2D(xn, yn, m, UN[:, c]) for c in range(nc)])
                        else:
                            vars2 = boundary_value(msh, x, y, UN, nbr, dirichlet)
                        sgn = 1 if iPhi == i else -1
                        return phiVal * psiVal * eqn.numericalFluxJacobian(vars_, vars2, x, y, nx, ny, sgn)

                    if nbr >= 0 or msh.bc[nbr][0]:
                        blocks[blk][np.ix_(comp + k, comp + j)] -= line_integral(msh, lambda x, y: bndj(x, y, nbr), a1, b1, 2 * deg)
                        blk += 1
                    blocks[d][np.ix_(comp + k, comp + j)] -= line_integral(msh, lambda x, y: bndj(x, y, i), a1, b1, 2 * deg)
    return blocks


def mass(msh, deg, nc):
    bs = (deg + 1) * (deg + 2) // 2
    m = deg + 1
    out = np.zeros((msh.np, nc * bs, nc * bs))
    for i in range(msh.np):
        for j in range(bs):
            phi = [0.0] * bs
            phi[j] = 1.0
            for k in range(j + 1):
                psi = [0.0] * bs
                psi[k] = 1.0

                def prod(x, y):
                    xx, yy = msh.local(i, x, y)
                    return Leg2D(xx, yy, m, phi) * Leg2D(xx, yy, m, psi)

                integ = polygon_integral(msh, prod, i, 2 * deg)
                for c in range(nc):
                    out[i, j + c * bs, k + c * bs] = integ
                    out[i, k + c * bs, j + c * bs] = integ
    return out


def block_ilu0(rowptr, colind, blocks):
    """BlockILU0::BlockILU0 with blocks[s] as (row, col) matrices; returns DD, L, DU in A's slots."""
    n = len(rowptr) - 1
    AD = {i: blocks[next(s for s in range(rowptr[i], rowptr[i + 1]) if colind[s] == i)].copy() for i in range(n)}
    AO = {s: blocks[s].copy() for s in range(rowptr[-1])}
    out = np.zeros_like(blocks)
    for i in range(n):
        D = np.linalg.inv(AD[i])
        out[next(s for s in range(rowptr[i], rowptr[i + 1]) if colind[s] == i)] = D
        for k in range(rowptr[i], rowptr[i + 1]):
            j = colind[k]
            if j > i:
                for k2 in range(rowptr[j], rowptr[j + 1]):
                    if colind[k2] == i:
                        AO[k2] = AO[k2] @ D
                        AD[j] = AD[j] - AO[k2] @ AO[k]
                        AO[k] = D @ AO[k]
                        break
    for s in range(rowptr[-1]):
        i = np.searchsorted(rowptr, s, side="right") - 1
        if colind[s] != i:
            out[s] = AO[s]
    return out


def gmres(rowptr, colind, blocks, F, b, x, m, tol, maxit):
    """BlockMatrix::gmres with BlockILU0::solve as the preconditioner."""
    n = len(rowptr) - 1
    bl = blocks.shape[1]

    def matvec(v):
        out = np.zeros_like(v)
        for i in range(n):
            for s in range(rowptr[i], rowptr[i + 1]):
                j = colind[s]
                out[i * bl:(i + 1) * bl] += blocks[s] @ v[j * bl:(j + 1) * bl]
        return out

    def pc(v):
        v = v.copy()
        for i in range(n):
            for s in range(rowptr[i], rowptr[i + 1]):
                j = colind[s]
                if j < i:
                    v[i * bl:(i + 1) * bl] -= F[s] @ v[j * bl:(j + 1) * bl]
        for i in range(n):
            s = next(s for s in range(rowptr[i], rowptr[i + 1]) if colind[s] == i)
            v[i * bl:(i + 1) * bl] = F[s] @ v[i * bl:(i + 1) * bl]
        for i in range(n - 1, -1, -1):
            for s in range(rowptr[i], rowptr[i + 1]):
                j = colind[s]
                if j > i:
                    v[i * bl:(i + 1) * bl] -= F[s] @ v[j * bl:(j + 1) * bl]
        return v

    N = n * bl
    r = pc(b.copy())
    nrm2b = np.linalg.norm(r)
    if maxit % m > 0:
        maxit = (maxit // m + 1) * m
    V = np.zeros((m + 1, N))
    H = np.zeros((m + 1, m))
    c = np.zeros(m)
    s = np.zeros(m)
    y = np.zeros(m + 1)
    i = 0
    for jj in range(maxit // m):
        r = pc(matvec(x) - b)
        beta = np.linalg.norm(r)
        V[0] = r / beta
        y[:] = 0
        y[0] = beta
        for i in range(m):
            if abs(y[i]) < tol * nrm2b:
                break
            w = pc(matvec(V[i]))
            h = V[:i + 1] @ w
            w = w - V[:i + 1].T @ h
            hnew = np.linalg.norm(w)
            V[i + 1] = w / hnew
            for k in range(i):
                tmp = c[k] * h[k] - s[k] * h[k + 1]
                h[k + 1] = s[k] * h[k] + c[k] * h[k + 1]
                h[k] = tmp
            rd = h[i]
            dd = math.sqrt(rd * rd + hnew * hnew)
            c[i] = rd / dd
            s[i] = -hnew / dd
            h[i] = dd
            H[:i + 1, i] = h
            y[i + 1] = s[i] * y[i]
            y[i] = c[i] * y[i]
        else:
            i = m
        yy = np.linalg.solve(np.triu(H[:i, :i]), y[:i]) if i > 0 else np.zeros(0)
        x = x - V[:i].T @ yy
        if abs(y[i]) < tol * nrm2b:
            break
    return x, m * jj + i


# ---- cases --------------------------------------------------------------------------
def mesh_arrays(m):
    return dict(v=m.v, pptr=m.pptr, pv=m.pv, p2p=m.p2p, bc=m.bc, tptr=m.tptr, txy=m.txy, bb=m.bb)


def run_case(name, m, eqn, deg, u, dirichlet, spec):
    import oracle as O

    msh = Msh(m)
    bs = (deg + 1) * (deg + 2) // 2
    nc = eqn.nc
    U = u.reshape(m.n_cells, nc, bs).transpose(0, 2, 1)   # (np, bs, nc) like arma::cube slices
    b = assemble(msh, eqn, U, deg, dirichlet)
    rowptr, colind = m.bsr_pattern()
    J = jacobian(msh, eqn, U, deg, dirichlet, rowptr)
    M = mass(msh, deg, nc)
    res = b.transpose(0, 2, 1).reshape(-1)
    # Newton matrix and its DILU / GMRES (blocks as (row, col))
    dt, a = 0.2, 0.7
    B = -dt * a * J
    for i in range(m.n_cells):
        B[rowptr[i]] += M[i]
    F = block_ilu0(rowptr, colind, B)
    x, iters = gmres(rowptr, colind, B, F, res, np.zeros_like(res), 10, 1e-10, 400)
    out = dict(mesh_arrays(m), deg=deg, u=u, residual=res, jacobian=J.transpose(0, 2, 1), mass=M.transpose(0, 2, 1),
               newton=B.transpose(0, 2, 1), dilu=F.transpose(0, 2, 1), gmres_x=x, gmres_iters=iters,
               rowptr=rowptr, colind=colind, kind=spec["kind"], beta=np.array(spec.get("beta", [0] * 6), float),
               trace=0)
    np.savez_compressed(os.path.join(HERE, f"literal_{name}.npz"), **out)
    print(name, m.n_cells, "cells: residual", np.abs(res).max(), "gmres iters", iters)


def main(which=None):
    rng = np.random.default_rng(11)
    if which == "p3":
        # 5) periodic 3x3 quads, constant advection p=3 (12-point triangle, 6-point line rules)
        m = periodic_rectangle(3, 3, 1.0, 1.0)
        u = rng.standard_normal(m.n_cells * 10)
        run_case("adv_quad3_p3", m, Advection((1.0, 0, 0, 0.5, 0, 0)), 3, u, lambda x, y: [0.0],
                 dict(kind=0, beta=[1.0, 0, 0, 0.5, 0, 0]))
        # 6) periodic 3x3 quads, Euler density wave p=3 (reference trace and reference Jacobian)
        u = setups.interp(m, 3, setups.sine_density(), 4)
        d = rng.standard_normal(u.shape) * 0.01
        d.reshape(-1, 4, 10)[:, :, 0] = 0
        run_case("euler_quad3_p3", m, Euler(1.4), 3, u + d, None, dict(kind=1))
        return
    # 1) periodic 3x3 quads (reference generator), constant advection p=2
    m = periodic_rectangle(3, 3, 1.0, 1.0)
    u = rng.standard_normal(m.n_cells * 6)
    run_case("adv_quad3_p2", m, Advection((1.0, 0, 0, 0.5, 0, 0)), 2, u, lambda x, y: [0.0],
             dict(kind=0, beta=[1.0, 0, 0, 0.5, 0, 0]))
    # 2) box Voronoi, reference advection velocity, Dirichlet zero, p=1
    m = box_voronoi(uniform_seeds(12, 1.0, 1.0, 5), 1.0, 1.0)
    u = rng.standard_normal(m.n_cells * 3)
    run_case("adv_box12_p1", m, Advection((-1.0, 0, 2.0, 1.0, -2.0, 0)), 1, u, lambda x, y: [0.0],
             dict(kind=0, beta=[-1.0, 0, 2.0, 1.0, -2.0, 0]))
    # 3) periodic Voronoi 4x4 on [0,10]^2, Euler vortex p=2 with perturbed modes (reference trace)
    m = periodic_voronoi(jittered_lattice_seeds(4, 4, 10.0, 10.0, 0.3, 7), 10.0, 10.0)
    u = setups.interp(m, 2, setups.vortex(), 4)
    d = rng.standard_normal(u.shape) * 0.02
    d.reshape(-1, 4, 6)[:, :, 0] = 0
    run_case("euler_vor16_p2", m, Euler(1.4), 2, u + d, None, dict(kind=1))
    # 4) periodic 3x3 quads, Euler p=1
    m = periodic_rectangle(3, 3, 1.0, 1.0)
    u = setups.interp(m, 1, setups.sine_density(), 4)
    d = rng.standard_normal(u.shape) * 0.01
    run_case("euler_quad3_p1", m, Euler(1.4), 1, u + d, None, dict(kind=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)
