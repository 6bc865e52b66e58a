// Exercises the C++ mirror of the reference API (include/cpolyg.hpp) the way the
// reference's drivers use theirs (examples/irk.cpp, examples/PaperGMRES.cpp), and
// checks each result against the FP64 oracle.  Prints PASS on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "../../include/cpolyg.hpp"
#include "../../oracle/oracle.h"

using namespace cpolyg::hip;

// periodic Nx x Ny quads on the unit square, the reference's ordering (src/Meshes.cpp:176-321)
static PolyMesh periodic_rectangle(int nx, int ny) {
    PolyMesh m;
    const double hx = 1.0 / nx, hy = 1.0 / ny;
    std::map<std::pair<int, int>, int> vid;
    auto add = [&](int ix, int iy) {
        auto key = std::make_pair(ix, iy);
        auto it = vid.find(key);
        if (it != vid.end()) return it->second;
        int id = (int)m.v.size() / 2;
        m.v.push_back(ix * hx);
        m.v.push_back(iy * hy);
        vid[key] = id;
        return id;
    };
    m.pptr.push_back(0);
    m.tptr.push_back(0);
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) {
            int a = add(ix, iy), b = add(ix + 1, iy), c = add(ix + 1, iy + 1), d = add(ix, iy + 1);
            m.pv.insert(m.pv.end(), {a, b, c, d});
            m.pptr.push_back((int)m.pv.size());
            const double x = ix * hx, y = iy * hy;
            m.txy.insert(m.txy.end(), {x, y, x + hx, y, x + hx, y + hy});
            m.txy.insert(m.txy.end(), {x, y, x + hx, y + hy, x, y + hy});
            m.tptr.push_back(m.tptr.back() + 2);
            m.bb.insert(m.bb.end(), {x, y, hx, hy});
        }
    int i = 0;
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++, i++) {
            auto per = [&](int p2, int a1, int b1, int a2, int b2) {
                m.bc.insert(m.bc.end(), {1, i, p2, a1, b1, a2, b2});
                m.p2p.push_back(-(int)(m.bc.size() / 7));
            };
            if (iy == 0) per(nx * (ny - 1) + ix, vid[{ix, 0}], vid[{ix + 1, 0}], vid[{ix, ny}], vid[{ix + 1, ny}]);
            else m.p2p.push_back(i - nx);
            if (ix == nx - 1) per(i - nx + 1, vid[{nx, iy}], vid[{nx, iy + 1}], vid[{0, iy}], vid[{0, iy + 1}]);
            else m.p2p.push_back(i + 1);
            if (iy == ny - 1) per(ix, vid[{ix, ny}], vid[{ix + 1, ny}], vid[{ix, 0}], vid[{ix + 1, 0}]);
            else m.p2p.push_back(i + nx);
            if (ix == 0) per(i + nx - 1, vid[{0, iy}], vid[{0, iy + 1}], vid[{nx, iy}], vid[{nx, iy + 1}]);
            else m.p2p.push_back(i - 1);
        }
    return m;
}

static double rel(const std::vector<double> &a, const std::vector<double> &b) {
    double d = 0, m = 0;
    for (size_t i = 0; i < a.size(); i++) {
        d = std::fmax(d, std::fabs(a[i] - b[i]));
        m = std::fmax(m, std::fabs(b[i]));
    }
    return d / (m > 0 ? m : 1);
}

int main() {
    const int N = 5, deg = 2, bs = 6;
    PolyMesh msh = periodic_rectangle(N, N);
    ConstantAdvection eqn(msh, 1.0, 0.5, deg);
    const int np = msh.np();
    // Gaussian bump coefficients (cell means only, plus a smooth linear part)
    std::vector<double> u0(np * bs, 0.0);
    for (int i = 0; i < np; i++) {
        const double cx = msh.bb[4 * i] + 0.5 * msh.bb[4 * i + 2], cy = msh.bb[4 * i + 1] + 0.5 * msh.bb[4 * i + 3];
        u0[i * bs] = std::exp(-20 * ((cx - .5) * (cx - .5) + (cy - .5) * (cy - .5)));
        u0[i * bs + 1] = 0.1 * std::sin(7 * cx);
        u0[i * bs + 3] = 0.1 * std::cos(5 * cy);
    }
    MeshFn u(np, deg, 1);
    u.a->upload(u0);

    o_mesh om{np, msh.v.data(), msh.pptr.data(), msh.pv.data(), msh.p2p.data(), (int)msh.bc.size() / 7,
              msh.bc.data(), msh.tptr.data(), msh.txy.data(), msh.bb.data()};
    o_eq oe{};
    oe.kind = O_ADVECTION;
    oe.beta[0] = 1.0;
    oe.beta[3] = 0.5;

    // Equation::assemble
    MeshFn b = eqn.assemble(u, 0.0);
    std::vector<double> bo(np * bs);
    o_assemble(&om, &oe, deg, u0.data(), 0.0, bo.data());
    const double e_res = rel(b.a->download(), bo);

    // Equation::jacobian, B *= -dt; B += M (through the stepper path below); matvec
    Jacobian J = eqn.jacobian(u, 0.0);
    std::vector<double> Jo((size_t)J.nb * 36);
    o_jacobian(&om, &oe, deg, u0.data(), 0.0, J.rowBlock.data(), Jo.data());
    const double e_jac = rel(J.download(), Jo);

    // BlockILU0 + Jacobian::solve on B = M - dt J
    std::vector<double> Mb((size_t)np * 36);
    o_mass(&om, deg, 1, Mb.data());
    const double dt = 0.05;
    std::vector<double> Bh = J.download();
    for (auto &x : Bh) x *= -dt;
    for (int i = 0; i < np; i++)
        for (int e = 0; e < 36; e++) Bh[(size_t)J.rowBlock[i] * 36 + e] += Mb[(size_t)i * 36 + e];
    J.upload(Bh);
    BlockILU0 pc(J);
    MeshFn x = J.solve(b, pc, 20, 1e-12, 2000);
    std::vector<double> xo(np * bs, 0.0), F(Bh.size());
    o_dilu_factor(np, 6, J.rowBlock.data(), J.colIndices.data(), Bh.data(), F.data());
    double tol = 1e-12;
    int it = 2000;
    o_gmres(np, 6, J.rowBlock.data(), J.colIndices.data(), Bh.data(), O_PC_DILU, F.data(), nullptr, bo.data(),
            xo.data(), 20, &tol, &it, O_GMRES_LEFT_REF);
    const double e_sol = rel(x.a->download(), xo);

    // BlockMatrix::jacobi with the block-Jacobi preconditioner of the same B
    BlockJacobi bj(J);
    MeshFn xj(np, deg, 1);
    xj.a->upload(std::vector<double>(np * bs, 0.0));
    double jtol = 1e-10;
    int jit = 1000;
    J.jacobi(b.data(), xj.data(), jtol, jit, bj);
    std::vector<double> LU(Bh.size()), xjo(np * bs, 0.0);
    std::vector<int> piv(np * bs);
    o_bjacobi_factor(np, 6, J.rowBlock.data(), J.colIndices.data(), Bh.data(), LU.data(), piv.data());
    double jtolo = 1e-10;
    int jito = 1000;
    o_jacobi_solve(np, 6, J.rowBlock.data(), J.colIndices.data(), Bh.data(), LU.data(), piv.data(), bo.data(),
                   xjo.data(), &jtolo, &jito);
    const double e_jacobi = rel(xj.a->download(), xjo);

    // DIRK::advance
    cpg_solver_params sp = reference_solver();
    sp.rtol = 1e-12;
    sp.maxit = 2000;
    sp.newton_tol = 1e-13;
    sp.newton_maxit = 10;
    DIRK dirk(3, eqn, sp);
    MeshFn u1 = dirk.advance(u, dt, 0.0);
    o_solver os{O_GMRES_LEFT_REF, 20, 1e-12, 2000, O_PC_DILU, 1e-13, 10};
    std::vector<double> uo(np * bs);
    o_stats ost{};
    o_dirk_step(&om, &oe, deg, 3, u0.data(), dt, 0.0, &os, uo.data(), &ost);
    const double e_dirk = rel(u1.a->download(), uo);

    // IRK::advance (coupled Newton system) and IRK::newAdvance (W form): both solve the
    // 2-stage Radau IIA step that the oracle takes
    IRK irk(2, eqn, sp);
    MeshFn ui = irk.advance(u, dt, 0.0);
    MeshFn uw = irk.newAdvance(u, dt, 0.0);
    std::vector<double> uio(np * bs);
    o_stats osi{};
    o_irk_step(&om, &oe, deg, 2, u0.data(), dt, 0.0, &os, uio.data(), &osi);
    const double e_irk = rel(ui.a->download(), uio), e_irkw = rel(uw.a->download(), uio);

    // RK2::advance and ForwardEuler::advance (explicit, one block mass solve per stage)
    RK2 rk2(eqn);
    ForwardEuler fe(eqn);
    std::vector<double> ur2o(np * bs), ufeo(np * bs);
    o_rk2_step(&om, &oe, deg, u0.data(), dt, 0.0, ur2o.data());
    o_forward_euler_step(&om, &oe, deg, u0.data(), dt, 0.0, ufeo.data());
    const double e_rk2 = rel(rk2.advance(u, dt, 0.0).a->download(), ur2o);
    const double e_fe = rel(fe.advance(u, dt, 0.0).a->download(), ufeo);

    // MeshFn::interp of a gaussian bump and MeshFn::L2Error against it, with the oracle's
    // restatement on the same points
    const std::vector<double> gpar = {0.5, 0.5, 20.0, 0.0, 0.0, 1.0, 1.0};
    MeshFn g = eqn.interp(CPG_FIELD_GAUSSIAN, gpar);
    const int Q = (deg + 1) * (deg + 1);
    std::vector<double> xy((size_t)np * Q * 2), gv((size_t)np * Q), go((size_t)np * bs);
    o_interp_points(&om, deg, xy.data());
    auto bump = [](double x, double y) {
        const double dx = x - 0.5 - std::rint(x - 0.5), dy = y - 0.5 - std::rint(y - 0.5);
        return std::exp(-20.0 * (dx * dx + dy * dy));
    };
    for (int e = 0; e < np * Q; e++) gv[e] = bump(xy[2 * e], xy[2 * e + 1]);
    o_interp_values(&om, deg, 1, gv.data(), go.data());
    const double e_interp = rel(g.a->download(), go);
    const double err = eqn.L2Error(g, CPG_FIELD_GAUSSIAN, gpar)[0];
    const int nq = o_quad_count(&om, 2 * deg);
    std::vector<double> qp((size_t)nq * 3), fq(nq);
    o_quad_points(&om, 2 * deg, qp.data());
    for (int e = 0; e < nq; e++) fq[e] = bump(qp[3 * e], qp[3 * e + 1]);
    double erro = 0.0;
    o_l2_error_values(&om, deg, 1, go.data(), fq.data(), &erro);
    const double e_l2 = std::abs(err - erro) / erro;

    // MeshFn arithmetic and MeshFn::L2Norm: (2 g + b - g) / 2 - g / 2 = b / 2, and the norm of
    // g carries its equation; a copy keeps its value when the original is updated in place
    MeshFn h = (2.0 * g + b - g) / 2.0 - g * 0.5;
    MeshFn gc = g;
    g += b;
    std::vector<double> hw = b.a->download(), gw = g.a->download(), gcw = gc.a->download();
    for (auto &v : hw) v *= 0.5;
    const double e_arith = rel(h.a->download(), hw) + rel(gcw, go);
    std::vector<double> gsum = go;
    for (size_t e = 0; e < gsum.size(); e++) gsum[e] += b.a->download()[e];
    const double e_inplace = rel(gw, gsum);
    // the L2 norm of the projected bump: sqrt of sum_i c_i^T M_i c_i with the oracle's mass blocks
    std::vector<double> Mg((size_t)np * bs * bs);
    o_mass(&om, deg, 1, Mg.data());
    double nrm = 0.0;
    for (int i = 0; i < np; i++)
        for (int r = 0; r < bs; r++)
            for (int c = 0; c < bs; c++)
                nrm += go[(size_t)i * bs + r] * Mg[(size_t)i * bs * bs + (size_t)c * bs + r] * go[(size_t)i * bs + c];
    const double e_norm = std::abs(gc.L2Norm()[0] - std::sqrt(nrm)) / std::sqrt(nrm);

    std::printf("residual %.3e jacobian %.3e gmres(it %d vs %d) %.3e dirk3 %.3e (newton %d vs %d) interp %.3e "
                "l2error %.3e irk2 %.3e irk2-w %.3e rk2 %.3e forward-euler %.3e jacobi(it %d vs %d) %.3e\n", e_res, e_jac,
                J.last_iterations, it, e_sol, e_dirk, dirk.last_stats.newton_iters, ost.newton_iters, e_interp, e_l2,
                e_irk, e_irkw, e_rk2, e_fe, jit, jito, e_jacobi);
    std::printf("meshfn arithmetic %.3e in-place %.3e l2norm %.3e\n", e_arith, e_inplace, e_norm);
    const bool ok = e_res < 1e-12 && e_jac < 1e-12 && e_sol < 1e-9 && e_dirk < 1e-10 &&
                    std::abs(J.last_iterations - it) <= 1 && e_interp < 1e-12 && e_l2 < 1e-9 && e_irk < 1e-10 &&
                    e_irkw < 1e-10 && e_rk2 < 1e-12 && e_fe < 1e-12 &&
                    jit == jito && jit < 1000 && e_jacobi < 1e-9 &&
                    e_arith < 1e-12 && e_inplace < 1e-14 && e_norm < 1e-12;
    std::printf(ok ? "PASS\n" : "FAIL\n");
    return ok ? 0 : 1;
}
