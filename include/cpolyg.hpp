// cpolyg.hpp -- C++ host API of the MI355X implicit DG library, mirroring the
// reference's classes (same names, argument meaning and ownership rules) over the
// C ABI in cpolyg.h.  Header-only; link with libcpolyg_hip.so.
//
//   reference (src/...)                      here (namespace cpolyg::hip)
//   BlockMatrix   BlockMatrix.h:10-77        BlockMatrix   (device BSR)
//   Preconditioner Preconditioners.h:10-15   Preconditioner (virtual solve(double*))
//   BlockJacobi / BlockILU0 :24-56           BlockJacobi / BlockILU0 (+ factor precision)
//   Jacobian      Jacobian.h:17-30           Jacobian : BlockMatrix
//   MeshFn        MeshFn.h:11-84             MeshFn (device coefficient vector)
//   MassMatrix    MassMatrix.h:9-27          MassMatrix
//   Equation      Equation.h:8-67            Equation / Euler / Advection / ConstantAdvection
//   TimeStepper   TimeIntegration.h:9-76     TimeStepper / DIRK / BackwardEuler / RK4 / RK2 /
//                                            ForwardEuler / IRK
//
// Differences a caller sees: vectors are device pointers; every failing call
// throws cpolyg::hip::Error carrying the C status; the GMRES restart and
// tolerances are parameters instead of the reference's constants.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cpolyg.h"

namespace cpolyg {
namespace hip {

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};

inline void check(int st) {
    if (st != CPG_OK) throw Error(st, std::string(cpg_status_string(st)) + ": " + cpg_last_error());
}

// ---- device memory -----------------------------------------------------------------
class DeviceArray {
  public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) : n_(n) { check(cpg_device_alloc(reinterpret_cast<void **>(&p_), n * sizeof(double))); }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DeviceArray &operator=(DeviceArray &&o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~DeviceArray() { if (p_) cpg_device_free(p_); }
    double *data() { return p_; }
    const double *data() const { return p_; }
    size_t size() const { return n_; }
    void upload(const std::vector<double> &h) { check(cpg_memcpy_h2d(p_, h.data(), h.size() * sizeof(double), nullptr)); check(cpg_stream_sync(nullptr)); }
    std::vector<double> download() const {
        std::vector<double> h(n_);
        check(cpg_memcpy_d2h(h.data(), p_, n_ * sizeof(double), nullptr));
        return h;
    }

  private:
    double *p_ = nullptr;
    size_t n_ = 0;
};

// ---- mesh --------------------------------------------------------------------------
// PolyMesh's public fields (src/PolyMesh.h:22-28), flattened.
struct PolyMesh {
    std::vector<double> v;                 // nv x 2
    std::vector<int> pptr, pv, p2p;        // polygons (CSR), neighbour per edge
    std::vector<int> bc;                   // nbc x 7: periodic,p1,p2,a1,b1,a2,b2
    std::vector<int> tptr;                 // sub-triangles per polygon
    std::vector<double> txy;               // ntri x 6
    std::vector<double> bb;                // np x 4 (may be empty: computed)
    int np() const { return (int)pptr.size() - 1; }

    cpg_mesh handle() const {
        if (!h_) {
            cpg_mesh h = nullptr;
            check(cpg_mesh_create(np(), (int)v.size() / 2, v.data(), pptr.data(), pv.data(), p2p.data(),
                                  (int)bc.size() / 7, bc.data(), tptr.data(), txy.data(), bb.empty() ? nullptr : bb.data(),
                                  &h));
            h_ = std::shared_ptr<cpg_mesh_s>(h, [](cpg_mesh m) { cpg_mesh_destroy(m); });
        }
        return h_.get();
    }

  private:
    mutable std::shared_ptr<cpg_mesh_s> h_;
};

// ---- MeshFn ------------------------------------------------------------------------
// arma::cube(bs, nc, np) coefficients, here a device vector u[i*bl + c*bs + k]
// Copies share the device vector; the arithmetic below never writes into a shared vector (+=
// and -= rebind to a new one), so a copy keeps its value as the reference's would.  `eq` is the
// equation whose mesh and basis the coefficients belong to (set by Equation's results, carried
// by the arithmetic); L2Norm needs it.
struct MeshFn {
    int deg = 0, nc = 1, np = 0;
    std::shared_ptr<DeviceArray> a;
    std::shared_ptr<cpg_equation_s> eq;
    MeshFn() = default;
    MeshFn(int a_np, int a_deg, int a_nc) : deg(a_deg), nc(a_nc), np(a_np),
        a(std::make_shared<DeviceArray>((size_t)a_np * a_nc * (a_deg + 1) * (a_deg + 2) / 2)) {}
    int bs() const { return (deg + 1) * (deg + 2) / 2; }
    size_t size() const { return (size_t)np * nc * bs(); }
    double *data() { return a->data(); }
    const double *data() const { return a->data(); }

    // MeshFn::L2Norm (src/MeshFn.cpp:299-312): one value per component
    std::vector<double> L2Norm() const {
        if (!eq) throw Error(CPG_ERR_ARG, "L2Norm: the coefficients carry no equation");
        std::vector<double> out(nc);
        check(cpg_l2norm(eq.get(), data(), out.data(), nullptr));
        return out;
    }
    // operator+, -, * (src/MeshFn.cpp:209-236), / and the in-place forms (src/MeshFn.h:59-79)
    MeshFn operator+(const MeshFn &o) const { return axpby(1.0, &o, 1.0); }
    MeshFn operator-(const MeshFn &o) const { return axpby(1.0, &o, -1.0); }
    MeshFn operator*(double s) const { return axpby(s, nullptr, 0.0); }
    MeshFn operator/(double s) const { return (*this) * (1.0 / s); }
    MeshFn &operator+=(const MeshFn &o) { return *this = *this + o; }
    MeshFn &operator-=(const MeshFn &o) { return *this = *this - o; }

  private:
    MeshFn axpby(double ca, const MeshFn *o, double cb) const {
        if (o && (o->deg != deg || o->nc != nc || o->np != np)) throw Error(CPG_ERR_ARG, "MeshFn shapes differ");
        MeshFn r(np, deg, nc);
        r.eq = eq ? eq : (o ? o->eq : nullptr);
        check(cpg_vec_axpby(size(), ca, data(), cb, o ? o->data() : nullptr, r.data(), nullptr));
        check(cpg_stream_sync(nullptr));
        return r;
    }
};
inline MeshFn operator*(double s, const MeshFn &fn) { return fn * s; }

// ---- BlockMatrix -------------------------------------------------------------------
class Preconditioner;

class BlockMatrix {
  public:
    int n_rows = 0, bl = 0, nb = 0;
    std::vector<int> colIndices, rowBlock;

    BlockMatrix() = default;
    BlockMatrix(int a_bl, const std::vector<int> &rowptr, const std::vector<int> &colind)
        : n_rows((int)rowptr.size() - 1), bl(a_bl), nb((int)colind.size()), colIndices(colind), rowBlock(rowptr) {
        cpg_bsr h = nullptr;
        check(cpg_bsr_create(n_rows, bl, nb, rowptr.data(), colind.data(), &h));
        adopt(h);
    }
    virtual ~BlockMatrix() = default;

    cpg_bsr handle() const { return h_.get(); }
    void upload(const std::vector<double> &blocks) { check(cpg_bsr_upload(handle(), blocks.data(), nullptr)); check(cpg_stream_sync(nullptr)); }
    std::vector<double> download() const {
        std::vector<double> out((size_t)nb * bl * bl);
        check(cpg_bsr_download(handle(), out.data(), nullptr));
        return out;
    }
    // y = A x (device vectors)
    void matvec(const double *x, double *y, void *stream = nullptr) const { check(cpg_bsr_matvec(handle(), x, y, stream)); }
    // reference semantics: overwrite b with A b (BlockMatrix::matvec(double*), src/BlockMatrix.cpp:28-49)
    void matvec(double *b) const {
        DeviceArray tmp((size_t)n_rows * bl);
        check(cpg_bsr_matvec(handle(), b, tmp.data(), nullptr));
        check(cpg_memcpy_d2d(b, tmp.data(), (size_t)n_rows * bl * sizeof(double), nullptr));
        check(cpg_stream_sync(nullptr));
    }
    BlockMatrix &operator*=(double s) {
        check(cpg_bsr_scale(handle(), s, nullptr));
        return *this;
    }
    // BlockMatrix::gmres (src/BlockMatrix.cpp:79-163): tol, maxIt in/out; x = initial guess in, solution out
    void gmres(const double *b, double *x, int m, double &tol, int &maxIt, Preconditioner &pc,
               int mode = CPG_GMRES_LEFT_REF);
    // BlockMatrix::jacobi (src/BlockMatrix.cpp:51-77): tol (absolute ||A x - b||_2), maxIt in/out
    void jacobi(const double *b, double *x, double &tol, int &maxIt, Preconditioner &pc);

  protected:
    void adopt(cpg_bsr h) { h_ = std::shared_ptr<cpg_bsr_s>(h, [](cpg_bsr b) { cpg_bsr_destroy(b); }); }
    std::shared_ptr<cpg_bsr_s> h_;
};

// ---- preconditioners ---------------------------------------------------------------
class Preconditioner {
  public:
    virtual ~Preconditioner() = default;
    // x <- M^-1 x in place (device vector)
    virtual void solve(double *x) = 0;
    virtual cpg_prec handle() const { return nullptr; }
};

struct NoPreconditioner : Preconditioner {
    void solve(double *) override {}
};

class FactoredPreconditioner : public Preconditioner {
  public:
    FactoredPreconditioner(BlockMatrix &A, int kind, int precision) {
        cpg_prec h = nullptr;
        check(cpg_prec_create(A.handle(), kind, precision, &h));
        h_ = std::shared_ptr<cpg_prec_s>(h, [](cpg_prec p) { cpg_prec_destroy(p); });
        check(cpg_prec_factor(h, A.handle(), nullptr));
    }
    void solve(double *x) override { check(cpg_prec_apply(h_.get(), x, nullptr)); }
    // out = M^-1 in on distinct device vectors (the reference's solve works in place only)
    void solve(const double *in, double *out) { check(cpg_prec_apply_to(h_.get(), in, out, nullptr)); }
    cpg_prec handle() const override { return h_.get(); }
    void refactor(BlockMatrix &A) { check(cpg_prec_factor(h_.get(), A.handle(), nullptr)); }

  private:
    std::shared_ptr<cpg_prec_s> h_;
};

// BlockJacobi (src/Preconditioners.cpp:8-47)
struct BlockJacobi : FactoredPreconditioner {
    explicit BlockJacobi(BlockMatrix &M, int precision = CPG_FP64) : FactoredPreconditioner(M, CPG_PC_BJACOBI, precision) {}
};

// BlockILU0 = block DILU (src/Preconditioners.cpp:49-158), level-scheduled
struct BlockILU0 : FactoredPreconditioner {
    explicit BlockILU0(BlockMatrix &A, int precision = CPG_FP64) : FactoredPreconditioner(A, CPG_PC_DILU, precision) {}
};

inline void BlockMatrix::gmres(const double *b, double *x, int m, double &tol, int &maxIt, Preconditioner &pc,
                               int mode) {
    cpg_krylov K = nullptr;
    check(cpg_krylov_create(n_rows, bl, m, mode, &K));
    std::unique_ptr<cpg_krylov_s, int (*)(cpg_krylov)> guard(K, cpg_krylov_destroy);
    check(cpg_gmres(K, handle(), pc.handle(), b, x, &tol, &maxIt, nullptr));
}

inline void BlockMatrix::jacobi(const double *b, double *x, double &tol, int &maxIt, Preconditioner &pc) {
    cpg_krylov K = nullptr;
    check(cpg_krylov_create(n_rows, bl, 1, CPG_GMRES_LEFT_REF, &K));
    std::unique_ptr<cpg_krylov_s, int (*)(cpg_krylov)> guard(K, cpg_krylov_destroy);
    check(cpg_jacobi(K, handle(), pc.handle(), b, x, &tol, &maxIt, nullptr));
}

// ---- equations ---------------------------------------------------------------------

class Equation;

// Jacobian (src/Jacobian.h): BlockMatrix with the mesh-adjacency pattern
// Jacobian::solve's choice of solver (src/Jacobian.h:9-13)
enum Solver { kGMRESSolver, kJacobiSolver };

class Jacobian : public BlockMatrix {
  public:
    int nc = 0, deg = 0;
    Jacobian() = default;
    Jacobian(cpg_equation E, int a_deg, int a_nc) : nc(a_nc), deg(a_deg) {
        int n, b, nnzb;
        check(cpg_equation_info(E, &n, &b, &nnzb));
        n_rows = n;
        bl = b;
        nb = nnzb;
        rowBlock.resize(n + 1);
        colIndices.resize(nnzb);
        check(cpg_jacobian_pattern(E, rowBlock.data(), colIndices.data()));
        cpg_bsr h = nullptr;
        check(cpg_jacobian_create(E, &h));
        adopt(h);
    }
    // Jacobian::solve (src/Jacobian.cpp:68-97): zero initial guess, GMRES(m)
    MeshFn solve(const MeshFn &b, Preconditioner &pc, int m = 20, double tol = 1e-14, int maxIt = 90000,
                 int mode = CPG_GMRES_LEFT_REF) {
        MeshFn x(b.np, b.deg, b.nc);
        x.eq = b.eq;
        std::vector<double> z(x.size(), 0.0);
        x.a->upload(z);
        gmres(b.data(), x.data(), m, tol, maxIt, pc, mode);
        last_tol = tol;
        last_iterations = maxIt;
        return x;
    }
    // Jacobian::solve(b, pc, s) (src/Jacobian.cpp:68-97): zero initial guess, tol 1e-14,
    // maxIt 90000, GMRES(20) or the Jacobi iteration
    MeshFn solve(const MeshFn &b, Preconditioner &pc, Solver s) {
        if (s == kGMRESSolver) return solve(b, pc);
        MeshFn x(b.np, b.deg, b.nc);
        x.eq = b.eq;
        std::vector<double> z(x.size(), 0.0);
        x.a->upload(z);
        double tol = 1e-14;
        int maxIt = 90000;
        jacobi(b.data(), x.data(), tol, maxIt, pc);
        last_tol = tol;
        last_iterations = maxIt;
        return x;
    }
    double last_tol = 0;
    int last_iterations = 0;
};

class MassMatrix;

class Equation {
  public:
    Equation(const PolyMesh &msh, const cpg_eq_params &p, int deg) : deg_(deg), nc_(p.kind == CPG_EULER ? 4 : 1) {
        cpg_equation h = nullptr;
        check(cpg_equation_create(msh.handle(), &p, deg, &h));
        h_ = std::shared_ptr<cpg_equation_s>(h, [](cpg_equation e) { cpg_equation_destroy(e); });
        np_ = msh.np();
    }
    virtual ~Equation() = default;
    int nc() const { return nc_; }
    int deg() const { return deg_; }
    cpg_equation handle() const { return h_.get(); }
    // Equation::assemble (src/Equation.cpp:152-191)
    virtual MeshFn assemble(const MeshFn &f, double t) {
        MeshFn b(np_, deg_, nc_);
        b.eq = h_;
        check(cpg_assemble_residual(handle(), f.data(), t, b.data(), nullptr));
        check(cpg_stream_sync(nullptr));
        return b;
    }
    // Equation::jacobian (src/Equation.cpp:193-279)
    virtual Jacobian jacobian(const MeshFn &f, double t) {
        Jacobian J(handle(), deg_, nc_);
        check(cpg_assemble_jacobian(handle(), f.data(), t, 1.0, 0, J.handle(), nullptr));
        check(cpg_stream_sync(nullptr));
        return J;
    }
    // (M + scale J(f, t)) v without assembling J: the matvec of the Newton matrix
    // `B = jacobian(f, t); B *= scale; B += M` (src/TimeIntegration.cpp:127-129)
    MeshFn newtonApply(const MeshFn &f, double t, double scale, const MeshFn &v) const {
        MeshFn y(np_, deg_, nc_);
        y.eq = h_;
        check(cpg_newton_apply(handle(), f.data(), t, scale, v.data(), y.data(), nullptr));
        check(cpg_stream_sync(nullptr));
        return y;
    }

    // MeshFn::interp (src/MeshFn.cpp:56-145) of a built-in field (CPG_FIELD_*, parameters as in
    // cpolyg.h) at time t: the least-squares fit at the bounding-box Gauss points
    MeshFn interp(int field, const std::vector<double> &par, double t = 0) const {
        std::vector<double> p(8, 0.0);
        for (size_t j = 0; j < par.size() && j < 8; j++) p[j] = par[j];
        MeshFn u(np_, deg_, nc_);
        u.eq = h_;
        check(cpg_interp_field(handle(), field, p.data(), t, u.data(), nullptr));
        check(cpg_stream_sync(nullptr));
        return u;
    }
    // MeshFn::L2Error (src/MeshFn.cpp:277-296) against the same fields, per component
    std::vector<double> L2Error(const MeshFn &f, int field, const std::vector<double> &par, double t = 0) const {
        std::vector<double> p(8, 0.0), err(nc_);
        for (size_t j = 0; j < par.size() && j < 8; j++) p[j] = par[j];
        check(cpg_l2_error_field(handle(), f.data(), field, p.data(), t, err.data(), nullptr));
        return err;
    }

  protected:
    int deg_, nc_, np_ = 0;
    std::shared_ptr<cpg_equation_s> h_;
};

inline cpg_eq_params euler_params(double gamma, int periodic_trace = 0, int jacobian = 0) {
    cpg_eq_params p{};
    p.kind = CPG_EULER;
    p.gamma = gamma;
    p.periodic_trace = periodic_trace;
    p.jacobian = jacobian;
    return p;
}

inline cpg_eq_params advection_params(double b0, double b1, double b2, double b3, double b4, double b5) {
    cpg_eq_params p{};
    p.kind = CPG_ADVECTION;
    double b[6] = {b0, b1, b2, b3, b4, b5};
    for (int j = 0; j < 6; j++) p.beta[j] = b[j];
    return p;
}

// eqns/Euler.h
struct Euler : Equation {
    Euler(const PolyMesh &m, double gamma, int deg, int periodic_trace = 0, int jacobian = 0)
        : Equation(m, euler_params(gamma, periodic_trace, jacobian), deg) {}
};

// eqns/Advection.h: default velocity (2y - 1, -2x + 1)
struct Advection : Equation {
    Advection(const PolyMesh &m, int deg) : Equation(m, advection_params(-1, 0, 2, 1, -2, 0), deg) {}

  protected:
    Advection(const PolyMesh &m, const cpg_eq_params &p, int deg) : Equation(m, p, deg) {}
};

struct ConstantAdvection : Advection {
    ConstantAdvection(const PolyMesh &m, double bx, double by, int deg)
        : Advection(m, advection_params(bx, 0, 0, by, 0, 0), deg) {}
};

// MassMatrix (src/MassMatrix.h): matvec only on the implicit path
class MassMatrix {
  public:
    explicit MassMatrix(const Equation &E) : E_(&E) {}
    MeshFn matvec(const MeshFn &fn) const {
        MeshFn y(fn.np, fn.deg, fn.nc);
        y.eq = fn.eq;
        check(cpg_mass_matvec(E_->handle(), fn.data(), y.data(), nullptr));
        check(cpg_stream_sync(nullptr));
        return y;
    }

  private:
    const Equation *E_;
};

// ---- time stepping -----------------------------------------------------------------
inline cpg_solver_params reference_solver() {
    cpg_solver_params s{};
    s.gmres_mode = CPG_GMRES_LEFT_REF;
    s.restart = 20;
    s.rtol = 1e-14;
    s.maxit = 90000;
    s.pc_kind = CPG_PC_DILU;
    s.precision = CPG_FP64;
    s.newton_tol = 5e-13;
    s.newton_maxit = 100;
    return s;
}

class TimeStepper {
  public:
    TimeStepper(Equation &eqn, int method, int nstages, const cpg_solver_params &sp) : eqn_(&eqn) {
        cpg_stepper h = nullptr;
        check(cpg_stepper_create(eqn.handle(), method, nstages, &sp, &h));
        h_ = std::shared_ptr<cpg_stepper_s>(h, [](cpg_stepper s) { cpg_stepper_destroy(s); });
    }
    virtual ~TimeStepper() = default;
    // TimeStepper::advance (src/TimeIntegration.h:16)
    virtual MeshFn advance(const MeshFn &u, double dt, double t = 0) {
        MeshFn out(u.np, u.deg, u.nc);
        out.eq = u.eq;
        check(cpg_stepper_advance(h_.get(), u.data(), dt, t, out.data(), &last_stats, nullptr));
        return out;
    }
    cpg_step_stats last_stats{};

  protected:
    Equation *eqn_;
    std::shared_ptr<cpg_stepper_s> h_;
};

struct DIRK : TimeStepper {
    DIRK(int nStages, Equation &eqn, const cpg_solver_params &sp = reference_solver())
        : TimeStepper(eqn, CPG_STEP_DIRK, nStages, sp) {}
};

struct BackwardEuler : TimeStepper {
    explicit BackwardEuler(Equation &eqn, const cpg_solver_params &sp = reference_solver())
        : TimeStepper(eqn, CPG_STEP_DIRK, 1, sp) {}
};

struct RK4 : TimeStepper {
    explicit RK4(Equation &eqn) : TimeStepper(eqn, CPG_STEP_RK4, 4, reference_solver()) {}
};

// RK2 (src/TimeIntegration.h:41-46): the midpoint stage with weights 1/2, 1/2, as the reference
struct RK2 : TimeStepper {
    explicit RK2(Equation &eqn) : TimeStepper(eqn, CPG_STEP_RK2, 2, reference_solver()) {}
};

// ForwardEuler (src/TimeIntegration.h:20-25)
struct ForwardEuler : TimeStepper {
    explicit ForwardEuler(Equation &eqn) : TimeStepper(eqn, CPG_STEP_FORWARD_EULER, 1, reference_solver()) {}
};

// IRK (src/TimeIntegration.h:64-76): Radau IIA with nStages = 2 or 3.  advance solves the
// stage-coupled Newton system (IRK::advance, src/TimeIntegration.cpp:153-281); newAdvance the
// W form in the unknowns w = A k (IRK::newAdvance, :283-395), on a second stepper made at its
// first call with the same parameters but flexible GMRES (the W form applies its matrix
// matrix-free).  The reference's constructor takes the mass matrix; here the equation's own is
// used, so the MassMatrix overload only keeps its call sites unchanged.
struct IRK : TimeStepper {
    IRK(int nStages, Equation &eqn, const cpg_solver_params &sp = reference_solver())
        : TimeStepper(eqn, CPG_STEP_IRK, nStages, sp), ns_(nStages), sp_(sp) {}
    IRK(int nStages, MassMatrix &, Equation &eqn, const cpg_solver_params &sp = reference_solver())
        : IRK(nStages, eqn, sp) {}
    MeshFn newAdvance(const MeshFn &u, double dt, double t = 0) {
        if (!w_) {
            cpg_solver_params w = sp_;
            w.gmres_mode = CPG_GMRES_FLEXIBLE;
            w_ = std::make_unique<TimeStepper>(*eqn_, CPG_STEP_IRK_W, ns_, w);
        }
        MeshFn out = w_->advance(u, dt, t);
        last_stats = w_->last_stats;
        return out;
    }

  private:
    int ns_;
    cpg_solver_params sp_;
    std::unique_ptr<TimeStepper> w_;
};

}  // namespace hip
}  // namespace cpolyg
