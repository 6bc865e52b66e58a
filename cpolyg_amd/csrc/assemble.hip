// DG residual and Jacobian assembly on polygon sub-triangulations.
//
// Reference: Equation::assemble / Equation::jacobian (src/Equation.cpp:152-279).
// The reference re-runs the whole quadrature for every test function (residual)
// and every (test, trial) pair (Jacobian).  Here each quadrature point is
// evaluated once per cell: state, flux and flux Jacobian are computed at the point
// and contracted against all test/trial basis values.
//
//   residual  b_i[c*bs+k] = sum_vol W (Fx_c d_x psi_k + Fy_c d_y psi_k)
//                          - sum_edges sum_q W psi_k Fhat_c(U-, U+, n)
//   Jacobian  J_ii[(c1,k),(c2,l)] = sum_vol W phi_l (Ax d_x psi_k + Ay d_y psi_k)_{c1c2}
//                                 - sum_edges sum_q W psi_k phi_l dFhat/dU-_{c1c2}
//             J_i,nb[(c1,k),(c2,l)] = - sum_q W psi_k phi^nb_l dFhat/dU+_{c1c2}
//
// Jacobian blocks are written straight into the BSR slots of the pattern of
// Jacobian::Jacobian (src/Jacobian.cpp:7-47): the diagonal block first, then one
// block per non-Dirichlet edge in edge order.  The Newton matrix
// scale*J + M (B *= -dt a; B += M, src/TimeIntegration.cpp:128-129) is produced in
// the same pass.
#include "cpg_internal.hpp"
#include "dg_device.hpp"

#include <cmath>
#include <cstring>
#include <mutex>

namespace cpg {

// triangle rules, values of src/Quadrature.cpp:126-171 (r, s, weight)
static const double TRI2[] = {0.16666666666667, 0.16666666666667, 0.33333333333333,
                              0.16666666666667, 0.66666666666667, 0.33333333333333,
                              0.66666666666667, 0.16666666666667, 0.33333333333333};
static const double TRI4[] = {0.44594849091597, 0.44594849091597, 0.22338158967801,
                              0.44594849091597, 0.10810301816807, 0.22338158967801,
                              0.10810301816807, 0.44594849091597, 0.22338158967801,
                              0.09157621350977, 0.09157621350977, 0.10995174365532,
                              0.09157621350977, 0.81684757298046, 0.10995174365532,
                              0.81684757298046, 0.09157621350977, 0.10995174365532};
static const double TRI6[] = {0.24928674517091, 0.24928674517091, 0.11678627572638,
                              0.24928674517091, 0.50142650965818, 0.11678627572638,
                              0.50142650965818, 0.24928674517091, 0.11678627572638,
                              0.06308901449150, 0.06308901449150, 0.05084490637021,
                              0.06308901449150, 0.87382197101700, 0.05084490637021,
                              0.87382197101700, 0.06308901449150, 0.05084490637021,
                              0.31035245103378, 0.63650249912140, 0.08285107561837,
                              0.63650249912140, 0.05314504984482, 0.08285107561837,
                              0.05314504984482, 0.31035245103378, 0.08285107561837,
                              0.63650249912140, 0.31035245103378, 0.08285107561837,
                              0.31035245103378, 0.05314504984482, 0.08285107561837,
                              0.05314504984482, 0.63650249912140, 0.08285107561837};
static const double TRI10[] = {0.33333333333333, 0.33333333333333, 0.09081799038275,
                               0.48557763338366, 0.48557763338366, 0.03672595775647,
                               0.48557763338366, 0.02884473323269, 0.03672595775647,
                               0.02884473323269, 0.48557763338366, 0.03672595775647,
                               0.10948157548504, 0.10948157548504, 0.04532105943553,
                               0.10948157548504, 0.78103684902993, 0.04532105943553,
                               0.78103684902993, 0.10948157548504, 0.04532105943553,
                               0.30793983876412, 0.55035294182100, 0.07275791684542,
                               0.55035294182100, 0.14170721941488, 0.07275791684542,
                               0.14170721941488, 0.30793983876412, 0.07275791684542,
                               0.55035294182100, 0.30793983876412, 0.07275791684542,
                               0.30793983876412, 0.14170721941488, 0.07275791684542,
                               0.14170721941488, 0.55035294182100, 0.07275791684542,
                               0.24667256063990, 0.72832390459741, 0.02832724253106,
                               0.72832390459741, 0.02500353476269, 0.02832724253106,
                               0.02500353476269, 0.24667256063990, 0.02832724253106,
                               0.72832390459741, 0.24667256063990, 0.02832724253106,
                               0.24667256063990, 0.02500353476269, 0.02832724253106,
                               0.02500353476269, 0.72832390459741, 0.02832724253106,
                               0.06680325101220, 0.92365593358750, 0.00942166696373,
                               0.92365593358750, 0.00954081540030, 0.00942166696373,
                               0.00954081540030, 0.06680325101220, 0.00942166696373,
                               0.92365593358750, 0.06680325101220, 0.00942166696373,
                               0.06680325101220, 0.00954081540030, 0.00942166696373,
                               0.00954081540030, 0.92365593358750, 0.00942166696373};

// edge record (host-computed geometry)
struct Edge {
    double x1, y1, x2, y2;  // endpoints (polygon edge pv[j] -> pv[j+1])
    double nx, ny;          // outward normal (centroid sign test, src/PolyMesh.cpp:248-273)
    double tx, ty;          // periodic translation v[a2] - v[a1] (src/PolyMesh.cpp:307-325)
    double len;
    int nb;                 // neighbour cell (interior or periodic partner), -1 Dirichlet
    int kind;               // 0 interior, 1 periodic, 2 Dirichlet
    int slot;               // BSR slot of the off-diagonal block, -1 if none
    int bslot;              // Dirichlet edges: index into the boundary-state table, else -1
};

struct Geo {
    const double4 *bb;      // x0, y0, w, h
    const int *vptr;        // volume points per cell
    const double4 *vpts;    // x, y, W
    const int *eptr;        // edges per cell
    const Edge *edges;
    const double *bcst;     // Dirichlet states at the edge points: [bslot][q][c] (k_bc_states)
};

// Dirichlet states of every boundary edge point at time t (BoundaryConditions::boundaryValue,
// src/BoundaryConditions.cpp:11-13).  Kept out of the cell kernels: the Dirichlet functors
// (exp, pow, atan2) would otherwise hold libm constants and parameter quotients in registers
// across their edge loops.
template <int DEG, int NC>
__global__ void k_bc_states(DevEq eq, const Edge *__restrict__ edges, const int *__restrict__ bedges, int nbe,
                            double t, double *__restrict__ bcst) {
    constexpr int NQL = Deg<DEG>::NQL;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nbe * NQL) return;
    const int b = e / NQL, q = e - b * NQL;
    const Edge E = edges[bedges[b]];
    const double *lr = line_rule<NQL>();
    const double s = 0.5 * (lr[2 * q] + 1.0);
    const double x = E.x1 + s * (E.x2 - E.x1), y = E.y1 + s * (E.y2 - E.y1);
    double U[NC];
    eval_bc<NC>(eq, x, y, t, U);
#pragma unroll
    for (int c = 0; c < NC; c++) bcst[(size_t)(E.bslot * NQL + q) * NC + c] = U[c];
}

// ---------------------------------------------------------------------------------------
template <int DEG, int NC>
__device__ inline void outer_state(const DevEq &eq, const Geo &g, const Edge &E, const double *__restrict__ u,
                                   double x, double y, int q, double *Up, double *phin) {
    constexpr int BS = Deg<DEG>::BS, BL = NC * BS, NQL = Deg<DEG>::NQL;
    if (E.kind == 2) {
#pragma unroll
        for (int c = 0; c < NC; c++) Up[c] = g.bcst[(size_t)(E.bslot * NQL + q) * NC + c];
#pragma unroll
        for (int k = 0; k < BS; k++) phin[k] = 0.0;
        return;
    }
    const double4 b = g.bb[E.nb];
    double xi, eta;
    if (E.kind == 0) {
        xi = 2.0 * (x - b.x) / b.z - 1.0;
        eta = 2.0 * (y - b.y) / b.w - 1.0;
        basis<DEG>(xi, eta, phin, nullptr, nullptr);
    } else {
        xi = 2.0 * (x + E.tx - b.x) / b.z - 1.0;
        eta = 2.0 * (y + E.ty - b.y) / b.w - 1.0;
        trace_basis<DEG>(eq.periodic_trace, xi, eta, phin);
    }
    eval_state<BS, NC>(u + (size_t)E.nb * BL, phin, Up);
}

// ---------------------------------------------------------------------------------------
// Cell operators: the residual b(u) (LIN = false, Equation::assemble, src/Equation.cpp:152-191)
// and the Newton operator applied without assembling it (LIN = true):
//   y = M v + scale * J(u) v,
// the product the BSR matvec forms with the matrix k_jacobian writes (scale * J + M;
// src/Jacobian.cpp:99-110 with src/BlockMatrix.cpp:28-49).  J(u) v is summed over the same
// quadrature points from the same per-point flux Jacobians that k_jacobian contracts
// (volume: A_x, A_y of fluxjac; edges: numfluxjac(+1) for the cell's own trace, and
// euler_offdiag_exact or numfluxjac(-1) for the neighbour's), applied to the traces of v
// instead of to basis functions, so the two agree to rounding in either Jacobian form.
//
// One wavefront per cell.  Lanes first take quadrature points: each evaluates the basis, the
// trace of its input and the pointwise flux (residual) or flux derivative (operator), and
// writes an entry (test-function values D[k], flux values G[c]) per point to an LDS record; a
// volume point gives two entries (d_xi phi_k with W sx F_x,c, then d_eta phi_k with W sy F_y,c,
// one record fill each), an edge point one (phi_k with -W Fhat_c).  Then the record is
// contracted: lane (jg, k) keeps NC accumulators out[c][k] over the entries j = jg mod JG (one
// LDS read of D and NC/2 16-byte reads of G per entry), and the JG partial sums of every output
// are added in a fixed order at the end.  The neighbours' coefficients of an edge chunk are
// staged in LDS with coalesced loads.
//
// LDS bandwidth, shared by the CU's waves, bounds much of this work, so the traces of the cell's
// own coefficients are formed from uniform (scalar) loads of x instead of LDS reads: the 40 (60
// at p=4) coefficients are the same for every lane.  This also frees the VGPRs that held LDS
// operands (p=3 operator 168 -> 156, p=4 240 -> 157).  Each edge chunk loads the lanes' edge records
// first, publishes the neighbour indices through LDS, and then issues the neighbours' coefficients
// and bounding boxes together, so the neighbour data is one round trip behind the records; the test
// values of an edge point go to the record as soon as they are formed.
//
// The operator (LIN) reads what depends on u alone from the point data k_linprep stored for
// this u (Euler; advection needs none), so its inner loops hold only traces of v: no divisions,
// square roots or traces of u per GMRES iteration.
template <int DEG, int NC, bool LIN>
struct CellOpLayout {
    static constexpr int BS = Deg<DEG>::BS, BL = NC * BS, NQL = Deg<DEG>::NQL;
    static constexpr int EPC = 64 / NQL;              // edges per chunk
    static constexpr int NE = 64;                     // record entries
    static constexpr int REC = NE * (BS + NC);        // D[NE][BS], then G[NE][NC]
    static constexpr int JG = 64 / BS;                // entry groups of the contraction
    static constexpr int SW = BL;                     // staged doubles per edge (the input's coefficients)
    static constexpr int STG = EPC * SW;
    static constexpr int SH = REC > JG * BL ? REC : JG * BL;
};

// acc[c] += sum over entries j = jg, jg + JG, ... < nent of D[j][k] G[j][c]
template <int BS, int NC, int JG>
__device__ inline void cellop_contract(const double *D, const double *G, int jg, int k, int nent, double (&acc)[NC]) {
    for (int j = jg; j < nent; j += JG) {
        const double dk = D[j * BS + k];
        const double *gj = G + j * NC;
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c] = fma(dk, gj[c], acc[c]);
    }
}

// Point data of the Euler Newton operator: nv volume points (LPV fields each) and nep edge
// points (LPE fields each), cell by cell and field-major inside a cell, so that a wavefront's
// loads are contiguous and a cell's data shares cache lines with its two neighbours in memory
// only at its ends (field-major over all points, every field of a cell began and ended inside
// lines that neighbouring cells, usually on other XCDs, fetched as well); read once per product
// with non-temporal loads, so that they do not evict the Krylov vectors that the preconditioner
// sweeps gather next.  Cell i with volume points [v0, v1) and edges [e0, e0 + ne):
struct LinData {
    const double *vol;   // vol[LPV v0 + f (v1 - v0) + (q - v0)] for volume point q of cell i
    const double *edg;   // edg[LPE NQL e0 + f ne NQL + (e - e0) NQL + k] for point k of edge e
    int nv, nep;
};

// out = b(x) (residual, LIN = false; x = u) or out = M xm + scale J(u) x (LIN = true; x = v, xm = v
// when null, the u-dependent part from `ld`; the coupled stages of an implicit Runge-Kutta step
// take the mass term and the Jacobian term from different vectors)
template <int DEG, int NC, bool LIN>
__global__ __launch_bounds__(64, DEG <= 3 ? 3 : 2) void k_cellop(DevEq eq, Geo g, const double *__restrict__ x, LinData ld,
                                                  double scale, const double *__restrict__ scale_dev,
                                                  const int *__restrict__ active, const double *__restrict__ mass,
                                                  const double *__restrict__ xm, double *__restrict__ out) {
    using L = CellOpLayout<DEG, NC, LIN>;
    if (active && !*active) return;   // a converged Krylov cycle replayed from a graph
    if (scale_dev) scale = *scale_dev;
    constexpr int BS = L::BS, BL = L::BL, NQL = L::NQL, SW = L::SW, JG = L::JG;
    __shared__ __align__(16) double sX[BL];   // the cell's coefficients of x
    __shared__ __align__(16) double sh[L::SH];
    __shared__ int sNbi[L::EPC];   // the neighbour of each edge of the current edge chunk
    __shared__ __align__(16) double stg[L::STG];   // the neighbours' coefficients of x for that chunk
    double *D = sh;
    double *G = sh + L::NE * BS;
    const int i = blockIdx.x, lane = threadIdx.x;
    // the cell's coefficients are the same for every lane: the traces read them through uniform
    // (scalar) loads, sX keeps a copy for the mass term
    const double *xc = x + (size_t)i * BL;
    for (int r = lane; r < BL; r += 64) sX[r] = x[(size_t)i * BL + r];
    const double4 bb = g.bb[i];
    const double sx = 2.0 / bb.z, sy = 2.0 / bb.w;
    const int jg = lane / BS, ck = lane - jg * BS;   // contraction role (lanes < JG*BS)
    const bool cl = jg < JG;
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    __syncthreads();

    const double *lr = line_rule<NQL>();
    const int e0 = g.eptr[i], ne = g.eptr[i + 1] - e0;

    // ---- volume points, 64 per chunk
    const int v0 = g.vptr[i], v1 = g.vptr[i + 1];
    for (int q0 = v0; q0 < v1; q0 += 64) {
        const int nq = min(64, v1 - q0);
        double deta[BS], Gy[NC];
        __syncthreads();   // readers of the previous record are done
        if (lane < nq) {
            const double4 P = g.vpts[q0 + lane];
            const double xi = 2.0 * (P.x - bb.x) / bb.z - 1.0, eta = 2.0 * (P.y - bb.y) / bb.w - 1.0;
            double phi[BS], dxi[BS], X[NC], Fx[NC];
            basis<DEG>(xi, eta, phi, dxi, deta);
#pragma unroll
            for (int k = 0; k < BS; k++) D[lane * BS + k] = dxi[k];
            eval_state<BS, NC>(xc, phi, X);
            if constexpr (LIN) {
                if constexpr (NC == 4) {
                    double p[LPV];
#pragma unroll
                    for (int f = 0; f < LPV; f++)
                        p[f] = __builtin_nontemporal_load(ld.vol + (size_t)LPV * v0 + (size_t)f * (v1 - v0) + (q0 - v0) + lane);
                    euler_jvp_p(eq.gamma, p, X, eq.jacobian == 0, Fx, Gy);
                } else {
                    double bx, by;
                    adv_beta(eq, P.x, P.y, bx, by);
                    Fx[0] = bx * X[0];
                    Gy[0] = by * X[0];
                }
            } else {
                flux<NC>(eq, X, P.x, P.y, Fx, Gy);
            }
            const double wx = P.z * sx, wy = P.z * sy;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                G[lane * NC + c] = wx * Fx[c];
                Gy[c] *= wy;
            }
        }
        __syncthreads();
        if (cl) cellop_contract<BS, NC, JG>(D, G, jg, ck, nq, acc);
        __syncthreads();
        if (lane < nq) {
#pragma unroll
            for (int k = 0; k < BS; k++) D[lane * BS + k] = deta[k];
#pragma unroll
            for (int c = 0; c < NC; c++) G[lane * NC + c] = Gy[c];
        }
        __syncthreads();
        if (cl) cellop_contract<BS, NC, JG>(D, G, jg, ck, nq, acc);
    }

    // ---- edge points, EPC edges per chunk
    for (int eb = 0; eb < ne; eb += L::EPC) {
        const int nec = min(L::EPC, ne - eb);
        const int np = nec * NQL;
        // the lane's edge record first, so that the neighbour loads below depend on one round trip
        const bool pt = lane < np;
        const int el = lane / NQL, q = lane - el * NQL;
        const Edge E = g.edges[e0 + eb + (pt ? el : 0)];
        // the neighbours' coefficients (neighbour indices from the lanes holding each edge's first
        // point, through LDS) and the lane's neighbour bounding box
        double4 nbb = make_double4(0.0, 0.0, 1.0, 1.0);
        if (pt && E.nb >= 0) nbb = g.bb[E.nb];
        if (pt && q == 0) sNbi[el] = E.nb;
        __syncthreads();   // readers of the previous chunk's staging and of the record are done
        {
            // every gather in flight before the first LDS write: one round trip per chunk (a loop
            // of load-then-store waited for each load in turn); Dirichlet edges (nb = -1) and idle
            // lanes read x[0] and store 0
            constexpr int GI = (L::STG + 63) / 64;
            double gv[GI];
#pragma unroll
            for (int t = 0; t < GI; t++) {
                const int idx = lane + 64 * t;
                const int el = idx / SW, r = idx - el * SW;
                const int nb = idx < nec * SW ? sNbi[el] : -1;
                const double v = x[nb >= 0 ? (size_t)nb * BL + r : 0];
                gv[t] = nb >= 0 ? v : 0.0;
            }
#pragma unroll
            for (int t = 0; t < GI; t++)
                if (lane + 64 * t < nec * SW) stg[lane + 64 * t] = gv[t];
        }
        __syncthreads();
        if (pt) {
            const double s = 0.5 * (lr[2 * q] + 1.0);
            const double px = E.x1 + s * (E.x2 - E.x1), py = E.y1 + s * (E.y2 - E.y1);
            const double W = 0.5 * E.len * lr[2 * q + 1];
            const double xi = 2.0 * (px - bb.x) / bb.z - 1.0, eta = 2.0 * (py - bb.y) / bb.w - 1.0;
            double Xm[NC], Xp[NC];
            {
                double phi[BS];
                basis<DEG>(xi, eta, phi, nullptr, nullptr);
#pragma unroll
                for (int k = 0; k < BS; k++) D[lane * BS + k] = phi[k];   // the entry's test values
                eval_state<BS, NC>(xc, phi, Xm);
            }
            if (E.kind == 2) {
                // Dirichlet: the outer state is data (k_bc_states), independent of u
#pragma unroll
                for (int c = 0; c < NC; c++) Xp[c] = LIN ? 0.0 : g.bcst[(size_t)(E.bslot * NQL + q) * NC + c];
            } else {
                double phin[BS];
                const double4 b = nbb;
                const double xo = E.kind == 0 ? px : px + E.tx, yo = E.kind == 0 ? py : py + E.ty;
                const double xin = 2.0 * (xo - b.x) / b.z - 1.0, etan = 2.0 * (yo - b.y) / b.w - 1.0;
                if (E.kind == 0) basis<DEG>(xin, etan, phin, nullptr, nullptr);
                else trace_basis<DEG>(eq.periodic_trace, xin, etan, phin);
                eval_state<BS, NC>(stg + el * SW, phin, Xp);
            }
            double F[NC];
            if constexpr (LIN) {
                if constexpr (NC == 4) {
                    double rec[LPE];
                    const double *ec = ld.edg + (size_t)LPE * NQL * e0 + (eb + el) * NQL + q;
#pragma unroll
                    for (int f = 0; f < LPE; f++) rec[f] = __builtin_nontemporal_load(ec + (size_t)f * ne * NQL);
                    euler_numflux_jvp_p(eq, rec, Xm, Xp, E.nx, E.ny, F);
                } else {
                    F[0] = adv_numflux_jvp(eq, Xm[0], Xp[0], px, py, E.nx, E.ny);
                }
            } else {
                numflux<NC>(eq, Xm, Xp, px, py, E.nx, E.ny, F);
            }
#pragma unroll
            for (int c = 0; c < NC; c++) G[lane * NC + c] = -W * F[c];
        }
        __syncthreads();
        if (cl) cellop_contract<BS, NC, JG>(D, G, jg, ck, np, acc);
    }

    // ---- sum the JG partial sums of every output in a fixed order
    __syncthreads();
    if (cl) {
#pragma unroll
        for (int c = 0; c < NC; c++) sh[jg * BL + c * BS + ck] = acc[c];
    }
    __syncthreads();
    if (lane < BL) {
        double r = 0.0;
#pragma unroll
        for (int q = 0; q < JG; q++) r += sh[q * BL + lane];
        if constexpr (LIN) {
            const int oc = lane / BS, ok = lane - oc * BS;
            const double *Mi = mass + (size_t)i * BS * BS;
            double m = 0.0;
            if (xm) {
                const double *xi = xm + (size_t)i * BL + oc * BS;
#pragma unroll
                for (int l = 0; l < BS; l++) m = fma(Mi[l * BS + ok], xi[l], m);
            } else {
#pragma unroll
                for (int l = 0; l < BS; l++) m = fma(Mi[l * BS + ok], sX[oc * BS + l], m);
            }
            r = fma(scale, r, m);
        }
        out[(size_t)i * BL + lane] = r;
    }
}

// Point data of the Euler Newton operator for the state u (the LinData layout): one
// wavefront per cell, lanes over quadrature points, u's traces from the cell's and the
// neighbours' coefficients (staged in LDS) or the Dirichlet states
template <int DEG>
__global__ __launch_bounds__(64) void k_linprep(DevEq eq, Geo g, const double *__restrict__ u, double *__restrict__ vol,
                                                double *__restrict__ edg, int nv, int nep) {
    constexpr int NC = 4, BS = Deg<DEG>::BS, BL = NC * BS, NQL = Deg<DEG>::NQL, EPC = 64 / NQL;
    __shared__ double sU[BL];
    __shared__ double stg[EPC * BL];
    const int i = blockIdx.x, lane = threadIdx.x;
    for (int r = lane; r < BL; r += 64) sU[r] = u[(size_t)i * BL + r];
    const double4 bb = g.bb[i];
    __syncthreads();
    const int v0 = g.vptr[i], v1 = g.vptr[i + 1];
    for (int q = v0 + lane; q < v1; q += 64) {
        const double4 P = g.vpts[q];
        const double xi = 2.0 * (P.x - bb.x) / bb.z - 1.0, eta = 2.0 * (P.y - bb.y) / bb.w - 1.0;
        double phi[BS], U[NC], p[LPV];
        basis<DEG>(xi, eta, phi, nullptr, nullptr);
        eval_state<BS, NC>(sU, phi, U);
        euler_prim4(eq.gamma, U, p);
#pragma unroll
        for (int f = 0; f < LPV; f++) vol[(size_t)LPV * v0 + (size_t)f * (v1 - v0) + (q - v0)] = p[f];
    }
    const double *lr = line_rule<NQL>();
    const int e0 = g.eptr[i], ne = g.eptr[i + 1] - e0;
    for (int eb = 0; eb < ne; eb += EPC) {
        const int nec = min(EPC, ne - eb);
        __syncthreads();
        for (int idx = lane; idx < nec * BL; idx += 64) {
            const int el = idx / BL, r = idx - el * BL;
            const int nb = g.edges[e0 + eb + el].nb;
            stg[idx] = nb >= 0 ? u[(size_t)nb * BL + r] : 0.0;
        }
        __syncthreads();
        if (lane < nec * NQL) {
            const int el = lane / NQL, q = lane - el * NQL;
            const Edge E = g.edges[e0 + eb + el];
            const double s = 0.5 * (lr[2 * q] + 1.0);
            const double px = E.x1 + s * (E.x2 - E.x1), py = E.y1 + s * (E.y2 - E.y1);
            const double xi = 2.0 * (px - bb.x) / bb.z - 1.0, eta = 2.0 * (py - bb.y) / bb.w - 1.0;
            double phi[BS], Um[NC], Up[NC], rec[LPE];
            basis<DEG>(xi, eta, phi, nullptr, nullptr);
            eval_state<BS, NC>(sU, phi, Um);
            if (E.kind == 2) {
#pragma unroll
                for (int c = 0; c < NC; c++) Up[c] = g.bcst[(size_t)(E.bslot * NQL + q) * NC + c];
            } else {
                const double4 b = g.bb[E.nb];
                const double xo = E.kind == 0 ? px : px + E.tx, yo = E.kind == 0 ? py : py + E.ty;
                const double xin = 2.0 * (xo - b.x) / b.z - 1.0, etan = 2.0 * (yo - b.y) / b.w - 1.0;
                double phin[BS];
                if (E.kind == 0) basis<DEG>(xin, etan, phin, nullptr, nullptr);
                else trace_basis<DEG>(eq.periodic_trace, xin, etan, phin);
                eval_state<BS, NC>(stg + el * BL, phin, Up);
            }
            euler_edge_prep(eq.gamma, Um, Up, E.nx, E.ny, rec);
            double *ec = edg + (size_t)LPE * NQL * e0 + (eb + el) * NQL + q;
#pragma unroll
            for (int f = 0; f < LPE; f++) ec[(size_t)f * ne * NQL] = rec[f];
        }
    }
}

// ---------------------------------------------------------------------------------------
// Jacobian: one workgroup of NC wavefronts per cell; wavefront c2 owns the trial-component
// column group c2 (columns c2*bs .. c2*bs+bs-1) of every block of the cell's row.
//
// Every block is a sum over quadrature points of rank-one products, i.e. a GEMM with the
// points as the contraction index:
//   volume   J_ii[(c1,k), l] += sum_q (Gx_k Ax[c1][c2] + Gy_k Ay[c1][c2])(q) * phi_l(q)
//   edge     J_ii[(c1,k), l] -= sum_q (W psi_k Jd[c1][c2])(q) * phi_l(q)
//            J_i,nb[(c1,k), l] = -scale sum_q (W psi_k Jo[c1][c2])(q) * phin_l(q)
// (Gx = W d_x psi, Gy = W d_y psi).  The products run on v_mfma_f64_16x16x4f64: the row
// index (c1,k) spans ceil(bl/16) 16-row tiles, the trial index l one 16-column tile, and each
// MFMA consumes four points.  Point data (basis values, scaled gradients, flux Jacobians) is
// evaluated once per point by one thread and staged in LDS in chunks of JQC points; the MFMA
// operands are formed from it on the fly (A = test factor x flux-Jacobian entry, B = trial
// basis value).  Results go from the accumulator tiles straight to the BSR slots.
typedef double jdx4 __attribute__((ext_vector_type(4)));
constexpr int JQC = 64;   // points per LDS chunk

template <int DEG, int NC>
struct JacLayout {
    static constexpr int BS = Deg<DEG>::BS, BL = NC * BS, NQL = Deg<DEG>::NQL;
    static constexpr int NC2 = NC * NC;
    static constexpr int MT = (BL + 15) / 16;          // 16-row tiles of a block
    static constexpr int NQLP = 4 * ((NQL + 3) / 4);   // edge points padded to whole MFMA k-steps
    static constexpr int EPC = JQC / NQLP;             // edges per chunk
    static_assert(BS <= 16, "trial basis must fit one 16-column tile");
    // volume point record: Gx[BS] Gy[BS] phi[16] Ax[NC2] Ay[NC2]
    static constexpr int V_GX = 0, V_GY = BS, V_PHI = 2 * BS, V_AX = 2 * BS + 16, V_AY = V_AX + NC2;
    static constexpr int VREC = V_AY + NC2;
    // edge point record: W psi[BS] phi[16] phin[16] Jd[NC2] Jo[NC2]
    static constexpr int E_WPHI = 0, E_PHI = BS, E_PHIN = BS + 16, E_JD = BS + 32, E_JO = E_JD + NC2;
    static constexpr int EREC = E_JO + NC2;
    static constexpr int REC = (VREC > EREC ? VREC : EREC) | 1;   // odd stride: fewer LDS bank conflicts
};

template <int DEG, int NC>
__device__ inline void jac_tiles_store(double *__restrict__ blk, const jdx4 (&acc)[JacLayout<DEG, NC>::MT],
                                       double scale, const double *__restrict__ massi, int c2) {
    using L = JacLayout<DEG, NC>;
    const int lane = threadIdx.x & 63;
    const int n = lane & 15, r0 = lane >> 4;
    if (n >= L::BS) return;
    double *col = blk + (size_t)(c2 * L::BS + n) * L::BL;
#pragma unroll
    for (int t = 0; t < L::MT; t++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int row = 16 * t + r0 + 4 * e;
            if (row < L::BL) {
                double v = scale * acc[t][e];
                if (massi && row / L::BS == c2) v += massi[n * L::BS + (row - c2 * L::BS)];
                col[row] = v;
            }
        }
}

template <int DEG, int NC>
__global__ __launch_bounds__(64 * NC, 2) void k_jacobian(DevEq eq, Geo g, const double *__restrict__ u, double t,
                                                      double scale, const double *__restrict__ cfac, int add_mass,
                                                      const double *__restrict__ mass,
                                                      const int *__restrict__ rowptr, double *__restrict__ J) {
    using L = JacLayout<DEG, NC>;
    constexpr int BS = L::BS, BL = L::BL, NQL = L::NQL, NC2 = L::NC2, MT = L::MT, REC = L::REC;
    constexpr int NT = 64 * NC;
    constexpr size_t BB = (size_t)BL * BL;
    __shared__ double rec[JQC * REC];
    __shared__ double sU[BL];
    const int i = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, c2 = tid >> 6;
    if (cfac) scale *= cfac[i];   // per-cell factor (batched solves: each instance has its own dt)
    for (int r = tid; r < BL; r += NT) sU[r] = u[(size_t)i * BL + r];
    const double4 bb = g.bb[i];
    const double sx = 2.0 / bb.z, sy = 2.0 / bb.w;
    // this lane's rows of the A operand: row m = 16 t + lane % 16 = (c1, k)
    int tk[MT], tc[MT];
    bool tv[MT];
#pragma unroll
    for (int tt = 0; tt < MT; tt++) {
        const int m = 16 * tt + (lane & 15);
        tv[tt] = m < BL;
        tc[tt] = tv[tt] ? m / BS : 0;
        tk[tt] = tv[tt] ? m - tc[tt] * BS : 0;
    }
    const int qsub = lane >> 4, ncol = lane & 15;
    jdx4 accD[MT];
#pragma unroll
    for (int tt = 0; tt < MT; tt++) accD[tt] = jdx4{0.0, 0.0, 0.0, 0.0};
    __syncthreads();

    // ---- volume terms
    const int v0 = g.vptr[i], v1 = g.vptr[i + 1];
    for (int q0 = v0; q0 < v1; q0 += JQC) {
        const int nq = min(JQC, v1 - q0);
        for (int p = tid; p < JQC; p += NT) {
            double *d = rec + p * REC;
            if (p < nq) {
                // results go to the LDS record as they are produced (short live ranges)
                const double4 P = g.vpts[q0 + p];
                double phi[BS], Um[NC];
                {
                    double dxi[BS], deta[BS];
                    const double xi = 2.0 * (P.x - bb.x) / bb.z - 1.0, eta = 2.0 * (P.y - bb.y) / bb.w - 1.0;
                    basis<DEG>(xi, eta, phi, dxi, deta);
#pragma unroll
                    for (int k = 0; k < BS; k++) {
                        d[L::V_GX + k] = P.z * dxi[k] * sx;
                        d[L::V_GY + k] = P.z * deta[k] * sy;
                    }
                }
#pragma unroll
                for (int k = 0; k < 16; k++) d[L::V_PHI + k] = k < BS ? phi[k] : 0.0;
                eval_state<BS, NC>(sU, phi, Um);
                fluxjac<NC>(eq, Um, P.x, P.y, d + L::V_AX, d + L::V_AY);
            } else {
#pragma unroll
                for (int k = 0; k < L::VREC; k++) d[k] = 0.0;
            }
        }
        __syncthreads();
        const int nks = (nq + 3) >> 2;
        for (int ks = 0; ks < nks; ks++) {
            const double *d = rec + (4 * ks + qsub) * REC;
            const double b = d[L::V_PHI + ncol];
#pragma unroll
            for (int tt = 0; tt < MT; tt++) {
                const int cc = tc[tt] * NC + c2;
                const double a = tv[tt] ? fma(d[L::V_GX + tk[tt]], d[L::V_AX + cc], d[L::V_GY + tk[tt]] * d[L::V_AY + cc])
                                        : 0.0;
                accD[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, accD[tt], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // ---- edge terms
    const double *lr = line_rule<NQL>();
    const int e0 = g.eptr[i], ne = g.eptr[i + 1] - e0;
    for (int eb = 0; eb < ne; eb += L::EPC) {
        const int nec = min(L::EPC, ne - eb);
        for (int p = tid; p < nec * L::NQLP; p += NT) {
            const int el = p / L::NQLP, q = p - el * L::NQLP;
            double *d = rec + p * REC;
            if (q < NQL) {
                const Edge E = g.edges[e0 + eb + el];
                const double s = 0.5 * (lr[2 * q] + 1.0);
                const double x = E.x1 + s * (E.x2 - E.x1), y = E.y1 + s * (E.y2 - E.y1);
                const double W = 0.5 * E.len * lr[2 * q + 1];
                const double xi = 2.0 * (x - bb.x) / bb.z - 1.0, eta = 2.0 * (y - bb.y) / bb.w - 1.0;
                double Um[NC], Up[NC];
                {
                    double phi[BS];
                    basis<DEG>(xi, eta, phi, nullptr, nullptr);
#pragma unroll
                    for (int k = 0; k < BS; k++) d[L::E_WPHI + k] = W * phi[k];
#pragma unroll
                    for (int k = 0; k < 16; k++) d[L::E_PHI + k] = k < BS ? phi[k] : 0.0;
                    eval_state<BS, NC>(sU, phi, Um);
                }
                {
                    double phin[BS];
                    outer_state<DEG, NC>(eq, g, E, u, x, y, q, Up, phin);
#pragma unroll
                    for (int k = 0; k < 16; k++) d[L::E_PHIN + k] = k < BS ? phin[k] : 0.0;
                }
                numfluxjac<NC>(eq, Um, Up, x, y, E.nx, E.ny, +1, d + L::E_JD);
                double *Jo = d + L::E_JO;
                if (E.kind != 2) {
                    if constexpr (NC == 4) {
                        if (eq.jacobian) euler_offdiag_exact(eq.gamma, Um, Up, E.nx, E.ny, Jo);
                        else numfluxjac<NC>(eq, Um, Up, x, y, E.nx, E.ny, -1, Jo);
                    } else {
                        numfluxjac<NC>(eq, Um, Up, x, y, E.nx, E.ny, -1, Jo);
                    }
                } else {
#pragma unroll
                    for (int c = 0; c < NC2; c++) Jo[c] = 0.0;
                }
            } else {
#pragma unroll
                for (int k = 0; k < L::EREC; k++) d[k] = 0.0;
            }
        }
        __syncthreads();
        for (int el = 0; el < nec; el++) {
            const int slot = g.edges[e0 + eb + el].slot;
            jdx4 accO[MT];
#pragma unroll
            for (int tt = 0; tt < MT; tt++) accO[tt] = jdx4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < L::NQLP / 4; ks++) {
                const double *d = rec + (el * L::NQLP + 4 * ks + qsub) * REC;
                const double bd = d[L::E_PHI + ncol];
                const double bo = d[L::E_PHIN + ncol];
#pragma unroll
                for (int tt = 0; tt < MT; tt++) {
                    const int cc = tc[tt] * NC + c2;
                    const double w = tv[tt] ? d[L::E_WPHI + tk[tt]] : 0.0;
                    accD[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(-w * d[L::E_JD + cc], bd, accD[tt], 0, 0, 0);
                    if (slot >= 0)
                        accO[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(w * d[L::E_JO + cc], bo, accO[tt], 0, 0, 0);
                }
            }
            if (slot >= 0) jac_tiles_store<DEG, NC>(J + (size_t)slot * BB, accO, -scale, nullptr, c2);
        }
        __syncthreads();
    }

    // ---- diagonal block: scale * J_ii (+ M_i on the diagonal component blocks)
    jac_tiles_store<DEG, NC>(J + (size_t)rowptr[i] * BB, accD, scale,
                             add_mass ? mass + (size_t)i * BS * BS : nullptr, c2);
}

// ---------------------------------------------------------------------------------------
// mass blocks M_i[k,l] = sum_q W phi_k phi_l  (MassMatrix::MassMatrix, src/MassMatrix.cpp:30-94)
template <int DEG>
__global__ __launch_bounds__(64) void k_mass(Geo g, double *__restrict__ M) {
    constexpr int BS = Deg<DEG>::BS;
    const int i = blockIdx.x;
    const double4 bb = g.bb[i];
    double acc[(BS * BS + 63) / 64];
#pragma unroll
    for (int j = 0; j < (BS * BS + 63) / 64; j++) acc[j] = 0.0;
    double phi[BS];
    for (int q = g.vptr[i]; q < g.vptr[i + 1]; q++) {
        const double4 P = g.vpts[q];
        const double xi = 2.0 * (P.x - bb.x) / bb.z - 1.0, eta = 2.0 * (P.y - bb.y) / bb.w - 1.0;
        basis<DEG>(xi, eta, phi, nullptr, nullptr);
#pragma unroll
        for (int j = 0; j < (BS * BS + 63) / 64; j++) {
            const int e = threadIdx.x + 64 * j;
            if (e < BS * BS) acc[j] += P.z * phi[e % BS] * phi[e / BS];
        }
    }
#pragma unroll
    for (int j = 0; j < (BS * BS + 63) / 64; j++) {
        const int e = threadIdx.x + 64 * j;
        if (e < BS * BS) M[(size_t)i * BS * BS + e] = acc[j];
    }
}

// y_i[c*bs+k] = sum_l M_i[k,l] x_i[c*bs+l]
template <int BS, int NC>
__global__ void k_mass_matvec(int n, const double *__restrict__ M, const double *__restrict__ x,
                              double *__restrict__ y) {
    constexpr int BL = NC * BS;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * BL) return;
    const size_t i = e / BL;
    const int r = (int)(e - i * BL);
    const int c = r / BS, k = r - c * BS;
    const double *Mi = M + i * BS * BS;
    const double *xi = x + i * BL + c * BS;
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < BS; l++) s = fma(Mi[l * BS + k], xi[l], s);
    y[e] = s;
}

// r_i = M_i k_i - dt b_i (or r = x when k is null); normc[i*NC + c] = r_c^T M_i r_c
template <int BS, int NC>
__global__ void k_newton_resid(int n, const double *__restrict__ M, const double *__restrict__ kv,
                               const double *__restrict__ b, double dt, const double *__restrict__ cfac,
                               double *__restrict__ r, double *__restrict__ normc) {
    constexpr int BL = NC * BS;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // (cell, component)
    if (e >= (size_t)n * NC) return;
    const size_t i = e / NC;
    const int c = (int)(e - i * NC);
    const double *Mi = M + i * BS * BS;
    if (cfac) dt *= cfac[i];
    double rv[BS];
    if (kv) {
        const double *ki = kv + i * BL + c * BS;
#pragma unroll
        for (int k = 0; k < BS; k++) {
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < BS; l++) s = fma(Mi[l * BS + k], ki[l], s);
            rv[k] = s - dt * b[i * BL + c * BS + k];
            r[i * BL + c * BS + k] = rv[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < BS; k++) rv[k] = b[i * BL + c * BS + k];
    }
    double q = 0.0;
#pragma unroll
    for (int k = 0; k < BS; k++) {
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < BS; l++) s = fma(Mi[l * BS + k], rv[l], s);
        q = fma(rv[k], s, q);
    }
    normc[e] = q;
}

// column sums of normc (n x NC) in two fixed-order stages
__global__ __launch_bounds__(256) void k_colsum_partial(const double *__restrict__ a, int n, int nc,
                                                        double *__restrict__ partial) {
    __shared__ double red[4][4];
    const int c = blockIdx.y;
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += a[(size_t)i * nc + c];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][0] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[(size_t)c * gridDim.x + blockIdx.x] = red[0][0] + red[1][0] + red[2][0] + red[3][0];
}

__global__ void k_colsum_final(const double *__restrict__ partial, int nb, int nc, double *__restrict__ out) {
    const int c = threadIdx.x;
    if (c >= nc) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += partial[(size_t)c * nb + b];
    out[c] = s;
}

// ---- analytic fields: MeshFn::interp (src/MeshFn.cpp:56-145) and MeshFn::L2Error
// (src/MeshFn.cpp:277-296) on the device.  The reference evaluates a host functor; the fields
// here are a fixed set (CPG_FIELD_*) evaluated in the kernels, or values supplied at the points.
struct FieldArgs {
    int kind;
    double par[8];
    double t;
};

// nearest periodic image of a difference (no wrap for a non-positive period)
__device__ inline double wrap_diff(double d, double L) { return L > 0.0 ? d - L * rint(d / L) : d; }

template <int NC>
__device__ inline void field_eval(const FieldArgs &f, double x, double y, double *out) {
    const double *p = f.par;
    const double t = f.t;
    if constexpr (NC == 1) {
        // gaussian bump moving with (vx, vy)
        const double dx = wrap_diff(x - p[0] - p[3] * t, p[5]), dy = wrap_diff(y - p[1] - p[4] * t, p[6]);
        out[0] = exp(-p[2] * (dx * dx + dy * dy));
    } else if (f.kind == CPG_FIELD_VORTEX) {
        // isentropic vortex (Yee, Sandham and Djomehri 1999) advected by the free stream
        const double beta = p[0], uinf = p[3], vinf = p[4], gam = p[5];
        const double dx = wrap_diff(x - p[1] - uinf * t, p[6]), dy = wrap_diff(y - p[2] - vinf * t, p[7]);
        const double ex = exp(0.5 * (1.0 - dx * dx - dy * dy));
        const double b2p = beta / (2.0 * M_PI);
        const double u = uinf - b2p * dy * ex, v = vinf + b2p * dx * ex;
        const double T = 1.0 - (gam - 1.0) * beta * beta / (8.0 * gam * M_PI * M_PI) * ex * ex;
        const double rho = pow(T, 1.0 / (gam - 1.0));
        out[0] = rho;
        out[1] = rho * u;
        out[2] = rho * v;
        out[3] = rho * T / (gam - 1.0) + 0.5 * rho * (u * u + v * v);
    } else {
        // density wave advected by a uniform flow
        const double ux = p[0], uy = p[1], pr = p[2], gam = p[3], amp = p[4];
        const double rho = 1.0 + amp * sin(2.0 * M_PI * (x - ux * t)) * sin(2.0 * M_PI * (y - uy * t));
        out[0] = rho;
        out[1] = rho * ux;
        out[2] = rho * uy;
        out[3] = pr / (gam - 1.0) + 0.5 * rho * (ux * ux + uy * uy);
    }
}

// one thread per cell: u_i = Gp vals_i, Gp the least-squares pseudo-inverse (bs x (p+1)^2) of the
// basis at the tensor Gauss points; vals from the field or from `vals` (n x Q x NC)
template <int DEG, int NC>
__global__ void k_interp(const double4 *__restrict__ bb, int n, FieldArgs f, const double *__restrict__ vals,
                         const double *__restrict__ nodes, const double *__restrict__ Gp, double *__restrict__ u) {
    constexpr int M = DEG + 1, Q = M * M, BS = Deg<DEG>::BS;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 B = bb[i];
    double acc[NC][BS];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int k = 0; k < BS; k++) acc[c][k] = 0.0;
    for (int qi = 0; qi < M; qi++)
        for (int qj = 0; qj < M; qj++) {
            const int q = qi * M + qj;
            double fv[NC];
            if (vals) {
#pragma unroll
                for (int c = 0; c < NC; c++) fv[c] = vals[((size_t)i * Q + q) * NC + c];
            } else {
                field_eval<NC>(f, (nodes[qi] + 1.0) * B.z / 2.0 + B.x, (nodes[qj] + 1.0) * B.w / 2.0 + B.y, fv);
            }
#pragma unroll
            for (int k = 0; k < BS; k++) {
                const double g = Gp[k * Q + q];
#pragma unroll
                for (int c = 0; c < NC; c++) acc[c][k] = fma(g, fv[c], acc[c][k]);
            }
        }
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int k = 0; k < BS; k++) u[(size_t)i * NC * BS + c * BS + k] = acc[c][k];
}

template <int DEG>
__global__ void k_interp_points(const double4 *__restrict__ bb, int n, const double *__restrict__ nodes,
                                double *__restrict__ xy) {
    constexpr int M = DEG + 1, Q = M * M;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * Q) return;
    const int i = e / Q, q = e - i * Q, qi = q / M, qj = q - qi * M;
    const double4 B = bb[i];
    xy[2 * (size_t)e] = (nodes[qi] + 1.0) * B.z / 2.0 + B.x;
    xy[2 * (size_t)e + 1] = (nodes[qj] + 1.0) * B.w / 2.0 + B.y;
}

// one thread per cell: err_i[c] = sum_q W_q (u_h(x_q) - f(x_q))_c^2 over the volume points
template <int DEG, int NC>
__global__ void k_l2_error(Geo g, int n, FieldArgs f, const double *__restrict__ u, double *__restrict__ err) {
    constexpr int BS = Deg<DEG>::BS;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 B = g.bb[i];
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = 0.0;
    for (int q = g.vptr[i]; q < g.vptr[i + 1]; q++) {
        const double4 P = g.vpts[q];
        double phi[BS];
        basis<DEG>(2.0 * (P.x - B.x) / B.z - 1.0, 2.0 * (P.y - B.y) / B.w - 1.0, phi, nullptr, nullptr);
        double fv[NC];
        field_eval<NC>(f, P.x, P.y, fv);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double uh = 0.0;
#pragma unroll
            for (int k = 0; k < BS; k++) uh = fma(u[(size_t)i * NC * BS + c * BS + k], phi[k], uh);
            const double d = uh - fv[c];
            acc[c] = fma(P.z * d, d, acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) err[(size_t)i * NC + c] = acc[c];
}

}  // namespace cpg

using namespace cpg;

// ---- handles -------------------------------------------------------------------------
struct cpg_mesh_s {
    int np = 0, nv = 0, nbc = 0;
    std::vector<double> v, txy, bb;
    std::vector<int> pptr, pv, p2p, bc, tptr;
};

struct cpg_equation_s {
    cpg_mesh_s *mesh = nullptr;
    DevEq eq{};
    int deg = 0, bs = 0, nc = 0, bl = 0, n = 0, nnzb = 0;
    std::vector<int> rowptr, colind;
    DevBuf<int> d_rowptr;
    DevBuf<double4> d_bb, d_vpts;
    DevBuf<int> d_vptr, d_eptr;
    DevBuf<Edge> d_edges;
    DevBuf<int> d_bedges;       // Dirichlet edges (indices into d_edges)
    DevBuf<double> bcst;        // their states at the edge points (k_bc_states)
    int nbe = 0;
    int nvpts = 0, nedges = 0;  // volume quadrature points, polygon edges
    DevBuf<double> lin_vol, lin_edg;   // point data of the Newton operator (k_linprep; Euler)
    bool prepared = false;
    DevBuf<double> mass;        // n x bs x bs
    DevBuf<double> normc, partial, sums;
    int npartial = 0;
    DevBuf<double> fit;         // interp: Gauss nodes (p+1), then the pseudo-inverse (bs x (p+1)^2)
    Geo geo() const { return Geo{d_bb.p, d_vptr.p, d_vpts.p, d_eptr.p, d_edges.p, bcst.p}; }
};

namespace cpg {

// Launch helpers shared with the stepper
// Dirichlet states at time t for the kernels that follow on the same stream
template <int DEG, int NC>
static void launch_bc_states(const cpg_equation_s &E, double t, hipStream_t s) {
    if (!E.nbe) return;
    const int cnt = E.nbe * Deg<DEG>::NQL;
    hipLaunchKernelGGL((k_bc_states<DEG, NC>), dim3(ceil_div(cnt, 256)), dim3(256), 0, s, E.eq, E.d_edges.p,
                       E.d_bedges.p, E.nbe, t, E.bcst.p);
    CPG_LAUNCH_CHECK();
}

template <int DEG, int NC>
static void launch_residual(const cpg_equation_s &E, const double *u, double t, double *b, hipStream_t s) {
    launch_bc_states<DEG, NC>(E, t, s);
    hipLaunchKernelGGL((k_cellop<DEG, NC, false>), dim3(E.n), dim3(64), 0, s, E.eq, E.geo(), u, LinData{}, 0.0,
                       nullptr, nullptr, nullptr, nullptr, b);
    CPG_LAUNCH_CHECK();
}

// point data of the Newton operator for u (Euler; advection has none), with the Dirichlet
// states already at the right time
// point data of the Newton operator for u (Euler; advection has none) into vol / edg (the
// equation's own buffers when null), with the Dirichlet states already at the right time
template <int DEG, int NC>
static void launch_newton_prepare(cpg_equation_s &E, const double *u, double *vol, double *edg, hipStream_t s) {
    if constexpr (NC == 4) {
        const size_t nep = (size_t)E.nedges * Deg<DEG>::NQL;
        if (!vol) {
            if (!E.lin_vol.p) {
                E.lin_vol.alloc((size_t)LPV * E.nvpts);
                E.lin_edg.alloc((size_t)LPE * nep);
            }
            vol = E.lin_vol.p;
            edg = E.lin_edg.p;
        }
        hipLaunchKernelGGL((k_linprep<DEG>), dim3(E.n), dim3(64), 0, s, E.eq, E.geo(), u, vol, edg, E.nvpts, (int)nep);
        CPG_LAUNCH_CHECK();
    }
}

// y = M xm + scale J(u) v (xm = v when null) for the u whose point data vol / edg hold (the
// equation's own, from its last prepare, when null); scale_dev (if non-null) overrides scale
// from device memory, so a captured launch can be replayed for another Newton matrix
template <int DEG, int NC>
static void launch_newton_apply(const cpg_equation_s &E, double scale, const double *scale_dev, const double *v,
                                const double *xm, double *y, const int *active, const double *vol, const double *edg,
                                hipStream_t s) {
    const LinData ld{vol ? vol : E.lin_vol.p, vol ? edg : E.lin_edg.p, E.nvpts, E.nedges * Deg<DEG>::NQL};
    launch_timed(k_cellop<DEG, NC, true>, dim3(E.n), dim3(64), 0, s, E.eq, E.geo(), v, ld, scale, scale_dev, active,
                 (const double *)E.mass.p, xm, y);
    CPG_LAUNCH_CHECK();
}

template <int DEG, int NC>
static void launch_bc_only(const cpg_equation_s &E, double t, hipStream_t s) {
    launch_bc_states<DEG, NC>(E, t, s);
}

template <int DEG, int NC>
static void launch_jacobian(const cpg_equation_s &E, const double *u, double t, double scale, const double *cfac,
                            int add_mass, double *J, hipStream_t s) {
    launch_bc_states<DEG, NC>(E, t, s);
    hipLaunchKernelGGL((k_jacobian<DEG, NC>), dim3(E.n), dim3(64 * NC), 0, s, E.eq, E.geo(), u, t, scale, cfac,
                       add_mass, E.mass.p, E.d_rowptr.p, J);
    CPG_LAUNCH_CHECK();
}

#define CPG_DISPATCH(E, FN, ...)                                                          \
    do {                                                                                  \
        if ((E).nc == 4) {                                                                \
            switch ((E).deg) {                                                            \
            case 1: FN<1, 4>(__VA_ARGS__); break;                                         \
            case 2: FN<2, 4>(__VA_ARGS__); break;                                         \
            case 3: FN<3, 4>(__VA_ARGS__); break;                                         \
            case 4: FN<4, 4>(__VA_ARGS__); break;                                         \
            default: CPG_CHECK(false, CPG_ERR_UNSUPPORTED, "degree %d", (E).deg);         \
            }                                                                             \
        } else {                                                                          \
            switch ((E).deg) {                                                            \
            case 1: FN<1, 1>(__VA_ARGS__); break;                                         \
            case 2: FN<2, 1>(__VA_ARGS__); break;                                         \
            case 3: FN<3, 1>(__VA_ARGS__); break;                                         \
            case 4: FN<4, 1>(__VA_ARGS__); break;                                         \
            default: CPG_CHECK(false, CPG_ERR_UNSUPPORTED, "degree %d", (E).deg);         \
            }                                                                             \
        }                                                                                 \
    } while (0)

void eq_residual(const cpg_equation_s &E, const double *u, double t, double *b, hipStream_t s) {
    CPG_DISPATCH(E, launch_residual, E, u, t, b, s);
}

// the u-dependent point data of the Newton operator at (u, t), into vol / edg (sized by
// eq_newton_data_size; the equation's own buffers when null)
void eq_newton_prepare(cpg_equation_s &E, const double *u, double t, hipStream_t s, double *vol, double *edg) {
    CPG_DISPATCH(E, launch_bc_only, E, t, s);
    CPG_DISPATCH(E, launch_newton_prepare, E, u, vol, edg, s);
}

// doubles of the operator's point data (volume, edges); 0 for advection
void eq_newton_data_size(const cpg_equation_s &E, size_t *nvol, size_t *nedg) {
    const int nql = 2 * E.deg <= 2 ? 2 : 2 * E.deg <= 4 ? 4 : 2 * E.deg <= 6 ? 6 : 10;
    *nvol = E.nc == 4 ? (size_t)LPV * E.nvpts : 0;
    *nedg = E.nc == 4 ? (size_t)LPE * E.nedges * nql : 0;
}

// y = M v + scale J(u, t) v without assembling J
void eq_newton_apply(cpg_equation_s &E, const double *u, double t, double scale, const double *v, double *y,
                     hipStream_t s) {
    eq_newton_prepare(E, u, t, s, nullptr, nullptr);
    E.prepared = true;
    CPG_DISPATCH(E, launch_newton_apply, E, scale, nullptr, v, nullptr, y, nullptr, nullptr, nullptr, s);
}

// the same for the (u, t) of the last eq_newton_prepare, with the scale in device memory (the
// Krylov operator of a Newton step)
void eq_newton_apply_dev(const cpg_equation_s &E, const double *scale_dev, const double *v, double *y,
                         const int *active, hipStream_t s, double scale, const double *xm, const double *vol,
                         const double *edg) {
    CPG_DISPATCH(E, launch_newton_apply, E, scale, scale_dev, v, xm, y, active, vol, edg, s);
}

void eq_jacobian(const cpg_equation_s &E, const double *u, double t, double scale, int add_mass, double *J,
                 hipStream_t s) {
    CPG_DISPATCH(E, launch_jacobian, E, u, t, scale, nullptr, add_mass, J, s);
}

void eq_jacobian_cells(const cpg_equation_s &E, const double *u, double t, double scale, const double *cfac,
                       int add_mass, double *J, hipStream_t s) {
    CPG_DISPATCH(E, launch_jacobian, E, u, t, scale, cfac, add_mass, J, s);
}

template <int BS, int NC>
static void launch_mass_matvec(const cpg_equation_s &E, const double *x, double *y, hipStream_t s) {
    const size_t cnt = (size_t)E.n * NC * BS;
    hipLaunchKernelGGL((k_mass_matvec<BS, NC>), dim3(ceil_div(cnt, 256)), dim3(256), 0, s, E.n, E.mass.p, x, y);
    CPG_LAUNCH_CHECK();
}

template <int BS, int NC>
static void launch_newton_resid(const cpg_equation_s &E, const double *kv, const double *b, double dt,
                                const double *cfac, double *r, hipStream_t s) {
    const size_t cnt = (size_t)E.n * NC;
    hipLaunchKernelGGL((k_newton_resid<BS, NC>), dim3(ceil_div(cnt, 128)), dim3(128), 0, s, E.n, E.mass.p, kv, b,
                       dt, cfac, r, E.normc.p);
    CPG_LAUNCH_CHECK();
}

#define CPG_DISPATCH_BS(E, FN, ...)                                                       \
    do {                                                                                  \
        const int bs_ = (E).bs;                                                           \
        if ((E).nc == 4) {                                                                \
            if (bs_ == 3) FN<3, 4>(__VA_ARGS__);                                          \
            else if (bs_ == 6) FN<6, 4>(__VA_ARGS__);                                     \
            else if (bs_ == 10) FN<10, 4>(__VA_ARGS__);                                   \
            else FN<15, 4>(__VA_ARGS__);                                                  \
        } else {                                                                          \
            if (bs_ == 3) FN<3, 1>(__VA_ARGS__);                                          \
            else if (bs_ == 6) FN<6, 1>(__VA_ARGS__);                                     \
            else if (bs_ == 10) FN<10, 1>(__VA_ARGS__);                                   \
            else FN<15, 1>(__VA_ARGS__);                                                  \
        }                                                                                 \
    } while (0)

void eq_mass_matvec(const cpg_equation_s &E, const double *x, double *y, hipStream_t s) {
    CPG_DISPATCH_BS(E, launch_mass_matvec, E, x, y, s);
}

// ---- MeshFn::L2Norm without allocation (cpg_l2norm_noalloc) ----------------------------
// The per-component sums sum_i u_ic^T M_i u_ic are reduced in two fixed-order stages through
// a file-scope device array instead of the equation's per-cell buffers: kL2Parts workgroups
// per component each take a grid-stride share of the cells and leave one partial sum, and
// k_colsum_final adds the kL2Parts partials of each component in order.  The array is the
// library's for the life of the process on every GPU it runs on; calls that overlap on one GPU
// share it and race.  Every call overwrites every entry it reads, so nothing is cleared.
constexpr int kL2Parts = 1024;                   // partial sums per component
constexpr int kL2MaxNc = 4;                      // components
constexpr int kL2Scratch = kL2Parts * kL2MaxNc + kL2MaxNc;   // partials, then the nc sums (32.8 KB)
static __device__ __align__(256) double scratch[kL2Scratch];

template <int BS, int NC>
__global__ __launch_bounds__(256) void k_l2norm_parts(int n, const double *__restrict__ M, const double *__restrict__ u,
                                                      double *__restrict__ parts) {
    constexpr int BL = NC * BS;
    __shared__ double red[4];
    const int c = blockIdx.y;
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kL2Parts * 256) {
        const double *Mi = M + (size_t)i * BS * BS;
        const double *ui = u + (size_t)i * BL + c * BS;
        double q = 0.0;
#pragma unroll
        for (int k = 0; k < BS; k++) {
            double t = 0.0;
#pragma unroll
            for (int l = 0; l < BS; l++) t = fma(Mi[l * BS + k], ui[l], t);
            q = fma(ui[k], t, q);
        }
        s += q;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) parts[(size_t)c * kL2Parts + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

template <int BS, int NC>
static void launch_l2norm_parts(const cpg_equation_s &E, const double *u, double *parts, hipStream_t s) {
    hipLaunchKernelGGL((k_l2norm_parts<BS, NC>), dim3(kL2Parts, NC), dim3(256), 0, s, E.n, E.mass.p, u, parts);
    CPG_LAUNCH_CHECK();
}

// device address of `scratch` on the current device, looked up once per device (the array's
// name in host code is not a device address)
static double *scratch_ptr() {
    static std::mutex mu;
    static std::vector<double *> ptrs;
    int dev = 0;
    CPG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
    if (!ptrs[dev]) {
        void *p = nullptr;
        CPG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        CPG_CHECK(p, CPG_ERR_HIP, "no device address for the L2-norm scratch array on device %d", dev);
        ptrs[dev] = static_cast<double *>(p);
    }
    return ptrs[dev];
}

// r = M k - dt b (k non-null) and returns max_c sqrt(r_c^T M r_c) summed over cells (synchronous);
// with k null, the norm of b itself.  out_nc (optional) receives the per-component norms.
double eq_newton_resid(const cpg_equation_s &E, const double *kv, const double *b, double dt, double *r,
                       double *out_nc, hipStream_t s) {
    CPG_DISPATCH_BS(E, launch_newton_resid, E, kv, b, dt, nullptr, r, s);
    hipLaunchKernelGGL(k_colsum_partial, dim3(E.npartial, E.nc), dim3(256), 0, s, E.normc.p, E.n, E.nc, E.partial.p);
    CPG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(64), 0, s, E.partial.p, E.npartial, E.nc, E.sums.p);
    CPG_LAUNCH_CHECK();
    double h[4] = {0, 0, 0, 0};
    CPG_HIP(hipMemcpyAsync(h, E.sums.p, sizeof(double) * E.nc, hipMemcpyDeviceToHost, s));
    CPG_HIP(hipStreamSynchronize(s));
    double mx = 0.0;
    for (int c = 0; c < E.nc; c++) {
        const double v = std::sqrt(h[c]);
        if (out_nc) out_nc[c] = v;
        if (!(v <= mx)) mx = v;  // NaN propagates
    }
    return mx;
}

// r = M k - dt cfac_i b per cell; the per-cell, per-component norms r_c^T M_i r_c stay on
// the device (eq_normc) for segmented reductions
void eq_newton_resid_cells(const cpg_equation_s &E, const double *kv, const double *b, double dt, const double *cfac,
                           double *r, hipStream_t s) {
    CPG_DISPATCH_BS(E, launch_newton_resid, E, kv, b, dt, cfac, r, s);
}

// MeshFn::interp's fitting data: the p+1 Gauss-Legendre nodes (ascending, Newton's method on
// P_{p+1}) and the pseudo-inverse (G^T G)^-1 G^T of the basis at the tensor points
// (G[i (p+1) + j][k] = phi_k(x_i, x_j)), the least-squares solve the reference does per cell
static void ensure_fit(cpg_equation_s &E) {
    if (E.fit.p) return;
    const int M = E.deg + 1, Q = M * M, BS = E.bs;
    std::vector<double> nodes(M);
    for (int i = 0; i < M; i++) {
        double x = std::cos(M_PI * (i + 0.75) / (M + 0.5));
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= M; k++) {
                const double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double dp = M == 1 ? 1.0 : M * (x * p1 - p0) / (x * x - 1.0);
            const double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        nodes[M - 1 - i] = x;
    }
    auto leg = [](int a, double x) {
        double p0 = 1.0, p1 = x;
        if (a == 0) return p0;
        for (int k = 2; k <= a; k++) {
            const double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        return p1;
    };
    std::vector<double> G((size_t)Q * BS);
    for (int qi = 0; qi < M; qi++)
        for (int qj = 0; qj < M; qj++) {
            int k = 0;
            for (int a = 0; a <= E.deg; a++)
                for (int b = 0; b <= E.deg - a; b++) G[(size_t)(qi * M + qj) * BS + k++] = leg(a, nodes[qi]) * leg(b, nodes[qj]);
        }
    // [G^T G | I] -> [I | (G^T G)^-1], partial pivoting
    std::vector<double> aug((size_t)BS * 2 * BS, 0.0);
    for (int r = 0; r < BS; r++) {
        for (int c = 0; c < BS; c++) {
            double v = 0.0;
            for (int q = 0; q < Q; q++) v += G[(size_t)q * BS + r] * G[(size_t)q * BS + c];
            aug[(size_t)r * 2 * BS + c] = v;
        }
        aug[(size_t)r * 2 * BS + BS + r] = 1.0;
    }
    for (int c = 0; c < BS; c++) {
        int pr = c;
        for (int r = c + 1; r < BS; r++)
            if (std::fabs(aug[(size_t)r * 2 * BS + c]) > std::fabs(aug[(size_t)pr * 2 * BS + c])) pr = r;
        for (int j = 0; j < 2 * BS; j++) std::swap(aug[(size_t)c * 2 * BS + j], aug[(size_t)pr * 2 * BS + j]);
        const double d = aug[(size_t)c * 2 * BS + c];
        CPG_CHECK(d != 0.0, CPG_ERR_SINGULAR, "interpolation normal matrix is singular");
        for (int j = 0; j < 2 * BS; j++) aug[(size_t)c * 2 * BS + j] /= d;
        for (int r = 0; r < BS; r++) {
            if (r == c) continue;
            const double f = aug[(size_t)r * 2 * BS + c];
            for (int j = 0; j < 2 * BS; j++) aug[(size_t)r * 2 * BS + j] -= f * aug[(size_t)c * 2 * BS + j];
        }
    }
    std::vector<double> fit(M + (size_t)BS * Q, 0.0);
    for (int i = 0; i < M; i++) fit[i] = nodes[i];
    for (int k = 0; k < BS; k++)
        for (int q = 0; q < Q; q++) {
            double v = 0.0;
            for (int l = 0; l < BS; l++) v += aug[(size_t)k * 2 * BS + BS + l] * G[(size_t)q * BS + l];
            fit[M + (size_t)k * Q + q] = v;
        }
    E.fit.upload(fit);
}

template <int DEG, int NC>
static void launch_interp(const cpg_equation_s &E, const FieldArgs &f, const double *vals, double *u, hipStream_t s) {
    hipLaunchKernelGGL((k_interp<DEG, NC>), dim3(ceil_div(E.n, 128)), dim3(128), 0, s, E.d_bb.p, E.n, f, vals,
                       E.fit.p, E.fit.p + DEG + 1, u);
    CPG_LAUNCH_CHECK();
}

template <int DEG, int NC>
static void launch_interp_points(const cpg_equation_s &E, double *xy, hipStream_t s) {
    const int cnt = E.n * (DEG + 1) * (DEG + 1);
    hipLaunchKernelGGL((k_interp_points<DEG>), dim3(ceil_div(cnt, 256)), dim3(256), 0, s, E.d_bb.p, E.n, E.fit.p, xy);
    CPG_LAUNCH_CHECK();
}

template <int DEG, int NC>
static void launch_l2_error(const cpg_equation_s &E, const FieldArgs &f, const double *u, hipStream_t s) {
    hipLaunchKernelGGL((k_l2_error<DEG, NC>), dim3(ceil_div(E.n, 128)), dim3(128), 0, s, E.geo(), E.n, f, u,
                       E.normc.p);
    CPG_LAUNCH_CHECK();
}

const double *eq_normc(const cpg_equation_s &E) { return E.normc.p; }

const double *equation_mass(const cpg_equation_s &E) { return E.mass.p; }
int equation_bs(const cpg_equation_s &E) { return E.bs; }

}  // namespace cpg

namespace {

template <int DEG>
void launch_mass(cpg_equation_s &E) {
    hipLaunchKernelGGL(k_mass<DEG>, dim3(E.n), dim3(64), 0, nullptr, E.geo(), E.mass.p);
    CPG_LAUNCH_CHECK();
}

void pick_tri(int deg2, const double **tri, int *nt, int *nl) {
    if (deg2 <= 2) { *tri = TRI2; *nt = 3; *nl = 2; }
    else if (deg2 <= 4) { *tri = TRI4; *nt = 6; *nl = 4; }
    else if (deg2 <= 6) { *tri = TRI6; *nt = 12; *nl = 6; }
    else { *tri = TRI10; *nt = 25; *nl = 10; }
}

}  // namespace

extern "C" {

int cpg_mesh_create(int np, int nv, const double *v, const int *pptr, const int *pv, const int *p2p, int nbc,
                    const int *bc, const int *tptr, const double *txy, const double *bb, cpg_mesh *out) {
    return guard([&] {
        CPG_CHECK(out && v && pptr && pv && p2p && tptr && txy && np > 0 && nv > 0, CPG_ERR_ARG, "bad mesh arguments");
        CPG_CHECK(nbc == 0 || bc, CPG_ERR_ARG, "null bc table");
        auto *M = new cpg_mesh_s;
        try {
            M->np = np;
            M->nv = nv;
            M->nbc = nbc;
            M->v.assign(v, v + 2 * nv);
            M->pptr.assign(pptr, pptr + np + 1);
            const int ne = pptr[np];
            M->pv.assign(pv, pv + ne);
            M->p2p.assign(p2p, p2p + ne);
            M->bc.assign(bc, bc + 7 * nbc);
            M->tptr.assign(tptr, tptr + np + 1);
            M->txy.assign(txy, txy + 6 * (size_t)tptr[np]);
            for (int e = 0; e < ne; e++) {
                CPG_CHECK(pv[e] >= 0 && pv[e] < nv, CPG_ERR_ARG, "vertex index out of range");
                CPG_CHECK(p2p[e] < np && (p2p[e] >= 0 || -1 - p2p[e] < nbc), CPG_ERR_ARG,
                          "neighbour id %d out of range", p2p[e]);
            }
            if (bb) {
                M->bb.assign(bb, bb + 4 * (size_t)np);
            } else {
                M->bb.resize(4 * (size_t)np);
                for (int i = 0; i < np; i++) {
                    double x0 = v[2 * pv[pptr[i]]], y0 = v[2 * pv[pptr[i]] + 1], x1 = x0, y1 = y0;
                    for (int j = pptr[i]; j < pptr[i + 1]; j++) {
                        x0 = std::min(x0, v[2 * pv[j]]); x1 = std::max(x1, v[2 * pv[j]]);
                        y0 = std::min(y0, v[2 * pv[j] + 1]); y1 = std::max(y1, v[2 * pv[j] + 1]);
                    }
                    M->bb[4 * i] = x0; M->bb[4 * i + 1] = y0; M->bb[4 * i + 2] = x1 - x0; M->bb[4 * i + 3] = y1 - y0;
                }
            }
        } catch (...) {
            delete M;
            throw;
        }
        *out = M;
    });
}

int cpg_mesh_destroy(cpg_mesh M) {
    return guard([&] { delete M; });
}

int cpg_equation_create(cpg_mesh M, const cpg_eq_params *p, int deg, cpg_equation *out) {
    return guard([&] {
        CPG_CHECK(M && p && out, CPG_ERR_ARG, "null argument");
        CPG_CHECK(deg >= 1 && deg <= 4, CPG_ERR_UNSUPPORTED, "degree %d not in 1..4", deg);
        CPG_CHECK(p->kind == CPG_ADVECTION || p->kind == CPG_EULER, CPG_ERR_ARG, "bad equation kind");
        auto *E = new cpg_equation_s;
        try {
            E->mesh = M;
            E->eq.kind = p->kind;
            for (int j = 0; j < 6; j++) E->eq.beta[j] = p->beta[j];
            E->eq.gamma = p->gamma;
            E->eq.bc_kind = p->bc_kind;
            for (int j = 0; j < 8; j++) E->eq.bc_par[j] = p->bc_par[j];
            E->eq.periodic_trace = p->periodic_trace;
            E->eq.jacobian = p->jacobian;
            E->deg = deg;
            E->bs = (deg + 1) * (deg + 2) / 2;
            E->nc = p->kind == CPG_EULER ? 4 : 1;
            E->bl = E->nc * E->bs;
            const int n = M->np;
            E->n = n;
            const auto &v = M->v;
            // Jacobian pattern (src/Jacobian.cpp:7-47) and edge records
            E->rowptr.assign(n + 1, 0);
            std::vector<Edge> edges;
            std::vector<int> eptr(n + 1, 0), bedges;
            for (int i = 0; i < n; i++) {
                E->rowptr[i] = (int)E->colind.size();
                E->colind.push_back(i);
                const int nv = M->pptr[i + 1] - M->pptr[i];
                double cx = 0, cy = 0;
                for (int j = 0; j < nv; j++) { cx += v[2 * M->pv[M->pptr[i] + j]]; cy += v[2 * M->pv[M->pptr[i] + j] + 1]; }
                cx /= nv;
                cy /= nv;
                for (int j = 0; j < nv; j++) {
                    const int a = M->pv[M->pptr[i] + j], b = M->pv[M->pptr[i] + (j + 1) % nv];
                    const int nb = M->p2p[M->pptr[i] + j];
                    Edge Ed{};
                    Ed.x1 = v[2 * a]; Ed.y1 = v[2 * a + 1]; Ed.x2 = v[2 * b]; Ed.y2 = v[2 * b + 1];
                    const double dx = v[2 * a] - v[2 * b], dy = v[2 * a + 1] - v[2 * b + 1];
                    const double len = std::sqrt(dx * dx + dy * dy);
                    const double sv = -dy * (v[2 * a] - cx) + dx * (v[2 * a + 1] - cy);
                    const int sg = (0.0 < sv) - (sv < 0.0);
                    Ed.nx = -sg * dy / len;
                    Ed.ny = sg * dx / len;
                    Ed.len = std::sqrt((Ed.x2 - Ed.x1) * (Ed.x2 - Ed.x1) + (Ed.y2 - Ed.y1) * (Ed.y2 - Ed.y1));
                    Ed.slot = -1;
                    Ed.bslot = -1;
                    if (nb >= 0) {
                        Ed.kind = 0;
                        Ed.nb = nb;
                    } else {
                        const int *row = &M->bc[7 * (-1 - nb)];
                        if (row[0]) {
                            Ed.kind = 1;
                            Ed.nb = row[2];
                            Ed.tx = v[2 * row[5]] - v[2 * row[3]];
                            Ed.ty = v[2 * row[5] + 1] - v[2 * row[3] + 1];
                        } else {
                            Ed.kind = 2;
                            Ed.nb = -1;
                            Ed.bslot = E->nbe++;
                            bedges.push_back((int)edges.size());
                        }
                    }
                    if (Ed.kind != 2) {
                        Ed.slot = (int)E->colind.size();
                        E->colind.push_back(Ed.nb);
                    }
                    edges.push_back(Ed);
                }
                eptr[i + 1] = (int)edges.size();
            }
            E->rowptr[n] = (int)E->colind.size();
            E->nnzb = E->rowptr[n];
            // volume quadrature points (Quadrature::polygonIntegral, src/Quadrature.cpp:13-57)
            const double *tr;
            int nt, nl;
            pick_tri(2 * deg, &tr, &nt, &nl);
            std::vector<int> vptr(n + 1, 0);
            std::vector<double4> vpts;
            for (int i = 0; i < n; i++) {
                for (int tt = M->tptr[i]; tt < M->tptr[i + 1]; tt++) {
                    const double *T = &M->txy[6 * (size_t)tt];
                    const double area = 0.5 * std::fabs(T[0] * T[3] + T[2] * T[5] + T[4] * T[1] - T[0] * T[5] -
                                                        T[4] * T[3] - T[2] * T[1]);
                    for (int q = 0; q < nt; q++) {
                        const double r = tr[3 * q], s = tr[3 * q + 1];
                        double4 P;
                        P.x = T[0] * (1 - r - s) + T[2] * r + T[4] * s;
                        P.y = T[1] * (1 - r - s) + T[3] * r + T[5] * s;
                        P.z = tr[3 * q + 2] * area;
                        P.w = 0.0;
                        vpts.push_back(P);
                    }
                }
                vptr[i + 1] = (int)vpts.size();
            }
            std::vector<double4> bbv(n);
            for (int i = 0; i < n; i++)
                bbv[i] = make_double4(M->bb[4 * i], M->bb[4 * i + 1], M->bb[4 * i + 2], M->bb[4 * i + 3]);
            E->d_rowptr.upload(E->rowptr);
            E->d_bb.upload(bbv);
            E->d_vptr.upload(vptr);
            E->d_vpts.upload(vpts);
            E->d_eptr.upload(eptr);
            E->d_edges.upload(edges);
            E->nvpts = (int)vpts.size();
            E->nedges = (int)edges.size();
            if (E->nbe) {
                const int nql = 2 * deg <= 2 ? 2 : 2 * deg <= 4 ? 4 : 2 * deg <= 6 ? 6 : 10;
                E->d_bedges.upload(bedges);
                E->bcst.alloc((size_t)E->nbe * nql * E->nc);
            }
            E->mass.alloc((size_t)n * E->bs * E->bs);
            switch (deg) {
            case 1: launch_mass<1>(*E); break;
            case 2: launch_mass<2>(*E); break;
            case 3: launch_mass<3>(*E); break;
            default: launch_mass<4>(*E); break;
            }
            E->normc.alloc((size_t)n * E->nc);
            E->npartial = std::max(1, std::min(1024, ceil_div(n, 2048)));
            E->partial.alloc((size_t)E->npartial * E->nc);
            E->sums.alloc(4);
            CPG_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete E;
            throw;
        }
        *out = E;
    });
}

int cpg_equation_destroy(cpg_equation E) {
    return guard([&] { delete E; });
}

int cpg_equation_info(cpg_equation E, int *n_cells, int *bl, int *nnzb) {
    return guard([&] {
        CPG_CHECK(E, CPG_ERR_ARG, "null equation");
        if (n_cells) *n_cells = E->n;
        if (bl) *bl = E->bl;
        if (nnzb) *nnzb = E->nnzb;
    });
}

int cpg_jacobian_pattern(cpg_equation E, int *rowptr, int *colind) {
    return guard([&] {
        CPG_CHECK(E && rowptr && colind, CPG_ERR_ARG, "null argument");
        std::memcpy(rowptr, E->rowptr.data(), sizeof(int) * (E->n + 1));
        std::memcpy(colind, E->colind.data(), sizeof(int) * E->nnzb);
    });
}

int cpg_jacobian_create(cpg_equation E, cpg_bsr *out) {
    if (!E || !out) {
        set_error("null argument");
        return CPG_ERR_ARG;
    }
    return cpg_bsr_create(E->n, E->bl, E->nnzb, E->rowptr.data(), E->colind.data(), out);
}

int cpg_assemble_residual(cpg_equation E, const double *u, double t, double *b, void *stream) {
    return guard([&] {
        CPG_CHECK(E && u && b, CPG_ERR_ARG, "null argument");
        CPG_CHECK(u != b, CPG_ERR_ARG, "u and b alias");
        eq_residual(*E, u, t, b, S(stream));
    });
}

int cpg_assemble_jacobian(cpg_equation E, const double *u, double t, double scale, int add_mass, cpg_bsr J,
                          void *stream) {
    return guard([&] {
        CPG_CHECK(E && u && J, CPG_ERR_ARG, "null argument");
        CPG_CHECK(J->n == E->n && J->bl == E->bl && J->nnzb == E->nnzb && J->h_colind == E->colind, CPG_ERR_ARG,
                  "matrix does not have the Jacobian pattern");
        eq_jacobian(*E, u, t, scale, add_mass, J->blocks.p, S(stream));
    });
}

int cpg_newton_apply(cpg_equation E, const double *u, double t, double scale, const double *v, double *y,
                     void *stream) {
    return guard([&] {
        CPG_CHECK(E && u && v && y, CPG_ERR_ARG, "null argument");
        CPG_CHECK(v != y && u != y, CPG_ERR_ARG, "y aliases u or v");
        eq_newton_apply(*E, u, t, scale, v, y, S(stream));
    });
}

int cpg_newton_prepare(cpg_equation E, const double *u, double t, void *stream) {
    return guard([&] {
        CPG_CHECK(E && u, CPG_ERR_ARG, "null argument");
        eq_newton_prepare(*E, u, t, S(stream), nullptr, nullptr);
        E->prepared = true;
    });
}

int cpg_newton_apply_prepared(cpg_equation E, double scale, const double *v, double *y, void *stream) {
    return guard([&] {
        CPG_CHECK(E && v && y && v != y, CPG_ERR_ARG, "bad argument");
        CPG_CHECK(E->prepared || E->nc != 4, CPG_ERR_ARG, "cpg_newton_apply_prepared before cpg_newton_prepare");
        eq_newton_apply_dev(*E, nullptr, v, y, nullptr, S(stream), scale, nullptr, nullptr, nullptr);
    });
}

int cpg_mass_matvec(cpg_equation E, const double *x, double *y, void *stream) {
    return guard([&] {
        CPG_CHECK(E && x && y && x != y, CPG_ERR_ARG, "bad argument");
        eq_mass_matvec(*E, x, y, S(stream));
    });
}

int cpg_mass_download(cpg_equation E, double *host_blocks) {
    return guard([&] {
        CPG_CHECK(E && host_blocks, CPG_ERR_ARG, "null argument");
        std::vector<double> m((size_t)E->n * E->bs * E->bs);
        CPG_HIP(hipMemcpy(m.data(), E->mass.p, m.size() * sizeof(double), hipMemcpyDeviceToHost));
        const int bs = E->bs, bl = E->bl;
        std::memset(host_blocks, 0, sizeof(double) * (size_t)E->n * bl * bl);
        for (int i = 0; i < E->n; i++)
            for (int c = 0; c < E->nc; c++)
                for (int l = 0; l < bs; l++)
                    for (int k = 0; k < bs; k++)
                        host_blocks[(size_t)i * bl * bl + (size_t)(c * bs + l) * bl + c * bs + k] =
                            m[(size_t)i * bs * bs + l * bs + k];
    });
}

int cpg_l2norm(cpg_equation E, const double *u, double *out, void *stream) {
    return guard([&] {
        CPG_CHECK(E && u && out, CPG_ERR_ARG, "null argument");
        eq_newton_resid(*E, nullptr, u, 0.0, nullptr, out, S(stream));
    });
}

int cpg_l2norm_noalloc(cpg_equation E, const double *u, double *out, void *stream) {
    return guard([&] {
        CPG_CHECK(E && u && out, CPG_ERR_ARG, "null argument");
        CPG_CHECK(E->nc >= 1 && E->nc <= kL2MaxNc, CPG_ERR_ARG,
                  "%d components: the allocation-free L2 norm holds partial sums for at most %d", E->nc, kL2MaxNc);
        hipStream_t s = S(stream);
        double *parts = scratch_ptr();
        double *sums = parts + (size_t)kL2Parts * kL2MaxNc;
        CPG_DISPATCH_BS(*E, launch_l2norm_parts, *E, u, parts, s);
        hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(64), 0, s, parts, kL2Parts, E->nc, sums);
        CPG_LAUNCH_CHECK();
        double h[kL2MaxNc] = {0, 0, 0, 0};
        CPG_HIP(hipMemcpyAsync(h, sums, sizeof(double) * E->nc, hipMemcpyDeviceToHost, s));
        CPG_HIP(hipStreamSynchronize(s));
        for (int c = 0; c < E->nc; c++) out[c] = std::sqrt(h[c]);
    });
}

static FieldArgs field_args(const cpg_equation_s &E, int field, const double *par, double t) {
    CPG_CHECK(field == CPG_FIELD_GAUSSIAN || field == CPG_FIELD_VORTEX || field == CPG_FIELD_SINE_DENSITY, CPG_ERR_ARG,
              "unknown field %d", field);
    CPG_CHECK((field == CPG_FIELD_GAUSSIAN) == (E.nc == 1), CPG_ERR_ARG, "field %d does not have %d components",
              field, E.nc);
    CPG_CHECK(par, CPG_ERR_ARG, "null parameters");
    FieldArgs f{};
    f.kind = field;
    const int np = field == CPG_FIELD_SINE_DENSITY ? 5 : field == CPG_FIELD_GAUSSIAN ? 7 : 8;
    for (int j = 0; j < np; j++) f.par[j] = par[j];
    f.t = t;
    return f;
}

int cpg_interp_field(cpg_equation E, int field, const double *par, double t, double *u, void *stream) {
    return guard([&] {
        CPG_CHECK(E && u, CPG_ERR_ARG, "null argument");
        const FieldArgs f = field_args(*E, field, par, t);
        ensure_fit(*E);
        CPG_DISPATCH(*E, launch_interp, *E, f, nullptr, u, S(stream));
    });
}

int cpg_interp_points(cpg_equation E, double *xy, void *stream) {
    return guard([&] {
        CPG_CHECK(E && xy, CPG_ERR_ARG, "null argument");
        ensure_fit(*E);
        CPG_DISPATCH(*E, launch_interp_points, *E, xy, S(stream));
    });
}

int cpg_interp_values(cpg_equation E, const double *vals, double *u, void *stream) {
    return guard([&] {
        CPG_CHECK(E && vals && u, CPG_ERR_ARG, "null argument");
        ensure_fit(*E);
        CPG_DISPATCH(*E, launch_interp, *E, FieldArgs{}, vals, u, S(stream));
    });
}

int cpg_l2_error_field(cpg_equation E, const double *u, int field, const double *par, double t, double *err,
                       void *stream) {
    return guard([&] {
        CPG_CHECK(E && u && err, CPG_ERR_ARG, "null argument");
        const FieldArgs f = field_args(*E, field, par, t);
        hipStream_t s = S(stream);
        CPG_DISPATCH(*E, launch_l2_error, *E, f, u, s);
        hipLaunchKernelGGL(k_colsum_partial, dim3(E->npartial, E->nc), dim3(256), 0, s, E->normc.p, E->n, E->nc,
                           E->partial.p);
        CPG_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(64), 0, s, E->partial.p, E->npartial, E->nc, E->sums.p);
        CPG_LAUNCH_CHECK();
        double h[4] = {0, 0, 0, 0};
        CPG_HIP(hipMemcpyAsync(h, E->sums.p, sizeof(double) * E->nc, hipMemcpyDeviceToHost, s));
        CPG_HIP(hipStreamSynchronize(s));
        for (int c = 0; c < E->nc; c++) err[c] = std::sqrt(h[c]);
    });
}

}  // extern "C"
