"""Benchmark: DIRK3 steps/s for compressible Euler, p=3, 65,536-cell periodic Voronoi mesh.

Workload (one "step" = one DIRK3 time step, i.e. three Newton solves on the device):
  isentropic vortex (strength 5, centre (5,5), free stream (1,1), gamma 1.4) on [0,10]^2,
  periodic; Voronoi mesh from a 256x256 lattice with uniform jitter 0.3h (seed 3),
  multicolour-ordered; Legendre p=3 (40x40 blocks); DIRK3 at CFL 20
  (dt = 20 h / ((2p+1) max(|v|+c)), h = 10/256) with Alexander's L-stable SDIRK3 tableau (the
  reference's 3-stage tableau, --dirk3 reference, is not A-stable: |R(-inf)| = 1.78, and at
  CFL 20 it amplifies the under-resolved modes every step until the state is non-physical);
  Newton to 5e-13 on the L2 residual norm; flexible GMRES(50) to 1e-10 with block DILU; exact
  periodic traces and exact Euler off-diagonal Jacobian blocks.
Solver configuration of the headline: the Newton operator M - dt a J(u) is applied at the
quadrature points (cpg_newton_prepare once per Newton matrix, cpg_newton_apply_prepared per
GMRES iteration; equal to the assembled matrix's product to rounding), and the assembled matrix
and its block DILU are refreshed once per step (--pc-refresh 2; the reference refreshes them at
every Newton iteration, --pc-refresh 0).  The Newton iteration and the GMRES tolerance are those
of the reference policy; the same number of Newton and GMRES iterations results.
The headline value is the mixed-precision path (by default DILU factors stored in BF16 with one
power-of-two scale per block and applied with FP32 partial sums, inside FP64 flexible GMRES); the
FP64 path (FP64 factors, same solver) is timed in the same run and reported beside it, with the
difference between the two solutions after every timed step, and both are timed again under the
reference refresh policy with the matrix-free and with the assembled (BSR matvec) operator.

Prints ONE JSON line.  Only one GPU is used: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DIRK steps/s, Euler p=3 65k cells: mixed-prec vs FP64 on GPU at 1e-10 resid"
HBM_PEAK_GBS = 8000.0
# diagonal coefficient a_ss of the two 3-stage DIRK tableaux (the Newton matrix is M - dt a_ss J)
A_SS = {"reference": 0.5 * (1 + 0.435866521508459), "l-stable": 0.43586652150845899942}


def traffic(kernel, cells, deg, precision):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc (profiles/hbm_traffic.json,
    written from the PMC summaries in profiles/), or None where no profile covers the config."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)["entries"]
    except (OSError, ValueError, KeyError):
        return None
    for e in table:
        if (e["kernel"], e["cells"], e["deg"], e["precision"]) == (kernel, cells, deg, precision):
            return e["bytes"]
    return None


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--nx", type=int, default=256, help="lattice size (cells = nx^2)")
    p.add_argument("--deg", type=int, default=3)
    p.add_argument("--cfl", type=float, default=20.0)
    p.add_argument("--dirk3", choices=["l-stable", "reference"], default="l-stable",
                   help="3-stage DIRK tableau: Alexander's L-stable SDIRK3, or the reference's (not A-stable)")
    p.add_argument("--restart", type=int, default=50)
    p.add_argument("--gmres-maxit", type=int, default=20000, help="GMRES iteration limit per Newton solve")
    p.add_argument("--mixed", choices=["fp32", "bf16", "bf16-d32", "inner-fp32", "inner-bf16"], default="bf16",
                   help="fp32/bf16: DILU factors stored in that precision inside FP64 FGMRES; bf16-d32: L and DU "
                        "blocks in BF16, diagonal inverses in FP32; inner-*: FP32 GMRES(--inner) with those factors "
                        "as the FGMRES preconditioner")
    p.add_argument("--inner", type=int, default=20)
    p.add_argument("--jacobian", type=int, default=1, help="1: exact off-diagonal Jacobian, 0: reference form")
    p.add_argument("--operator", choices=["matrix-free", "assembled"], default="matrix-free",
                   help="GMRES operator: the Newton matrix applied at the quadrature points, or the BSR matvec of "
                        "the assembled matrix")
    p.add_argument("--trace", type=int, default=1, help="periodic trace: 1 exact, 0 reference (truncated)")
    p.add_argument("--pc-refresh", type=int, default=2, choices=[0, 1, 2, 3],
                   help="0: Newton matrix and preconditioner at every Newton iteration (reference); 1: once per stage; "
                        "2: once per step; 3: kept across steps, renewed when GMRES slows (matrix-free operator "
                        "only)")
    p.add_argument("--dilu-apply", choices=["single", "levels"], default="single",
                   help="block DILU apply of FP64 vectors: one launch per apply (rows wait on device-side "
                        "completion counters) or one launch per colour level")
    p.add_argument("--no-fp64", action="store_true")
    p.add_argument("--no-compare", dest="compare_operators", action="store_false",
                   help="skip the runs under the reference refresh policy (matrix-free and assembled operator)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-nx", type=int, default=64,
                   help="lattice size of the single-thread CPU-baseline sample (0: none)")
    p.add_argument("--workload", choices=["dirk3-euler", "batch", "pmc", "irk", "rk4-advection"],
                   default="dirk3-euler",
                   help="batch: 64 independent implicit advection solves advanced as one batch (prints its own "
                        "line); pmc: the short kernel sequence profiled with rocprofv3 --pmc for HBM traffic; "
                        "rk4-advection: explicit RK4 linear advection, 1,024 cells, p=2, 100 steps (its own line)")
    p.add_argument("--rk4-steps", type=int, default=100)
    p.add_argument("--batch-stages", type=int, default=2, help="batch: 1 backward Euler, 2 SDIRK2")
    p.add_argument("--batch-count", type=int, default=64)
    p.add_argument("--no-sequential", action="store_true", help="batch: skip the one-stepper-per-instance timing")
    p.add_argument("--batch-mode", choices=["persistent", "graph"], default="persistent",
                   help="batch: the timed solver (one cooperative launch per batched solve, or hipGraph-replayed "
                        "launches per phase); the other one is timed beside it unless --no-batch-compare")
    p.add_argument("--no-batch-compare", action="store_true")
    p.add_argument("--batch-team-detail", action="store_true",
                   help="batch, persistent: per instance [cells, degree, team size, team ms, Arnoldi steps]")
    p.add_argument("--batch-wg-per-cu", type=int, default=0,
                   help="batch, persistent: workgroups per CU (0: autotuned over the first solves)")
    p.add_argument("--irk-nx", type=int, default=512, help="irk: lattice size (cells = nx^2)")
    p.add_argument("--irk-cfl", type=float, default=10.0)
    p.add_argument("--irk-restart", type=int, default=60)
    p.add_argument("--irk-pc", default="bjacobi,dilu,wform",
                   help="irk: preconditioners, comma-separated (bjacobi: BF16 and FP64 factors; dilu: BF16 and FP32; "
                        "wform: the W form with stage-block-diagonal DILU, BF16)")
    return p.parse_args()


def workload(nx, deg):
    from cpolyg_amd import setups
    from cpolyg_amd.mesh import jittered_lattice_seeds, periodic_voronoi

    mesh = periodic_voronoi(jittered_lattice_seeds(nx, nx, 10.0, 10.0, 0.3, 3), 10.0, 10.0)
    u0 = setups.interp(mesh, deg, setups.vortex(beta=5.0, x0=5.0, y0=5.0, uinf=1.0, vinf=1.0, gamma=1.4), 4)
    return mesh, u0


def operator_bytes(mesh, deg, nc):
    """Algorithmic bytes of one matrix-free Newton-operator product (k_cellop<DEG, NC, true>):
    the point data of the Newton matrix (Euler: 4 doubles per volume quadrature point, 17 per edge
    quadrature point of every cell's edges), the volume points (x, y, weight, pad: 32 B each),
    the 88-byte edge records, the cell bounding boxes and point/edge offsets, v in and y out
    (N doubles each) and the mass blocks (bs^2 doubles per cell).  The neighbours' coefficients
    of v that the edge points read are counted once, in v."""
    bs = (deg + 1) * (deg + 2) // 2
    n = mesh.n_cells
    nqt = 3 if 2 * deg <= 2 else 6 if 2 * deg <= 4 else 12 if 2 * deg <= 6 else 25
    nql = 2 if 2 * deg <= 2 else 4 if 2 * deg <= 4 else 6 if 2 * deg <= 6 else 10
    nv = int(mesh.tptr[-1]) * nqt
    ne = len(mesh.pv)
    point = (4 * nv + 17 * ne * nql) * 8 if nc == 4 else 0
    return point + 32 * nv + 88 * ne + n * (32 + 8) + 2 * n * nc * bs * 8 + n * bs * bs * 8


def algorithmic_bytes(n, nnzb, bl, elem, scales, diag32=False):
    """matvec: blocks + colind + rowptr + x + y;  DILU apply: factor blocks (+ scales) + x in + x out
    (diag32: the n diagonal inverses in FP32, the other blocks at elem bytes with scales)."""
    N = n * bl
    matvec = nnzb * bl * bl * 8 + nnzb * 4 + (n + 1) * 4 + 2 * N * 8
    if diag32:
        apply = (nnzb - n) * bl * bl * elem + (nnzb - n) * 4 + n * bl * bl * 4 + 2 * N * 8
    else:
        apply = nnzb * bl * bl * elem + (nnzb * 4 if scales else 0) + 2 * N * 8
    return matvec, apply


def run_path(E, u0, dt, steps, warmup, solver, keep=False, label=""):
    """Warmup, then `steps` timed DIRK3 steps (the production path: captured GMRES cycles),
    then one more step with HIP events around every operator and preconditioner application
    (eager launches; not part of the timed region's wall time).  keep: a device copy of the
    state after every timed step (one 21 MB device-to-device copy per step at 65,536 cells)."""
    import torch
    from cpolyg_amd import hip

    S = hip.Stepper(E, 0, 3, solver)
    stream = torch.cuda.current_stream()
    u = torch.from_numpy(u0).cuda()
    nxt = torch.empty_like(u)
    t = 0.0
    for _ in range(warmup):
        S.advance(u, dt, t, nxt, stream)
        u, nxt = nxt, u
        t += dt
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    states = []
    for _ in range(steps):
        st = S.advance(u, dt, t, nxt, stream)
        stats.append(st.as_dict())
        u, nxt = nxt, u
        if keep:
            states.append(u.clone())
        t += dt
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"# {label}: {1e3 * wall / max(steps, 1):.1f} ms/step, "
          f"{sum(x['gmres_iters'] for x in stats) / max(steps, 1):.1f} GMRES iterations/step", file=sys.stderr,
          flush=True)
    S.enable_kernel_timing(True)
    S.advance(u, dt, t, nxt, stream)
    kt = S.kernel_timing()
    S.enable_kernel_timing(False)
    torch.cuda.synchronize()
    if keep:
        stats[-1]["states"] = states
    return u.cpu().numpy(), wall, stats, kt, u, t


def standalone_rates(E, u, t, scale, reps=100, single_launch=True):
    """Sustained HIP-event timing on torch's current stream (the stream the library launches on):
    `reps` back-to-back launches between two events, the way the solver issues them, of the BSR
    matvec, the matrix-free Newton apply and the out-of-place DILU apply z = M^-1 x in each factor
    precision, on the Newton matrix M + scale J(u, t).  These agree with rocprofv3's dispatch
    timestamps of the same launches to about 1%."""
    import torch
    from cpolyg_amd import hip
    from cpolyg_amd.spec import PC_DILU, PREC_BF16, PREC_BF16_D32, PREC_FP32, PREC_FP64

    stream = torch.cuda.current_stream()
    J = E.jacobian_matrix()
    E.jacobian(u, t, J, scale=scale, add_mass=True, stream=stream)
    x = torch.randn(E.n * E.bl, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    zs = [torch.empty_like(x) for _ in range(4)]

    def sustained(fn, n=reps):
        for k in range(3):
            fn(k)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for k in range(n):
            fn(k)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    E.newton_prepare(u, t, stream)
    out = {"bsr_matvec_ms": sustained(lambda k: J.matvec(x, y, stream), 30),
           "newton_prepare_ms": sustained(lambda k: E.newton_prepare(u, t, stream), 10),
           "newton_apply_ms": sustained(lambda k: E.newton_apply_prepared(scale, x, y, stream))}
    E.newton_prepare(u, t, stream)
    for name, prec in (("bf16", PREC_BF16), ("bf16-d32", PREC_BF16_D32), ("fp32", PREC_FP32), ("fp64", PREC_FP64)):
        P = hip.Prec(J, PC_DILU, prec)
        P.set_flow(int(single_launch))
        P.factor(J, stream)
        out[f"dilu_apply_{name}_ms"] = sustained(lambda k: P.apply_to(x, zs[k % 4], stream))
        out["dilu_levels"] = P.levels()
        del P
    return out


def cpu_phase_times(nx, deg, cfl, counts, a_ss, threads, gmres_its):
    """The oracle's per-phase times on the workload at nx^2 cells with `threads` host threads:
    one residual, one Jacobian, one DILU factor of the Newton matrix, and one flexible GMRES cycle
    of exactly `gmres_its` iterations (restart = gmres_its, unreachable tolerance: the oracle rounds
    maxit up to whole restart cycles), i.e. the iterations of one Newton solve of the GPU run with
    their Gram-Schmidt lengths (seconds each; the GMRES figure per iteration)."""
    import oracle as O
    from cpolyg_amd import setups
    from cpolyg_amd.spec import EquationSpec

    mesh, u = workload(nx, deg)
    orc = O.Oracle(mesh)
    spec = EquationSpec.euler(periodic_trace=counts["trace"], jacobian=counts["jacobian"])
    dt = setups.dg_time_step(cfl, 10.0 / nx, deg, setups.euler_max_speed(u, deg))
    O.set_threads(threads)
    try:
        t0 = time.perf_counter()
        b = orc.assemble(spec, deg, u, 0.0)
        t_res = time.perf_counter() - t0
        t0 = time.perf_counter()
        B = orc.jacobian(spec, deg, u, 0.0)
        t_jac = time.perf_counter() - t0
        B *= -dt * a_ss
        B[orc.rowptr[:-1]] += orc.mass(deg, 4)
        t0 = time.perf_counter()
        O.dilu_factor(orc.rowptr, orc.colind, B)
        t_fac = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.gmres(orc.rowptr, orc.colind, B, b, m=gmres_its, tol=1e-30, maxit=gmres_its, pc_kind=2, mode=1)
        t_gm = (time.perf_counter() - t0 - t_fac) / gmres_its
    finally:
        O.set_threads(1)
    return mesh.n_cells, t_res, t_jac, t_fac, t_gm


def cpu_baseline(nx, deg, cfl, counts, n_target, a_ss, threads, sample_nx):
    """The oracle (C restatement of the reference algorithm) timed per phase on the host and
    composed with the per-step phase counts of the GPU run under the reference's refresh policy
    (one Jacobian and one DILU factor per Newton iteration, a residual per Newton iteration and
    per stage, the GMRES iterations of the GPU run).  Main figure: all `threads` host threads on
    the full workload (nx^2 cells, measured directly).  Beside it: one thread on a sample_nx^2-cell
    sample of the same workload, scaled linearly in cells to nx^2."""
    def compose(n, t_res, t_jac, t_fac, t_gm, scale=1.0):
        return scale * (counts["residuals"] * t_res + counts["newton"] * (t_jac + t_fac) + counts["gmres"] * t_gm)

    it = max(1, int(round(counts["gmres"] / max(counts["newton"], 1.0))))   # iterations per Newton solve
    n, t_res, t_jac, t_fac, t_gm = cpu_phase_times(nx, deg, cfl, counts, a_ss, threads, it)
    step = compose(n, t_res, t_jac, t_fac, t_gm)
    out = {"value": 1.0 / step, "unit": "DIRK steps/s", "cores": threads, "kind": "port",
           "sample": (f"oracle with {threads} threads on the full workload ({n} cells, measured directly): 1 residual "
                      f"{t_res:.2f}s, 1 Jacobian {t_jac:.2f}s, 1 DILU factor (level-parallel) {t_fac:.2f}s, {it} "
                      f"flexible GMRES iterations (one solve's worth) {t_gm:.3f}s each; composed with the GPU run's per-step counts "
                      f"({counts['residuals']:.1f} residuals, {counts['newton']:.1f} Jacobians+factors as the "
                      f"reference refreshes them, {counts['gmres']:.1f} GMRES iterations)"),
           "s_per_step": step, **host_cpu()}
    if sample_nx:
        n1, r1, j1, f1, g1 = cpu_phase_times(sample_nx, deg, cfl, counts, a_ss, 1, it)
        step1 = compose(n1, r1, j1, f1, g1, n_target / n1)
        out["single_core"] = {"value": 1.0 / step1, "s_per_step": step1, "cores": 1,
                              "sample": (f"oracle on 1 thread at {n1} cells: residual {r1:.2f}s, Jacobian {j1:.2f}s, "
                                         f"DILU factor {f1:.2f}s, FGMRES iteration {g1:.3f}s; scaled x{n_target / n1:g} "
                                         f"to {n_target} cells")}
    return out


def bench_rk4(args):
    """Linear advection with constant velocity (1, 1) on the periodic unit square, Voronoi mesh of
    1,024 cells from uniform random seeds (seed 1), p = 2, u0 = exp(-100 |x - (1/2, 1/2)|^2), RK4
    at dt = 0.1 h_min (h_min the smallest sqrt(cell area)), --rk4-steps steps in FP64 with exact
    periodic traces (RK4::advance, src/TimeIntegration.cpp:14-22, as driven by
    examples/ExpAdv.cpp:66-83).  Times the steps (one warm-up run of the same length first),
    reports the L2 error against the advected exact Gaussian, and times the oracle's RK4 step on
    the host (one core) on the same mesh and state."""
    import numpy as np
    import torch

    from cpolyg_amd import hip
    from cpolyg_amd.mesh import periodic_voronoi, uniform_seeds
    from cpolyg_amd.spec import FIELD_GAUSSIAN, EquationSpec, SolverSpec

    deg, nsteps = 2, args.rk4_steps
    mesh = periodic_voronoi(uniform_seeds(1024, 1.0, 1.0, 1), 1.0, 1.0)
    spec = EquationSpec.constant_advection(1.0, 1.0, periodic_trace=1)
    E = hip.Equation(hip.Mesh(mesh), spec, deg)
    dt = 0.1 * mesh.h_min()
    par = (0.5, 0.5, 100.0, 1.0, 1.0, 1.0, 1.0)
    u0 = torch.zeros(E.n * E.bl, dtype=torch.float64, device="cuda")
    E.interp_field(FIELD_GAUSSIAN, par, 0.0, u0)
    S = hip.Stepper(E, 2, 4, SolverSpec())
    a, b = u0.clone(), torch.empty_like(u0)
    for k in range(nsteps):                    # warm-up run
        S.advance(a, dt, k * dt, b)
        a, b = b, a
    a.copy_(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(nsteps):
        S.advance(a, dt, k * dt, b)
        a, b = b, a
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    err0 = float(E.l2_error_field(u0, FIELD_GAUSSIAN, par, 0.0)[0])
    err = float(E.l2_error_field(a, FIELD_GAUSSIAN, par, nsteps * dt)[0])
    out = {"metric": "RK4 steps/s, linear advection p=2, 1,024-cell periodic Voronoi", "value": nsteps / wall,
           "unit": "RK4 steps/s", "n_gpus": 1, "steps": nsteps, "ms_per_step": 1e3 * wall / nsteps,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"constant advection (1,1), periodic unit square, Voronoi {mesh.n_cells} cells "
                                  f"(uniform seeds), p={deg}, RK4 dt = 0.1 h_min = {dt:.3e}, {nsteps} steps, exact "
                                  f"periodic traces", "cells": mesh.n_cells, "deg": deg, "dt": dt},
           "l2_error_initial_projection": err0, "l2_error_final": err}
    if not args.no_cpu:
        import oracle as O
        orc = O.Oracle(mesh)
        uc = u0.cpu().numpy()
        k_cpu = min(nsteps, 20)
        t0 = time.perf_counter()
        for k in range(k_cpu):
            uc = orc.rk4_step(spec, deg, uc, dt, k * dt)
        t_cpu = (time.perf_counter() - t0) / k_cpu
        out["cpu_baseline"] = {"value": 1.0 / t_cpu, "unit": "RK4 steps/s", "cores": 1, "kind": "port",
                               "sample": f"oracle RK4 step (o_rk4_step) on the same mesh and state, {k_cpu} steps",
                               **host_cpu()}
    print(json.dumps(out))


def host_cpu():
    """CPU model and logical core count of the host (reported with every CPU number)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def bench_batch(args):
    """Batched implicit advection: 64 instances (Voronoi meshes of the unit square with
    log-uniform 500..20,000 cells, degrees 1..4, Gaussian bumps, rotating field, Dirichlet
    inflow), CFL 5 each, SDIRK2 (or backward Euler), FP32 block DILU in FP64 flexible
    GMRES(30) to 1e-10, one batch per step.  The timed solver is the persistent one (every
    restart cycle of every instance in one cooperative launch per batched solve) or the
    launch-per-phase one with hipGraph-captured restart cycles; the other is timed beside it,
    and the same instances stepped one stepper at a time (same solver) for comparison."""
    import numpy as np
    import torch

    from cpolyg_amd import hip, setups
    from cpolyg_amd.spec import EquationSpec, GMRES_FLEXIBLE, PC_DILU, PREC_FP32, SolverSpec

    t0 = time.perf_counter()
    classes, meta = setups.batch_advection(args.batch_count, seed=4)
    t_setup = time.perf_counter() - t0
    spec = EquationSpec.rotating_advection()
    solver = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=30, rtol=1e-10, maxit=6000, pc_kind=PC_DILU,
                        precision=PREC_FP32, newton_tol=1e-12, newton_maxit=10)
    eqs = [(hip.Equation(hip.Mesh(um), spec, p), cp) for p, um, cp, _, _, _ in classes]
    dts = [d for _, _, d in meta]
    modes = {"persistent": hip.Batch.PERSISTENT, "graph": hip.Batch.GRAPH}

    def run(mode_name):
        t0 = time.perf_counter()
        B = hip.Batch(eqs, args.batch_stages, solver, mode=modes[mode_name])
        t_create = time.perf_counter() - t0
        if mode_name == "persistent" and args.batch_wg_per_cu:
            B.set_teams(args.batch_wg_per_cu)
        us = [torch.from_numpy(u0).cuda() for _, _, _, u0, _, _ in classes]
        outs = [torch.zeros_like(u) for u in us]
        t = 0.0
        # warmup: at least the persistent solver's autotuner trials (up to 4 candidates, two solves per step);
        # both modes take the same steps, so that their final states compare
        for _ in range(max(args.warmup, 3)):
            B.advance(us, dts, t, outs)
            us, outs = outs, us
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = []
        st = None
        for _ in range(args.steps):
            st, it = B.advance(us, dts, t, outs)
            its.append(it)
            us, outs = outs, us
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        final = [u.cpu().numpy() for u in us]
        return B, wall, np.asarray(its), st, t_create, final

    B, wall, its, st, t_create, final = run(args.batch_mode)
    cells = sum(n for n, _, _ in meta)
    dofs = sum(n * (p + 1) * (p + 2) // 2 for n, p, _ in meta)
    solver_desc = ("one cooperative launch per batched solve (persistent teams)" if args.batch_mode == "persistent"
                   else "hipGraph-replayed restart cycles")
    out = {"metric": "batched implicit steps: 64 advection solves, instance-steps/s", "value": len(meta) / wall,
           "unit": "instance-steps/s", "ms_per_batch_step": wall * 1e3, "steps": args.steps, "warmup": args.warmup,
           "config": {"workload": f"{len(meta)} rotating-field advection solves, Voronoi unit squares, "
                                  f"{min(n for n, _, _ in meta)}..{max(n for n, _, _ in meta)} cells, p 1..4, "
                                  f"{'SDIRK2' if args.batch_stages == 2 else 'backward Euler'} at CFL 5, FP32 DILU in "
                                  f"FP64 FGMRES(30) rtol 1e-10, {solver_desc}",
                      "instances": len(meta), "cells": cells, "dofs": dofs, "mode": args.batch_mode,
                      "classes": [{"deg": p, "instances": len(cp) - 1, "cells": int(um.n_cells)}
                                  for p, um, cp, _, _, _ in classes]},
           "gmres_steps_per_instance_step": {"min": int(its.min()), "max": int(its.max()), "mean": float(its.mean())},
           "newton_iters_per_step": st.newton_iters, "setup_s": t_setup, "create_s": t_create}
    if args.batch_mode == "persistent":
        tm = B.teams()
        out["teams"] = {"grid": tm["grid"], "wg_per_cu": tm["wg_per_cu"],
                        "team_wg": {"min": int(tm["team_wg"].min()), "max": int(tm["team_wg"].max())},
                        "team_ms_last_solve": {"min": float(tm["team_ms"].min()), "max": float(tm["team_ms"].max()),
                                               "mean": float(tm["team_ms"].mean())}}
        if args.batch_team_detail:
            out["teams"]["detail"] = [[int(n), int(d), int(w), round(float(ms), 3), int(i)] for (n, d, _), w, ms, i in
                                      zip(meta, tm["team_wg"], tm["team_ms"], its[-1])]
    else:
        out["graph"] = B.info()
    del B
    if not args.no_batch_compare:
        other = "graph" if args.batch_mode == "persistent" else "persistent"
        B2, wall2, its2, _, _, final2 = run(other)
        diff = max(float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)) for a, b in zip(final, final2))
        out[other] = {"ms_per_batch_step": wall2 * 1e3, "value": len(meta) / wall2,
                      "gmres_steps_per_instance_step_mean": float(its2.mean()),
                      "max_rel_diff_vs_timed": diff}
        out[f"speedup_vs_{other}"] = wall2 / wall
        del B2
    if not args.no_sequential:
        # the same instances one stepper at a time (same solver, same kernels, one launch per operation
        # per instance)
        seq = 0.0
        for c, (p, um, cp, u0, ids, meshes) in enumerate(classes):
            bs = (p + 1) * (p + 2) // 2
            for k, m in enumerate(meshes):
                q = sum(len(cc[2]) - 1 for cc in classes[:c]) + k
                E = hip.Equation(hip.Mesh(m), spec, p)
                S = hip.Stepper(E, 0, args.batch_stages, solver)
                ui = torch.from_numpy(u0[cp[k] * bs:cp[k + 1] * bs]).cuda()
                uo = torch.zeros_like(ui)
                S.advance(ui, dts[q], 0.0, uo)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                S.advance(ui, dts[q], 0.0, uo)
                torch.cuda.synchronize()
                seq += time.perf_counter() - t1
        out["sequential_ms_per_step"] = seq * 1e3
        out["batch_speedup_vs_sequential"] = seq / wall
    print(json.dumps(out))


def bench_irk(args):
    """Radau IIA (3 stages) for compressible Euler, density 1 + 0.2 sin(2 pi x) sin(2 pi y), velocity
    (1, 0.5), pressure 1, periodic unit square, Voronoi mesh of a jittered nx^2 lattice (seed 5), p=3.
    The stage-coupled Newton system (3 x 40 unknowns per cell) is applied matrix-free and
    preconditioned on its cell-interleaved form (120 x 120 blocks) by block Jacobi (factors in BF16
    and FP64) and by block DILU formed from the three stage Jacobians (BF16 and FP32: FP64 DILU
    factors of 262,144 cells are 211 GB); FP64 flexible GMRES(--irk-restart) to 1e-10; Newton to
    1e-10 (the reference's 5e-13 is an absolute L2 bound set for its small meshes)."""
    import numpy as np
    import torch

    from cpolyg_amd import hip, setups
    from cpolyg_amd.mesh import jittered_lattice_seeds, periodic_voronoi
    from cpolyg_amd.spec import (EquationSpec, GMRES_FLEXIBLE, OP_MATRIX_FREE, PC_BJACOBI, PC_DILU, PREC_BF16,
                                 PREC_FP32, PREC_FP64, SolverSpec)

    nx, deg = args.irk_nx, 3
    t0 = time.perf_counter()
    mesh = periodic_voronoi(jittered_lattice_seeds(nx, nx, 1.0, 1.0, 0.3, 5), 1.0, 1.0)
    u0 = setups.interp(mesh, deg, setups.sine_density(), 4)
    t_setup = time.perf_counter() - t0
    n = mesh.n_cells
    dt = setups.dg_time_step(args.irk_cfl, 1.0 / nx, deg, setups.euler_max_speed(u0, deg))
    E = hip.Equation(hip.Mesh(mesh), EquationSpec.euler(periodic_trace=1, jacobian=1), deg)
    runs = []
    for pc in args.irk_pc.split(","):
        if pc == "bjacobi":
            runs += [("bjacobi_bf16", PC_BJACOBI, PREC_BF16), ("bjacobi_fp64", PC_BJACOBI, PREC_FP64)]
        elif pc == "dilu":
            runs += [("dilu_bf16", PC_DILU, PREC_BF16), ("dilu_fp32", PC_DILU, PREC_FP32)]
        elif pc == "wform":
            # IRK::newAdvance's form: stage-block-diagonal DILU (40 x 40 blocks)
            runs += [("wform_dilu_bf16", PC_DILU, PREC_BF16)]
        else:
            raise SystemExit(f"unknown --irk-pc entry {pc!r}")
    res = {}
    free0, total = torch.cuda.mem_get_info()
    for name, pc, prec in runs:
        sp = SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=args.irk_restart, rtol=1e-10, maxit=6000,
                        pc_kind=pc, newton_tol=1e-10, newton_maxit=10, precision=prec,
                        newton_operator=OP_MATRIX_FREE, pc_refresh=args.pc_refresh)
        try:
            S = hip.Stepper(E, 3 if name.startswith("wform") else 1, 3, sp)
        except hip.CpgError as e:   # e.g. the factors of this precision do not fit
            res[name] = {"error": str(e)}
            print(f"# {name}: {e}", file=sys.stderr, flush=True)
            continue
        free1, _ = torch.cuda.mem_get_info()
        u = torch.from_numpy(u0).cuda()
        nxt = torch.empty_like(u)
        t = 0.0
        for _ in range(args.warmup):
            S.advance(u, dt, t, nxt)
            u, nxt = nxt, u
            t += dt
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sts = []
        for _ in range(args.steps):
            sts.append(S.advance(u, dt, t, nxt).as_dict())
            u, nxt = nxt, u
            t += dt
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        res[name] = {"ms_per_step": wall * 1e3, "steps_per_s": 1.0 / wall,
                     "newton_iters_per_step": sum(x["newton_iters"] for x in sts) / args.steps,
                     "gmres_iters_per_step": sum(x["gmres_iters"] for x in sts) / args.steps,
                     "phase_ms_per_step": {k: sum(x[k] for x in sts) / args.steps
                                           for k in ("ms_residual", "ms_jacobian", "ms_factor", "ms_gmres")},
                     "last_resid": sts[-1]["last_resid"],
                     "stepper_device_bytes": free0 - free1, "u": u.cpu().numpy()}
        print(f"# {name}: {res[name]['ms_per_step']:.1f} ms/step, {res[name]['gmres_iters_per_step']:.1f} "
              f"GMRES iterations", file=sys.stderr, flush=True)
        del S, u, nxt
        torch.cuda.synchronize()
    # every run against the first FP64-factor one (or the first run)
    done = [k for k in res if "u" in res[k]]
    base = next((k for k in done if k.endswith("fp64")), done[0])
    ub = res[base]["u"]
    for k in done:
        res[k]["max_rel_diff_vs_" + base] = float(np.max(np.abs(res[k]["u"] - ub)) / np.max(np.abs(ub)))
    for k in done:
        del res[k]["u"]
    head = next((k for k in ("dilu_bf16", "bjacobi_bf16") if k in done), base)
    out = {"metric": "Radau IIA (3-stage) steps/s, Euler p=3, 262k cells, coupled system matrix-free",
           "value": res[head]["steps_per_s"], "unit": "IRK steps/s", "value_run": head, "steps": args.steps,
           "warmup": args.warmup,
           "config": {"workload": f"Euler density wave, periodic Voronoi {n} cells ({nx}^2 jittered lattice), p={deg}, "
                                  f"Radau IIA 3 stages at CFL {args.irk_cfl:g}, matrix-free coupled operator, "
                                  f"cell-interleaved 120x120 coupled blocks, FGMRES({args.irk_restart}) rtol 1e-10",
                      "cells": n, "dt": dt, "pc_refresh": args.pc_refresh, "colours": mesh.n_colours},
           "runs": res, "gpu_memory_total_bytes": total, "setup_s": t_setup}
    print(json.dumps(out))


def pmc_workload(args):
    """Kernel sequence of the HBM-traffic profile: a 1 GiB device copy of known size (calibrates
    FETCH_SIZE / WRITE_SIZE), three FP64 BSR matvecs of the headline Newton matrix, two DILU applies
    with factors in the --mixed precision (BF16 or FP32), the first in one launch, the second in one
    launch per level, then the matrix-free Newton operator of the same matrix: one prepare (point
    data) and three products.  Meant to run under rocprofv3 --pmc; prints nothing but a completion
    line."""
    import torch

    from cpolyg_amd import hip, setups
    from cpolyg_amd.spec import EquationSpec, PC_DILU, PREC_BF16, PREC_BF16_D32, PREC_FP32

    mesh, u0 = workload(args.nx, args.deg)
    E = hip.Equation(hip.Mesh(mesh), EquationSpec.euler(periodic_trace=args.trace, jacobian=args.jacobian), args.deg)
    J = E.jacobian_matrix()
    u = torch.from_numpy(u0).cuda()
    dt = setups.dg_time_step(args.cfl, 10.0 / args.nx, args.deg, setups.euler_max_speed(u0, args.deg))
    E.jacobian(u, 0.0, J, scale=-dt * A_SS[args.dirk3], add_mass=True)
    P = hip.Prec(J, PC_DILU, {"bf16": PREC_BF16, "bf16-d32": PREC_BF16_D32}.get(args.mixed, PREC_FP32))
    P.factor(J)
    x = torch.randn(E.n * E.bl, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    for _ in range(3):
        J.matvec(x, y)
    P.apply(x)          # one launch (k_dilu_flow)
    P.set_flow(0)
    P.apply(x)          # one launch per level (k_rowstream)
    E.newton_prepare(u, 0.0)
    for _ in range(3):
        E.newton_apply_prepared(-dt * A_SS[args.dirk3], x, y)
    torch.cuda.synchronize()
    print("pmc workload done", flush=True)


def main():
    args = parse()
    if args.workload == "pmc":
        pmc_workload(args)
        return
    if args.workload == "irk":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        bench_irk(args)
        return
    if args.workload == "rk4-advection":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        bench_rk4(args)
        return
    if args.workload == "batch":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        bench_batch(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path is measured on a single GPU "
              f"and has no multi-GPU run.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from cpolyg_amd import hip, setups
    from cpolyg_amd.spec import (DIRK3_L_STABLE, DIRK3_REFERENCE, EquationSpec, GMRES_FLEXIBLE, OP_ASSEMBLED,
                                 OP_MATRIX_FREE, PC_DILU, PC_INNER_GMRES, PREC_BF16, PREC_BF16_D32, PREC_FP32, PREC_FP64,
                                 SolverSpec)

    deg = args.deg
    t_setup = time.perf_counter()
    mesh, u0 = workload(args.nx, deg)
    n = mesh.n_cells
    print(f"# mesh {n} cells, {mesh.n_colours} colours, initial state: {time.perf_counter() - t_setup:.1f} s",
          file=sys.stderr, flush=True)
    dt = setups.dg_time_step(args.cfl, 10.0 / args.nx, deg, setups.euler_max_speed(u0, deg))
    spec = EquationSpec.euler(periodic_trace=args.trace, jacobian=args.jacobian)
    E = hip.Equation(hip.Mesh(mesh), spec, deg)
    bl, nnzb = E.bl, E.nnzb

    dirk3 = DIRK3_L_STABLE if args.dirk3 == "l-stable" else DIRK3_REFERENCE

    def solver(prec, pc=PC_DILU, op=None, refresh=None):
        # each stage's Newton iteration starts from the previous stage's increment and halves an
        # update that would leave the physical states (the reference starts from zero; at CFL 20 its
        # first update on the third stage can overshoot to negative pressure)
        return SolverSpec(gmres_mode=GMRES_FLEXIBLE, restart=args.restart, rtol=1e-10, maxit=args.gmres_maxit,
                          pc_kind=pc, newton_tol=5e-13, newton_maxit=100, precision=prec, inner_iters=args.inner,
                          newton_guess=1, newton_backtrack=1, dirk_tableau=dirk3,
                          pc_level_launches=int(args.dilu_apply == "levels"),
                          pc_refresh=args.pc_refresh if refresh is None else refresh,
                          newton_operator=(OP_MATRIX_FREE if args.operator == "matrix-free" else OP_ASSEMBLED)
                          if op is None else op)

    mixed_prec = {"bf16": PREC_BF16, "bf16-d32": PREC_BF16_D32}.get(args.mixed.replace("inner-", ""), PREC_FP32)
    mixed_pc = PC_INNER_GMRES if args.mixed.startswith("inner") else PC_DILU
    steps = args.steps
    a_ss = A_SS[args.dirk3]
    u_mix, wall_mix, st_mix, kt, u_last, t_last = run_path(E, u0, dt, steps, args.warmup,
                                                           solver(mixed_prec, mixed_pc), keep=not args.no_fp64,
                                                           label=f"mixed {args.mixed}")
    res = {}
    if not args.no_fp64:
        u64, wall64, st64, kt64, _, _ = run_path(E, u0, dt, steps, args.warmup, solver(PREC_FP64), keep=True,
                                              label="fp64")
        res["fp64_steps_per_s"] = steps / wall64
        res["fp64_ms_per_step"] = 1e3 * wall64 / steps
        res["speedup_mixed_vs_fp64"] = wall64 / wall_mix
        res["mixed_vs_fp64_max_rel_diff"] = float(np.max(np.abs(u_mix - u64)) / np.max(np.abs(u64)))
        # the mixed-precision solution against the FP64 path's after every timed step (both
        # start from the same state and take the same dt): max |u_mixed - u_fp64| / max |u_fp64|
        sm, s64 = st_mix[-1].pop("states"), st64[-1].pop("states")
        res["mixed_vs_fp64_max_rel_diff_per_step"] = [float((a - b).abs().max() / b.abs().max())
                                                      for a, b in zip(sm, s64)]
        del sm, s64
        res["fp64_gmres_iters_per_step"] = sum(s["gmres_iters"] for s in st64) / steps
        res["fp64_newton_iters_per_step"] = sum(s["newton_iters"] for s in st64) / steps
        res["fp64_phase_ms_per_step"] = {k: sum(s[k] for s in st64) / steps for k in
                                         ("ms_residual", "ms_jacobian", "ms_factor", "ms_gmres")}
    if args.compare_operators:
        # the same pair of paths under the reference's refresh policy (Newton matrix and factor at
        # every Newton iteration), with the matrix-free and with the assembled (BSR) operator
        for key, op in (("reference_refresh_matrix_free", OP_MATRIX_FREE), ("reference_refresh_assembled", OP_ASSEMBLED)):
            if key.endswith("matrix_free") and args.pc_refresh == 0 and args.operator == "matrix-free":
                continue
            cmp_ = {"operator": "matrix-free" if op == OP_MATRIX_FREE else "assembled", "pc_refresh": 0}
            try:
                ua, wa, sta, _, _, _ = run_path(E, u0, dt, steps, args.warmup, solver(mixed_prec, mixed_pc, op, 0),
                                                label=f"{key} mixed")
                cmp_["mixed_ms_per_step"] = 1e3 * wa / steps
                cmp_["mixed_max_rel_diff_vs_headline"] = float(np.max(np.abs(ua - u_mix)) / np.max(np.abs(u_mix)))
                cmp_["gmres_iters_per_step"] = sum(x["gmres_iters"] for x in sta) / steps
                if not args.no_fp64:
                    _, wa64, _, _, _, _ = run_path(E, u0, dt, steps, args.warmup, solver(PREC_FP64, PC_DILU, op, 0),
                                                   label=f"{key} fp64")
                    cmp_["fp64_ms_per_step"] = 1e3 * wa64 / steps
                    cmp_["speedup_mixed_vs_fp64"] = wa64 / wa
            except hip.CpgError as e:   # a comparison leg that fails is reported, not fatal
                cmp_["error"] = str(e)
                print(f"# {key}: {e}", file=sys.stderr, flush=True)
            res[key] = cmp_
    # per-launch rates on the last Newton matrix of the mixed run (stage-3 scale)
    rates = standalone_rates(E, u_last, t_last, -dt * a_ss, single_launch=args.dilu_apply == "single")
    elem = {PREC_BF16: 2, PREC_BF16_D32: 2, PREC_FP32: 4}[mixed_prec]
    b_mv, b_ap = algorithmic_bytes(n, nnzb, bl, elem, mixed_prec != PREC_FP32, mixed_prec == PREC_BF16_D32)
    b_ap32 = algorithmic_bytes(n, nnzb, bl, 4, False)[1]
    b_ap64 = algorithmic_bytes(n, nnzb, bl, 8, False)[1]
    nm, msm, na, msa = kt
    op_ms = msm / max(nm, 1)
    ap_ms = msa / max(na, 1)
    if mixed_pc != PC_DILU:
        res["inner_solve"] = {"iterations": args.inner, "factors": args.mixed, "avg_ms": ap_ms, "launches": na}
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9   # noqa: E731
    b_op = operator_bytes(mesh, deg, 4)
    mv_gbs = gbs(b_mv, rates["bsr_matvec_ms"])
    newton = sum(s["newton_iters"] for s in st_mix) / steps
    gm = sum(s["gmres_iters"] for s in st_mix) / steps
    if mixed_pc == PC_DILU:
        # dominant kernel of the mixed step: the DILU apply (one launch, or one row-stream launch per
        # colour level and sweep, forward level 0 and backward level 0 excepted), at its sustained
        # rate on the timed run's last Newton matrix (the in-solver event figure is reported beside it:
        # inside the solver's timing step HIP events read 7-9% below rocprofv3's dispatch timestamps
        # of the same launches, while the sustained figure agrees with them)
        ap_sus = rates[f"dilu_apply_{args.mixed}_ms"]
        ap_gbs = gbs(b_ap, ap_sus)
        lev = rates["dilu_levels"]
        if mixed_prec == PREC_BF16_D32 and args.dilu_apply == "single":
            shape = (f"two launches: the rows of {lev - 1} forward levels (level 0 a no-op), then of {lev} backward "
                     f"levels, each backward row forming h = D^-1 y from the FP32 inverse and x = h - DU x from the "
                     f"BF16 stream; rows wait on completion counters of the rows they read")
        elif mixed_prec == PREC_BF16_D32:
            shape = (f"{3 * lev - 1} row-stream launches: {lev - 1} forward levels (level 0 a no-op), then per "
                     f"backward level the FP32 D^-1 products and the BF16 DU stream")
        elif args.dilu_apply == "single":
            shape = (f"one k_dilu_flow launch: the rows of {lev - 1} forward levels (level 0 a no-op), the last "
                     f"finishing its rows' backward step, then of {lev - 1} backward levels, each row waiting on "
                     f"completion counters of the rows it reads")
        else:
            shape = (f"{2 * lev - 2} row-stream launches: {lev} colour levels x 2 sweeps, forward level 0 a no-op, "
                     f"the last forward level finishing its rows' backward step")
        roof = {"bound": "hbm", "kernel": f"dilu_apply ({args.mixed} factors, {bl}x{bl} blocks, {shape})",
                "achieved": ap_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ap_gbs / HBM_PEAK_GBS,
                "traffic": traffic("dilu_apply", n, deg, args.mixed), "algorithmic_bytes": b_ap, "avg_ms": ap_sus,
                "avg_ms_source": "100 back-to-back out-of-place applies, HIP events at both ends",
                "avg_ms_in_solver": ap_ms, "frac_in_solver": gbs(b_ap, ap_ms) / HBM_PEAK_GBS, "launches": na}
    else:
        roof = {"bound": "hbm", "kernel": "bsr_matvec (FP64, 40x40 blocks)", "achieved": mv_gbs,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": mv_gbs / HBM_PEAK_GBS,
                "traffic": traffic("bsr_matvec", n, deg, "fp64"), "algorithmic_bytes": b_mv,
                "avg_ms": rates["bsr_matvec_ms"]}
    headline = args.nx == 256 and deg == 3 and args.cfl == 20.0
    out = {
        "metric": METRIC if headline else (f"DIRK steps/s, Euler p={deg} {n} cells CFL {args.cfl:g}: mixed-prec vs "
                                           f"FP64 on GPU at 1e-10 resid"),
        "value": steps / wall_mix,
        "unit": "DIRK steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_mix / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f64 outer / {args.mixed} preconditioner factors",
        "data": "synthetic",
        "config": {"workload": f"Euler isentropic vortex, periodic Voronoi {n} cells ({args.nx}^2 jittered "
                               f"lattice), p={deg} ({bl}x{bl} blocks), DIRK3 "
                               f"({'Alexander L-stable SDIRK3' if args.dirk3 == 'l-stable' else 'reference tableau'}) "
                               f"CFL {args.cfl:g}, FGMRES({args.restart}) "
                               f"rtol 1e-10, block DILU, {args.operator} Newton operator, preconditioner refreshed "
                               f"{['every Newton iteration', 'once per stage', 'once per step', 'kept across steps'][args.pc_refresh]}",
                   "cells": n, "deg": deg, "pc_refresh": args.pc_refresh, "dirk3_tableau": args.dirk3,
                   "block": bl, "nnzb": nnzb, "dt": dt, "cfl": args.cfl, "mixed": args.mixed,
                   "operator": args.operator, "jacobian": "exact" if args.jacobian else "reference",
                   "periodic_trace": "exact" if args.trace else "reference", "colours": mesh.n_colours},
        "roofline": roof,
        "in_solver_ms": {"operator_avg_ms": op_ms, "operator_launches": nm, "operator": args.operator,
                         "prec_apply_avg_ms": ap_ms, "prec_applies": na},
        "bsr_matvec": {"algorithmic_bytes": b_mv, "avg_ms": rates["bsr_matvec_ms"], "achieved_gbs": mv_gbs,
                       "frac": mv_gbs / HBM_PEAK_GBS, "traffic": traffic("bsr_matvec", n, deg, "fp64")},
        "newton_operator_ms": {"prepare_per_newton_matrix": rates["newton_prepare_ms"],
                               "apply_per_product": rates["newton_apply_ms"]},
        # the GMRES operator of the headline (replaces BlockMatrix::matvec): sustained per-product
        # rate (and the in-solver event figure) against its algorithmic bytes
        "newton_operator_roofline": {"bound": "hbm", "kernel": f"k_cellop<{deg},4,true> (matrix-free Newton operator)",
                                     "algorithmic_bytes": b_op, "avg_ms": rates["newton_apply_ms"],
                                     "achieved": gbs(b_op, rates["newton_apply_ms"]), "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": gbs(b_op, rates["newton_apply_ms"]) / HBM_PEAK_GBS,
                                     "avg_ms_in_solver": op_ms, "frac_in_solver": gbs(b_op, op_ms) / HBM_PEAK_GBS,
                                     "traffic": traffic("newton_apply", n, deg, "fp64")},
        "dilu_apply_standalone": {p: {"avg_ms": rates[f"dilu_apply_{p}_ms"], "algorithmic_bytes": b,
                                      "achieved_gbs": gbs(b, rates[f"dilu_apply_{p}_ms"]),
                                      "frac": gbs(b, rates[f"dilu_apply_{p}_ms"]) / HBM_PEAK_GBS}
                                  for p, b in (("bf16", algorithmic_bytes(n, nnzb, bl, 2, True)[1]),
                                               ("bf16-d32", algorithmic_bytes(n, nnzb, bl, 2, True, True)[1]),
                                               ("fp32", b_ap32), ("fp64", b_ap64))},
        "newton_iters_per_step": newton,
        "gmres_iters_per_step": gm,
        "phase_ms_per_step": {k: sum(s[k] for s in st_mix) / steps for k in
                              ("ms_residual", "ms_jacobian", "ms_factor", "ms_gmres")},
        **res,
    }
    if not args.no_cpu:
        counts = {"residuals": newton + 3, "newton": newton, "gmres": gm, "trace": args.trace,
                  "jacobian": args.jacobian}
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
        out["cpu_baseline"] = cpu_baseline(args.nx, deg, args.cfl, counts, n, a_ss, min(threads, 16), args.cpu_nx)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
