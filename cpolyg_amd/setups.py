"""Initial data for the workloads.

:func:`interp` is ``MeshFn::interp`` (``src/MeshFn.cpp:56-145``): a least-squares fit
of the Legendre basis to samples on the ``(p+1) x (p+1)`` Gauss-Legendre tensor grid
of each cell's bounding box, vectorised over cells.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh import PolyMesh
from .spec import bs_of, isentropic_vortex


def legendre_basis(deg: int, xi: np.ndarray, eta: np.ndarray) -> np.ndarray:
    """phi_k(xi, eta) = P_a(xi) P_b(eta), k over a = 0..p, b = 0..p-a
    (``src/Legendre.h:20-35``).  Returns ``(..., bs)``."""
    xi = np.asarray(xi, np.float64)
    eta = np.asarray(eta, np.float64)
    Px = [np.ones_like(xi), xi]
    Py = [np.ones_like(eta), eta]
    for n in range(1, deg):
        Px.append(((2 * n + 1) * xi * Px[n] - n * Px[n - 1]) / (n + 1))
        Py.append(((2 * n + 1) * eta * Py[n] - n * Py[n - 1]) / (n + 1))
    out = []
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            out.append(Px[a] * Py[b])
    return np.stack(out, axis=-1)


def interp(mesh: PolyMesh, deg: int, fn, nc: int) -> np.ndarray:
    """Project ``fn(x, y) -> (..., nc)`` onto the DG space; returns the flat
    coefficient vector ``u[i*bl + c*bs + k]``."""
    bs = bs_of(deg)
    q, _ = np.polynomial.legendre.leggauss(deg + 1)
    qi, qj = np.meshgrid(q, q, indexing="ij")
    G = legendre_basis(deg, qi.ravel(), qj.ravel())           # (Q, bs)
    bb = mesh.bb
    x = (qi.ravel()[None, :] + 1) * bb[:, 2:3] / 2.0 + bb[:, 0:1]
    y = (qj.ravel()[None, :] + 1) * bb[:, 3:4] / 2.0 + bb[:, 1:2]
    vals = np.asarray(fn(x, y), np.float64)
    if nc == 1 and vals.ndim == 2:
        vals = vals[..., None]
    coef = np.linalg.lstsq(G, vals.transpose(1, 0, 2).reshape(G.shape[0], -1), rcond=None)[0]
    coef = coef.reshape(bs, mesh.n_cells, nc).transpose(1, 2, 0)  # (np, nc, bs)
    return np.ascontiguousarray(coef.reshape(-1))


def gaussian(cx=0.5, cy=0.5, a=100.0):
    return lambda x, y: np.exp(-a * ((x - cx) ** 2 + (y - cy) ** 2))


def vortex(beta=5.0, x0=5.0, y0=5.0, uinf=1.0, vinf=1.0, gamma=1.4, t=0.0):
    return lambda x, y: isentropic_vortex(x, y, t, beta, x0, y0, uinf, vinf, gamma)


def sine_density(ux=1.0, uy=0.5, p=1.0, gamma=1.4, amp=0.2):
    """rho = 1 + amp sin(2 pi x) sin(2 pi y), velocity (ux, uy), pressure p."""
    def f(x, y):
        rho = 1 + amp * np.sin(2 * math.pi * x) * np.sin(2 * math.pi * y)
        return np.stack([rho, rho * ux, rho * uy, p / (gamma - 1) + 0.5 * rho * (ux * ux + uy * uy)], axis=-1)
    return f


def dg_time_step(cfl: float, h: float, deg: int, max_speed: float) -> float:
    """dt = CFL h / ((2p+1) max(|v|+c)): the usual DG CFL number, in units of the explicit
    stability limit of degree p.  The reference fixes dt in each driver and has no CFL
    convention (e.g. examples/PaperEuler.cpp:68: dt = h*cfl)."""
    return cfl * h / ((2 * deg + 1) * max_speed)


def euler_max_speed(u: np.ndarray, deg: int, gamma: float = 1.4) -> float:
    """max(|v| + c) over the cell means (coefficient 0 of each component)."""
    bs = bs_of(deg)
    U = u.reshape(-1, 4, bs)[:, :, 0]
    rho = U[:, 0]
    vx = U[:, 1] / rho
    vy = U[:, 2] / rho
    P = (gamma - 1) * (U[:, 3] - 0.5 * rho * (vx * vx + vy * vy))
    c = np.sqrt(gamma * P / rho)
    return float(np.max(np.hypot(vx, vy) + c))


def batch_advection(count: int = 64, seed: int = 4, nmin: int = 500, nmax: int = 20000, degs=(1, 2, 3, 4),
                    cfl: float = 5.0, width: float = 100.0):
    """Independent advection problems for a batched implicit step: instance q has a Voronoi
    mesh of the unit square (Dirichlet boundary, zero inflow data) from uniform random seeds,
    its cell count log-uniform in [nmin, nmax], a degree drawn from ``degs``, a Gaussian bump
    ``exp(-width |x - c|^2)`` with centre c uniform in [0.3, 0.7]^2, and the rotating field
    ``(-(y - 1/2), x - 1/2)``; dt_q = dg_time_step(cfl, 1/sqrt(n_q), p_q, max |beta|).

    Returns a list of classes ``(deg, union mesh, cell_ptr, u0, instance ids, instance meshes)``
    (one per degree, instances in id order inside a class) and the per-instance ``(n, deg, dt)``
    in class order."""
    from .mesh import box_voronoi, uniform_seeds, union_mesh

    rng = np.random.default_rng(seed)
    inst = []
    for q in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        p = int(degs[int(rng.integers(0, len(degs)))])
        cx, cy = rng.uniform(0.3, 0.7, 2)
        inst.append((q, n, p, float(cx), float(cy)))
    vmax = math.sqrt(0.5)   # max |(-(y-1/2), x-1/2)| on the unit square
    classes, meta = [], []
    for p in sorted({i[2] for i in inst}):
        members = [i for i in inst if i[2] == p]
        meshes, us, ids = [], [], []
        for q, n, _, cx, cy in members:
            m = box_voronoi(uniform_seeds(n, 1.0, 1.0, 1000 * seed + q), 1.0, 1.0)
            meshes.append(m)
            us.append(interp(m, p, gaussian(cx, cy, width), 1))
            ids.append(q)
            meta.append((m.n_cells, p, dg_time_step(cfl, 1.0 / math.sqrt(m.n_cells), p, vmax)))
        um, cptr = union_mesh(meshes)
        classes.append((p, um, cptr, np.concatenate(us), ids, meshes))
    return classes, meta
