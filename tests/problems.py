"""Shared test problems (meshes, states and oracle-built matrices)."""
from __future__ import annotations

import functools

import numpy as np

from cpolyg_amd import setups
from cpolyg_amd.mesh import jittered_lattice_seeds, periodic_voronoi, periodic_rectangle, box_voronoi, uniform_seeds
from cpolyg_amd.spec import EquationSpec, bs_of


@functools.lru_cache(maxsize=None)
def vortex_mesh(nx=8, seed=1):
    return periodic_voronoi(jittered_lattice_seeds(nx, nx, 10.0, 10.0, 0.3, seed), 10.0, 10.0)


@functools.lru_cache(maxsize=None)
def unit_periodic_mesh(nx=8, seed=2):
    return periodic_voronoi(jittered_lattice_seeds(nx, nx, 1.0, 1.0, 0.3, seed), 1.0, 1.0)


@functools.lru_cache(maxsize=None)
def box_mesh(n=60, seed=3):
    return box_voronoi(uniform_seeds(n, 1.0, 1.0, seed), 1.0, 1.0)


def vortex_state(mesh, deg, perturb=0.0, seed=0):
    u = setups.interp(mesh, deg, setups.vortex(), 4)
    if perturb:
        bs = bs_of(deg)
        rng = np.random.default_rng(seed)
        d = rng.standard_normal(u.shape) * perturb
        d.reshape(-1, 4, bs)[:, :, 0] = 0.0          # keep cell means (positivity)
        u = u + d
    return u


def newton_matrix(orc, spec, deg, u, dt=0.3, a=0.7179332607542295):
    """B = M - dt*a*J(u) in the Jacobian pattern (what a DIRK stage factors)."""
    J = orc.jacobian(spec, deg, u)
    M = orc.mass(deg, spec.nc)
    B = -dt * a * J
    B[orc.rowptr[:-1]] += M
    return B


def dense(rowptr, colind, blocks):
    """Dense matrix from BSR blocks stored column-major: blocks[s, c, r] = entry (r, c)."""
    n = len(rowptr) - 1
    bl = blocks.shape[1]
    D = np.zeros((n * bl, n * bl))
    for i in range(n):
        for s in range(rowptr[i], rowptr[i + 1]):
            j = colind[s]
            D[i * bl:(i + 1) * bl, j * bl:(j + 1) * bl] += blocks[s].T
    return D


def rel(a, b):
    """max |a - b| / max |b|"""
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
