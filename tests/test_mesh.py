"""Mesh pipeline (CPU): periodic / box Voronoi, the reference's periodic quads and the
multicolour renumbering."""
import numpy as np
import pytest

from cpolyg_amd.mesh import (box_voronoi, jittered_lattice_seeds, periodic_rectangle, periodic_voronoi,
                             uniform_seeds, PolyMesh)


@pytest.mark.parametrize("nx,L,jit", [(8, 10.0, 0.3), (16, 1.0, 0.3), (5, 1.0, 0.0)])
def test_periodic_voronoi(nx, L, jit):
    m = periodic_voronoi(jittered_lattice_seeds(nx, nx, L, L, jit, 4), L, L)
    m.check()
    assert m.n_cells == nx * nx
    assert abs(m.areas().sum() - L * L) < 1e-10 * L * L
    # periodic translations are lattice shifts and partners point back
    for row in m.bc:
        assert row[0] == 1
        t = m.v[row[5]] - m.v[row[3]]
        t2 = m.v[row[6]] - m.v[row[4]]
        assert np.allclose(t, t2, atol=1e-9 * L)
        assert np.allclose(np.round(t / L), t / L, atol=1e-9)
        assert np.abs(np.round(t / L)).sum() >= 1
    # each cell's polygon is counter-clockwise and convex
    for i in range(m.n_cells):
        p = m.v[m.polygon(i)]
        e = np.roll(p, -1, 0) - p
        cross = e[:, 0] * np.roll(e, -1, 0)[:, 1] - e[:, 1] * np.roll(e, -1, 0)[:, 0]
        assert np.all(cross > -1e-12)


def test_uniform_periodic_voronoi_1024():
    m = periodic_voronoi(uniform_seeds(1024, 1.0, 1.0, 1), 1.0, 1.0)
    m.check()
    assert abs(m.areas().sum() - 1.0) < 1e-10


def test_box_voronoi_boundary_on_box():
    m = box_voronoi(uniform_seeds(200, 1.0, 1.0, 2), 1.0, 1.0)
    m.check()
    assert abs(m.areas().sum() - 1.0) < 1e-10
    for i in range(m.n_cells):
        for e, k in enumerate(m.neighbours(i)):
            if k < 0:
                a = m.v[m.pv[m.pptr[i] + e]]
                b = m.v[m.pv[m.pptr[i] + (e + 1) % (m.pptr[i + 1] - m.pptr[i])]]
                on = [np.isclose(a[0], 0) and np.isclose(b[0], 0), np.isclose(a[0], 1) and np.isclose(b[0], 1),
                      np.isclose(a[1], 0) and np.isclose(b[1], 0), np.isclose(a[1], 1) and np.isclose(b[1], 1)]
                assert any(on)


def test_reference_periodic_rectangle_ordering():
    # src/Meshes.cpp:176-321: cell 0's edges are bottom (periodic id -1, partner Nx(Ny-1)),
    # right (cell 1), top (cell Nx), left (periodic id -2, partner Nx-1)
    m = periodic_rectangle(4, 4, 1.0, 1.0)
    m.check()
    assert list(m.neighbours(0)) == [-1, 1, 4, -2]
    assert list(m.bc[0][:3]) == [1, 0, 12]
    assert list(m.bc[1][:3]) == [1, 0, 3]
    rp, ci = m.bsr_pattern()
    assert list(ci[rp[0]:rp[1]]) == [0, 12, 1, 4, 3]   # diagonal first, then edge order


def test_multicolour_ordering_property():
    m = periodic_voronoi(jittered_lattice_seeds(16, 16, 1.0, 1.0, 0.3, 6), 1.0, 1.0)
    assert m.colour is not None and 4 <= m.n_colours <= 8
    m.check()   # neighbours i < j  <=>  colour(i) < colour(j)
    assert np.all(np.diff(m.colour) >= 0)


def test_save_load_roundtrip(tmp_path):
    m = periodic_voronoi(jittered_lattice_seeds(6, 6, 1.0, 1.0, 0.3, 9), 1.0, 1.0)
    m.save(str(tmp_path / "m.npz"))
    m2 = PolyMesh.load(str(tmp_path / "m.npz"))
    for k, v in m.arrays().items():
        np.testing.assert_array_equal(v, m2.arrays()[k])
    np.testing.assert_array_equal(m.colour_ptr, m2.colour_ptr)
