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


@pytest.mark.parametrize("order", ["dsatur", "greedy"])
def test_multicolour_is_proper_and_balanced(order):
    """Both colourings are proper; the default one has at most 5 classes of near-equal size on a
    jittered Voronoi lattice (each class is one launch per DILU sweep)."""
    base = periodic_voronoi(jittered_lattice_seeds(24, 24, 1.0, 1.0, 0.3, 2), 1.0, 1.0, colour=False)
    m = base.multicolour(order)
    m.check()
    adj = m.adjacency()
    assert all(m.colour[i] != m.colour[j] for i in range(m.n_cells) for j in adj[i])
    sizes = np.diff(m.colour_ptr)
    assert sizes.sum() == base.n_cells
    if order == "dsatur":
        assert m.n_colours <= 5
        assert sizes.max() - sizes.min() <= max(2, sizes.max() // 50)
    # a renumbering of the same cells
    np.testing.assert_allclose(np.sort(m.areas()), np.sort(base.areas()))


def test_save_load_roundtrip(tmp_path):
    m = periodic_voronoi(jittered_lattice_seeds(6, 6, 1.0, 1.0, 0.3, 9), 1.0, 1.0)
    m.save(str(tmp_path / "m.npz"))
    m2 = PolyMesh.load(str(tmp_path / "m.npz"))
    for k, v in m.arrays().items():
        np.testing.assert_array_equal(v, m2.arrays()[k])
    np.testing.assert_array_equal(m.colour_ptr, m2.colour_ptr)


def test_union_mesh_is_disjoint_union():
    """Each part of a union keeps its neighbours, boundary data and areas, renumbered by its
    cell offset; the Jacobian pattern is block diagonal by part (what the batched solver
    assembles and factors as one system per polynomial degree)."""
    from cpolyg_amd.mesh import box_voronoi, periodic_voronoi, jittered_lattice_seeds, uniform_seeds, union_mesh
    parts = [box_voronoi(uniform_seeds(n, 1.0, 1.0, s), 1.0, 1.0) for n, s in ((30, 1), (55, 2))]
    parts.append(periodic_voronoi(jittered_lattice_seeds(4, 4, 1.0, 1.0, 0.3, 5), 1.0, 1.0))
    u, cp = union_mesh(parts)
    u.check()
    assert list(cp) == [0, 30, 85, 101]
    rp, ci = u.bsr_pattern()
    for q, m in enumerate(parts):
        r2, c2 = m.bsr_pattern()
        np.testing.assert_array_equal(rp[cp[q]:cp[q + 1] + 1] - rp[cp[q]], r2)
        cols = ci[rp[cp[q]]:rp[cp[q + 1]]]
        assert cols.min() >= cp[q] and cols.max() < cp[q + 1]
        np.testing.assert_array_equal(cols - cp[q], c2)
        np.testing.assert_allclose(u.areas()[cp[q]:cp[q + 1]], m.areas(), rtol=0, atol=0)
        np.testing.assert_array_equal(u.bb[cp[q]:cp[q + 1]], m.bb)


def test_batch_workload_shapes():
    from cpolyg_amd import setups
    classes, meta = setups.batch_advection(count=9, seed=5, nmin=20, nmax=80)
    assert sum(len(cp) - 1 for _, _, cp, _, _, _ in classes) == 9 == len(meta)
    for p, um, cp, u0, ids, meshes in classes:
        assert [mm.n_cells for mm in meshes] == list(np.diff(cp))
        assert um.n_cells == cp[-1] and len(u0) == um.n_cells * (p + 1) * (p + 2) // 2
        assert ids == sorted(ids)
    assert all(20 <= n <= 80 and dt > 0 for n, _, dt in meta)
