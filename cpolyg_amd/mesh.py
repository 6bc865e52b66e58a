"""Polygonal meshes for the implicit DG path.

A :class:`PolyMesh` carries the same public fields as cpolyg's ``PolyMesh``
(``src/PolyMesh.h:22-28``), flattened into arrays so they can cross a C ABI:

* ``v``      -- vertex coordinates, ``(nv, 2)`` float64            (``v``)
* ``pptr/pv`` -- polygon vertex lists in CSR form, counter-clockwise (``p``)
* ``p2p``    -- neighbour across edge ``pv[j] -> pv[j+1]`` of each polygon,
  same CSR as ``pv``; ``>= 0`` is a cell, ``< 0`` a boundary id   (``p2p``)
* ``bc``     -- ``(nbc, 7)`` int32 rows ``periodic, p1, p2, a1, b1, a2, b2``;
  row ``k`` describes boundary id ``-1-k``                        (``bc``,
  ``PolyMesh::BoundaryInfo`` ``src/PolyMesh.h:15-20``)
* ``tptr/txy`` -- sub-triangles of each polygon, ``(ntri, 6)`` corner
  coordinates ``x1 y1 x2 y2 x3 y3``                               (``tri[i].p``,
  ``src/Triangulation.h:15``)
* ``bb``     -- bounding boxes ``x0, y0, w, h`` (``computebb``, ``src/PolyMesh.cpp:70-104``)

Generators:

* :func:`periodic_voronoi` -- periodic Voronoi diagram in O(n log n) (Qhull on the
  seeds plus the periodic images that can touch the domain).  The reference builds
  only non-periodic diagrams through voro++ with an O(n^2) neighbour search
  (``src/PolyMesh.cpp:12-67,162-218``).
* :func:`box_voronoi` -- Voronoi diagram clipped to a rectangle (seeds mirrored
  across the sides), Dirichlet boundary id ``-1`` like ``computep2p``.
* :func:`periodic_rectangle` -- the reference's periodic quad mesh
  (``src/Meshes.cpp:176-321``), same vertex, edge and boundary-id order.
* :func:`jittered_lattice_seeds`, :func:`uniform_seeds` -- seed generators.

Every mesh can be renumbered into colour-major order (:meth:`PolyMesh.multicolour`)
so that for every pair of neighbours ``i < j`` implies ``colour(i) < colour(j)``:
the block DILU of ``src/Preconditioners.cpp:49-123`` taken in natural row order is
then exactly a colour-level schedule.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

__all__ = [
    "PolyMesh",
    "periodic_voronoi",
    "box_voronoi",
    "periodic_rectangle",
    "quad_rectangle",
    "jittered_lattice_seeds",
    "uniform_seeds",
    "union_mesh",
]


def _first_fit(adj: list[list[int]]) -> np.ndarray:
    """Greedy first-fit colouring in cell order."""
    colour = np.full(len(adj), -1, np.int64)
    for i in range(len(adj)):
        used = {colour[j] for j in adj[i] if colour[j] >= 0}
        c = 0
        while c in used:
            c += 1
        colour[i] = c
    return colour


def _dsatur(adj: list[list[int]]) -> np.ndarray:
    """Saturation-degree colouring (Brelaz 1979): repeatedly colour, with its smallest free
    colour, the uncoloured cell with the most distinct colours among its neighbours (ties:
    most neighbours, then lowest index)."""
    import heapq

    n = len(adj)
    colour = np.full(n, -1, np.int64)
    sat: list[set] = [set() for _ in range(n)]
    heap = [(0, -len(adj[i]), i) for i in range(n)]
    heapq.heapify(heap)
    while heap:
        s, d, v = heapq.heappop(heap)
        if colour[v] >= 0 or -s != len(sat[v]):
            continue   # stale entry
        c = 0
        while c in sat[v]:
            c += 1
        colour[v] = c
        for j in adj[v]:
            if colour[j] < 0 and c not in sat[j]:
                sat[j].add(c)
                heapq.heappush(heap, (-len(sat[j]), -len(adj[j]), j))
    return colour


def _balance(adj: list[list[int]], colour: np.ndarray, passes: int = 6) -> np.ndarray:
    """Move cells out of classes above the mean size into the smallest class that none of
    their neighbours holds, until no move is possible (the colouring stays proper)."""
    colour = colour.copy()
    nc = int(colour.max()) + 1 if len(colour) else 0
    if nc < 2:
        return colour
    sizes = np.bincount(colour, minlength=nc)
    mean = len(colour) / nc
    for _ in range(passes):
        moved = 0
        for v in range(len(colour)):
            c0 = colour[v]
            if sizes[c0] <= mean:
                continue
            used = {colour[j] for j in adj[v]}
            best = -1
            for c in range(nc):
                if c not in used and sizes[c] < sizes[c0] - 1 and (best < 0 or sizes[c] < sizes[best]):
                    best = c
            if best >= 0:
                sizes[c0] -= 1
                sizes[best] += 1
                colour[v] = best
                moved += 1
        if not moved:
            break
    return colour


@dataclasses.dataclass
class PolyMesh:
    v: np.ndarray
    pptr: np.ndarray
    pv: np.ndarray
    p2p: np.ndarray
    bc: np.ndarray
    tptr: np.ndarray
    txy: np.ndarray
    bb: np.ndarray
    width: float = 1.0
    height: float = 1.0
    colour: np.ndarray | None = None
    colour_ptr: np.ndarray | None = None

    @property
    def np_(self) -> int:
        return len(self.pptr) - 1

    @property
    def n_cells(self) -> int:
        return len(self.pptr) - 1

    @property
    def n_colours(self) -> int:
        return 0 if self.colour_ptr is None else len(self.colour_ptr) - 1

    def polygon(self, i: int) -> np.ndarray:
        return self.pv[self.pptr[i]:self.pptr[i + 1]]

    def neighbours(self, i: int) -> np.ndarray:
        return self.p2p[self.pptr[i]:self.pptr[i + 1]]

    def cell_neighbour(self, i: int, e: int) -> int:
        """Cell index across edge ``e`` of cell ``i`` (periodic resolved) or -1."""
        k = int(self.p2p[self.pptr[i] + e])
        if k >= 0:
            return k
        row = self.bc[-1 - k]
        return int(row[2]) if row[0] else -1

    def adjacency(self) -> list[list[int]]:
        adj = []
        for i in range(self.n_cells):
            nb = []
            for e in range(self.pptr[i + 1] - self.pptr[i]):
                j = self.cell_neighbour(i, e)
                if j >= 0:
                    nb.append(j)
            adj.append(nb)
        return adj

    def bsr_pattern(self) -> tuple[np.ndarray, np.ndarray]:
        """Block sparsity of the Jacobian (``Jacobian::Jacobian``, ``src/Jacobian.cpp:7-47``):
        the diagonal block first, then one block per non-Dirichlet edge in edge order."""
        rowptr = np.zeros(self.n_cells + 1, np.int32)
        cols = []
        for i in range(self.n_cells):
            cols.append(i)
            for e in range(self.pptr[i + 1] - self.pptr[i]):
                j = self.cell_neighbour(i, e)
                if j >= 0:
                    cols.append(j)
            rowptr[i + 1] = len(cols)
        return rowptr, np.asarray(cols, np.int32)

    def h_min(self) -> float:
        """Smallest sqrt(area) over the cells."""
        return float(np.sqrt(self.areas().min()))

    def areas(self) -> np.ndarray:
        t = self.txy
        a = 0.5 * np.abs(t[:, 0] * t[:, 3] + t[:, 2] * t[:, 5] + t[:, 4] * t[:, 1]
                         - t[:, 0] * t[:, 5] - t[:, 4] * t[:, 3] - t[:, 2] * t[:, 1])
        return np.add.reduceat(a, self.tptr[:-1])

    def multicolour(self, order: str = "dsatur") -> "PolyMesh":
        """Return a copy renumbered colour-major.

        ``order="dsatur"`` (default): saturation-degree colouring (Brelaz), then a balancing
        pass that moves cells from the largest classes into smaller ones their neighbours
        leave free.  Every class is one level, hence one launch, of each DILU sweep and of
        the factor, so few classes of equal size keep every launch large (a 256^2 jittered
        Voronoi lattice: 5 classes of ~13.1k cells, against 6 from first-fit, two of them
        nearly empty).  ``order="greedy"``: first-fit in cell order.

        Cells keep their relative order inside a colour, so a lattice-ordered
        seed set keeps its locality inside each colour.
        """
        n = self.n_cells
        adj = self.adjacency()
        if order == "greedy":
            colour = _first_fit(adj)
        elif order == "dsatur":
            colour = _balance(adj, _dsatur(adj))
        else:
            raise ValueError(f"unknown colouring order {order!r}")
        perm = np.argsort(colour, kind="stable")  # new -> old
        m = self.permute(perm)
        m.colour = colour[perm].astype(np.int32)
        nc = int(colour.max()) + 1
        m.colour_ptr = np.searchsorted(m.colour, np.arange(nc + 1)).astype(np.int32)
        return m

    def permute(self, perm: np.ndarray) -> "PolyMesh":
        """Renumber cells: new cell ``k`` is old cell ``perm[k]``."""
        n = self.n_cells
        perm = np.asarray(perm, np.int64)
        inv = np.empty(n, np.int64)
        inv[perm] = np.arange(n)
        cnt = np.diff(self.pptr)[perm]
        pptr = np.zeros(n + 1, np.int32)
        pptr[1:] = np.cumsum(cnt)
        idx = np.concatenate([np.arange(self.pptr[o], self.pptr[o + 1]) for o in perm]) if n else np.zeros(0, np.int64)
        pv = self.pv[idx].copy()
        p2p = self.p2p[idx].copy()
        pos = p2p >= 0
        p2p[pos] = inv[p2p[pos]]
        bc = self.bc.copy()
        if len(bc):
            per = bc[:, 0] != 0
            bc[per, 1] = inv[bc[per, 1]]
            bc[per, 2] = inv[bc[per, 2]]
        tcnt = np.diff(self.tptr)[perm]
        tptr = np.zeros(n + 1, np.int32)
        tptr[1:] = np.cumsum(tcnt)
        tidx = np.concatenate([np.arange(self.tptr[o], self.tptr[o + 1]) for o in perm]) if n else np.zeros(0, np.int64)
        return PolyMesh(self.v.copy(), pptr, pv.astype(np.int32), p2p.astype(np.int32), bc.astype(np.int32),
                        tptr, self.txy[tidx].copy(), self.bb[perm].copy(), self.width, self.height)

    def check(self) -> None:
        """Structural checks: neighbour symmetry, boundary rows, positive areas."""
        n = self.n_cells
        assert np.all(self.areas() > 0)
        for i in range(n):
            nbs = [self.cell_neighbour(i, e) for e in range(self.pptr[i + 1] - self.pptr[i])]
            real = [j for j in nbs if j >= 0]
            assert len(real) == len(set(real)), f"cell {i} touches a neighbour twice"
            assert i not in real, f"cell {i} is its own neighbour"
            for j in real:
                assert i in [self.cell_neighbour(j, e) for e in range(self.pptr[j + 1] - self.pptr[j])]
        if self.colour is not None:
            for i in range(n):
                for e in range(self.pptr[i + 1] - self.pptr[i]):
                    j = self.cell_neighbour(i, e)
                    if j >= 0:
                        assert self.colour[i] != self.colour[j]
                        assert (i < j) == (self.colour[i] < self.colour[j])

    def arrays(self) -> dict:
        return dict(v=self.v, pptr=self.pptr, pv=self.pv, p2p=self.p2p, bc=self.bc,
                    tptr=self.tptr, txy=self.txy, bb=self.bb)

    def save(self, path: str) -> None:
        extra = {}
        if self.colour is not None:
            extra = dict(colour=self.colour, colour_ptr=self.colour_ptr)
        np.savez_compressed(path, width=self.width, height=self.height, **self.arrays(), **extra)

    @staticmethod
    def load(path: str) -> "PolyMesh":
        z = np.load(path)
        m = PolyMesh(z["v"], z["pptr"], z["pv"], z["p2p"], z["bc"], z["tptr"], z["txy"], z["bb"],
                     float(z["width"]), float(z["height"]))
        if "colour" in z:
            m.colour, m.colour_ptr = z["colour"], z["colour_ptr"]
        return m


def union_mesh(meshes: list[PolyMesh]) -> tuple[PolyMesh, np.ndarray]:
    """Disjoint union of meshes: one mesh whose connected components are the given meshes,
    cells numbered mesh by mesh.  Returns the union and ``cell_ptr`` (mesh ``q`` owns cells
    ``cell_ptr[q] .. cell_ptr[q+1]-1``).  Boundary ids of mesh ``q`` are shifted past those of
    the meshes before it and periodic rows are renumbered, so every cell keeps exactly the
    neighbours and boundary data it had.  Each part keeps its own cell order (e.g. colour-major),
    so the level schedule of a block DILU on the union is the union of the parts' schedules.
    The batched solver assembles and factors one such union per polynomial degree."""
    if not meshes:
        raise ValueError("empty mesh list")
    vs, pptrs, pvs, p2ps, bcs, tptrs, txys, bbs = [], [], [], [], [], [], [], []
    cell_ptr = np.zeros(len(meshes) + 1, np.int32)
    nv = nc = nb = nt = 0
    for q, m in enumerate(meshes):
        vs.append(m.v)
        pptrs.append(m.pptr[:-1] + len(np.concatenate(pvs)) if pvs else m.pptr[:-1])
        pvs.append(m.pv + nv)
        p2p = m.p2p.astype(np.int64).copy()
        p2p[p2p >= 0] += nc
        p2p[p2p < 0] -= nb
        p2ps.append(p2p)
        bc = m.bc.astype(np.int64).copy()
        if len(bc):
            per = bc[:, 0] != 0
            bc[per, 1:3] += nc
            bc[per, 3:7] += nv
        bcs.append(bc)
        tptrs.append(m.tptr[:-1] + nt)
        txys.append(m.txy)
        bbs.append(m.bb)
        nv += len(m.v)
        nc += m.n_cells
        nb += len(m.bc)
        nt += len(m.txy)
        cell_ptr[q + 1] = nc
    pv = np.concatenate(pvs).astype(np.int32)
    pptr = np.concatenate(pptrs + [np.asarray([len(pv)])]).astype(np.int32)
    tptr = np.concatenate(tptrs + [np.asarray([nt])]).astype(np.int32)
    u = PolyMesh(np.concatenate(vs), pptr, pv, np.concatenate(p2ps).astype(np.int32),
                 np.concatenate(bcs).astype(np.int32).reshape(-1, 7), tptr, np.concatenate(txys),
                 np.concatenate(bbs), max(m.width for m in meshes), max(m.height for m in meshes))
    return u, cell_ptr


# --------------------------------------------------------------------------------------
# helpers

def _bbox(v: np.ndarray, pptr: np.ndarray, pv: np.ndarray) -> np.ndarray:
    """``PolyMesh::computebb`` (``src/PolyMesh.cpp:70-104``): x0, y0, w, h."""
    n = len(pptr) - 1
    xs = v[pv, 0]
    ys = v[pv, 1]
    x0 = np.minimum.reduceat(xs, pptr[:-1])
    y0 = np.minimum.reduceat(ys, pptr[:-1])
    x1 = np.maximum.reduceat(xs, pptr[:-1])
    y1 = np.maximum.reduceat(ys, pptr[:-1])
    bb = np.stack([x0, y0, x1 - x0, y1 - y0], axis=1)
    assert bb.shape == (n, 4)
    return bb


def _fan(v: np.ndarray, pptr: np.ndarray, pv: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Fan sub-triangulation from each polygon's first vertex (convex cells)."""
    n = len(pptr) - 1
    nv = np.diff(pptr)
    tcnt = nv - 2
    tptr = np.zeros(n + 1, np.int32)
    tptr[1:] = np.cumsum(tcnt)
    txy = np.empty((int(tptr[-1]), 6))
    t = 0
    for i in range(n):
        poly = pv[pptr[i]:pptr[i + 1]]
        a = v[poly[0]]
        for k in range(1, len(poly) - 1):
            b = v[poly[k]]
            c = v[poly[k + 1]]
            txy[t] = (a[0], a[1], b[0], b[1], c[0], c[1])
            t += 1
    return tptr, txy


def _ccw(points: np.ndarray, centre: np.ndarray) -> np.ndarray:
    ang = np.arctan2(points[:, 1] - centre[1], points[:, 0] - centre[0])
    return np.argsort(ang, kind="stable")


# --------------------------------------------------------------------------------------
# seed generators

def jittered_lattice_seeds(nx: int, ny: int, width: float, height: float, jitter: float,
                           seed: int) -> np.ndarray:
    """Seeds ``((i+1/2)h + xi, (j+1/2)h + eta)``, ``xi, eta ~ U[-jitter*h, jitter*h]``,
    row-major (x fastest), from ``numpy.random.default_rng(seed)``."""
    rng = np.random.default_rng(seed)
    hx = width / nx
    hy = height / ny
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    x = (ii.ravel() + 0.5) * hx + rng.uniform(-jitter * hx, jitter * hx, nx * ny)
    y = (jj.ravel() + 0.5) * hy + rng.uniform(-jitter * hy, jitter * hy, nx * ny)
    return np.stack([x, y], axis=1)


def uniform_seeds(n: int, width: float, height: float, seed: int) -> np.ndarray:
    """``n`` seeds uniform in ``[0,width) x [0,height)``, sorted along a coarse
    row-major bucket order so that cell numbering keeps some locality."""
    rng = np.random.default_rng(seed)
    pts = np.stack([rng.uniform(0, width, n), rng.uniform(0, height, n)], axis=1)
    nb = max(1, int(math.sqrt(n)))
    key = np.floor(pts[:, 1] / height * nb) * nb + np.floor(pts[:, 0] / width * nb)
    return pts[np.argsort(key, kind="stable")]


# --------------------------------------------------------------------------------------
# Voronoi meshes

def _voronoi_cells(aug: np.ndarray, n: int):
    from scipy.spatial import Voronoi

    vor = Voronoi(aug)
    ridge = {}
    for (p, q), (va, vb) in zip(vor.ridge_points, vor.ridge_vertices):
        if va < 0 or vb < 0:
            continue
        key = (va, vb) if va < vb else (vb, va)
        ridge[key] = (int(p), int(q))
    cells = []
    for i in range(n):
        reg = vor.regions[vor.point_region[i]]
        if len(reg) == 0 or -1 in reg:
            raise RuntimeError(f"unbounded Voronoi region for seed {i}")
        reg = np.asarray(reg)
        order = _ccw(vor.vertices[reg], aug[i])
        reg = reg[order]
        edges = []
        for k in range(len(reg)):
            va, vb = int(reg[k]), int(reg[(k + 1) % len(reg)])
            key = (va, vb) if va < vb else (vb, va)
            p, q = ridge[key]
            other = q if p == i else p
            edges.append((va, vb, other))
        cells.append(edges)
    return vor, cells


def periodic_voronoi(seeds: np.ndarray, width: float, height: float, colour: bool = True) -> PolyMesh:
    """Periodic Voronoi mesh on ``[0,width) x [0,height)``.

    Edges across the periodic boundary get their own negative boundary id with a
    ``BoundaryInfo`` row ``(1, p1=i, p2=j, a1, b1, a2, b2)``: ``v[a2] - v[a1]`` is the
    translation from cell ``i``'s side of the edge into cell ``j``'s frame, as used by
    ``PolyMesh::getPeriodicCoordinates`` (``src/PolyMesh.cpp:307-325``).
    """
    seeds = np.asarray(seeds, np.float64)
    n = len(seeds)
    h = math.sqrt(width * height / n)
    margin = min(max(width, height), 5.0 * h)
    while True:
        imgs = [seeds]
        src = [np.arange(n)]
        shifts = [np.zeros((n, 2), np.int64)]
        for sx in (-1, 0, 1):
            for sy in (-1, 0, 1):
                if sx == 0 and sy == 0:
                    continue
                p = seeds + np.array([sx * width, sy * height])
                keep = ((p[:, 0] > -margin) & (p[:, 0] < width + margin)
                        & (p[:, 1] > -margin) & (p[:, 1] < height + margin))
                imgs.append(p[keep])
                src.append(np.nonzero(keep)[0])
                shifts.append(np.tile([sx, sy], (int(keep.sum()), 1)))
        aug = np.concatenate(imgs)
        srcs = np.concatenate(src)
        shs = np.concatenate(shifts)
        try:
            vor, cells = _voronoi_cells(aug, n)
        except RuntimeError:
            margin *= 2
            continue
        rmax = 0.0
        for i, edges in enumerate(cells):
            pts = vor.vertices[[e[0] for e in edges]]
            rmax = max(rmax, float(np.max(np.hypot(pts[:, 0] - seeds[i, 0], pts[:, 1] - seeds[i, 1]))))
        if 2.0 * rmax < margin or margin >= 2 * max(width, height):
            break
        margin = 2.5 * rmax
    # vertices used by central cells
    vmap = {}
    vlist = []
    for edges in cells:
        for va, _, _ in edges:
            if va not in vmap:
                vmap[va] = len(vlist)
                vlist.append(vor.vertices[va])
    v = np.asarray(vlist, np.float64)
    pptr = np.zeros(n + 1, np.int32)
    pv = []
    p2p = []
    bc = []
    pending = []  # (row index in bc, i, j, shift, va, vb)
    for i, edges in enumerate(cells):
        for va, vb, other in edges:
            pv.append(vmap[va])
            if other < n:
                p2p.append(other)
            else:
                j = int(srcs[other])
                s = shs[other]
                bc.append([1, i, j, vmap[va], vmap[vb], -1, -1])
                pending.append((len(bc) - 1, j, s))
                p2p.append(-len(bc))
        pptr[i + 1] = len(pv)
    pv = np.asarray(pv, np.int32)
    bc = np.asarray(bc, np.int32).reshape(-1, 7)
    tol = 1e-7 * h
    for row, j, s in pending:
        tvec = -np.array([s[0] * width, s[1] * height])
        poly = pv[pptr[j]:pptr[j + 1]]
        for col_src, col_dst in ((3, 5), (4, 6)):
            target = v[bc[row, col_src]] + tvec
            d = np.hypot(v[poly, 0] - target[0], v[poly, 1] - target[1])
            k = int(np.argmin(d))
            if d[k] > tol:
                raise RuntimeError("periodic vertex image not found")
            bc[row, col_dst] = poly[k]
    p2p = np.asarray(p2p, np.int32)
    tptr, txy = _fan(v, pptr, pv)
    m = PolyMesh(v, pptr, pv, p2p, bc, tptr, txy, _bbox(v, pptr, pv), width, height)
    return m.multicolour() if colour else m


def box_voronoi(seeds: np.ndarray, width: float, height: float, colour: bool = True) -> PolyMesh:
    """Voronoi mesh clipped to ``[0,width] x [0,height]``; every exterior edge gets the
    Dirichlet boundary id ``-1`` (``computep2p``, ``src/PolyMesh.cpp:12-67``)."""
    seeds = np.asarray(seeds, np.float64)
    n = len(seeds)
    mir = [seeds,
           np.stack([-seeds[:, 0], seeds[:, 1]], 1),
           np.stack([2 * width - seeds[:, 0], seeds[:, 1]], 1),
           np.stack([seeds[:, 0], -seeds[:, 1]], 1),
           np.stack([seeds[:, 0], 2 * height - seeds[:, 1]], 1)]
    aug = np.concatenate(mir)
    vor, cells = _voronoi_cells(aug, n)
    vmap = {}
    vlist = []
    pptr = np.zeros(n + 1, np.int32)
    pv = []
    p2p = []
    for i, edges in enumerate(cells):
        for va, vb, other in edges:
            if va not in vmap:
                vmap[va] = len(vlist)
                vlist.append(vor.vertices[va])
            pv.append(vmap[va])
            p2p.append(other if other < n else -1)
        pptr[i + 1] = len(pv)
    v = np.asarray(vlist, np.float64)
    pv = np.asarray(pv, np.int32)
    p2p = np.asarray(p2p, np.int32)
    bc = np.asarray([[0, 0, 0, 0, 0, 0, 0]], np.int32)
    tptr, txy = _fan(v, pptr, pv)
    m = PolyMesh(v, pptr, pv, p2p, bc, tptr, txy, _bbox(v, pptr, pv), width, height)
    return m.multicolour() if colour else m


# --------------------------------------------------------------------------------------
# structured meshes with the reference's ordering

def periodic_rectangle(nx: int, ny: int, width: float, height: float, colour: bool = False) -> PolyMesh:
    """``periodicRectangle(Nx, Ny, width, height)`` (``src/Meshes.cpp:176-321``).

    Same vertex numbering (first-seen order of ``addVertex``), polygon order
    (row-major), edge order (bottom, right, top, left) and boundary-id order
    (``-1, -2, ...`` as periodic edges are met).  Like the reference,
    ``hy = width/Ny``.  Sub-triangles: two per quad, fan from the first vertex.
    """
    hx = width / nx
    hy = width / ny
    vid = {}
    v = []

    def add(x, y):
        key = (round(x / hx), round(y / hy))
        if key not in vid:
            vid[key] = len(v)
            v.append((x, y))
        return vid[key]

    pv = []
    for iy in range(ny):
        for ix in range(nx):
            x = ix * hx
            y = iy * hy
            pv += [add(x, y), add(x + hx, y), add(x + hx, y + hy), add(x, y + hy)]

    def find(x, y):
        return vid[(round(x / hx), round(y / hy))]

    p2p = []
    bc = []
    i = 0
    for iy in range(ny):
        for ix in range(nx):
            if iy == 0:
                bc.append([1, i, nx * (ny - 1) + ix, find(ix * hx, 0), find((ix + 1) * hx, 0),
                           find(ix * hx, ny * hy), find((ix + 1) * hx, ny * hy)])
                p2p.append(-len(bc))
            else:
                p2p.append(i - nx)
            if ix == nx - 1:
                bc.append([1, i, i - nx + 1, find(nx * hx, iy * hy), find(nx * hx, (iy + 1) * hy),
                           find(0, iy * hy), find(0, (iy + 1) * hy)])
                p2p.append(-len(bc))
            else:
                p2p.append(i + 1)
            if iy == ny - 1:
                bc.append([1, i, ix, find(ix * hx, ny * hy), find((ix + 1) * hx, ny * hy),
                           find(ix * hx, 0), find((ix + 1) * hx, 0)])
                p2p.append(-len(bc))
            else:
                p2p.append(i + nx)
            if ix == 0:
                bc.append([1, i, i + nx - 1, find(0, iy * hy), find(0, (iy + 1) * hy),
                           find(nx * hx, iy * hy), find(nx * hx, (iy + 1) * hy)])
                p2p.append(-len(bc))
            else:
                p2p.append(i - 1)
            i += 1
    n = nx * ny
    pptr = np.arange(0, 4 * n + 1, 4, dtype=np.int32)
    v = np.asarray(v, np.float64)
    pv = np.asarray(pv, np.int32)
    tptr, txy = _fan(v, pptr, pv)
    m = PolyMesh(v, pptr, pv, np.asarray(p2p, np.int32), np.asarray(bc, np.int32), tptr, txy,
                 _bbox(v, pptr, pv), width, height)
    return m.multicolour() if colour else m


def quad_rectangle(nx: int, ny: int, width: float, height: float, colour: bool = False) -> PolyMesh:
    """Axis-aligned quads with Dirichlet boundary id ``-1`` on the exterior."""
    hx = width / nx
    hy = height / ny
    vid = {}
    v = []

    def add(ix, iy):
        if (ix, iy) not in vid:
            vid[(ix, iy)] = len(v)
            v.append((ix * hx, iy * hy))
        return vid[(ix, iy)]

    pv = []
    p2p = []
    for iy in range(ny):
        for ix in range(nx):
            i = iy * nx + ix
            pv += [add(ix, iy), add(ix + 1, iy), add(ix + 1, iy + 1), add(ix, iy + 1)]
            p2p += [i - nx if iy > 0 else -1, i + 1 if ix < nx - 1 else -1,
                    i + nx if iy < ny - 1 else -1, i - 1 if ix > 0 else -1]
    n = nx * ny
    pptr = np.arange(0, 4 * n + 1, 4, dtype=np.int32)
    v = np.asarray(v, np.float64)
    pv = np.asarray(pv, np.int32)
    tptr, txy = _fan(v, pptr, pv)
    m = PolyMesh(v, pptr, pv, np.asarray(p2p, np.int32), np.asarray([[0] * 7], np.int32), tptr, txy,
                 _bbox(v, pptr, pv), width, height)
    return m.multicolour() if colour else m
