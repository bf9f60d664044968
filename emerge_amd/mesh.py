"""Tetrahedral mesh topology for the FEM hot path.

`Mesh3D` holds the arrays the reference's ``Mesh3D._pre_update`` builds
(src/emerge/_emerge/mesh3d.py:386-498) and that fix the global DOF numbering,
hence the bit-exact sparsity pattern:

* ``edges (2, nE)`` / ``tris (3, nF)``: unique vertex tuples taken from each
  tet's *sorted* vertices, stored in lexicographic order (mesh3d.py:386-407).
* ``tet_to_edge (6, T)``: local vertex pairs (1,2),(1,3),(1,4),(2,3),(4,2),(3,4)
  in the tet's own vertex order (mesh3d.py:427).
* ``tet_to_tri (4, T)``: local triples (1,2,3),(1,3,4),(1,4,2),(2,3,4)
  (mesh3d.py:431).
* ``tri_to_edge (3, nF)``: (i1,i2),(i2,i3),(i1,i3) of the sorted triangle
  (mesh3d.py:475-484).
* ``tri_to_tet (2, nF)``: adjacent tets in increasing tet order, ``-1234`` for a
  missing second tet (mesh3d.py:449-453, ``_MISSING_ID`` at :26).
* ``edge_lengths (nE,)``: ``sqrt(sum((n[:,e0]-n[:,e1])**2))`` (mesh3d.py:497).

The reference derives these with Python sets and dicts per tet; here the same
orderings come from sort/unique/searchsorted over packed integer keys, which is
O(T log T) and vectorised.  gmsh is not available, so meshes come from
`kuhn_box` (hexahedra split into 6 positively oriented Kuhn tets, optional
seeded interior jitter).
"""
from __future__ import annotations

import numpy as np

MISSING_ID = -1234

# tet-local vertex pairs / triples (0-based) in the reference's order
TET_EDGE_LOCAL = np.array([[0, 0, 0, 1, 3, 2], [1, 2, 3, 2, 1, 3]], dtype=np.int64)
TET_TRI_LOCAL = np.array([[0, 0, 0, 1], [1, 2, 3, 2], [2, 3, 1, 3]], dtype=np.int64)


def _pair_key(a: np.ndarray, b: np.ndarray, nv: int) -> np.ndarray:
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    return lo * nv + hi


def _triple_key(a: np.ndarray, b: np.ndarray, c: np.ndarray, nv: int) -> np.ndarray:
    s = np.sort(np.stack([a, b, c]), axis=0)
    return (s[0] * nv + s[1]) * nv + s[2]


class Mesh3D:
    """Nodes, tets and the derived edge/triangle topology of a tet mesh."""

    def __init__(self, nodes: np.ndarray, tets: np.ndarray):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.int64)
        if nodes.ndim != 2 or nodes.shape[0] != 3:
            raise ValueError("nodes must have shape (3, V)")
        if tets.ndim != 2 or tets.shape[0] != 4:
            raise ValueError("tets must have shape (4, T)")
        nv = nodes.shape[1]
        if nv >= 2_000_000:
            raise ValueError("vertex count too large for packed 64-bit triangle keys")
        self.nodes = nodes
        self.tets = tets
        self._build(nv)

    # ------------------------------------------------------------------ build
    def _build(self, nv: int) -> None:
        tets = self.tets
        st = np.sort(tets, axis=0)
        pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
        ekeys = np.concatenate([st[a] * nv + st[b] for a, b in pairs])
        ekeys = np.unique(ekeys)
        self._edge_keys = ekeys
        self.edges = np.stack([ekeys // nv, ekeys % nv])

        triples = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
        tkeys = np.concatenate([(st[a] * nv + st[b]) * nv + st[c] for a, b, c in triples])
        tkeys = np.unique(tkeys)
        self._tri_keys = tkeys
        self.tris = np.stack([tkeys // (nv * nv), (tkeys // nv) % nv, tkeys % nv])

        t2e = np.empty((6, tets.shape[1]), dtype=np.int64)
        for k in range(6):
            key = _pair_key(tets[TET_EDGE_LOCAL[0, k]], tets[TET_EDGE_LOCAL[1, k]], nv)
            t2e[k] = np.searchsorted(ekeys, key)
        self.tet_to_edge = t2e

        t2t = np.empty((4, tets.shape[1]), dtype=np.int64)
        for k in range(4):
            key = _triple_key(tets[TET_TRI_LOCAL[0, k]], tets[TET_TRI_LOCAL[1, k]],
                              tets[TET_TRI_LOCAL[2, k]], nv)
            t2t[k] = np.searchsorted(tkeys, key)
        self.tet_to_tri = t2t

        i1, i2, i3 = self.tris
        self.tri_to_edge = np.stack([
            np.searchsorted(ekeys, i1 * nv + i2),
            np.searchsorted(ekeys, i2 * nv + i3),
            np.searchsorted(ekeys, i1 * nv + i3),
        ])

        ntri = self.tris.shape[1]
        tri_ids = t2t.T.reshape(-1)  # tet-major, local face minor
        tet_ids = np.repeat(np.arange(tets.shape[1]), 4)
        order = np.argsort(tri_ids, kind="stable")
        tri_sorted = tri_ids[order]
        tet_sorted = tet_ids[order]
        first = np.ones(tri_sorted.shape[0], dtype=bool)
        first[1:] = tri_sorted[1:] != tri_sorted[:-1]
        tri_to_tet = np.full((2, ntri), MISSING_ID, dtype=np.int64)
        tri_to_tet[0, tri_sorted[first]] = tet_sorted[first]
        tri_to_tet[1, tri_sorted[~first]] = tet_sorted[~first]
        counts = np.bincount(tri_ids, minlength=ntri)
        if counts.max(initial=0) > 2:
            raise ValueError("non-manifold mesh: a triangle is shared by more than two tets")
        self.tri_to_tet = tri_to_tet

        d = self.nodes[:, self.edges[0]] - self.nodes[:, self.edges[1]]
        self.edge_lengths = np.sqrt(np.sum(d ** 2, axis=0))

    @classmethod
    def on_device(cls, nodes: np.ndarray, tets: np.ndarray, ctx) -> "Mesh3D":
        """The same topology built on the GPU (``emg_topology_build``: sort-
        unique keys and binary searches), bit-exact with ``_build``."""
        import ctypes

        from ._lib import ptr
        obj = cls.__new__(cls)
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.int64)
        if nodes.ndim != 2 or nodes.shape[0] != 3 or tets.ndim != 2 or tets.shape[0] != 4:
            raise ValueError("nodes must have shape (3, V) and tets (4, T)")
        nv, T = nodes.shape[1], tets.shape[1]
        obj.nodes, obj.tets = nodes, tets
        t32 = np.ascontiguousarray(tets, dtype=np.int32)
        nE, nF = ctypes.c_int64(), ctypes.c_int64()
        ctx.call("emg_topology_build", nv, ptr(nodes), T, ptr(t32), ctypes.byref(nE), ctypes.byref(nF))
        nE, nF = nE.value, nF.value
        edges = np.empty((2, nE), dtype=np.int32)
        tris = np.empty((3, nF), dtype=np.int32)
        t2e = np.empty((6, T), dtype=np.int32)
        t2t = np.empty((4, T), dtype=np.int32)
        r2e = np.empty((3, nF), dtype=np.int32)
        r2t = np.empty((2, nF), dtype=np.int32)
        lengths = np.empty(nE, dtype=np.float64)
        ctx.call("emg_topology_get", ptr(edges), ptr(tris), ptr(t2e), ptr(t2t), ptr(r2e), ptr(r2t), ptr(lengths))
        obj.edges = edges.astype(np.int64)
        obj.tris = tris.astype(np.int64)
        obj.tet_to_edge = t2e.astype(np.int64)
        obj.tet_to_tri = t2t.astype(np.int64)
        obj.tri_to_edge = r2e.astype(np.int64)
        obj.tri_to_tet = r2t.astype(np.int64)
        obj.edge_lengths = lengths
        obj._edge_keys = obj.edges[0] * nv + obj.edges[1]
        obj._tri_keys = (obj.tris[0] * nv + obj.tris[1]) * nv + obj.tris[2]
        return obj

    # ------------------------------------------------------------- properties
    @property
    def n_nodes(self) -> int:
        return self.nodes.shape[1]

    @property
    def n_tets(self) -> int:
        return self.tets.shape[1]

    @property
    def n_edges(self) -> int:
        return self.edges.shape[1]

    @property
    def n_tris(self) -> int:
        return self.tris.shape[1]

    # --------------------------------------------------------------- lookups
    def get_edge(self, i1: int, i2: int) -> int:
        nv = self.n_nodes
        key = min(i1, i2) * nv + max(i1, i2)
        idx = int(np.searchsorted(self._edge_keys, key))
        if idx >= self._edge_keys.shape[0] or self._edge_keys[idx] != key:
            raise ValueError(f"There is no edge with indices {i1}, {i2}")
        return idx

    def edge_ids(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        """Vectorised ``get_edge``: edge ids of the vertex pairs (a, b)."""
        key = _pair_key(np.asarray(a, dtype=np.int64), np.asarray(b, dtype=np.int64), self.n_nodes)
        idx = np.searchsorted(self._edge_keys, key)
        if np.any(idx >= self._edge_keys.shape[0]) or np.any(self._edge_keys[np.minimum(idx, self._edge_keys.shape[0] - 1)] != key):
            raise ValueError("edge_ids: a vertex pair is not an edge of the mesh")
        return idx

    def get_tri(self, i1: int, i2: int, i3: int) -> int:
        nv = self.n_nodes
        a, b, c = sorted((int(i1), int(i2), int(i3)))
        key = (a * nv + b) * nv + c
        idx = int(np.searchsorted(self._tri_keys, key))
        if idx >= self._tri_keys.shape[0] or self._tri_keys[idx] != key:
            raise ValueError(f"There is no triangle with indices {a}, {b}, {c}")
        return idx

    def boundary_tris(self) -> np.ndarray:
        """Triangles with a single adjacent tet (the exterior surface)."""
        return np.nonzero(self.tri_to_tet[1] == MISSING_ID)[0]

    def plane_tris(self, axis: int, value: float, tol: float = 1e-9) -> np.ndarray:
        """Boundary triangles lying on the plane ``x[axis] == value``."""
        bt = self.boundary_tris()
        coords = self.nodes[axis][self.tris[:, bt]]
        on = np.all(np.abs(coords - value) <= tol, axis=0)
        return bt[on]

    def tris_in_box(self, lo, hi, tol: float = 1e-9) -> np.ndarray:
        """All triangles (interior ones included) whose three vertices lie in
        the closed axis-aligned box [lo, hi]; a degenerate box (lo == hi on
        one axis) selects a planar patch, e.g. an internal PEC sheet."""
        lo = np.asarray(lo, dtype=np.float64)[:, None, None]
        hi = np.asarray(hi, dtype=np.float64)[:, None, None]
        c = self.nodes[:, self.tris]  # (3, 3, nF)
        inside = np.all((c >= lo - tol) & (c <= hi + tol), axis=(0, 1))
        return np.nonzero(inside)[0]

    def tri_nodes(self, tri_ids: np.ndarray) -> np.ndarray:
        return np.unique(self.tris[:, tri_ids].reshape(-1))

    def get_face_tets(self, tri_ids: np.ndarray) -> np.ndarray:
        """Tets sharing a vertex with the given triangles (mesh3d.py:267-272)."""
        nodes = self.tri_nodes(tri_ids)
        mask = np.zeros(self.n_nodes, dtype=bool)
        mask[nodes] = True
        return np.nonzero(mask[self.tets].any(axis=0))[0]


def kuhn_grid(xs: np.ndarray, ys: np.ndarray, zs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Tensor-product grid with coordinate lines ``xs``, ``ys``, ``zs``
    (strictly increasing), each cell split into 6 Kuhn tets.

    Vertex ``(i, j, k)`` has id ``i + nx1*(j + ny1*k)`` (``nx1 = len(xs)``);
    the 6 tets of a cell are consecutive, cells ordered x fastest.  Every tet
    is positively oriented.  Returns ``(nodes (3,V), tets (4,T))``.
    """
    xs, ys, zs = (np.asarray(a, dtype=np.float64) for a in (xs, ys, zs))
    for a in (xs, ys, zs):
        if a.ndim != 1 or a.size < 2 or np.any(np.diff(a) <= 0):
            raise ValueError("kuhn_grid: coordinate lines must be strictly increasing")
    nx, ny, nz = xs.size - 1, ys.size - 1, zs.size - 1
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij")
    ii = ii.transpose(2, 1, 0).reshape(-1)
    jj = jj.transpose(2, 1, 0).reshape(-1)
    kk = kk.transpose(2, 1, 0).reshape(-1)
    nodes = np.stack([xs[ii], ys[jj], zs[kk]])

    def vid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.transpose(2, 1, 0).reshape(-1)
    cj = cj.transpose(2, 1, 0).reshape(-1)
    ck = ck.transpose(2, 1, 0).reshape(-1)
    unit = np.eye(3, dtype=np.int64)
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    parity = [1, -1, -1, 1, 1, -1]
    tets = []
    for perm, sgn in zip(perms, parity):
        p0 = np.zeros(3, dtype=np.int64)
        p1 = p0 + unit[perm[0]]
        p2 = p1 + unit[perm[1]]
        p3 = np.ones(3, dtype=np.int64)
        v = [vid(ci + p[0], cj + p[1], ck + p[2]) for p in (p0, p1, p2, p3)]
        if sgn < 0:
            v[1], v[2] = v[2], v[1]
        tets.append(np.stack(v))
    # cell-major, Kuhn-permutation minor: the 6 tets of one cell are consecutive
    tets = np.stack(tets, axis=2).reshape(4, -1)
    return nodes, tets


def kuhn_box(nx: int, ny: int, nz: int,
             size: tuple[float, float, float] = (1.0, 1.0, 1.0),
             origin: tuple[float, float, float] = (0.0, 0.0, 0.0),
             jitter: float = 0.0, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Box of nx*ny*nz hexahedra, each split into 6 Kuhn tets.

    Vertex ``(i, j, k)`` has id ``i + (nx+1)*(j + (ny+1)*k)``.  Interior vertices
    move by ``U[-jitter, jitter] * h`` per axis (PCG64 stream ``seed``); a Kuhn
    tet's smallest altitude is h/sqrt(2), so ``jitter < 0.2`` cannot invert one.
    Every tet is positively oriented.  Returns ``(nodes (3,V), tets (4,T))``.
    """
    hx, hy, hz = size[0] / nx, size[1] / ny, size[2] / nz
    nodes, tets = kuhn_grid(origin[0] + np.arange(nx + 1) * hx, origin[1] + np.arange(ny + 1) * hy,
                            origin[2] + np.arange(nz + 1) * hz)
    if jitter > 0.0:
        ii = np.tile(np.arange(nx + 1), (ny + 1) * (nz + 1))
        jj = np.tile(np.repeat(np.arange(ny + 1), nx + 1), nz + 1)
        kk = np.repeat(np.arange(nz + 1), (nx + 1) * (ny + 1))
        rng = np.random.default_rng(seed)
        shift = rng.uniform(-jitter, jitter, size=nodes.shape)
        shift *= np.array([hx, hy, hz])[:, None]
        interior = (ii > 0) & (ii < nx) & (jj > 0) & (jj < ny) & (kk > 0) & (kk < nz)
        nodes[:, interior] += shift[:, interior]
    return nodes, tets
