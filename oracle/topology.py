"""Oracle: mesh topology and Nedelec-2 DOF maps, restated with sets and dicts.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Pinned by: exact integer rules (sorted tuples, fixed local orders) evaluated by
brute force, plus the closed-form Kuhn counts checked in tests.  Pure-Python
loops: use on small meshes only.
"""
from __future__ import annotations

from collections import defaultdict

import numpy as np

MISSING_ID = -1234


def pre_update(nodes: np.ndarray, tets: np.ndarray) -> dict:
    """Restates ``Mesh3D._pre_update`` (E/mesh3d.py:386-498) without gmsh.

    edges/tris: unique sorted tuples from each tet's sorted vertices, sorted
    lexicographically (:386-407); tet_to_edge pairs (1,2),(1,3),(1,4),(2,3),
    (4,2),(3,4) (:427); tet_to_tri triples (1,2,3),(1,3,4),(1,4,2),(2,3,4)
    (:431); tri_to_tet in tet order (:449-453); tri_to_edge (i1i2, i2i3, i1i3)
    (:475-484); edge lengths (:497).
    """
    T = tets.shape[1]
    edgeset, triset = set(), set()
    for it in range(T):
        i1, i2, i3, i4 = sorted(int(x) for x in tets[:, it])
        edgeset.update([(i1, i2), (i1, i3), (i1, i4), (i2, i3), (i2, i4), (i3, i4)])
        triset.update([(i1, i2, i3), (i1, i2, i4), (i1, i3, i4), (i2, i3, i4)])
    edges = np.array(sorted(edgeset)).T
    tris = np.array(sorted(triset)).T
    inv_edges = {(int(edges[0, i]), int(edges[1, i])): i for i in range(edges.shape[1])}
    inv_tris = {tuple(int(v) for v in tris[:, i]): i for i in range(tris.shape[1])}

    def get_edge(a, b):
        return inv_edges[(min(int(a), int(b)), max(int(a), int(b)))]

    def get_tri(a, b, c):
        return inv_tris[tuple(sorted((int(a), int(b), int(c))))]

    tet_to_edge = np.zeros((6, T), dtype=np.int64)
    tet_to_tri = np.zeros((4, T), dtype=np.int64)
    tri_to_tet_l = defaultdict(list)
    for it in range(T):
        tv = tets[:, it]
        tet_to_edge[:, it] = [get_edge(tv[i - 1], tv[j - 1])
                              for i, j in zip([1, 1, 1, 2, 4, 3], [2, 3, 4, 3, 2, 4])]
        tet_to_tri[:, it] = [get_tri(tv[i - 1], tv[j - 1], tv[k - 1])
                             for i, j, k in zip([1, 1, 1, 2], [2, 3, 4, 3], [3, 4, 2, 4])]
        for k in range(4):
            tri_to_tet_l[int(tet_to_tri[k, it])].append(it)
    ntri = tris.shape[1]
    tri_to_tet = np.zeros((2, ntri), dtype=np.int64) + MISSING_ID
    for itri in range(ntri):
        lst = tri_to_tet_l[itri]
        tri_to_tet[:len(lst), itri] = lst
    tri_to_edge = np.zeros((3, ntri), dtype=np.int64)
    for itri in range(ntri):
        i1, i2, i3 = tris[:, itri]
        tri_to_edge[:, itri] = [get_edge(i1, i2), get_edge(i2, i3), get_edge(i1, i3)]
    edge_lengths = np.sqrt(np.sum((nodes[:, edges[0, :]] - nodes[:, edges[1, :]]) ** 2, axis=0))
    return dict(edges=edges, tris=tris, tet_to_edge=tet_to_edge, tet_to_tri=tet_to_tri,
                tri_to_tet=tri_to_tet, tri_to_edge=tri_to_edge, edge_lengths=edge_lengths)


def nedelec2_maps(topo: dict, T: int) -> dict:
    """Restates ``Nedelec2.__init__`` DOF maps (E/elements/nedelec2.py:42-58)."""
    ne = topo["edges"].shape[1]
    nf = topo["tris"].shape[1]
    t2f = np.zeros((20, T), dtype=np.int64)
    t2f[:6, :] = topo["tet_to_edge"]
    t2f[6:10, :] = topo["tet_to_tri"] + ne
    t2f[10:16, :] = topo["tet_to_edge"] + (nf + ne)
    t2f[16:20, :] = topo["tet_to_tri"] + (nf + 2 * ne)
    e2f = np.zeros((2, ne), dtype=np.int64)
    e2f[0, :] = np.arange(ne)
    e2f[1, :] = np.arange(ne) + nf + ne
    r2f = np.zeros((8, nf), dtype=np.int64)
    r2f[:3, :] = topo["tri_to_edge"]
    r2f[3, :] = np.arange(nf) + ne
    r2f[4:7, :] = topo["tri_to_edge"] + ne + nf
    r2f[7, :] = np.arange(nf) + 2 * ne + nf
    return dict(tet_to_field=t2f, edge_to_field=e2f, tri_to_field=r2f, n_field=2 * ne + 2 * nf)


def empty_tri_rowcol(tri_to_field: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Restates ``FEMBasis.empty_tri_rowcol`` (E/elements/femdata.py:105-121)."""
    N, nf = 8, tri_to_field.shape[1]
    N2 = N * N
    rows = np.empty(nf * N2, dtype=np.int64)
    cols = np.empty(nf * N2, dtype=np.int64)
    for itri in range(nf):
        p = itri * N2
        idx = tri_to_field[:, itri]
        for ii in range(N):
            rows[p + N * ii:p + N * (ii + 1)] = idx[ii]
            cols[p + ii:p + N2:N] = idx[ii]
    return rows, cols
