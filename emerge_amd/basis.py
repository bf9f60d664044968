"""Second-order Nedelec DOF numbering.

Mirrors ``Nedelec2.__init__`` (src/emerge/_emerge/elements/nedelec2.py:28-66):
global DOFs are laid out in four blocks ``[edge 1st | face 1st | edge 2nd |
face 2nd]`` so that

* ``tet_to_field (20, T) = [tet_to_edge; tet_to_tri + nE;
  tet_to_edge + nE + nF; tet_to_tri + 2nE + nF]`` (nedelec2.py:42-46),
* ``tri_to_field (8, nF) = [tri_to_edge; tri + nE; tri_to_edge + nE + nF;
  tri + 2nE + nF]`` (nedelec2.py:53-58),
* ``edge_to_field (2, nE) = [edge; edge + nE + nF]`` (nedelec2.py:48-51).
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh3D


class Nedelec2:
    n_tet_dofs = 20
    n_tri_dofs = 8

    def __init__(self, mesh: Mesh3D):
        self.mesh = mesh
        ne, nf = mesh.n_edges, mesh.n_tris
        self.n_edges = ne
        self.n_tris = nf
        self.n_tets = mesh.n_tets
        self.n_field = 2 * ne + 2 * nf

        t2f = np.empty((20, mesh.n_tets), dtype=np.int64)
        t2f[:6] = mesh.tet_to_edge
        t2f[6:10] = mesh.tet_to_tri + ne
        t2f[10:16] = mesh.tet_to_edge + (nf + ne)
        t2f[16:20] = mesh.tet_to_tri + (nf + 2 * ne)
        self.tet_to_field = t2f

        e = np.arange(ne)
        self.edge_to_field = np.stack([e, e + nf + ne])

        tr = np.arange(nf)
        r2f = np.empty((8, nf), dtype=np.int64)
        r2f[:3] = mesh.tri_to_edge
        r2f[3] = tr + ne
        r2f[4:7] = mesh.tri_to_edge + ne + nf
        r2f[7] = tr + 2 * ne + nf
        self.tri_to_field = r2f

    @classmethod
    def on_device(cls, mesh: Mesh3D, ctx) -> "Nedelec2":
        """The same DOF maps built on the GPU (``emg_dof_maps``) from the
        topology that ``Mesh3D.on_device(nodes, tets, ctx)`` left in `ctx`."""
        from ._lib import ptr
        obj = cls.__new__(cls)
        obj.mesh = mesh
        ne, nf, T = mesh.n_edges, mesh.n_tris, mesh.n_tets
        obj.n_edges, obj.n_tris, obj.n_tets = ne, nf, T
        obj.n_field = 2 * ne + 2 * nf
        t2f = np.empty((20, T), dtype=np.int32)
        r2f = np.empty((8, nf), dtype=np.int32)
        e2f = np.empty((2, ne), dtype=np.int32)
        ctx.call("emg_dof_maps", ptr(t2f), ptr(r2f), ptr(e2f))
        obj.tet_to_field = t2f.astype(np.int64)
        obj.tri_to_field = r2f.astype(np.int64)
        obj.edge_to_field = e2f.astype(np.int64)
        return obj

    def tri_rowcol(self, tri_ids: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Row/col index of each 8x8 entry of the given triangles, row-major
        per triangle (femdata.py:105-121)."""
        idx = self.tri_to_field[:, tri_ids]  # (8, n)
        rows = np.repeat(idx.T, 8, axis=1).reshape(-1)
        cols = np.tile(idx.T, (1, 8)).reshape(-1)
        return rows, cols


def pec_keep_ids(ctx, pec_tri_ids: np.ndarray) -> np.ndarray:
    """Ascending DOFs kept by the PEC elimination (ASM/assembler.py:274-307,
    406-407) on the topology in `ctx` (``emg_pec_keep``): every DOF except
    both DOFs of each edge of the PEC triangles and their 8 face DOFs."""
    import ctypes

    from ._lib import ptr
    tri = np.ascontiguousarray(pec_tri_ids, dtype=np.int32)
    n = ctypes.c_int64()
    ctx.call("emg_pec_keep", tri.shape[0], ptr(tri) if tri.size else None, ctypes.byref(n), None)
    out = np.empty(n.value, dtype=np.int32)
    ctx.call("emg_pec_keep", tri.shape[0], ptr(tri) if tri.size else None, ctypes.byref(n), ptr(out))
    return out.astype(np.int64)
