"""Far field by the Stratton-Chu integral on the GPU.

Mirrors ``stratton_chu(Ein, Hin, mesh, theta, phi, k0) -> (Eout, Hout, wns)``
(MW/sc.py:135-170, ``MW/`` = src/emerge/_emerge/physics/microwave/) with the
sum (``stratton_chu_ff``, :41-132) in ``emg_farfield``.  The surface is
described as the reference's ``SurfaceMesh`` describes it
(E/mesh3d.py:690-716, 820-885): samples at the edge centres of the surface
triangles, each triangle adding n A / 3 to the weighted normal of its three
edges; unit normals point away from ``origin`` (``outward_normal``,
E/mth/optimized.py:252-274) and, without an origin, away from the tet each
triangle belongs to.  The caller samples E and H at ``edge_centers``.
"""
from __future__ import annotations

import numpy as np

from ._lib import Context, ptr


class FarFieldSurface:
    """Edge-centre samples of the triangles `tri_ids` of a Mesh3D."""

    def __init__(self, mesh, tri_ids: np.ndarray, origin=None):
        tri_ids = np.asarray(tri_ids)
        tris = mesh.tris[:, tri_ids]
        X = mesh.nodes[:, tris]                                     # (3, 3 vert, n)
        cr = np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], axis=0)
        length = np.sqrt(cr[0] ** 2 + cr[1] ** 2 + cr[2] ** 2)
        normals = cr / length
        self.areas = length / 2
        centres = X.mean(axis=1)
        o = (mesh.nodes[:, np.unique(tris)].mean(axis=1) if origin is None
             else np.asarray(origin, dtype=np.float64))
        normals = normals * np.where(((centres - o[:, None]) * normals).sum(axis=0) < 0, -1.0, 1.0)
        if origin is None:
            tets = np.max(mesh.tri_to_tet[:, tri_ids], axis=0)
            tc = mesh.nodes[:, mesh.tets[:, tets]].mean(axis=1)
            normals = normals * np.sign(((centres - tc) * normals).sum(axis=0))
        self.normals = normals
        # edges (x < y) of every triangle: pairs (0,1), (1,2), (0,2)
        pairs = np.concatenate([np.sort(tris[[0, 1]], axis=0), np.sort(tris[[1, 2]], axis=0),
                                np.sort(tris[[0, 2]], axis=0)], axis=1)
        self.edges, inv = np.unique(pairs, axis=1, return_inverse=True)
        self.tri_to_edge = inv.reshape(3, -1)
        w = normals * self.areas / 3
        wn = np.zeros((3, self.edges.shape[1]))
        for k in range(3):
            wn[k] = np.bincount(self.tri_to_edge[0], w[k], self.edges.shape[1]) \
                + np.bincount(self.tri_to_edge[1], w[k], self.edges.shape[1]) \
                + np.bincount(self.tri_to_edge[2], w[k], self.edges.shape[1])
        self.weighted_normals = wn
        self.edge_centers = (mesh.nodes[:, self.edges[0]] + mesh.nodes[:, self.edges[1]]) / 2

    @property
    def exyz(self):
        return self.edge_centers[0], self.edge_centers[1], self.edge_centers[2]


def stratton_chu(Ein, Hin, surface: FarFieldSurface, theta, phi, k0: float, ctx: Context | None = None):
    """Far-field E and H (complex (3, N)) in the directions (theta, phi)
    from E and H (complex (3, M)) at the surface's edge centres; returns
    (Eout, Hout, weighted normals) as the reference does."""
    E = np.ascontiguousarray(Ein, dtype=np.complex128)
    H = np.ascontiguousarray(Hin, dtype=np.complex128)
    M = surface.edge_centers.shape[1]
    if E.shape != (3, M) or H.shape != (3, M):
        raise ValueError("Ein and Hin must have shape (3, n_samples)")
    th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
    ph = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
    if th.shape != ph.shape or th.ndim != 1:
        raise ValueError("theta and phi must be 1-D arrays of one length")
    pos = np.ascontiguousarray(surface.edge_centers, dtype=np.float64)
    wn = np.ascontiguousarray(surface.weighted_normals, dtype=np.float64)
    ctx = ctx if ctx is not None else Context(0)
    N = th.shape[0]
    Eout = np.empty((3, N), dtype=np.complex128)
    Hout = np.empty((3, N), dtype=np.complex128)
    ctx.call("emg_farfield", M, ptr(E), ptr(H), ptr(pos), ptr(wn), N, ptr(th), ptr(ph), float(k0),
             ptr(Eout), ptr(Hout))
    return Eout, Hout, wn


class _Mirrored:
    """Samples and weighted normals of a surface mirrored in the planes
    x = 0, y = 0 and/or z = 0 (SurfaceMesh.flip, E/mesh3d.py:760-787)."""

    def __init__(self, surface, axes: str):
        self.edge_centers = surface.edge_centers.copy()
        self.weighted_normals = surface.weighted_normals.copy()
        for a in axes:
            k = "xyz".index(a)
            self.edge_centers[k] = -self.edge_centers[k]
            self.weighted_normals[k] = -self.weighted_normals[k]


def _signs(syms) -> tuple[np.ndarray, np.ndarray]:
    """Component signs of E and H for the symmetry planes `syms` ('Ex' = E
    symmetry in x = 0, 'Hy' = H symmetry in y = 0, ...), as _EHSign
    (MW/microwave_data.py:253-303): 'E' flips all of E, 'H' all of H, and
    the axis flips that component of both."""
    se = np.ones(3)
    sh = np.ones(3)
    for s in syms:
        f, c = s[0], s[1].lower()
        if f == "E":
            se = -se
        elif f == "H":
            sh = -sh
        k = "xyz".index(c)
        se[k] = -se[k]
        sh[k] = -sh[k]
    return se, sh


def farfield_symmetric(Ein, Hin, surface, theta, phi, k0: float, syms=None, ctx: Context | None = None):
    """``stratton_chu`` plus the mirror images of the surface for symmetry
    planes (MW/microwave_data.py:790-861): with planes s1..sn every non-empty
    combination of them adds the far field of the mirrored surface with the
    component signs of that combination."""
    Eff, Hff, _ = stratton_chu(Ein, Hin, surface, theta, phi, k0, ctx=ctx)
    syms = list(syms or [])
    if len(syms) > 3:
        raise ValueError("at most three symmetry planes")
    from itertools import combinations
    for r in range(1, len(syms) + 1):
        for combo in combinations(syms, r):
            se, sh = _signs(combo)
            surf = _Mirrored(surface, "".join(s[1].lower() for s in combo))
            E2, H2, _ = stratton_chu(Ein * se[:, None], Hin * sh[:, None], surf, theta, phi, k0, ctx=ctx)
            Eff = Eff + E2
            Hff = Hff + H2
    return Eff, Hff


def arc_on_plane(ref_dir, normal, angle_range_deg, num_points: int = 100):
    """(theta, phi) of unit directions along an arc in the plane with normal
    `normal`, angle 0 along `ref_dir` projected into the plane and 90
    degrees along normal x ref (arc_on_plane, MW/microwave_data.py:42-91)."""
    n = np.asarray(normal, dtype=np.float64)
    n = n / np.linalg.norm(n)
    d = np.asarray(ref_dir, dtype=np.float64)
    d = d - np.dot(d, n) * n
    if np.linalg.norm(d) < 1e-12:
        raise ValueError("Reference direction is parallel to the normal vector.")
    e1 = d / np.linalg.norm(d)
    e2 = np.cross(n, e1)
    a = np.deg2rad(np.linspace(angle_range_deg[0], angle_range_deg[1], num_points))
    u = np.outer(np.cos(a), e1) + np.outer(np.sin(a), e2)
    return np.arccos(u[:, 2]), np.arctan2(u[:, 1], u[:, 0])
