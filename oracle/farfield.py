"""Oracle: Stratton-Chu far field from surface fields, restated.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates MW/sc.py:41-170 (``MW/`` = src/emerge/_emerge/physics/microwave/)
and the surface-mesh geometry it reads (E/mesh3d.py:690-716, 820-885;
``outward_normal`` E/mth/optimized.py:252-274):

* samples at the edge centres of the surface triangles; every triangle adds
  n A / 3 to the weighted normal of each of its three edges;
* samples with |E| <= 1e-3 max |E| are left out (``LR``);
* Eout(r) = Q sum_j r x [(n x E)_j - Z0 r x (n x H)_j] exp(j k0 r . x_j),
  Q = -j k0 / (4 pi), r = (sin t cos p, sin t sin p, cos t);
  Hout = r x Eout / Z0.

Pinned in tests/test_farfield.py by the analytic far field of a Hertzian
dipole (exact near fields sampled on a closed cube) and by the extinction
theorem (a plane wave through a closed surface radiates nothing outside).
"""
from __future__ import annotations

import numpy as np

Z0 = 376.73031366857   # E/const.py:3
LR = 0.001             # MW/sc.py:25


def surface_geometry(nodes: np.ndarray, tris: np.ndarray, origin=None, tri_tets=None, tet_centers=None):
    """(edge_centers (3, nE), weighted normals (3, nE), edges (2, nE)) of the
    triangles `tris` (3, n): unit normals from (p2 - p1) x (p3 - p1), signed
    to point away from `origin`; with origin None, the face's mean vertex is
    the origin and each normal is then flipped to point away from the centre
    of its tet (`tri_tets`, `tet_centers`), as E/mesh3d.py:696-715."""
    P = nodes[:, tris]                                   # (3, 3 vert, n)
    c = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], axis=0)
    nrm = np.sqrt((c ** 2).sum(axis=0))
    n = c / nrm
    area = nrm / 2
    tc = P.mean(axis=1)
    if origin is None:
        used = np.unique(tris)
        o = nodes[:, used].mean(axis=1)
    else:
        o = np.asarray(origin, dtype=np.float64)
    n = n * np.where(((tc - o[:, None]) * n).sum(axis=0) < 0, -1.0, 1.0)
    if origin is None:
        align = tc - tet_centers[:, tri_tets]
        n = n * np.sign((align * n).sum(axis=0))
    pairs = np.concatenate([np.sort(tris[[0, 1]], axis=0), np.sort(tris[[1, 2]], axis=0),
                            np.sort(tris[[0, 2]], axis=0)], axis=1)
    edges, inv = np.unique(pairs, axis=1, return_inverse=True)
    inv = inv.reshape(-1)
    wn = np.zeros((3, edges.shape[1]))
    w = np.tile(n * area / 3, (1, 3))
    for k in range(3):
        np.add.at(wn[k], inv, w[k])
    centers = (nodes[:, edges[0]] + nodes[:, edges[1]]) / 2
    return centers, wn, edges


def stratton_chu_ff(Ein: np.ndarray, Hin: np.ndarray, pos: np.ndarray, wn: np.ndarray, theta: np.ndarray,
                    phi: np.ndarray, k0: float, chunk: int = 256):
    """Far-field E and H (3, N) complex for directions (theta, phi)."""
    Emag = np.sqrt(np.abs(Ein[0]) ** 2 + np.abs(Ein[1]) ** 2 + np.abs(Ein[2]) ** 2)
    keep = Emag > Emag.max() * LR
    E, H, x, n = Ein[:, keep], Hin[:, keep], pos[:, keep], wn[:, keep]
    nxE = np.cross(n, E, axis=0)
    nxH = np.cross(n, H, axis=0)
    r = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])   # (3, N)
    acc = np.zeros((3, theta.shape[0]), dtype=np.complex128)
    for s in range(0, x.shape[1], chunk):
        G = np.exp(1j * k0 * (r.T @ x[:, s:s + chunk]))                     # (N, m)
        a = nxE[:, s:s + chunk]
        b = nxH[:, s:s + chunk]
        # r x (n x H) per (direction, sample)
        rxb = np.stack([r[1][:, None] * b[2][None] - r[2][:, None] * b[1][None],
                        r[2][:, None] * b[0][None] - r[0][:, None] * b[2][None],
                        r[0][:, None] * b[1][None] - r[1][:, None] * b[0][None]])  # (3, N, m)
        ie = (a[:, None, :] - Z0 * rxb) * G[None]
        acc += ie.sum(axis=2)
    Q = -1j * k0 / (4 * np.pi)
    Eout = Q * np.cross(r, acc, axis=0)
    Hout = np.cross(r, Eout, axis=0) / Z0
    return Eout, Hout
