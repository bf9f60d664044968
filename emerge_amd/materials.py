"""Material tensors of the sweep that the reference derives per tet.

* ``lossy_permittivity``: loss tangent and conductivity folded into eps_r
  (ASM/assembler.py:218, 240); a frequency-dependent material.
* ``pml_layer_tensors``: the coordinate-dependent perfectly matched layer of
  ``pmlbox`` (E/geo/pmlbox.py:25-121): per axis with a layer direction
  d != 0 a complex stretch s = 1 - j (depth / thickness)^exponent deltamax,
  depth measured from the inner face of the layer; eps and mu become
  diag(sy sz / sx, sz sx / sy, sx sy / sz) times the background values.
  The reference evaluates such materials at the tet centres
  (MW/microwave_3d.py:908-928, centres from E/mesh3d.py:366); the tensors do
  not depend on the frequency.
"""
from __future__ import annotations

import numpy as np

from .sweep import lossy_permittivity  # noqa: F401  (re-exported)


def tet_centers(mesh) -> np.ndarray:
    """(3, T) mean of the four vertices of every tet."""
    n = mesh.nodes
    t = mesh.tets
    return (n[:, t[0]] + n[:, t[1]] + n[:, t[2]] + n[:, t[3]]) / 4


def pml_stretch(xyz: np.ndarray, center, dims, direction, thickness: float, exponent: float = 1.5,
                deltamax: float = 8.0) -> np.ndarray:
    """(3, n) complex stretch factors (sx, sy, sz) at points `xyz` (3, n) of
    the layer that borders the box of centre `center` and size `dims` on the
    side `direction` (each component -1, 0 or +1; several non-zero
    components make an edge or corner layer).  Axes with a zero component
    have s = 1."""
    xyz = np.asarray(xyz, dtype=np.float64)
    s = np.ones((3, xyz.shape[1]), dtype=np.complex128)
    for a in range(3):
        d = direction[a]
        if d == 0:
            continue
        depth = d * (xyz[a] - center[a] - d * dims[a] / 2) / thickness
        if np.any(depth < -1e-12) or np.any(depth > 1 + 1e-12):
            raise ValueError("points outside the PML layer along axis %d" % a)
        depth = np.clip(depth, 0.0, 1.0)
        s[a] = 1 - 1j * depth ** exponent * deltamax
    return s


def pml_layer_tensors(mesh, tet_ids: np.ndarray, center, dims, direction, thickness: float,
                      exponent: float = 1.5, deltamax: float = 8.0, er: complex = 1.0, ur: complex = 1.0):
    """(3, 3, n) complex eps_r and mu_r of the layer's tets `tet_ids`,
    evaluated at their centres."""
    c = tet_centers(mesh)[:, np.asarray(tet_ids)]
    sx, sy, sz = pml_stretch(c, center, dims, direction, thickness, exponent, deltamax)
    lam = np.stack([sy * sz / sx, sz * sx / sy, sx * sy / sz])
    n = c.shape[1]
    eps = np.zeros((3, 3, n), dtype=np.complex128)
    mu = np.zeros((3, 3, n), dtype=np.complex128)
    for k in range(3):
        eps[k, k] = er * lam[k]
        mu[k, k] = ur * lam[k]
    return eps, mu
