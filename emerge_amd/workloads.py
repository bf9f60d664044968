"""Synthetic workloads of the named benchmark shapes (gmsh is not available).

* ``cube_assembly``: unit cube split into n^3 hexahedra x 6 Kuhn tets,
  interior vertices jittered by U[-0.15h, 0.15h] per axis (PCG64 ``seed``),
  per-tet eps_r = U[1,10] * (1 - j tan_d) with tan_d ~ U[0, 0.02], mu_r = 1.
  n = 60 gives 1,296,000 tets and 8,316,360 complex DOFs.
* ``waveguide``: air-filled WR-90 (22.86 x 10.16 mm) of length 40 mm as a
  Kuhn box (11 x 5 x 18 cells by default), ports on z = 0 and z = L.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .basis import Nedelec2
from .mesh import Mesh3D, kuhn_box


@dataclass
class VolumeProblem:
    mesh: Mesh3D
    basis: Nedelec2
    eps_r: np.ndarray      # (T,) complex
    inv_mu_r: np.ndarray   # (T,) complex
    meta: dict = field(default_factory=dict)


def cube_assembly(n: int = 60, seed: int = 0, jitter: float = 0.15) -> VolumeProblem:
    nodes, tets = kuhn_box(n, n, n, jitter=jitter, seed=seed)
    mesh = Mesh3D(nodes, tets)
    basis = Nedelec2(mesh)
    rng = np.random.default_rng(seed + 1)
    T = mesh.n_tets
    er = rng.uniform(1.0, 10.0, T)
    tand = rng.uniform(0.0, 0.02, T)
    eps = er * (1 - 1j * tand)
    inv_mu = np.ones(T, dtype=np.complex128)
    return VolumeProblem(mesh, basis, eps.astype(np.complex128), inv_mu,
                         meta={"workload": f"cube_{n}^3_kuhn_jitter{jitter}", "n": n, "seed": seed})


WR90_A = 22.86e-3
WR90_B = 10.16e-3


def waveguide(cells=(11, 5, 18), length: float = 40e-3, jitter: float = 0.0, seed: int = 0) -> VolumeProblem:
    nodes, tets = kuhn_box(*cells, size=(WR90_A, WR90_B, length), jitter=jitter, seed=seed)
    mesh = Mesh3D(nodes, tets)
    basis = Nedelec2(mesh)
    T = mesh.n_tets
    return VolumeProblem(mesh, basis, np.ones(T, dtype=np.complex128), np.ones(T, dtype=np.complex128),
                         meta={"workload": "wr90_waveguide", "cells": cells, "length": length,
                               "a": WR90_A, "b": WR90_B})


def waveguide_bcs(problem: VolumeProblem) -> list:
    """PEC walls and TE10 RectangularWaveguide ports on z = 0 and z = L
    (local frame origin at the face centre, x along the broad side)."""
    from . import bc
    m = problem.mesh
    L = problem.meta["length"]
    a, b = problem.meta["a"], problem.meta["b"]
    t1 = m.plane_tris(2, 0.0)
    t2 = m.plane_tris(2, L)
    pec = np.setdiff1d(m.boundary_tris(), np.concatenate([t1, t2]))
    cs1 = bc.CoordinateSystem([1, 0, 0], [0, 1, 0], [0, 0, 1], [a / 2, b / 2, 0.0])
    cs2 = bc.CoordinateSystem([1, 0, 0], [0, 1, 0], [0, 0, 1], [a / 2, b / 2, L])
    return [bc.PEC(pec), bc.RectangularWaveguide(t1, 1, cs1, (a, b)),
            bc.RectangularWaveguide(t2, 2, cs2, (a, b))]
