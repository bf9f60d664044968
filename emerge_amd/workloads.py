"""Synthetic workloads of the named benchmark shapes (gmsh is not available).

* ``cube_assembly``: unit cube split into n^3 hexahedra x 6 Kuhn tets,
  interior vertices jittered by U[-0.15h, 0.15h] per axis (PCG64 ``seed``),
  per-tet eps_r = U[1,10] * (1 - j tan_d) with tan_d ~ U[0, 0.02], mu_r = 1.
  n = 60 gives 1,296,000 tets and 8,316,360 complex DOFs.
* ``waveguide``: air-filled WR-90 (22.86 x 10.16 mm) of length 40 mm as a
  Kuhn box (11 x 5 x 18 cells by default), ports on z = 0 and z = L.
* ``waveguide_family``: independent jittered WR-90 sections of uneven
  length (DOF counts drawn from a seed), solved side by side.
* ``cavity``: closed PEC box 60 x 40 x 30 mm with a centred dielectric block
  20 x 20 x 10 mm (eps_r = 9.8, tan_d = 1e-4) and a probe fed by a lumped
  port from the floor.
* ``patch_antenna``: substrate slab 60 x 60 x 1.52 mm (eps_r = 3.38,
  tan_d = 0.0027) on a PEC ground, 30 x 40 mm PEC patch, probe-style lumped
  port under the patch at an inset from its edge, air box a quarter
  wavelength (at 2 GHz) away truncated by a first-order absorbing boundary.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .basis import Nedelec2
from .mesh import Mesh3D, kuhn_box, kuhn_grid


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


def waveguide_family(count: int = 16, n_range=(20_000, 120_000), seed: int = 0,
                     jitter: float = 0.15) -> list[VolumeProblem]:
    """``count`` WR-90 sections on the 11 x 5 cross-section grid of
    ``waveguide`` with the same axial spacing (40/18 mm), their DOF counts
    drawn U[n_range] from PCG64 ``seed``, interior vertices jittered by
    U[-jitter, jitter] h (stream ``seed + 1 + i``)."""
    rng = np.random.default_rng(seed)
    targets = rng.uniform(n_range[0], n_range[1], count)
    dz = 40e-3 / 18
    out = []
    for i, tgt in enumerate(targets):
        nz = 2
        while _ndof_estimate(11, 5, nz) < tgt:
            nz += 1
        L = nz * dz
        nodes, tets = kuhn_box(11, 5, nz, size=(WR90_A, WR90_B, L), jitter=jitter, seed=seed + 1 + i)
        mesh = Mesh3D(nodes, tets)
        basis = Nedelec2(mesh)
        T = mesh.n_tets
        out.append(VolumeProblem(mesh, basis, np.ones(T, dtype=np.complex128), np.ones(T, dtype=np.complex128),
                                 meta={"workload": "wr90_waveguide", "cells": (11, 5, nz), "length": L,
                                       "a": WR90_A, "b": WR90_B, "target_dofs": float(tgt)}))
    return out


def _ndof_estimate(nx: int, ny: int, nz: int) -> int:
    """2 nE + 2 nF of an nx x ny x nz Kuhn box (closed form: a Kuhn cell
    adds 7 edges and 12 faces, boundary layers fewer)."""
    V = (nx + 1) * (ny + 1) * (nz + 1)
    # edges: axis edges + 3 face diagonals per cell face family + 1 body diagonal per cell
    nE = (nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
          + nx * ny * (nz + 1) + nx * (ny + 1) * nz + (nx + 1) * ny * nz + nx * ny * nz)
    T = 6 * nx * ny * nz
    # Euler characteristic of a ball: V - E + F - T = 1
    nF = 1 - V + nE + T
    return 2 * nE + 2 * nF


def graded_lines(a: float, b: float, h_a: float, h_b: float, ratio: float = 1.2) -> np.ndarray:
    """Coordinate lines from a to b whose spacing starts near h_a at a and
    grows (or shrinks) geometrically by at most ``ratio`` per cell towards
    h_b, capped at max(h_a, h_b); the last line is b exactly."""
    L = b - a
    hs = []
    h = h_a
    tot = 0.0
    cap = max(h_a, h_b)
    while tot < L - 1e-12:
        hs.append(h)
        tot += h
        h = min(cap, h * ratio) if h_b >= h_a else max(h_b, h / ratio)
    hs = np.array(hs) * (L / tot)
    return a + np.concatenate([[0.0], np.cumsum(hs)])


def _uniform(a: float, b: float, h: float) -> np.ndarray:
    n = max(1, int(round((b - a) / h)))
    return np.linspace(a, b, n + 1)


def _join(*segments) -> np.ndarray:
    out = [segments[0]]
    for seg in segments[1:]:
        if abs(seg[0] - out[-1][-1]) > 1e-12:
            raise ValueError("grid segments do not meet")
        out.append(seg[1:])
    return np.concatenate(out)


def cavity(cells=(48, 32, 24), size=(60e-3, 40e-3, 30e-3)) -> VolumeProblem:
    """Closed PEC cavity with a centred dielectric block and a probe-fed
    lumped port (one cell wide, from the floor to a PEC probe strip).

    Uniform Kuhn grid (48 x 32 x 24 cells at h = 1.25 mm by default:
    221,184 tets).  Block: centre +- (10, 10, 5) mm, eps_r = 9.8 (1 - 1e-4 j).
    Probe: the plane x = size_x / 6, one cell wide in y starting at the
    centre line; lumped port over the first two cell layers above the floor
    (field along +z, 50 ohm, 1 W), PEC strip above it up to a third of the
    height."""
    nx, ny, nz = cells
    X, Y, Z = size
    nodes, tets = kuhn_box(nx, ny, nz, size=size)
    mesh = Mesh3D(nodes, tets)
    basis = Nedelec2(mesh)
    c = mesh.nodes[:, mesh.tets].mean(axis=1)
    blk = (np.abs(c[0] - X / 2) < 10e-3) & (np.abs(c[1] - Y / 2) < 10e-3) & (np.abs(c[2] - Z / 2) < 5e-3)
    eps = np.where(blk, 9.8 * (1 - 1e-4j), 1.0 + 0j).astype(np.complex128)
    hx, hy, hz = X / nx, Y / ny, Z / nz
    xp = round(nx / 6) * hx
    y0 = round(ny / 2) * hy
    zp = 2 * hz
    zt = max(round(nz / 3), 3) * hz
    meta = {"workload": f"cavity_{nx}x{ny}x{nz}_kuhn", "cells": cells, "size": size,
            "port": dict(x=xp, y=(y0, y0 + hy), z=(0.0, zp)), "probe": dict(x=xp, y=(y0, y0 + hy), z=(zp, zt))}
    return VolumeProblem(mesh, basis, eps, np.ones(mesh.n_tets, dtype=np.complex128), meta=meta)


def cavity_bcs(problem: VolumeProblem) -> list:
    from . import bc
    m, pm, pr = problem.mesh, problem.meta["port"], problem.meta["probe"]
    port = m.tris_in_box((pm["x"], pm["y"][0], pm["z"][0]), (pm["x"], pm["y"][1], pm["z"][1]))
    probe = m.tris_in_box((pr["x"], pr["y"][0], pr["z"][0]), (pr["x"], pr["y"][1], pr["z"][1]))
    pec = np.union1d(m.boundary_tris(), probe)
    w, h = pm["y"][1] - pm["y"][0], pm["z"][1] - pm["z"][0]
    return [bc.PEC(pec), bc.LumpedPort(port, 1, w, h, (0, 0, 1))]


def patch_antenna(h_slab: float = 1.25e-3, n_sub: int = 4, h_air: float = 5e-3, ratio: float = 1.3,
                  air_gap: float = 37.5e-3, inset: float = 6.25e-3) -> VolumeProblem:
    """Microstrip-patch-like slab in an air box (graded Kuhn grid).

    Slab 60 x 60 x 1.52 mm centred on the z axis, ground plane z = 0 (under
    the slab only), patch 30 x 40 mm (x by y) on top, eps_r = 3.38
    (1 - 0.0027 j).  Lumped port: the plane y = 0, one cell wide in x from
    the patch edge x = -15 mm + ``inset``, from the ground to the patch (field
    along +z, 50 ohm, 1 W).  The air box extends ``air_gap`` beyond the slab
    on every side (a quarter wavelength at 2 GHz).  Spacing ``h_slab`` in x
    and y over the slab, ``n_sub`` layers across the substrate, growing by
    ``ratio`` per cell to ``h_air`` in the air.  The defaults give about one
    million tets."""
    t = 1.52e-3
    half = 30e-3
    xs_in = _uniform(-half, half, h_slab)
    xs = _join(-graded_lines(half, half + air_gap, h_slab, h_air, ratio)[::-1], xs_in,
               graded_lines(half, half + air_gap, h_slab, h_air, ratio))
    hz = t / n_sub
    zs = _join(-graded_lines(0.0, air_gap, hz, h_air, ratio)[::-1], np.linspace(0.0, t, n_sub + 1),
               graded_lines(t, t + air_gap, hz, h_air, ratio))
    for v in (-15e-3, 15e-3, -20e-3, 20e-3, -15e-3 + inset, 0.0):
        if np.min(np.abs(xs - v)) > 1e-9:
            raise ValueError(f"patch_antenna: no grid line at {v}")
    nodes, tets = kuhn_grid(xs, xs, zs)
    mesh = Mesh3D(nodes, tets)
    basis = Nedelec2(mesh)
    c = mesh.nodes[:, mesh.tets].mean(axis=1)
    sub = (np.abs(c[0]) < half) & (np.abs(c[1]) < half) & (c[2] > 0) & (c[2] < t)
    eps = np.where(sub, 3.38 * (1 - 0.0027j), 1.0 + 0j).astype(np.complex128)
    xf = -15e-3 + inset
    i = int(np.argmin(np.abs(xs - xf)))
    meta = {"workload": "patch_antenna_graded_kuhn", "t": t, "half": half, "air_gap": air_gap,
            "port": dict(x=(xs[i], xs[i + 1]), y=0.0, z=(0.0, t)),
            "grid": (xs.size - 1, xs.size - 1, zs.size - 1)}
    return VolumeProblem(mesh, basis, eps, np.ones(mesh.n_tets, dtype=np.complex128), meta=meta)


def patch_bcs(problem: VolumeProblem) -> list:
    """PEC ground (z = 0 under the slab) and patch, first-order absorbing
    boundary on the air box, lumped port (ground to patch)."""
    from . import bc
    m, meta = problem.mesh, problem.meta
    t, half, pm = meta["t"], meta["half"], meta["port"]
    ground = m.tris_in_box((-half, -half, 0.0), (half, half, 0.0))
    patch = m.tris_in_box((-15e-3, -20e-3, t), (15e-3, 20e-3, t))
    port = m.tris_in_box((pm["x"][0], pm["y"], pm["z"][0]), (pm["x"][1], pm["y"], pm["z"][1]))
    w, h = pm["x"][1] - pm["x"][0], pm["z"][1] - pm["z"][0]
    return [bc.PEC(np.union1d(ground, patch)), bc.AbsorbingBoundary(m.boundary_tris(), order=1),
            bc.LumpedPort(port, 1, w, h, (0, 0, 1))]
