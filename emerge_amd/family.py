"""Independent meshes of uneven size solved together as one block-diagonal
system, so that every Krylov kernel is one launch over all of them.

The reference solves the jobs of a sweep one by one or in parallel worker
processes/threads (MW/microwave_3d.py:773-861); here the meshes of a
family are placed side by side (disjoint vertex sets, shifted along x so
that their positions do not overlap) and assembled as one mesh.  Its system
is the direct sum of the meshes' systems, its CSR pattern the union of
theirs, and one batched sweep covers all of them.  Port p of every mesh is
excited in the same system ("group" p), so a system's right-hand side is
the sum of the meshes' own right-hand sides and its solution restricted to
mesh m is mesh m's solution for its port p.

By default (``segmented=True``) the reduced rows are grouped by mesh and the
COCG solve keeps one set of Krylov scalars per mesh (emg_solver_set_segments):
every mesh follows the recurrence it would follow alone, with its own stop
test on ||b_m - A_m x_m|| / ||b_m||, while each kernel launch still covers
all meshes.  With ``segmented=False`` the scalars are global over the union,
and the union system is solved to tol * min_m ||b_m|| / ||b||, which bounds
every mesh's own relative residual by tol.
"""
from __future__ import annotations

import numpy as np

from . import bc as bcs
from .basis import Nedelec2
from .mesh import Mesh3D
from .sweep import FrequencySweep


def _shifted_bc(b, dx: float, tri_off: int, port_offset: int):
    """Copy of boundary condition `b` of one mesh on the union: triangle ids
    offset, port frames shifted by dx along x, port numbers offset."""
    t = np.asarray(b.tri_ids, dtype=np.int64) + tri_off
    if isinstance(b, bcs.PEC):
        return bcs.PEC(t)
    if isinstance(b, bcs.AbsorbingBoundary):
        return bcs.AbsorbingBoundary(t, order=b.order, abctype=b.abctype)
    if isinstance(b, bcs.RectangularWaveguide):
        cs = bcs.CoordinateSystem(b.cs.xax, b.cs.yax, b.cs.zax, b.cs.origin + np.array([dx, 0.0, 0.0]))
        out = bcs.RectangularWaveguide(t, b.port_number + port_offset, cs, b.dims, mode=b.mode, er=b.er,
                                       power=b.power)
        out._polarization = b._polarization
        return out
    if isinstance(b, bcs.LumpedPort):
        return bcs.LumpedPort(t, b.port_number + port_offset, b.width, b.height, b.direction, power=b.power,
                              Z0_port=b.Z0)
    raise TypeError(f"unsupported boundary condition {type(b).__name__}")


def dof_segments(mesh: Mesh3D, basis: Nedelec2, vert_off) -> np.ndarray:
    """Mesh index of every field DOF of a union mesh whose meshes own the
    vertex ranges [vert_off[m], vert_off[m+1]): the mesh of the DOF's edge or
    triangle (all its vertices lie in one mesh)."""
    voff = np.asarray(vert_off, dtype=np.int64)
    seg = np.full(basis.n_field, -1, dtype=np.int64)
    e_mesh = np.searchsorted(voff, mesh.edges[0], side="right") - 1
    t_mesh = np.searchsorted(voff, mesh.tris[0], side="right") - 1
    if np.any(e_mesh != np.searchsorted(voff, mesh.edges[1], side="right") - 1):
        raise ValueError("an edge joins two meshes")
    seg[basis.edge_to_field] = e_mesh[None, :]
    seg[basis.tri_to_field] = t_mesh[None, :]
    if np.any(seg < 0):
        raise ValueError("a field DOF without a mesh")
    return seg


class FamilySweep:
    """A family of independent meshes (each with isotropic scalar materials
    and its own boundary conditions, the same number of ports each) solved as
    one block-diagonal device system."""

    def __init__(self, problems: list, boundary_conditions: list, device: int = 0, gap: float | None = None,
                 segmented: bool = True):
        if len(problems) != len(boundary_conditions) or not problems:
            raise ValueError("one list of boundary conditions per problem")
        nports = {sum(1 for b in bl if isinstance(b, bcs.PortBC)) for bl in boundary_conditions}
        if len(nports) != 1:
            raise ValueError("every mesh of a family needs the same number of ports")
        self.n_ports = nports.pop()
        self.n_meshes = len(problems)
        nodes, tets, eps, inu, all_bcs = [], [], [], [], []
        x0 = 0.0
        voff = toff = 0
        self.vert_off, self.tri_off, self.tet_off, self.x_shift = [], [], [], []
        for p, bl in zip(problems, boundary_conditions):
            m = p.mesh
            lo, hi = m.nodes.min(axis=1), m.nodes.max(axis=1)
            g = gap if gap is not None else 0.25 * float(hi[0] - lo[0])
            dx = x0 - lo[0]
            shifted = m.nodes.copy()
            shifted[0] += dx
            nodes.append(shifted)
            tets.append(m.tets + voff)
            eps.append(np.asarray(p.eps_r, dtype=np.complex128))
            inu.append(np.asarray(p.inv_mu_r, dtype=np.complex128))
            self.vert_off.append(voff)
            self.tri_off.append(toff)
            self.tet_off.append(sum(q.shape[1] for q in tets[:-1]))
            self.x_shift.append(dx)
            pn = len(self.vert_off) - 1
            all_bcs += [_shifted_bc(b, dx, toff, pn * self.n_ports) for b in bl]
            voff += m.n_nodes
            toff += m.n_tris
            x0 = hi[0] + dx + g
        self.problems = problems
        self.mesh = Mesh3D(np.concatenate(nodes, axis=1), np.concatenate(tets, axis=1))
        # the union's triangles are the meshes' triangles in mesh order
        # (vertex ids of mesh m all precede those of mesh m + 1)
        for p, o, vo in zip(problems, self.tri_off, self.vert_off):
            nt = p.mesh.n_tris
            if not np.array_equal(self.mesh.tris[:, o:o + nt], p.mesh.tris + vo):
                raise AssertionError("union triangle order differs from the meshes' order")
        self.basis = Nedelec2(self.mesh)
        self.segmented = bool(segmented)
        seg = None
        if self.segmented:
            seg = dof_segments(self.mesh, self.basis, self.vert_off)
        self.sweep = FrequencySweep(self.mesh, self.basis, np.concatenate(eps), np.concatenate(inu), all_bcs,
                                    device=device, dof_segments=seg)
        P = self.n_ports
        # group g = port g of every mesh; sweep.ports are sorted by port number
        self.groups = [[m * P + g for m in range(self.n_meshes)] for g in range(P)]
        self.blocks = [list(range(m * P, (m + 1) * P)) for m in range(self.n_meshes)]
        ports = self.sweep.ports
        for grp in self.groups:
            for k in (50.0, 250.0):
                f0 = ports[grp[0]].uinc_factor(k)
                if any(abs(ports[j].uinc_factor(k) - f0) > 1e-12 * abs(f0) for j in grp):
                    raise ValueError("the ports excited together must share their excitation factor")
        F = self.sweep.forcing()
        nrm = np.linalg.norm(F, axis=1)
        # min_m ||b_m|| / ||b|| per group (the factor uinc_factor(k0) is
        # common to a group, so the forcing shapes' norms give the ratio)
        self.tol_scale = 1.0 if self.segmented else min(
            float(nrm[grp].min() / np.sqrt(np.sum(nrm[grp] ** 2))) for grp in self.groups)

    @property
    def ctx(self):
        return self.sweep.ctx

    def solve(self, freqs, tol: float = 1e-8, **kw):
        """Every (frequency, port group) system of the union.  Returns
        (S, relres_bound, result): S[m] is mesh m's S-matrices
        (nfreq, P, P), relres_bound (nfreq, P) bounds every mesh's own
        relative residual, result the union's SweepResult."""
        res = self.sweep.solve(freqs, tol=tol * self.tol_scale, groups=self.groups, blocks=self.blocks, **kw)
        P = self.n_ports
        S = [res.S[:, m * P:(m + 1) * P, m * P:(m + 1) * P] for m in range(self.n_meshes)]
        return S, res.relres / self.tol_scale, res

    def mesh_field(self, full: np.ndarray, m: int) -> np.ndarray:
        """Mesh m's own DOF vector from a union solution (its DOF numbering:
        edges, faces, edges, faces of that mesh)."""
        p = self.problems[m]
        f = p.basis
        out = np.zeros(f.n_field, dtype=np.complex128)
        ub = self.basis
        eo = int(self.mesh.edge_ids(p.mesh.edges[0, :1] + self.vert_off[m], p.mesh.edges[1, :1] + self.vert_off[m])[0])
        to = self.tri_off[m]
        ne, nf = f.n_edges, f.n_tris
        e = np.arange(ne)
        t = np.arange(nf)
        out[e] = full[ub.edge_to_field[0, eo + e]]
        out[ne + t] = full[ub.tri_to_field[3, to + t]]
        out[ne + nf + e] = full[ub.edge_to_field[1, eo + e]]
        out[2 * ne + nf + t] = full[ub.tri_to_field[7, to + t]]
        return out
