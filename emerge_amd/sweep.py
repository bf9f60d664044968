"""Frequency sweep on the GPU: PEC elimination, Robin/port terms, batched solve.

Host side of the path that replaces, per frequency, ``assemble_freq_matrix``
(ASM/assembler.py:189-429) and ``SolveRoutine.solve`` (E/solver.py:1241-1326)
inside ``Microwave3D.run_sweep`` (MW/microwave_3d.py:773-794).  Because E, B
and every surface matrix share one pattern and only scalars depend on k0
(K = E - k0^2 B + sum_bc gamma_bc(k0) S_bc, ASM/assembler.py:255, 349-350),
the volume matrices are assembled once and all frequency points (times all
excited ports, MW/simjob.py:82-108) are solved in batches on the device.

Host work here is integer bookkeeping done once per mesh (which DOFs a PEC
face removes, which triangle entries land in which reduced row); all
floating-point work (element and surface integrals, their sums, the solves)
runs in libemerge_hip.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import bc as bcs
from . import gradient
from ._lib import Context, ptr
from .assembler import MeshOnDevice

C0 = 299792458
EPS0 = 8.854187818814e-12


PECLIM = 1e7   # settings.mw_3d_peclim (E/settings.py:24): bulk conductors above it are PEC


def conductor_split(cond, peclim: float = PECLIM):
    """Per-tet classification of a conductivity (T,) or (3, 3, T) as the
    reference makes it (ASM/assembler.py:240-288, on cond[0, 0]): tets above
    `peclim` are perfect conductors (all their DOFs eliminated), the others
    with cond > 0 make the permittivity frequency-dependent.  Returns
    (pec_tets bool (T,), lossy conductivity with the PEC tets zeroed)."""
    c = np.asarray(cond)
    c00 = (c if c.ndim == 1 else c[0, 0]).real
    pec = c00 > peclim
    lossy = np.array(c, dtype=np.complex128, copy=True)
    if lossy.ndim == 1:
        lossy[pec] = 0
    else:
        lossy[:, :, pec] = 0
    return pec, lossy


def lossy_permittivity(eps_r, tand, cond, freq: float) -> np.ndarray:
    """Complex relative permittivity with loss tangent and conductivity (S/m)
    at `freq`: ``eps_r (1 - j tand) - j cond / (w eps0)``, w = 2 pi freq
    (ASM/assembler.py:218, 240).  Arrays broadcast (scalars (T,) or tensors
    (3, 3, T))."""
    w0 = 2 * np.pi * freq
    return np.asarray(eps_r) * (1 - 1j * np.asarray(tand)) - 1j * np.asarray(cond) / (w0 * EPS0)


# gaus_quad_tri(4) (E/mth/optimized.py:31-32, 87-118): weights and
# barycentric coordinates in generation order
_DUN_W = np.array([0.223381589678011] * 3 + [0.109951743655322] * 3)
_DUN_A = (0.108103018168070, 0.445948490915965)
_DUN_B = (0.816847572980459, 0.091576213509771)
_DUN_L = np.array([[_DUN_A[0], _DUN_A[1], _DUN_A[1]], [_DUN_A[1], _DUN_A[1], _DUN_A[0]],
                   [_DUN_A[1], _DUN_A[0], _DUN_A[1]],
                   [_DUN_B[0], _DUN_B[1], _DUN_B[1]], [_DUN_B[1], _DUN_B[1], _DUN_B[0]],
                   [_DUN_B[1], _DUN_B[0], _DUN_B[1]]])


def dunavant_points(nodes: np.ndarray, tris: np.ndarray, tri_ids: np.ndarray) -> np.ndarray:
    """Quadrature points (n_tri, 6, 3) of gaus_quad_tri(4) on the triangles."""
    v = nodes[:, tris[:, tri_ids]]  # (3 coords, 3 verts, n)
    return np.einsum("qk,ckn->nqc", _DUN_L, v)


@dataclass
class SweepResult:
    freqs: np.ndarray
    ports: list
    fields: dict            # (freq index, port number) -> full solution vector (if kept)
    relres: np.ndarray      # (nfreq, nports)
    iters: np.ndarray
    precision: str
    S: np.ndarray | None = None   # (nfreq, nports, nports), rows = receiving port
    timing: dict = field(default_factory=dict)
    fp64_fallback: np.ndarray | None = None  # (nfreq, nports) solved again in FP64 from zero
    fp64_escalated: int = 0   # mixed systems finished with complex128 inner solves (warm)


def apply_s_caps(Sp: np.ndarray, cap_single: bool = True, cap_col: bool = True, recip: bool = False):
    """Post-corrections of MW/microwave_3d.py:1223-1235 (defaults of
    E/settings.py:25-27): reciprocity average, column power cap (divides by
    the column power), single-entry cap."""
    Sp = Sp.copy()
    n = Sp.shape[0]
    if recip:
        Sp = (Sp + Sp.T) / 2
    if cap_col:
        for j in range(n):
            Sp[:, j] = Sp[:, j] / max(1.0, np.sum(np.abs(Sp[:, j]) ** 2))
    if cap_single:
        for i in range(n):
            for j in range(n):
                Sp[i, j] = Sp[i, j] / max(1.0, np.abs(Sp[i, j]))
    return Sp


def _morton_order(pos: np.ndarray, ids: np.ndarray, h: float) -> np.ndarray:
    """Permutation sorting points (3, n) along a Morton curve of cells of
    size h (ties by id)."""
    lo = pos.min(axis=1, keepdims=True)
    q = np.clip(np.floor((pos - lo) / max(h, 1e-300)), 0, 1023).astype(np.int64)
    key = np.zeros(pos.shape[1], dtype=np.int64)
    for b in range(10):
        for d in range(3):
            key |= ((q[d] >> b) & 1) << (3 * b + d)
    return np.lexsort((ids, key))


class FrequencySweep:
    """Device-resident E, B, surface terms and forcing of one mesh."""

    def __init__(self, mesh, basis, eps_r, inv_mu_r, boundary_conditions: list,
                 device: int = 0, ctx: Context | None = None, dof_order: str = "spatial", mu_r=None,
                 cond=None, peclim: float = PECLIM, dof_segments=None):
        """Materials per tet: eps_r a complex scalar (T,) or tensor (3, 3, T);
        inv_mu_r complex scalars 1/mu_r (T,), or mu_r a tensor (3, 3, T) (the
        reference's ``ur``, inverted on the device) in its place.  Any of them
        may be a callable of the frequency in Hz returning such an array:
        the sweep is then frequency-dependent and every frequency is
        re-assembled before its systems are solved, as the reference does
        for frequency-dependent or conductive materials
        (ASM/assembler.py:222-252; conductivity as ``eps_r - 1j*cond/(w eps0)``,
        ``lossy_permittivity``).  cond: per-tet conductivity in S/m, (T,) or
        (3, 3, T), as the reference's ``cond`` array: tets above `peclim` are
        perfect conductors (their DOFs are eliminated with the PEC faces, and
        their permittivity plays no part), the other conducting tets add
        -j cond / (w eps0) to eps_r at every frequency (the sweep is then
        frequency-dependent; ASM/assembler.py:240-288).  dof_segments: a
        segment id per field DOF (N,) for a mesh made of independent parts
        (no matrix entry couples two segments): the reduced rows are grouped
        by segment and the COCG solves run one Krylov recurrence per segment
        (``emg_solver_set_segments``), each stopped on its own residual."""
        self.conductor_tets = None
        if cond is not None:
            pec_t, lossy = conductor_split(cond, peclim)
            self.conductor_tets = np.nonzero(pec_t)[0]
            if pec_t.any():
                boundary_conditions = list(boundary_conditions) + [
                    bcs.PEC(np.unique(mesh.tet_to_tri[:, pec_t].reshape(-1)))]
            if np.any(lossy != 0):
                base = eps_r

                def eps_r(freq, base=base, sig=lossy):
                    e = np.asarray(base(freq) if callable(base) else base, dtype=np.complex128)
                    if e.ndim != sig.ndim:
                        eye = np.eye(3)[:, :, None]
                        if e.ndim == 1:
                            e = e[None, None, :] * eye
                        else:
                            sig = sig[None, None, :] * eye
                    return e - 1j * sig / (2 * np.pi * freq * EPS0)
        if dof_order not in ("spatial", "natural"):
            raise ValueError("dof_order must be 'spatial' or 'natural'")
        if (inv_mu_r is None) == (mu_r is None):
            raise ValueError("give exactly one of inv_mu_r (scalars) and mu_r (tensors)")
        self.dof_order = dof_order
        self.dof_segments = None if dof_segments is None else np.asarray(dof_segments, dtype=np.int64)
        if self.dof_segments is not None and self.dof_segments.shape != (basis.n_field,):
            raise ValueError("dof_segments needs one segment id per field DOF")
        self.mesh = mesh
        self.basis = basis
        self.ctx = ctx if ctx is not None else Context(device)
        self.dm = MeshOnDevice(self.ctx, mesh.nodes, mesh.tets, basis.tet_to_field, basis.n_field)
        self._materials = (eps_r, inv_mu_r, mu_r)
        self.frequency_dependent = any(callable(v) for v in self._materials)
        self._mat_freq = None     # frequency whose materials are assembled (frequency-dependent sweeps)
        if self.frequency_dependent:
            # vacuum placeholder for the reduction below; every solve assembles
            # the materials of its own frequencies first
            one = np.ones(mesh.n_tets, dtype=np.complex128)
            self._assemble_materials(one, one, None)
        else:
            self._assemble_materials(eps_r, inv_mu_r, mu_r)
        self._aux = None          # host arrays of the auxiliary space, built on first use
        self.vector_aux = True    # vector-nodal functions in the auxiliary space (set_vector_aux)
        self.progress = None      # optional callable (systems done, systems) after each block
        self._dvol = None         # its volume diagonal for the assembled materials
        self._aux_on = False
        self._escalated = 0       # systems finished with complex128 inner solves in the last solve
        self.bcs = list(boundary_conditions)
        self.pec = [b for b in self.bcs if isinstance(b, bcs.PEC)]
        self.robin = [b for b in self.bcs if isinstance(b, bcs.RobinBC)]
        self.ports = sorted([b for b in self.robin if b.is_port], key=lambda p: p.port_number)
        self._setup()

    # -------------------------------------------------------------- materials
    def _assemble_materials(self, eps_r, inv_mu_r, mu_r) -> None:
        """Upload one material set and assemble E and B: the isotropic kernel
        for scalar eps_r and inv_mu_r, the tensor kernel otherwise (scalars
        become multiples of the identity, 1/inv_mu_r for mu_r)."""
        T = self.mesh.n_tets
        e = np.asarray(eps_r, dtype=np.complex128)
        if e.shape not in ((T,), (3, 3, T)):
            raise ValueError("eps_r must have shape (n_tets,) or (3, 3, n_tets)")
        eye = np.eye(3)[:, :, None]
        if mu_r is None and e.ndim == 1:
            self.dm.assemble(e, np.asarray(inv_mu_r, dtype=np.complex128))
            self._imu = np.asarray(inv_mu_r, dtype=np.complex128)
        else:
            et = e if e.ndim == 3 else e[None, None, :] * eye
            if mu_r is None:
                inu = np.asarray(inv_mu_r, dtype=np.complex128)
                if inu.shape != (T,):
                    raise ValueError("inv_mu_r must have shape (n_tets,); give mu_r for tensors")
                mt = (1.0 / inu)[None, None, :] * eye
            else:
                mt = np.asarray(mu_r, dtype=np.complex128)
                mt = mt if mt.ndim == 3 else mt[None, None, :] * eye
            self.ctx.material_set_tensor(et, mt)
            self.ctx.assemble()
            self._imu = np.moveaxis(np.linalg.inv(np.moveaxis(mt, 2, 0)), 0, 2)
        self.eps_r = e
        self._dvol = None

    def materials_at(self, freq: float):
        """(eps_r, inv_mu_r, mu_r) at `freq` (callables evaluated)."""
        return tuple(v(freq) if callable(v) else v for v in self._materials)

    def _load_frequency(self, freq: float) -> None:
        """Assemble the materials of `freq` into the device system (frequency-
        dependent sweeps): new E and B values on the same pattern, gathered
        into the reduced system; surface terms, forcing and ports stay."""
        if self._mat_freq == freq:
            return
        self._assemble_materials(*self.materials_at(freq))
        self.ctx.call("emg_system_refresh")
        self._mat_freq = freq
        if self._aux_on:          # the gradient-space diagonal follows eps
            self._aux_on = False
            self._set_aux(True)

    def interpolate(self, field: np.ndarray, xs, ys, zs, freq: float, mu_r=None):
        """E and H of a full-length solution (``SweepResult.fields``) at the
        points (xs, ys, zs), as MWField.interpolate (MW/microwave_data.py:
        402-430): E interpolated in the last tet containing each point, H =
        curl E / (-j w mu0 mu_r) with the owner tet's scalar mu_r (default:
        the sweep's isotropic 1 / inv_mu_r; pass per-tet scalars for tensor
        or frequency-dependent media).  NaN outside the mesh.  Returns
        (E, H), complex (3, P)."""
        pts = np.ascontiguousarray(np.vstack([np.ravel(xs), np.ravel(ys), np.ravel(zs)]), dtype=np.float64)
        P = pts.shape[1]
        x = np.ascontiguousarray(field, dtype=np.complex128)
        if x.shape != (self.basis.n_field,):
            raise ValueError("field must hold every DOF (n_field,)")
        E = np.empty((3, P), dtype=np.complex128)
        C = np.empty((3, P), dtype=np.complex128)
        own = np.empty(P, dtype=np.int32)
        self.ctx.call("emg_field_points", ptr(x), P, ptr(pts), ptr(E), ptr(C), ptr(own))
        if mu_r is None:
            inu = self.materials_at(freq)[1]
            if inu is None or np.ndim(inu) != 1:
                raise ValueError("pass mu_r (per-tet scalars) for tensor permeability")
            mu_r = 1.0 / np.asarray(inu, dtype=np.complex128)
        mu_r = np.asarray(mu_r, dtype=np.complex128)
        mu0 = 1 / (C0 * C0 * EPS0)
        ok = own >= 0
        H = np.full((3, P), np.nan + 1j * np.nan)
        H[:, ok] = C[:, ok] / (-1j * 2 * np.pi * freq * mu0 * mu_r[own[ok]])[None, :]
        return E, H

    def farfield(self, field: np.ndarray, freq: float, tri_ids, theta, phi, origin=None, syms=None, mu_r=None):
        """Far field of a solution over the boundary triangles `tri_ids`
        (MWField.farfield, MW/microwave_data.py:790-861): E and H sampled at
        the surface's edge centres on the device, the Stratton-Chu sum on the
        device, mirror images for the symmetry planes `syms`.  Returns
        (Eff, Hff, Ptot) with Ptot = sum_j Re(S_j . wn_j), S = E x conj(H) / 2
        the time-averaged Poynting vector at the samples."""
        from .farfield import FarFieldSurface, farfield_symmetric
        surf = FarFieldSurface(self.mesh, tri_ids, origin=origin)
        E, H = self.interpolate(field, *surf.exyz, freq=freq, mu_r=mu_r)
        k0 = 2 * np.pi * freq / C0
        Eff, Hff = farfield_symmetric(E, H, surf, theta, phi, k0, syms=syms, ctx=self.ctx)
        S = 0.5 * np.cross(E, np.conj(H), axis=0)
        Ptot = float(np.sum(S * surf.weighted_normals).real)
        return Eff, Hff, Ptot

    def farfield_2d(self, field: np.ndarray, freq: float, tri_ids, ref_direction, plane_normal,
                    ang_range=(-180, 180), Npoints: int = 201, origin=None, syms=None, mu_r=None):
        """Far field along a cut (MWField.farfield_2d, MW/microwave_data.py:
        730-756): returns (angles in radians, Eff (3, N), Hff (3, N), Ptot)."""
        from .farfield import arc_on_plane
        theta, phi = arc_on_plane(ref_direction, plane_normal, ang_range, Npoints)
        Eff, Hff, Ptot = self.farfield(field, freq, tri_ids, theta, phi, origin=origin, syms=syms, mu_r=mu_r)
        return np.linspace(*ang_range, Npoints) * np.pi / 180, Eff, Hff, Ptot

    def farfield_3d(self, field: np.ndarray, freq: float, tri_ids, thetas=None, phis=None, origin=None,
                    syms=None, mu_r=None):
        """Far field on a theta x phi grid (MWField.farfield_3d,
        MW/microwave_data.py:758-788; defaults 0..pi in 91 and -pi..pi in
        181 steps): returns (T, P, Eff (3, *T.shape), Hff, Ptot), T and P from
        meshgrid(thetas, phis)."""
        thetas = np.linspace(0, np.pi, 91) if thetas is None else np.asarray(thetas)
        phis = np.linspace(-np.pi, np.pi, 181) if phis is None else np.asarray(phis)
        T, P = np.meshgrid(thetas, phis)
        Eff, Hff, Ptot = self.farfield(field, freq, tri_ids, T.ravel(), P.ravel(), origin=origin, syms=syms,
                                       mu_r=mu_r)
        return T, P, Eff.reshape((3,) + T.shape), Hff.reshape((3,) + T.shape), Ptot

    # ------------------------------------------------------------------ setup
    def pec_dofs(self) -> np.ndarray:
        """DOFs removed by PEC faces: both DOFs of every edge of a PEC triangle
        and all 8 DOFs of the triangle (ASM/assembler.py:290-307)."""
        f = self.basis
        ids = [np.zeros(0, dtype=np.int64)]
        for p in self.pec:
            t = p.tri_ids
            ids.append(f.edge_to_field[:, self.mesh.tri_to_edge[:, t].reshape(-1)].reshape(-1))
            ids.append(f.tri_to_field[:, t].reshape(-1))
        return np.unique(np.concatenate(ids))

    def spatial_order(self, ids: np.ndarray) -> np.ndarray:
        """`ids` reordered along a Morton curve of the DOF positions (edge
        midpoints, face centroids; both DOFs of an entity stay adjacent), so
        that consecutive reduced rows -- and each XCD's contiguous share of
        them -- couple to nearby columns."""
        m = self.mesh
        ec = 0.5 * (m.nodes[:, m.edges[0]] + m.nodes[:, m.edges[1]])
        fc = (m.nodes[:, m.tris[0]] + m.nodes[:, m.tris[1]] + m.nodes[:, m.tris[2]]) / 3.0
        pos = np.concatenate([ec, fc, ec, fc], axis=1)[:, ids]
        return ids[_morton_order(pos, ids, float(np.median(m.edge_lengths)))]

    def _setup(self) -> None:
        f, m = self.basis, self.mesh
        N = f.n_field
        keep_mask = np.ones(N, dtype=bool)
        keep_mask[self.pec_dofs()] = False
        self.solve_ids = np.nonzero(keep_mask)[0]
        if self.dof_order == "spatial":
            self.solve_ids = self.spatial_order(self.solve_ids)
        if self.dof_segments is not None:
            # segment-major rows, the order above kept inside each segment
            self.solve_ids = self.solve_ids[np.argsort(self.dof_segments[self.solve_ids], kind="stable")]
        red = np.full(N, -1, dtype=np.int64)
        red[self.solve_ids] = np.arange(self.solve_ids.shape[0])
        self.red = red
        ids32 = np.ascontiguousarray(self.solve_ids, dtype=np.int32)
        self.ctx.call("emg_system_reduce", ids32.shape[0], ptr(ids32))
        self.n_red = ids32.shape[0]
        self.segment_rows = None
        if self.dof_segments is not None:
            sid = self.dof_segments[self.solve_ids]
            starts = np.concatenate([[0], np.nonzero(np.diff(sid))[0] + 1, [sid.shape[0]]]).astype(np.int64)
            self.segment_rows = np.ascontiguousarray(starts)
            self.ctx.call("emg_solver_set_segments", int(starts.shape[0] - 1), ptr(self.segment_rows))

        # surface terms on the device, one per Robin BC plus one per
        # second-order ABC correction: (kind, bc); the solve scales term i by
        # gamma_i(k0)
        self.surface = [("robin", b) for b in self.robin]
        self.surface += [("abc2", b) for b in self.robin if getattr(b, "order", 1) == 2]
        tri_off = []
        for kind, b in self.surface:
            tv = np.ascontiguousarray(m.tris[:, b.tri_ids], dtype=np.int32)
            bid = ctypes.c_int32()
            off = ctypes.c_int64()
            if kind == "abc2":
                self.ctx.call("emg_bc_add_abc2", tv.shape[1], ptr(tv), ctypes.byref(bid), ctypes.byref(off))
            else:
                u = None
                if b.excited:
                    pts = dunavant_points(m.nodes, m.tris, b.tri_ids)  # (n,6,3)
                    shp = b.uinc_shape(pts[..., 0].reshape(-1), pts[..., 1].reshape(-1), pts[..., 2].reshape(-1))
                    u = np.ascontiguousarray(np.asarray(shp, dtype=np.complex128).reshape(3, -1).T)  # (n*6, 3)
                self.ctx.call("emg_bc_add", tv.shape[1], ptr(tv), ptr(u) if u is not None else None,
                              ctypes.byref(bid), ctypes.byref(off))
            tri_off.append(off.value)
        self.tri_off = tri_off

        # surface entries (row, col, term) over reduced rows, contributions in tri order
        rows_l, cols_l, bcs_l, con_l = [], [], [], []
        for ib, (kind, b) in enumerate(self.surface):
            t = b.tri_ids
            dof = f.tri_to_field[:, t]  # (8, n)
            rr = red[dof]
            n = t.shape[0]
            g = tri_off[ib] + np.arange(n)
            r = np.repeat(rr.T, 8, axis=1).reshape(-1)        # entry (tri, a, b) row dof[a]
            c = np.tile(rr.T, (1, 8)).reshape(-1)             # col dof[b]
            con = (g[:, None] * 64 + np.arange(64)[None, :]).reshape(-1)
            ok = (r >= 0) & (c >= 0)
            rows_l.append(r[ok]); cols_l.append(c[ok]); con_l.append(con[ok])
            bcs_l.append(np.full(ok.sum(), ib))
        if rows_l:
            r = np.concatenate(rows_l); c = np.concatenate(cols_l)
            bb = np.concatenate(bcs_l); con = np.concatenate(con_l)
        else:
            r = c = bb = con = np.zeros(0, dtype=np.int64)
        order = np.lexsort((con, bb, c, r))
        r, c, bb, con = r[order], c[order], bb[order], con[order]
        key = (r * (self.n_red + 1) + c) * (len(self.surface) + 1) + bb
        head = np.ones(key.shape[0], dtype=bool)
        head[1:] = key[1:] != key[:-1]
        starts = np.nonzero(head)[0]
        seg = np.append(starts, key.shape[0]).astype(np.int32)
        er, ec, eb = r[starts], c[starts], bb[starts]
        s_ptr = np.zeros(self.n_red + 1, dtype=np.int32)
        np.add.at(s_ptr, er + 1, 1)
        s_ptr = np.cumsum(s_ptr).astype(np.int32)
        a_col = np.ascontiguousarray(ec, dtype=np.int32)
        a_bc = np.ascontiguousarray(eb, dtype=np.int32)
        a_con = np.ascontiguousarray(con, dtype=np.int32)
        self.ctx.call("emg_surface_set", a_col.shape[0], ptr(s_ptr), ptr(a_col), ptr(a_bc), ptr(seg),
                      a_con.shape[0], ptr(a_con))
        self.n_sentries = a_col.shape[0]
        self.s_ptr, self.s_col, self.s_bc = s_ptr, a_col, a_bc

        # forcing vectors: one per port, rows p*n_red + r
        frows, fcon = [], []
        for ip, p in enumerate(self.ports):
            ib = self.robin.index(p)  # Robin terms come first in self.surface
            t = p.tri_ids
            rr = red[f.tri_to_field[:, t]].T  # (n, 8)
            g = tri_off[ib] + np.arange(t.shape[0])
            con = (g[:, None] * 8 + np.arange(8)[None, :]).reshape(-1)
            rr = rr.reshape(-1)
            ok = rr >= 0
            frows.append(ip * self.n_red + rr[ok]); fcon.append(con[ok])
        if self.ports:
            fr = np.concatenate(frows); fc = np.concatenate(fcon)
            order = np.lexsort((fc, fr))
            fr, fc = fr[order], fc[order]
            head = np.ones(fr.shape[0], dtype=bool)
            head[1:] = fr[1:] != fr[:-1]
            st = np.nonzero(head)[0]
            fseg = np.append(st, fr.shape[0]).astype(np.int32)
            frow = np.ascontiguousarray(fr[st], dtype=np.int32)
            fcon_a = np.ascontiguousarray(fc, dtype=np.int32)
            self.ctx.call("emg_forcing_set", len(self.ports), frow.shape[0], ptr(frow), ptr(fseg),
                          fcon_a.shape[0], ptr(fcon_a))

        # S-parameter probes: Dunavant points of the port faces for mode-overlap
        # ports, voltage-line points for lumped ports
        self.port_norms = None
        if self.ports:
            counts, owners, P_, C_, W_ = [], [], [], [], []
            for p in self.ports:
                if getattr(p, "v_integration", False):
                    own, pts, vec, w = self.voltage_line_probe(p)
                else:
                    t = p.tri_ids
                    pts = dunavant_points(m.nodes, m.tris, t)  # (n, 6, 3)
                    own = np.repeat(m.tri_to_tet[0, t], 6)
                    shp = np.asarray(p.mode_shape(pts[..., 0].reshape(-1), pts[..., 1].reshape(-1),
                                                  pts[..., 2].reshape(-1)), dtype=np.complex128)
                    vec = shp.reshape(3, -1).T
                    v = m.nodes[:, m.tris[:, t]]
                    cr = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], axis=0)
                    area = np.sqrt(cr[0] ** 2 + cr[1] ** 2 + cr[2] ** 2) / 2  # calc_area, E/mth/optimized.py:276-291
                    w = (_DUN_W[None, :] * area[:, None]).reshape(-1)
                    pts = pts.reshape(-1, 3)
                counts.append(pts.shape[0])
                owners.append(own)
                P_.append(pts)
                C_.append(vec)
                W_.append(w)
            cnt = np.ascontiguousarray(counts, dtype=np.int64)
            own = np.ascontiguousarray(np.concatenate(owners), dtype=np.int32)
            pts = np.ascontiguousarray(np.concatenate(P_), dtype=np.float64)
            vec = np.ascontiguousarray(np.concatenate(C_), dtype=np.complex128)
            w = np.ascontiguousarray(np.concatenate(W_), dtype=np.float64)
            self.ctx.call("emg_port_points_set", len(self.ports), ptr(cnt), ptr(own), ptr(pts), ptr(vec), ptr(w))
            norms = np.empty(len(self.ports))
            self.ctx.call("emg_port_norms_get", ptr(norms))
            self.port_norms = norms

    def aux_space(self):
        """Auxiliary space of the COCG preconditioner
        ``D^-1 + A W D_a^-1 A^T`` (Hiptmair-Xu type, with diagonal solves):
        A = [G | P] restricted to the reduced rows, with G the discrete
        gradient of the second-order nodal functions and P the vector-nodal
        interpolation of lam_a e_d (``emerge_amd.gradient``), both restricted
        to the functions that vanish on PEC (vertices and edges of PEC
        triangles dropped), columns in Morton order.  D_a per column:
        dconst - k0^2 dvol + sum_s gamma_s dsurf[s] from exact integrals (G:
        dconst 0, int eps |grad phi|^2, int_S |grad_t phi|^2; P: the curl-curl
        int (grad lam_a x e_d) . mu^-1 (grad lam_a x e_d), int lam_a^2 eps_dd,
        int_S lam_a^2 |e_d,t|^2; second-order ABC corrections contribute
        none); W = 2 on the vertex hats, 1 elsewhere.  With the vector-nodal
        part off (``self.vector_aux = False``) A = G, W = 1 (the plain
        gradient-space correction).  Returns (A csr, dvol, dsurf, dconst, W)."""
        m = self.mesh
        V, nE = m.n_nodes, m.n_edges
        vec = bool(self.vector_aux)
        if self._aux is None or self._aux[5] != vec:
            f = self.basis
            G = gradient.discrete_gradient(m, f)
            drop = np.zeros(G.shape[1], dtype=bool)
            pec_v = np.zeros(V, dtype=bool)
            for p in self.pec:
                pv = np.unique(m.tris[:, p.tri_ids])
                drop[pv] = True
                pec_v[pv] = True
                drop[V + np.unique(m.tri_to_edge[:, p.tri_ids])] = True
            mid = 0.5 * (m.nodes[:, m.edges[0]] + m.nodes[:, m.edges[1]])
            pos = np.concatenate([m.nodes, mid], axis=1)
            A = G
            if vec:
                P = gradient.vector_interpolation(m, f)
                A = sp.hstack([G, P], format="csr")
                drop = np.concatenate([drop, np.tile(pec_v, 3)])
                pos = np.concatenate([pos, m.nodes, m.nodes, m.nodes], axis=1)
            cols = np.nonzero(~drop)[0]
            cols = cols[_morton_order(pos[:, cols], cols, float(np.median(m.edge_lengths)))]
            Ar = A[self.solve_ids][:, cols].tocsr()
            Ar.sort_indices()
            nodal = cols < V + nE
            dsurf = np.zeros((len(self.surface), cols.shape[0]))
            for i, (kind, b) in enumerate(self.surface):
                if kind == "robin":
                    ds = gradient.nodal_surface_diagonal(m, b.tri_ids)
                    if vec:
                        ds = np.concatenate([ds, gradient.vector_surface_diagonal(m, b.tri_ids)])
                    dsurf[i] = ds[cols]
            wt = np.where(cols < V, 2.0, 1.0) if vec else np.ones(cols.shape[0])
            self._aux = (Ar, cols, dsurf, nodal, wt, vec)
        Ar, cols, dsurf, nodal, wt, vec = self._aux
        if self._dvol is None:
            dv = gradient.nodal_volume_diagonal(m, self.eps_r)
            dc = np.zeros(dv.shape[0], dtype=np.complex128)
            if vec:
                curl, mass = gradient.vector_volume_diagonals(m, self.eps_r, self._imu)
                dv = np.concatenate([dv, mass])
                dc = np.concatenate([dc, curl])
            self._dvol = (dv[cols], dc[cols])
        dvol, dconst = self._dvol
        return Ar, dvol, dsurf, dconst, wt

    def set_vector_aux(self, on: bool) -> None:
        """Keep (True, the default) or leave out the vector-nodal functions of
        the auxiliary space; the device copy is rebuilt on the next solve."""
        if bool(on) != self.vector_aux:
            self.vector_aux = bool(on)
            self._dvol = None
            if self._aux_on:
                self._aux_on = False
                self._set_aux(True)

    def _set_aux(self, on: bool) -> None:
        if on == self._aux_on:
            return
        if not on:
            self.ctx.call("emg_aux_set", 0, None, None, None, None, None, None, None, 0, None)
            self._aux_on = False
            return
        Gr, dvol, dsurf, dconst, wt = self.aux_space()
        Gt = Gr.T.tocsr()
        Gt.sort_indices()
        gp, gi = np.ascontiguousarray(Gr.indptr, np.int32), np.ascontiguousarray(Gr.indices, np.int32)
        gv = np.ascontiguousarray(Gr.data, np.float64)
        tp, ti = np.ascontiguousarray(Gt.indptr, np.int32), np.ascontiguousarray(Gt.indices, np.int32)
        tv = np.ascontiguousarray(Gt.data, np.float64)
        dv = np.ascontiguousarray(dvol, np.complex128)
        ds = np.ascontiguousarray(dsurf, np.float64)
        self.ctx.call("emg_aux_set", Gr.shape[1], ptr(gp), ptr(gi), ptr(gv), ptr(tp), ptr(ti), ptr(tv), ptr(dv),
                      ds.shape[0], ptr(ds) if ds.size else None)
        dc = np.ascontiguousarray(dconst, np.complex128)
        w = np.ascontiguousarray(wt, np.float64)
        self.ctx.call("emg_aux_set_terms", ptr(dc), ptr(w))
        self._aux_on = True

    def voltage_line_probe(self, port, npts: int = 21):
        """Probe points of a voltage-line port.  Lines (MW/microwave_3d.py:
        336-372): from every port-face node with the smallest coordinate along
        the port direction, `npts` equally spaced points over the port height.
        Line integral (E/coord.py:24-61, 149-153): the composite 3-point rule on
        the samples of E . step is, per section (y0, y1, y2), the combination
        (y0 + 4 y1 + y2) / 3; the port voltage is the mean over the lines.
        Each point belongs to the last face-adjacent tet that contains it
        (E/elements/ned2_interp.py:141-301 tolerances)."""
        m = self.mesh
        d = port.direction
        ids = np.unique(m.tris[:, port.tri_ids])
        P = m.nodes[:, ids]
        dot = P[0] * d[0] + P[1] * d[1] + P[2] * d[2]
        starts = P[:, dot == dot.min()]
        nl = starts.shape[1]
        lines = [np.stack([np.linspace(s0[k], s0[k] + d[k] * port.height, npts) for k in range(3)], axis=1)
                 for s0 in starts.T]
        pts = np.concatenate(lines)                                  # (nl*npts, 3)
        coef = np.zeros(npts)
        for i in range(0, npts - 2, 2):
            coef[i:i + 3] += (1.0 / 3.0, 4.0 / 3.0, 1.0 / 3.0)
        w = np.tile(coef, nl) / nl
        step = np.concatenate([np.repeat((ln[1] - ln[0])[None, :], npts, axis=0) for ln in lines])
        vec = step.astype(np.complex128)
        tets = m.get_face_tets(port.tri_ids)
        owner = np.full(pts.shape[0], -1, dtype=np.int64)
        X = pts.T
        for it in tets:
            v = m.nodes[:, m.tets[:, it]]
            Bl = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=1)
            loc = np.linalg.solve(Bl, X - v[:, 0:1])
            inside = ((loc[0] + loc[1] + loc[2]) <= 1.00000001) & (loc[0] >= -1e-6) & (loc[1] >= -1e-6) \
                & (loc[2] >= -1e-6)
            owner[inside] = it
        if np.any(owner < 0):
            raise ValueError(f"port {port.port_number}: voltage line leaves the face-adjacent tets")
        return owner, pts, vec, w

    # -------------------------------------------------------------- accessors
    def reduced_pattern(self):
        nnz = ctypes.c_int64()
        self.ctx.call("emg_system_pattern_get", None, None, ctypes.byref(nnz))
        ip = np.empty(self.n_red + 1, dtype=np.int32)
        ix = np.empty(nnz.value, dtype=np.int32)
        self.ctx.call("emg_system_pattern_get", ptr(ip), ptr(ix), ctypes.byref(nnz))
        return ip, ix

    def forcing(self) -> np.ndarray:
        out = np.empty((len(self.ports), self.n_red), dtype=np.complex128)
        self.ctx.call("emg_forcing_get", ptr(out))
        return out

    # ------------------------------------------------------------------ solve
    def batch_coefficients(self, freqs: np.ndarray, groups=None):
        """k2, gamma [batch][bc] and rhs coefficients [batch][port] for the
        (frequency, excitation) pairs in frequency-major order; an excitation
        is one port, or with `groups` (lists of port indices) every port of
        a group at once."""
        freqs = np.asarray(freqs, dtype=np.float64)
        nb = len(self.surface)
        npt = len(self.ports)
        groups = [[j] for j in range(npt)] if groups is None else groups
        ng = len(groups)
        k0 = 2 * np.pi * freqs / C0
        k2 = np.repeat(k0 ** 2, ng)
        gam = np.zeros((freqs.shape[0] * ng, max(nb, 1)), dtype=np.complex128)
        coef = np.zeros((freqs.shape[0] * ng, npt), dtype=np.complex128)
        for i, k in enumerate(k0):
            g = np.array([b.abc2_coefficient(k) if kind == "abc2" else b.get_gamma(k)
                          for kind, b in self.surface], dtype=np.complex128)
            for jg, grp in enumerate(groups):
                row = i * ng + jg
                gam[row, :nb] = g
                for j in grp:
                    coef[row, j] = self.ports[j].uinc_factor(k)
        return np.ascontiguousarray(k2), np.ascontiguousarray(gam[:, :nb] if nb else gam), np.ascontiguousarray(coef)

    def s_matrix(self, overlaps: np.ndarray, freqs: np.ndarray, caps: bool = True, groups=None,
                 blocks=None) -> np.ndarray:
        """S from the device overlaps, receiving port q, active port j
        (MW/microwave_3d.py:1193-1211, 1250-1303; MW/sparam.py:71-137).
        Mode-overlap port: field_p = (conj(q_m) O - [q==j] |q_m|^2 N_q) / 2Z,
        mode_p = |q_m|^2 N_q / 2Z, S_qj = field_p / mode_p / sqrt(P_j).
        Voltage-line port (O = mean line voltage V, V_q = sqrt(2 P_q Z0_q)):
        S_qj = ((V - [q==j] V_q) / V_q) / sqrt(P_j).  With port `groups`
        excited together and independent `blocks` of ports (one port of each
        group per block), the active port seen by q is the one of q's block;
        entries between blocks stay 0 and the caps act per block."""
        npt = len(self.ports)
        groups = [[j] for j in range(npt)] if groups is None else groups
        blocks = [list(range(npt))] if blocks is None else blocks
        block_of = np.empty(npt, dtype=np.int64)
        for ib, bl in enumerate(blocks):
            block_of[bl] = ib
        ng = len(groups)
        S = np.zeros((freqs.shape[0], npt, npt), dtype=np.complex128)
        for i, f in enumerate(freqs):
            k0 = 2 * np.pi * f / C0
            for jg, grp in enumerate(groups):
                row = i * ng + jg
                for q, prb in enumerate(self.ports):
                    act_in = [j for j in grp if block_of[j] == block_of[q]]
                    if len(act_in) != 1:
                        continue
                    j = act_in[0]
                    act = self.ports[j]
                    if getattr(prb, "v_integration", False):
                        V = overlaps[row, q]
                        a = prb.voltage
                        b = V - a if q == j else V
                        S[i, q, j] = (b / a) / np.sqrt(act.power)
                        continue
                    qm = prb.mode_factor(k0)
                    Zm = prb.Zmode(k0)
                    O = overlaps[row, q]
                    N = self.port_norms[q]
                    fp = (np.conj(qm) * O - (1.0 if q == j else 0.0) * abs(qm) ** 2 * N) / (2 * Zm)
                    mp = abs(qm) ** 2 * N / (2 * Zm)
                    S[i, q, j] = (fp / mp) / np.sqrt(act.power)
            if caps:
                for bl in blocks:
                    S[i][np.ix_(bl, bl)] = apply_s_caps(S[i][np.ix_(bl, bl)])
        return S

    def needs_safe32(self, freqs) -> bool:
        """True when (k0_min h_min)^2 < 0.05: on near-gradient fields E z
        cancels to O(k0^2 h^2) of its terms, so FP32 stiffness values (relative
        error 6e-8) would perturb the operator by 6e-8 / (k0 h)^2 there."""
        k0 = 2 * np.pi * float(np.min(freqs)) / C0
        h = float(np.min(self.mesh.edge_lengths))
        return (k0 * h) ** 2 < 0.05

    def choose_precision(self, freqs) -> str:
        """Inner-solve precision for precision='auto': 'fp64' where
        ``needs_safe32`` finds the mesh fine against the wavelength (there the
        complex64 inner solves need the FP64 stiffness and accumulation and
        about 20-30% more iterations, which eat their bandwidth advantage:
        measured 1.20 s mixed, 1.21 s FP64 and 1.13 s 'm32' on the 221k-tet
        cavity, 76.8 s, 61.2 s and 64.7 s for 8 points of the one-million-tet
        patch antenna), 'mixed' otherwise (1.23x the FP64 batched sweep and
        9.5x the plain FP64 path on the 201-point WR-90 sweep)."""
        return "fp64" if self.needs_safe32(freqs) else "mixed"

    @staticmethod
    def spmv_lane_options(F: int, precision: str, safe32: bool) -> list:
        """Entries-in-flight values of the COCG SpMV that batch width F can
        use (the device's layout rule: V = 8 complex64 or 4 complex128 (or
        FP64-accumulated) systems per lane, F / V lanes per entry, GR in
        {4, 8, 16, 32} with V <= GR <= 64 V / F); 0 stands for the
        default."""
        vm = 8 if precision == "mixed" and not safe32 else 4
        v = min(F, vm)
        gw = 64 // (F // v)
        return [0] + [g for g in (4, 8, 16, 32) if v <= g <= gw]

    def autotune(self, freqs, precision: str = "mixed", method: str = "cocg", precond: str = "aux",
                 candidates=(4, 8, 16, 32, 64), iters: int = 200, idr_s: int = 4, inner_rel32: float = 1e-2,
                 lanes: bool = True) -> tuple:
        """Frequency-block width F and SpMV lane layout (entries of a row in
        flight, hence rows per wavefront) with the least device time per
        system-iteration on this mesh: each candidate solves the first F
        (frequency, port) systems of `freqs` for at most `iters` inner
        iterations (one refinement cycle, complex128 continuation off), timed
        after a warm-up run; time is divided by the iterations actually run.
        Lane layouts are tried for COCG only (the other methods use other
        kernels).  The result (F, entries in flight; 0 = default) is cached
        per (precision, method, precond, safe mode)."""
        import time as _time
        safe = bool(getattr(self, "_safe32", False))
        key = (precision, method, precond, safe)
        cache = self.__dict__.setdefault("_fblock_tuned", {})
        if key in cache:
            return cache[key]
        freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
        k2, gam, coef = self.batch_coefficients(freqs)
        nb = k2.shape[0]
        self._set_aux(method == "cocg" and precond == "aux")
        self.ctx.call("emg_solver_config", {"gmres": 0, "idr": 1, "cocg": 2}[method], int(idr_s), float(inner_rel32))
        prec = {"fp64": 0, "mixed": 1, "m32": 2}[precision]
        best, best_cfg = np.inf, (1, 0)
        timings = {}
        self.ctx.call("emg_solver_set_escalation", 0)
        try:
            for F in candidates:
                if F > 32 and method != "cocg":
                    continue  # blocks of 64 are COCG only
                opts = self.spmv_lane_options(F, precision, safe) if (lanes and method == "cocg") else [0]
                rows = np.arange(F) % nb
                run = dict(fblock=F, tol=1e-300, restart=iters if method != "gmres" else min(iters, 100),
                           max_cycles=1, keep_fields=False)
                args = (k2[rows], gam[rows] if gam.size else gam, coef[rows], prec)
                for g in opts:
                    self.ctx.call("emg_solver_set_spmv_lanes", int(g))
                    self._run_rows(*args, **run)
                    t0 = _time.perf_counter()
                    _, _, it, _ = self._run_rows(*args, **run)
                    dt = _time.perf_counter() - t0
                    per = dt / max(int(it.sum()), 1)
                    timings[(F, g)] = per
                    if per < best:
                        best, best_cfg = per, (F, g)
        finally:
            self.ctx.call("emg_solver_set_escalation", 1)
            self.ctx.call("emg_solver_set_spmv_lanes", 0)
        cache[key] = best_cfg
        self.tuning = dict(cache=dict(cache), seconds_per_system_iteration=timings)
        return best_cfg

    def autotune_fblock(self, freqs, precision: str = "mixed", method: str = "cocg", precond: str = "aux",
                        candidates=(4, 8, 16, 32), iters: int = 200, idr_s: int = 4,
                        inner_rel32: float = 1e-2) -> int:
        """The frequency-block width of ``autotune`` (lane layout included in
        the search)."""
        return self.autotune(freqs, precision=precision, method=method, precond=precond, candidates=candidates,
                             iters=iters, idr_s=idr_s, inner_rel32=inner_rel32)[0]

    def solve(self, freqs, precision: str = "mixed", tol: float = 1e-8, fblock=32, restart: int | None = None,
              max_cycles: int = 200, keep_fields: bool = True, method: str = "cocg", idr_s: int = 4,
              inner_rel32: float | None = None, precond: str = "aux", fallback: bool = True,
              safe32="auto", groups=None, blocks=None) -> SweepResult:
        """Solve every (frequency, active port) system.  method 'gmres' =
        GMRES(restart); 'idr' = IDR(idr_s) with a budget of `restart`
        matvecs per FP64 refinement cycle; 'cocg' = COCG on the symmetric
        part inside FP64 refinement, at most `restart` iterations per cycle,
        each cycle reducing the residual by `inner_rel32` (defaults per
        method: restart 100 / 30000 / 20000, inner_rel32 1e-5 / 1e-4 / 1e-2
        for gmres / idr / cocg).  precond (COCG
        only): 'aux' = Jacobi plus the auxiliary gradient-space correction
        (``aux_space``), 'jacobi' = Jacobi alone.  fallback (mixed only):
        systems that miss tol are solved again in FP64
        (``SweepResult.fp64_fallback`` marks them).  fblock = systems per
        device batch (1..32; up to 64 with COCG) or 'auto' (``autotune_fblock``).  safe32
        (mixed COCG): keep the real stiffness in FP64 and accumulate the
        SpMV and the gradient-space correction in FP64; 'auto' turns it on
        when ``needs_safe32`` says the mesh is fine against the wavelength.
        precision: 'fp64', 'mixed' (complex64 inner solves), 'm32'
        (complex128 inner solves whose SpMV reads the symmetric mass as
        complex64: 20 instead of 28 bytes per entry, the FP64 refinement
        removes the perturbation) or 'auto' (``choose_precision``).  groups / blocks: excite the ports of each
        group together (one system per frequency and group) and extract S per
        block of independent ports (``s_matrix``); fields are then keyed by
        (frequency index, group index)."""
        freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
        if precision == "auto":
            precision = self.choose_precision(freqs)
        if precision not in ("fp64", "mixed", "m32"):
            raise ValueError("precision must be 'fp64', 'mixed', 'm32' or 'auto'")
        if precision == "m32" and method != "cocg":
            raise ValueError("precision 'm32' (complex64 mass values) is a COCG mode")
        if method not in ("gmres", "idr", "cocg"):
            raise ValueError("method must be 'gmres', 'idr' or 'cocg'")
        if precond not in ("aux", "jacobi"):
            raise ValueError("precond must be 'aux' or 'jacobi'")
        if restart is None:
            restart = {"gmres": 100, "idr": 30000, "cocg": 20000}[method]
        if inner_rel32 is None:
            inner_rel32 = {"gmres": 1e-5, "idr": 1e-4, "cocg": 1e-2}[method]
        if safe32 == "auto":
            safe32 = self.needs_safe32(freqs)
        self.ctx.call("emg_solver_set_safe32", int(bool(safe32)))
        self._safe32 = bool(safe32)
        npt = len(self.ports) if groups is None else len(groups)
        if self.frequency_dependent:
            # one frequency per device batch: each has its own E and B
            fblock = max(1, npt)
        lanes = 0
        if not self.frequency_dependent and fblock == "auto":
            fblock, lanes = self.autotune(freqs, precision=precision, method=method, precond=precond,
                                          idr_s=idr_s, inner_rel32=inner_rel32)
        self.ctx.call("emg_solver_set_spmv_lanes", int(lanes))
        self.ctx.call("emg_solver_config", {"gmres": 0, "idr": 1, "cocg": 2}[method], int(idr_s),
                      float(inner_rel32))
        k2, gam, coef = self.batch_coefficients(freqs, groups)
        self._escalated = 0
        run = dict(fblock=fblock, tol=tol, restart=restart, max_cycles=max_cycles, keep_fields=keep_fields)
        prec = {"fp64": 0, "mixed": 1, "m32": 2}[precision]
        aux = method == "cocg" and precond == "aux"
        if not self.frequency_dependent:
            self._set_aux(aux)
            x, relres, iters, ov, fell_back = self._solve_rows(k2, gam, coef, prec, fallback, run)
        else:
            parts = []
            for i, fq in enumerate(freqs):
                rows = np.arange(i * npt, (i + 1) * npt)
                self._load_frequency(float(fq))
                self._set_aux(aux)
                parts.append(self._solve_rows(k2[rows], gam[rows] if gam.size else gam, coef[rows], prec,
                                              fallback, run))
            x, relres, iters, ov, fell_back = (
                None if parts[0][j] is None else np.concatenate([q[j] for q in parts]) for j in range(5))
        fields = {}
        if x is not None:
            keys = [p.port_number for p in self.ports] if groups is None else list(range(npt))
            for i in range(freqs.shape[0]):
                for j, key in enumerate(keys):
                    full = np.zeros(self.basis.n_field, dtype=np.complex128)
                    full[self.solve_ids] = x[i * npt + j]
                    fields[(i, key)] = full
        S = self.s_matrix(ov, freqs, groups=groups, blocks=blocks) if ov is not None else None
        return SweepResult(freqs, self.ports, fields, relres.reshape(-1, npt), iters.reshape(-1, npt), precision,
                           S=S, fp64_fallback=fell_back.reshape(-1, npt), fp64_escalated=self._escalated)

    def _solve_rows(self, k2, gam, coef, prec, fallback, run):
        """`_run_rows`, then (precision 'mixed' or 'm32', fallback on) the
        rows the refinement gave up solved again in FP64."""
        x, relres, iters, ov = self._run_rows(k2, gam, coef, prec, **run)
        fell_back = np.zeros(k2.shape[0], dtype=bool)
        if prec != 0 and fallback:
            bad = np.nonzero(~(relres <= run["tol"]))[0]
            if bad.size:
                xb, rb, ib, ob = self._run_rows(k2[bad], gam[bad] if gam.size else gam, coef[bad], 0, **run)
                relres[bad] = rb
                iters[bad] += ib
                if x is not None:
                    x[bad] = xb
                if ov is not None:
                    ov[bad] = ob
                fell_back[bad] = True
        return x, relres, iters, ov, fell_back

    def _run_rows(self, k2, gam, coef, prec, fblock, tol, restart, max_cycles, keep_fields):
        """emg_sweep_solve over the given (frequency, port) rows: one call, or
        with ``self.progress`` (a callable taking (rows done, rows)) one call
        per block of `fblock` rows with a report after each (the device
        splits the rows into the same blocks either way)."""
        nbatch = k2.shape[0]
        cb = getattr(self, "progress", None)
        if cb is not None and nbatch > int(fblock):
            parts = []
            for a in range(0, nbatch, int(fblock)):
                b = min(nbatch, a + int(fblock))
                parts.append(self._run_rows_once(k2[a:b], gam[a:b] if gam.size else gam, coef[a:b], prec,
                                                 fblock, tol, restart, max_cycles, keep_fields))
                cb(b, nbatch)
            cat = lambda i: None if parts[0][i] is None else np.concatenate([q[i] for q in parts])  # noqa: E731
            return cat(0), cat(1), cat(2), cat(3)
        return self._run_rows_once(k2, gam, coef, prec, fblock, tol, restart, max_cycles, keep_fields)

    def _run_rows_once(self, k2, gam, coef, prec, fblock, tol, restart, max_cycles, keep_fields):
        """One emg_sweep_solve over the given (frequency, port) rows."""
        nbatch = k2.shape[0]
        k2 = np.ascontiguousarray(k2)
        gam = np.ascontiguousarray(gam)
        coef = np.ascontiguousarray(coef)
        x = np.empty((nbatch, self.n_red), dtype=np.complex128) if keep_fields else None
        relres = np.empty(nbatch)
        iters = np.empty(nbatch, dtype=np.int64)
        ov = np.empty((nbatch, len(self.ports)), dtype=np.complex128) if self.port_norms is not None else None
        self.ctx.call("emg_sweep_solve", nbatch, ptr(k2), ptr(gam) if gam.size else None, ptr(coef), prec,
                      int(fblock), float(tol), int(restart), int(max_cycles),
                      ptr(x) if x is not None else None, ptr(relres), ptr(iters),
                      ptr(ov) if ov is not None else None)
        esc = ctypes.c_int64()
        self.ctx.call("emg_sweep_escalated", ctypes.byref(esc))
        self._escalated += esc.value
        return x, relres, iters, ov
