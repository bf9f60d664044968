"""ModalPort: a port whose mode comes from a boundary-mode analysis of its face.

Counterpart of ``ModalPort`` / ``PortMode`` (MW/microwave_bc.py:348-372,
469-710) and ``MicrowaveSolver.modal_analysis`` (MW/microwave_3d.py:474-652),
``MW/`` = src/emerge/_emerge/physics/microwave/.  The face's mixed
second-order Nedelec / Lagrange element matrices come from the device
(``emg_port_mode_elements``); the port's pencil (a few thousand unknowns at
most) is solved once, on the host with LAPACK (``direct=True``, as the
reference's ``SolverLAPACK``) or ARPACK shift-invert (``direct=False``, as
``SolverARPACK``).  The mode then enters the frequency sweep exactly like an
analytic port mode: ``get_gamma``, ``uinc_shape`` / ``uinc_factor`` for the
forcing and ``mode_shape`` / ``mode_factor`` for the S-parameter overlaps.

Not carried over: automatic TEM conductor detection (a TEM port's impedance
needs ``set_integration_line``) and per-frequency re-analysis of ports with
mixed materials.  Alignment vectors sample the face at its triangle centres
(the reference samples a 5 x 5 grid of the face's gmsh parametrisation).
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.linalg
import scipy.sparse
import scipy.sparse.linalg

from . import bc as bcs
from ._lib import ptr

C0 = bcs.C0
MU0 = bcs.MU0
EPS0 = bcs.EPS0
PECLIM = 1e7
_EDGES = ((0, 1), (1, 2), (0, 2))

# full-precision 6-point Dunavant rule (the power integral's gaus_quad_tri(4))
_QW = np.array([0.223381589678011] * 3 + [0.109951743655322] * 3)
_QA = (0.108103018168070, 0.445948490915965)
_QB = (0.816847572980459, 0.091576213509771)
_QL = np.array([[_QA[0], _QA[1], _QA[1]], [_QA[1], _QA[1], _QA[0]], [_QA[1], _QA[0], _QA[1]],
                [_QB[0], _QB[1], _QB[1]], [_QB[1], _QB[1], _QB[0]], [_QB[1], _QB[0], _QB[1]]])


class PortFace:
    """The port's triangles as a 2-D mesh in the port frame, with the mixed
    DOF map of E/elements/nedleg2.py:117-160 (transverse: edge x3, face,
    edge x3, face; longitudinal: vertex x3, edge x3 per triangle).  Node
    order: increasing global id (E/mesh3d.py:717-745); edges (i1,i2),
    (i2,i3), (i1,i3) of each triangle, numbered in lexicographic order (the
    numbering does not enter any result)."""

    def __init__(self, mesh, tri_ids: np.ndarray, cs: bcs.CoordinateSystem):
        self.tri_ids = np.asarray(tri_ids, dtype=np.int64)
        self.cs = cs
        t = np.sort(mesh.tris[:, self.tri_ids], axis=0)
        self.node_ids, inv = np.unique(t, return_inverse=True)
        self.tris = inv.reshape(3, -1)
        nn, nt = self.node_ids.size, self.tris.shape[1]
        keys = np.stack([self.tris[p] * nn + self.tris[q] for p, q in _EDGES])  # (3, nt), p < q
        ek, t2e = np.unique(keys, return_inverse=True)
        self.edges = np.array([ek // nn, ek % nn])
        self.tri_to_edge = t2e.reshape(3, nt)
        ne = self.edges.shape[1]
        self.n_xy = 2 * ne + 2 * nt
        self.n_field = self.n_xy + nn + ne
        t2f = np.zeros((14, nt), dtype=np.int64)
        t2f[0:3] = self.tri_to_edge
        t2f[3] = np.arange(nt) + ne
        t2f[4:7] = self.tri_to_edge + ne + nt
        t2f[7] = np.arange(nt) + 2 * ne + nt
        t2f[8:11] = self.tris + self.n_xy
        t2f[11:14] = self.tri_to_edge + self.n_xy + nn
        self.tri_to_field = t2f
        e2f = np.zeros((5, ne), dtype=np.int64)
        e2f[0] = np.arange(ne)
        e2f[1] = np.arange(ne) + nt + ne
        e2f[2:4] = self.edges + self.n_xy
        e2f[4] = np.arange(ne) + self.n_xy + nn
        self.edge_to_field = e2f
        x, y, _ = cs.in_local_cs(*mesh.nodes[:, self.node_ids])
        self.xy = np.array([x, y])
        self._ek = ek                                      # sorted edge keys: edge e has key ek[e]
        # barycentric coefficients per triangle: lam_i = a_i + b_i x + c_i y
        v = self.xy[:, self.tris]                          # (2, 3, nt)
        x1, x2, x3 = v[0]
        y1, y2, y3 = v[1]
        sA = 0.5 * ((x1 - x3) * (y2 - y1) - (x1 - x2) * (y3 - y1))
        s = np.sign(sA) / (2 * np.abs(sA))
        self.area = np.abs(sA)
        self.a = np.array([x2 * y3 - y2 * x3, x3 * y1 - y3 * x1, x1 * y2 - y1 * x2]) * s
        self.b = np.array([y2 - y3, y3 - y1, y1 - y2]) * s
        self.c = np.array([x3 - x2, x1 - x3, x2 - x1]) * s

    def shared_edges(self, tris3d: np.ndarray) -> np.ndarray:
        """Face edges that are edges of the given 3-D triangles (3, k)."""
        nn = self.node_ids.size
        pos = np.minimum(np.searchsorted(self.node_ids, tris3d), nn - 1)
        hit = self.node_ids[pos] == tris3d
        found = []
        for p, q in ((0, 1), (1, 2), (0, 2)):
            ok = hit[p] & hit[q]
            keys = np.minimum(pos[p][ok], pos[q][ok]) * nn + np.maximum(pos[p][ok], pos[q][ok])
            i = np.minimum(np.searchsorted(self._ek, keys), self._ek.size - 1)
            found.append(i[self._ek[i] == keys])   # edge e has key _ek[e]
        return np.unique(np.concatenate(found))

    def vertex_xy(self) -> np.ndarray:
        """(nt, 3, 2) local vertex coordinates, vertices in increasing id."""
        return np.ascontiguousarray(np.moveaxis(self.xy[:, self.tris], 0, 2).transpose(1, 0, 2))

    # ------------------------------------------------------------ evaluation
    def locate(self, x, y):
        """Owner triangle and barycentrics (3, m) of local points (first
        containing triangle; -1 outside the face)."""
        m = x.size
        owner = np.full(m, -1, dtype=np.int64)
        lam = np.zeros((3, m))
        step = max(1, 4_000_000 // max(1, self.tris.shape[1]))
        for s in range(0, m, step):
            xs, ys = x[s:s + step], y[s:s + step]
            L = self.a[:, :, None] + self.b[:, :, None] * xs[None, None] + self.c[:, :, None] * ys[None, None]
            inside = np.all(L >= -1e-6, axis=0)                    # (nt, k)
            hit = inside.any(axis=0)
            first = np.argmax(inside, axis=0)
            owner[s:s + step] = np.where(hit, first, -1)
            lam[:, s:s + step] = L[:, first, np.arange(xs.size)]
        return owner, lam

    def basis(self, t, L):
        """Transverse functions (8, 2, m), their curls (8, m), Lagrange
        functions (6, m) and gradients (6, 2, m) at barycentrics L (3, m) of
        triangles t (m,)."""
        b, c = self.b[:, t], self.c[:, t]
        G = np.stack([b, c], axis=1)                              # (3, 2, m)

        def w(i, j):
            return L[i] * G[j] - L[j] * G[i], 2 * (b[i] * c[j] - c[i] * b[j])

        def scaled(f, gf, wv, wc):
            return f * wv, f * wc + (gf[0] * wv[1] - gf[1] * wv[0])

        m = L.shape[1]
        N = np.zeros((8, 2, m))
        Cu = np.zeros((8, m))
        for k, (p, q) in enumerate(_EDGES):
            wv, wc = w(p, q)
            N[k], Cu[k] = wv, wc
            N[k + 4], Cu[k + 4] = scaled(L[p] - L[q], G[p] - G[q], wv, wc)
        for k, (f1, i1, j1, f2, i2, j2) in ((3, (1, 0, 2, 0, 1, 2)), (7, (2, 1, 0, 1, 2, 0))):
            v1, c1 = scaled(L[f1], G[f1], *w(i1, j1))
            v2, c2 = scaled(L[f2], G[f2], *w(i2, j2))
            N[k], Cu[k] = v1 + v2, c1 + c2
        P = np.zeros((6, m))
        GP = np.zeros((6, 2, m))
        for k in range(3):
            P[k] = L[k] * (2 * L[k] - 1)
            GP[k] = G[k] * (4 * L[k] - 1)
        for k, (p, q) in enumerate(_EDGES):
            P[k + 3] = 4 * L[p] * L[q]
            GP[k + 3] = 4 * (G[p] * L[q] + G[q] * L[p])
        return N, Cu, P, GP

    def fields(self, field, x, y, k0=None, beta=None, urinv_diag=None):
        """E (and H when k0 is given) in local components (3, m) at local
        points; zero outside the face.  E as E/elements/ned2_interp.py:617-686;
        H = j/(k0 c0 mu0) diag(mu_r^-1) (d_y Ez - j beta Ey, -d_x Ez + j beta
        Ex, curl_z Et), the reference's convention (:688-800)."""
        owner, lam = self.locate(np.asarray(x, float), np.asarray(y, float))
        ok = owner >= 0
        E = np.zeros((3, owner.size), dtype=np.complex128)
        H = np.zeros((3, owner.size), dtype=np.complex128) if k0 is not None else None
        if not ok.any():
            return E, H
        t = owner[ok]
        N, Cu, P, GP = self.basis(t, lam[:, ok])
        e = field[self.tri_to_field[:, t]]                         # (14, m)
        Et = np.einsum("im,icm->cm", e[:8], N)
        E[:2, ok] = Et
        E[2, ok] = np.einsum("im,im->m", e[8:], P)
        if k0 is not None:
            gz = np.einsum("im,icm->cm", e[8:], GP)
            cz = np.einsum("im,im->m", e[:8], Cu)
            jb = 1j * beta
            f = 1j / (k0 * C0 * MU0)
            d = urinv_diag[:, t]
            H[0, ok] = f * (gz[1] - jb * Et[1]) * d[0]
            H[1, ok] = f * (-gz[0] + jb * Et[0]) * d[1]
            H[2, ok] = f * cz * d[2]
        return E, H

    def power_flux(self, field, k0, beta, urinv_diag) -> float:
        """0.5 Re (E x H*) . z over the face (MW/port_functions.py:21-36)."""
        px = (_QL @ self.xy[0, self.tris]).T.reshape(-1)          # (nt*6,)
        py = (_QL @ self.xy[1, self.tris]).T.reshape(-1)
        E, H = self.fields(field, px, py, k0, beta, urinv_diag)
        S = 0.5 * np.real(E[0] * np.conj(H[1]) - E[1] * np.conj(H[0])).reshape(-1, 6)
        return float(np.sum(self.area * (S @ _QW)))


class PortMode:
    """One mode of a port (MW/microwave_bc.py:348-372)."""

    def __init__(self, face: PortFace, field, k0, beta, freq, urinv_diag):
        self.face = face
        self.modefield = field
        self.k0 = k0
        self.beta = beta
        self.freq = freq
        self.neff = beta / k0
        self.energy = float(np.mean(np.abs(field) ** 2))
        self.norm_factor = 1.0
        self.polarity = 1.0
        self.Z0 = 50.0
        self.modetype = "TEM"
        self._urinv = urinv_diag

    def set_power(self, power) -> None:
        self.norm_factor = np.sqrt(1 / np.abs(power))

    def E_global(self, x, y, z):
        xl, yl, _ = self.face.cs.in_local_cs(x, y, z)
        E, _ = self.face.fields(self.modefield, np.atleast_1d(xl), np.atleast_1d(yl))
        return np.array(self.face.cs.in_global_basis(E[0], E[1], E[2]))


class ModalPort(bcs.PortBC):
    """Port on the triangles `tri_ids` whose mode is found by
    ``modal_analysis`` (MW/microwave_bc.py:469-710).  `cs` is the port
    frame (local z along the face normal); without one a frame is built from
    the face plane (origin at the centroid, z the mean normal)."""

    def __init__(self, tri_ids, port_number: int, cs: bcs.CoordinateSystem | None = None,
                 power: float = 1.0, modetype: str | None = None, mixed_materials: bool = False):
        super().__init__(tri_ids, port_number, power)
        self.cs = cs
        self.forced_modetype = modetype
        self.mixed_materials = mixed_materials
        self.modes: dict[float, list[PortMode]] = {}
        self.selected_mode = 0
        self._first_k0 = None
        self._last_k0 = None
        self.vintline: list[np.ndarray] = []
        self.alignment_vectors: list[np.ndarray] = []
        self.initialized = False

    # mode bookkeeping (MW/microwave_bc.py:597-710)
    def set_integration_line(self, c1, c2, N: int = 21) -> None:
        """Voltage line (global points) for the impedance of a TEM mode."""
        self.vintline.append(np.linspace(np.asarray(c1, float), np.asarray(c2, float), N))

    @property
    def nmodes(self) -> int:
        if self._last_k0 is None:
            raise ValueError("ModalPort is not properly configured. No modes are defined.")
        return len(self.modes[self._last_k0])

    def get_mode(self, k0: float, i: int | None = None) -> PortMode:
        if not self.modes:
            raise ValueError("ModalPort has no modes: run modal_analysis first")
        options = self.modes[min(self.modes, key=lambda k: abs(k - k0))]
        if i is None:
            i = min(len(options) - 1, self.selected_mode)
        return options[i]

    def modetype(self, k0):
        return self.get_mode(k0).modetype

    def portZ0(self, k0):
        return self.get_mode(k0).Z0

    def get_beta(self, k0):
        mode = self.get_mode(k0)
        if self.forced_modetype == "TEM":
            return mode.beta / mode.k0 * k0
        freq = k0 * C0 / (2 * np.pi)
        return np.sqrt(mode.beta ** 2 + k0 ** 2 * (1 - (mode.freq / freq) ** 2))

    def get_gamma(self, k0):
        return 1j * self.get_beta(k0)

    def mode_shape(self, x, y, z):
        """Normalised mode field without the _qmode(k0) factor (3, n)."""
        mode = self.get_mode(self._first_k0 if self._first_k0 is not None else 0.0)
        return mode.norm_factor * mode.polarity * mode.E_global(x, y, z)

    def mode_factor(self, k0):
        return self._qmode(k0)

    def port_mode_3d_global(self, x, y, z, k0):
        return self.mode_factor(k0) * self.mode_shape(x, y, z)

    def uinc_shape(self, x, y, z):
        return self.mode_shape(x, y, z)

    def uinc_factor(self, k0):
        return -2 * 1j * self.get_beta(k0) * self._qmode(k0)

    def add_mode(self, mode: PortMode) -> PortMode | None:
        """MW/microwave_bc.py:666-702: modes of negligible energy are dropped;
        polarity relative to the last mode of the first analysis."""
        if mode.energy < 1e-4:
            return None
        self.modes.setdefault(mode.k0, []).append(mode)
        self.initialized = True
        self._last_k0 = mode.k0
        if self._first_k0 is None:
            self._first_k0 = mode.k0
        else:
            ref = self.get_mode(self._first_k0, -1).modefield
            mode.polarity = float(np.sign(np.sum(mode.modefield * ref).real))
        return mode

    def align_modes(self, *axes) -> None:
        """Mode order and polarity by alignment vectors (MW/microwave_bc.py:561-575):
        mode i is the one whose field has the largest |sum E . axes[i]| over
        the face samples, with that sum made positive."""
        self.alignment_vectors = [np.asarray(a, dtype=float) for a in axes]

    def _face_samples(self, mode: PortMode):
        """Sample points of the face: its triangle centroids (global)."""
        f = mode.face
        cx, cy = f.xy[0, f.tris].mean(axis=0), f.xy[1, f.tris].mean(axis=0)
        o, ex, ey = f.cs.origin, f.cs.xax, f.cs.yax
        return [o[k] + ex[k] * cx + ey[k] * cy for k in range(3)]

    def sort_modes(self) -> None:
        """MW/microwave_bc.py:614-639: by the alignment vectors when given,
        else by decreasing beta."""
        for k0, modes in self.modes.items():
            if self.alignment_vectors:
                x, y, z = self._face_samples(modes[0])
                picked = []
                for ax in self.alignment_vectors:
                    ints = [np.sum(np.einsum("cn,c->n", m.E_global(x, y, z), ax)) for m in modes]
                    i = int(np.argmax(np.abs(ints)))
                    modes[i].polarity = float(np.sign(np.real(ints[i])))
                    picked.append(modes[i])
                self.modes[k0] = picked
            else:
                self.modes[k0] = sorted(modes, key=lambda m: np.real(m.beta), reverse=True)


def _face_frame(mesh, tri_ids) -> bcs.CoordinateSystem:
    """Frame of a port given without one: z along the face's normal pointing
    out of the volume (away from the tet behind each triangle), x and y as
    Axis.construct_cs builds them (E/cs.py:106-118: x-axis unless z is within
    0.999 of it, then y), origin at the face's centroid (the origin only
    shifts the local coordinates the modes are solved and evaluated in)."""
    v = mesh.nodes[:, mesh.tris[:, tri_ids]]                     # (3, 3, n)
    nt = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], axis=0)  # (3, n)
    tet = mesh.tets[:, mesh.tri_to_tet[0, tri_ids]]              # (4, n)
    inside = mesh.nodes[:, tet].mean(axis=1) - v.mean(axis=1)    # centroid of the tet minus the triangle's
    nt = nt * -np.sign(np.einsum("in,in->n", nt, inside))
    z = nt.sum(axis=1)
    z = z / np.linalg.norm(z)
    ax = np.array([1.0, 0, 0]) if abs(z[0]) <= 0.999 else np.array([0, 1.0, 0])
    y = np.cross(z, ax)
    x = -np.cross(z, y)
    return bcs.CoordinateSystem(x, y, z, v.reshape(3, -1).mean(axis=1))


def _per_tri_tensor(values, tets) -> np.ndarray:
    """(3, 3, n) complex tensors of per-tet scalars or tensors."""
    v = np.asarray(values, dtype=np.complex128)
    if v.ndim == 1:
        return v[tets][None, None, :] * np.eye(3)[:, :, None]
    return v[:, :, tets]


def modal_analysis(ctx, mesh, port: ModalPort, eps_r, inv_mu_r=None, boundary_conditions=(), freq: float = 1e9,
                   nmodes: int = 6, direct: bool = True, TEM: bool = False, target_kz: float | None = None,
                   target_neff: float | None = None, mu_r=None, cond=None) -> list[PortMode]:
    """Modes of `port` at `freq` (MicrowaveSolver.modal_analysis,
    MW/microwave_3d.py:474-652, with the materials of the tet behind each face
    triangle, E/physics/microwave/microwave_3d.py:533-541).  eps_r: (T,) or
    (3, 3, T) complex; inv_mu_r (T,) or mu_r (3, 3, T); cond (T,) S/m.
    Conductor DOFs: edges shared with PEC faces and triangles above the
    perfect-conductor limit (ASM/assembler.py:154-183)."""
    if port.cs is None:
        port.cs = _face_frame(mesh, port.tri_ids)
    face = PortFace(mesh, port.tri_ids, port.cs)
    k0 = 2 * np.pi * freq / C0
    tets = mesh.tri_to_tet[0, face.tri_ids]
    er = _per_tri_tensor(eps_r(freq) if callable(eps_r) else eps_r, tets)
    if mu_r is not None:
        ur = _per_tri_tensor(mu_r, tets)
    else:
        ur = _per_tri_tensor(1.0 / np.asarray(inv_mu_r, dtype=np.complex128), tets)
    sig = np.zeros(face.tri_ids.size) if cond is None else np.asarray(cond, float)[tets]
    er = er - 1j * (sig / (k0 * C0 * EPS0))[None, None, :] * np.eye(3)[:, :, None]
    urinv = np.moveaxis(np.linalg.inv(np.moveaxis(ur, 2, 0)), 0, 2)
    nt = face.tris.shape[1]
    xy = face.vertex_xy()
    e9 = np.ascontiguousarray(np.moveaxis(er, 2, 0))             # (nt, 3, 3)
    u9 = np.ascontiguousarray(np.moveaxis(urinv, 2, 0))
    Ae = np.empty((nt, 14, 14), dtype=np.complex128)
    Be = np.empty((nt, 14, 14), dtype=np.complex128)
    ctx.call("emg_port_mode_elements", nt, ptr(xy), ptr(e9), ptr(u9), ctypes.c_double(k0), ptr(Ae), ptr(Be))
    rows = np.repeat(face.tri_to_field.T, 14, axis=1).reshape(-1)
    cols = np.tile(face.tri_to_field.T, (1, 14)).reshape(-1)
    n = face.n_field
    A = scipy.sparse.csr_matrix((Ae.reshape(-1), (rows, cols)), shape=(n, n))
    B = scipy.sparse.csr_matrix((Be.reshape(-1), (rows, cols)), shape=(n, n))

    pec = set()
    for it in np.nonzero(sig > PECLIM)[0]:
        pec.update(int(i) for i in face.tri_to_field[:, it])
    for b in boundary_conditions:
        if isinstance(b, bcs.PEC) and len(b.tri_ids):
            for e in face.shared_edges(mesh.tris[:, b.tri_ids]):
                pec.update(int(i) for i in face.edge_to_field[:, e])
    solve_ids = np.array(sorted(set(range(n)) - pec), dtype=np.int64)

    ermean = float(np.mean(er[0, 0].real))
    urmean = float(np.mean(ur[0, 0].real))
    if target_neff is not None:
        target_kz = k0 * target_neff
    if target_kz is None:
        target_kz = ermean * urmean * (1.1 if (TEM or port.forced_modetype == "TEM") else 0.7) * k0
    As = A[solve_ids][:, solve_ids]
    Bs = B[solve_ids][:, solve_ids]
    if direct:
        # SolverLAPACK.eig (E/solver.py:790-808) and filter_real_modes (:192-227)
        lam, vec = scipy.linalg.eig(As.toarray(), Bs.toarray(), check_finite=False)
        ext = target_kz ** 2 * 2 * 2 * 2
        keep = (-lam <= ext) & (-lam >= 1)
        lam, vec = lam[keep], vec[:, keep]
        order = np.argsort(np.abs(np.sqrt(-lam)))
        lam, vec = lam[order], vec[:, order]
    else:
        # SolverARPACK.eig (E/solver.py:816-832)
        lam, vec = scipy.sparse.linalg.eigs(As.tocsc(), k=min(nmodes, As.shape[0] - 2), M=Bs.tocsc(),
                                            sigma=-(target_kz ** 2), which="LM")

    ermax = np.max(er.real)
    urmax = np.max(ur.real)
    urinv_diag = np.array([urinv[0, 0], urinv[1, 1], urinv[2, 2]])
    added = []
    for i in range(min(nmodes, lam.size)):
        E = np.zeros(n, dtype=np.complex128)
        E[solve_ids] = vec[:, i]
        E = E * np.exp(-1j * np.angle(E[np.argmax(np.abs(E))]))
        beta = min(k0 * np.sqrt(ermax * urmax), np.emath.sqrt(-lam[i]))
        if i < len(port.alignment_vectors):
            # MW/microwave_3d.py:595-602: the mean field over the triangle
            # centres points along alignment vector i
            probe = PortMode(face, E, k0, beta, freq, urinv_diag)
            Ec = probe.E_global(*port._face_samples(probe)).mean(axis=1)
            if np.real(np.dot(port.alignment_vectors[i], Ec)) < 0:
                E = -E
        P = face.power_flux(E, k0, beta, urinv_diag)
        mode = port.add_mode(PortMode(face, E, k0, beta, freq, urinv_diag))
        if mode is None:
            continue
        ez = np.max(np.abs(E[face.n_xy:]))
        exy = np.max(np.abs(E[:face.n_xy]))
        if port.forced_modetype == "TEM" or TEM:
            mode.modetype = "TEM"
            if port.vintline:
                line = port.vintline[0]
                mid = 0.5 * (line[1:] + line[:-1])
                dl = line[1:] - line[:-1]
                Em = mode.E_global(mid[:, 0], mid[:, 1], mid[:, 2])
                voltage = complex(np.sum(Em.T * dl))
                if voltage.real < 0:
                    mode.polarity = -mode.polarity
                mode.Z0 = abs(voltage ** 2 / (2 * P))
        elif ez / exy < 1e-1 or port.forced_modetype == "TE":
            mode.modetype = "TE"
        else:
            mode.modetype = "TM"
        mode.set_power(P * port._qmode(k0) ** 2)
        added.append(mode)
    port.sort_modes()
    return added
