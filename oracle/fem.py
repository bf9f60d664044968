"""Oracle: frequency-point assembly, direct solve and S-parameter extraction.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates (``ASM/`` = src/emerge/_emerge/physics/microwave/assembly/,
``MW/`` = src/emerge/_emerge/physics/microwave/, ``E/`` = src/emerge/_emerge/):
* ``Assembler.assemble_freq_matrix`` for isotropic, non-conducting materials
  without periodic BCs (ASM/assembler.py:218-429);
* ``SimJob.iter_Ab`` (MW/simjob.py:82-108);
* ``SolveRoutine.solve`` with ``SolverSuperLU`` (E/solver.py:1241-1326,
  535-572);
* ``ned2_tet_interp`` (E/elements/ned2_interp.py:141-301);
* ``surface_integral`` (E/mth/integrals.py:28-92), ``sparam_field_power`` /
  ``sparam_mode_power`` (MW/sparam.py:71-137) and the S-matrix loop with its
  caps (MW/microwave_3d.py:1167-1303).
Pinned by the closed forms of an empty rectangular waveguide (tests) and by
SuperLU itself (a library routine at the version the reference pins).
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.linalg import splu

from . import robin
from .curlcurl import local_mapping
from .quadrature import tet_coefficients

C0 = 299792458


def assemble_freq_matrix(mesh, maps: dict, E, B, pec_tris, robin_bcs, frequency: float, cond=None,
                         peclim: float = 1e7, tet_to_field=None):
    """Returns (K, b, port_vectors, solve_ids).  ``robin_bcs`` are oracle.ports
    objects; ports (with a ``port_number``) are excited.  ``cond`` (3, 3, T):
    every tet with cond[0, 0] > peclim is a conductor whose 20 DOFs are
    eliminated (ASM/assembler.py:280-288; settings.mw_3d_peclim = 1e7)."""
    K0 = 2 * np.pi * frequency / C0
    K = (E - B * (K0**2)).tocsr()
    N = E.shape[0]
    b = np.zeros(N, dtype=np.complex128)
    port_vectors = {bc.port_number: np.zeros(N, dtype=np.complex128)
                    for bc in robin_bcs if hasattr(bc, "port_number")}
    pec_ids = []
    if cond is not None:
        for itet in range(mesh.n_tets):
            if cond[0, 0, itet] > peclim:
                pec_ids.extend(tet_to_field[:, itet])
    tri_ids = np.asarray(pec_tris, dtype=np.int64)
    for ii in list(mesh.tri_to_edge[:, tri_ids].flatten()):
        pec_ids.extend(list(maps["edge_to_field"][:, ii]))
    for ii in tri_ids:
        pec_ids.extend(list(maps["tri_to_field"][:, ii]))
    if robin_bcs:
        pts = robin.gaus_quad_tri(4)
        Bempty = np.zeros(mesh.n_tris * 64, dtype=np.complex128)
        for bc in robin_bcs:
            gamma = bc.get_gamma(K0)
            if hasattr(bc, "port_number"):
                Bempty, b_p = robin.assemble_robin_bc_excited(
                    mesh.nodes, mesh.tris, maps["tri_to_field"], N, Bempty, bc.tri_ids,
                    lambda x, y, z, bc=bc: bc.get_Uinc(x, y, z, K0), gamma, pts)
                port_vectors[bc.port_number] += b_p
            else:
                Bempty = robin.assemble_robin_bc(mesh.nodes, mesh.tris, Bempty, bc.tri_ids, gamma)
            # second-order ABC correction (ASM/assembler.py:342-348)
            if getattr(bc, "order", 1) == 2:
                from . import abc2
                c2 = bc.o2coeffs[bc.abctype][1]
                Bempty = Bempty + abc2.abc_order_2_matrix(mesh.nodes, mesh.tris, mesh.n_tris, bc.tri_ids,
                                                          1j * c2 / K0)
        from .topology import empty_tri_rowcol
        rows, cols = empty_tri_rowcol(maps["tri_to_field"])
        ids = np.argwhere(Bempty != 0)[:, 0]
        Bp = csr_matrix((Bempty[ids], (rows[ids], cols[ids])), shape=(N, N))
        K = K + Bp
    pec = set(int(i) for i in pec_ids)
    solve_ids = np.array([i for i in range(N) if i not in pec])
    return K, b, port_vectors, solve_ids


def solve_superlu(A, b, solve_ids, lu=None):
    """SolveRoutine.solve + SolverSuperLU: DOF elimination, splu with the
    reference's options, scatter back.  Returns (x, lu)."""
    A = A.tocsc()
    Asel = A[np.ix_(solve_ids, solve_ids)]
    bsel = b[solve_ids]
    if lu is None:
        lu = splu(Asel, permc_spec="MMD_AT_PLUS_A", relax=0, diag_pivot_thresh=0.001,
                  options=dict(SymmetricMode=True, Equil=False, IterRefine="SINGLE"))
    x = lu.solve(bsel)
    sol = np.zeros(A.shape[0], dtype=np.complex128)
    sol[solve_ids] = x
    return sol, lu


def solve_all_ports(K, b, port_vectors, solve_ids) -> dict:
    """SimJob.iter_Ab: one solve per port with b + port vector, factorization
    reused after the first."""
    fields = {}
    lu = None
    for key, mode in port_vectors.items():
        fields[key], lu = solve_superlu(K, b + mode, solve_ids, lu)
    return fields


def ned2_tet_interp(coords, solution, mesh, tet_to_field, tetids):
    """E/elements/ned2_interp.py:141-301: field at points, each point owned by
    the last listed tet that contains it (with the reference's tolerances)."""
    n = coords.shape[1]
    xs, ys, zs = coords
    Ex = np.zeros(n, dtype=np.complex128)
    Ey = np.zeros_like(Ex)
    Ez = np.zeros_like(Ex)
    setnan = np.zeros(n)
    assigned = np.zeros(n, dtype=np.int64) - 1
    nodes, tets = mesh.nodes, mesh.tets
    for itet in tetids:
        v = nodes[:, tets[:, itet]]
        blocal = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=1)
        loc = np.linalg.inv(blocal) @ (coords - v[:, 0:1])
        inside = ((loc[0] + loc[1] + loc[2]) <= 1.00000001) & (loc[0] >= -1e-6) & (loc[1] >= -1e-6) \
            & (loc[2] >= -1e-6)
        assigned[inside] = itet
    nE = mesh.n_edges
    for itet in tetids:
        inside = assigned == itet
        if not inside.any():
            continue
        x, y, z = xs[inside], ys[inside], zs[inside]
        xv, yv, zv = nodes[0, tets[:, itet]], nodes[1, tets[:, itet]], nodes[2, tets[:, itet]]
        a_s, b_s, c_s, d_s, V = tet_coefficients(xv, yv, zv)
        g_edges = mesh.edges[:, tet_to_field[:6, itet]]
        g_tris = mesh.tris[:, tet_to_field[6:10, itet] - nE]
        le = local_mapping(tets[:, itet][:, None], g_edges[:, :, None])[:, :, 0]
        lt = local_mapping(tets[:, itet][:, None], g_tris[:, :, None])[:, :, 0]
        Et = solution[tet_to_field[:, itet]]
        V1 = 216 * V**3
        ex = np.zeros(x.shape, dtype=np.complex128)
        ey = np.zeros_like(ex)
        ez = np.zeros_like(ex)
        for ie in range(6):
            e1, e2 = le[:, ie]
            F1 = a_s[e1] + b_s[e1]*x + c_s[e1]*y + d_s[e1]*z
            F2 = a_s[e2] + b_s[e2]*x + c_s[e2]*y + d_s[e2]*z
            L = np.sqrt((xv[e1]-xv[e2])**2 + (yv[e1]-yv[e2])**2 + (zv[e1]-zv[e2])**2)
            amp = L*(Et[ie]*F1 + Et[ie + 10]*F2)
            ex += amp*(b_s[e1]*F2 - b_s[e2]*F1)/V1
            ey += amp*(c_s[e1]*F2 - c_s[e2]*F1)/V1
            ez += amp*(d_s[e1]*F2 - d_s[e2]*F1)/V1
        for it in range(4):
            i1, i2, i3 = lt[:, it]
            L1 = np.sqrt((xv[i1]-xv[i3])**2 + (yv[i1]-yv[i3])**2 + (zv[i1]-zv[i3])**2)
            L2 = np.sqrt((xv[i1]-xv[i2])**2 + (yv[i1]-yv[i2])**2 + (zv[i1]-zv[i2])**2)
            F1 = a_s[i1] + b_s[i1]*x + c_s[i1]*y + d_s[i1]*z
            F2 = a_s[i2] + b_s[i2]*x + c_s[i2]*y + d_s[i2]*z
            F3 = a_s[i3] + b_s[i3]*x + c_s[i3]*y + d_s[i3]*z
            Q1 = Et[it + 6]*L1*F2
            Q2 = Et[it + 16]*L2*F3
            ex += (-Q1*(b_s[i1]*F3 - b_s[i3]*F1) + Q2*(b_s[i1]*F2 - b_s[i2]*F1))/V1
            ey += (-Q1*(c_s[i1]*F3 - c_s[i3]*F1) + Q2*(c_s[i1]*F2 - c_s[i2]*F1))/V1
            ez += (-Q1*(d_s[i1]*F3 - d_s[i3]*F1) + Q2*(d_s[i1]*F2 - d_s[i2]*F1))/V1
        Ex[inside] += ex
        Ey[inside] += ey
        Ez[inside] += ez
        setnan[inside] += 1
    Ex[setnan == 0] = np.nan
    Ey[setnan == 0] = np.nan
    Ez[setnan == 0] = np.nan
    return Ex, Ey, Ez


def ned2_tet_interp_curl(coords, solution, mesh, tet_to_field, tetids, c):
    """E/elements/ned2_interp.py:302-507: c[tet] times the curl of the field
    at points, each point owned by the last listed tet that contains it (the
    same search as ``ned2_tet_interp``).  With F_i = 6V lam_i (cofactors
    a_i + b_i x + c_i y + d_i z, gradients g_i) the reference's expanded
    expressions are the analytic curl of its interpolation, written here in
    product form: edge (e1, e2):  -L (m1 F1 + m2 F2) (g1 x g2) / (72 V^3);
    face (i1, i2, i3):  [-m1 L1 (g2 x (F3 g1 - F1 g3) + 2 F2 g3 x g1)
    + m2 L2 (g3 x (F2 g1 - F1 g2) + 2 F3 g2 x g1)] / (216 V^3).
    Pinned by central differences of ``ned2_tet_interp`` (exact for the
    quadratic field up to rounding)."""
    n = coords.shape[1]
    out = np.zeros((3, n), dtype=np.complex128)
    found = np.zeros(n, dtype=bool)
    assigned = np.zeros(n, dtype=np.int64) - 1
    nodes, tets = mesh.nodes, mesh.tets
    for itet in tetids:
        v = nodes[:, tets[:, itet]]
        blocal = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=1)
        loc = np.linalg.inv(blocal) @ (coords - v[:, 0:1])
        inside = ((loc[0] + loc[1] + loc[2]) <= 1.00000001) & (loc[0] >= -1e-6) & (loc[1] >= -1e-6) \
            & (loc[2] >= -1e-6)
        assigned[inside] = itet
    nE = mesh.n_edges
    for itet in tetids:
        inside = assigned == itet
        if not inside.any():
            continue
        P = coords[:, inside]
        xv, yv, zv = nodes[0, tets[:, itet]], nodes[1, tets[:, itet]], nodes[2, tets[:, itet]]
        a_s, b_s, c_s, d_s, V = tet_coefficients(xv, yv, zv)
        g = np.stack([b_s, c_s, d_s])                                  # (3, 4)
        Fv = a_s[:, None] + g.T @ P                                    # (4, m)
        le = local_mapping(tets[:, itet][:, None], mesh.edges[:, tet_to_field[:6, itet]][:, :, None])[:, :, 0]
        lt = local_mapping(tets[:, itet][:, None], mesh.tris[:, tet_to_field[6:10, itet] - nE][:, :, None])[:, :, 0]
        Et = solution[tet_to_field[:, itet]]
        V1 = 216 * V ** 3
        V2 = 72 * V ** 3
        acc = np.zeros((3, P.shape[1]), dtype=np.complex128)

        def dist(p, q):
            return np.sqrt((xv[p] - xv[q]) ** 2 + (yv[p] - yv[q]) ** 2 + (zv[p] - zv[q]) ** 2)

        for ie in range(6):
            e1, e2 = le[:, ie]
            L = dist(e1, e2)
            phi = L * (Et[ie] * Fv[e1] + Et[ie + 10] * Fv[e2])
            acc -= np.cross(g[:, e1], g[:, e2])[:, None] * phi[None, :] / V2
        for it in range(4):
            i1, i2, i3 = lt[:, it]
            L1, L2 = dist(i1, i3), dist(i1, i2)
            g1, g2, g3 = g[:, i1][:, None], g[:, i2][:, None], g[:, i3][:, None]
            F1, F2, F3 = Fv[i1][None, :], Fv[i2][None, :], Fv[i3][None, :]
            t1 = np.cross(g2, F3 * g1 - F1 * g3, axis=0) + 2 * F2 * np.cross(g3, g1, axis=0)
            t2 = np.cross(g3, F2 * g1 - F1 * g2, axis=0) + 2 * F3 * np.cross(g2, g1, axis=0)
            acc += (-Et[it + 6] * L1 * t1 + Et[it + 16] * L2 * t2) / V1
        out[:, inside] = acc * c[itet]
        found[inside] = True
    out[:, ~found] = np.nan
    return out


def surface_integral(nodes, tri_vertices, function, constants=None, gq_order=4):
    """E/mth/integrals.py:28-92."""
    if constants is None:
        constants = np.ones(tri_vertices.shape[1])
    DPTs = robin.gaus_quad_tri(gq_order)
    n = tri_vertices.shape[1]
    xall = np.zeros((DPTs.shape[1], n))
    yall = np.zeros_like(xall)
    zall = np.zeros_like(xall)
    for it in range(n):
        vid = tri_vertices[:, it]
        x1, x2, x3 = nodes[0, vid]
        y1, y2, y3 = nodes[1, vid]
        z1, z2, z3 = nodes[2, vid]
        xall[:, it] = x1*DPTs[1] + x2*DPTs[2] + x3*DPTs[3]
        yall[:, it] = y1*DPTs[1] + y2*DPTs[2] + y3*DPTs[3]
        zall[:, it] = z1*DPTs[1] + z2*DPTs[2] + z3*DPTs[3]
    f = np.asarray(function(xall.flatten(), yall.flatten(), zall.flatten())).reshape(xall.shape)
    tot = 0j
    for it in range(n):
        vid = tri_vertices[:, it]
        v1, v2, v3 = nodes[:, vid[0]], nodes[:, vid[1]], nodes[:, vid[2]]
        av = np.cross(v2 - v1, v3 - v1)
        A = np.sqrt(av[0]**2 + av[1]**2 + av[2]**2) / 2
        tot = tot + constants[it] * np.sum(DPTs[0]*f[:, it]) * A
    return tot


def gauss3_composite(x, y):
    """E/coord.py:24-61: per section (y0, y1, y2) the quadratic through the
    samples evaluated at the 3 Gauss nodes of [-1, 1] and weighted."""
    x = np.asarray(x, dtype=np.complex128)
    y = np.asarray(y, dtype=np.complex128)
    if x.size % 2 == 0:
        raise ValueError("Number of samples must be odd")
    xi = np.sqrt(3 / 5)
    nodes = np.array([-xi, 0.0, +xi])
    weights = np.array([5 / 9, 8 / 9, 5 / 9])
    total = 0.0
    for i in range(0, x.size - 2, 2):
        y0, y1, y2 = y[i:i + 3]
        a = y0 * 0.5 - y1 + 0.5 * y2
        b = -y0 * 0.5 + 0.5 * y2
        c = y1
        total += np.dot(weights, a * nodes ** 2 + b * nodes + c)
    return total


def voltage_lines(mesh, port, npts=21):
    """MW/microwave_3d.py:336-372 (Line.from_points, E/coord.py:100-117): one
    line per port-face node at the minimum coordinate along the port
    direction, from that node over the port height."""
    points = np.unique(mesh.tris[:, port.tri_ids].ravel())
    xs, ys, zs = mesh.nodes[0, points], mesh.nodes[1, points], mesh.nodes[2, points]
    ax = port.direction
    dot = xs * ax[0] + ys * ax[1] + zs * ax[2]
    start_id = np.argwhere(dot == np.min(dot)).flatten()
    lines = []
    for x, y, z in zip(xs[start_id], ys[start_id], zs[start_id]):
        start = np.array([x, y, z])
        end = start + ax * port.height
        lines.append((np.linspace(start[0], end[0], npts), np.linspace(start[1], end[1], npts),
                      np.linspace(start[2], end[2], npts)))
    return lines


def line_integral(line, fieldf):
    """Line.line_integral (E/coord.py:63-84, 149-153)."""
    xs, ys, zs = line
    dxs, dys, dzs = xs[1:] - xs[:-1], ys[1:] - ys[:-1], zs[1:] - zs[:-1]
    dl = np.sqrt(dxs ** 2 + dys ** 2 + dzs ** 2)
    ll = np.concatenate((np.array([0.0]), np.cumsum(dl)))
    Ex, Ey, Ez = fieldf(xs, ys, zs)
    return gauss3_composite(ll, Ex * dxs[0] + Ey * dys[0] + Ez * dzs[0])


def s_parameters(mesh, tet_to_field, fields: dict, ports: list, k0: float,
                 cap_single=True, cap_col=True, recip=False) -> np.ndarray:
    """MW/microwave_3d.py:1167-1235 with _compute_s_data (:1250-1303):
    waveguide (mode-power) ports and voltage-line (lumped) ports; mu_r = 1 on
    the port faces."""
    numbers = [p.port_number for p in ports]
    Sp = np.zeros((len(ports), len(ports)), dtype=np.complex128)
    for ja, act in enumerate(ports):
        act.active = True
        sol = fields[act.port_number]
        for ip, bc in enumerate(ports):
            port_tets = mesh.get_face_tets(bc.tri_ids)
            tv = mesh.tris[:, bc.tri_ids]

            def fieldf(x, y, z):
                return np.array(ned2_tet_interp(np.array([x, y, z]), sol, mesh, tet_to_field, port_tets))

            if getattr(bc, "v_integration", False):
                V = sum(line_integral(ln, fieldf) for ln in voltage_lines(mesh, bc)) / len(voltage_lines(mesh, bc))
                a = bc.voltage
                b = V - bc.voltage if bc.active else V
                a_sig = a * np.sqrt(1 / (2 * bc.Z0) + 0j)
                b_sig = b * np.sqrt(1 / (2 * bc.Z0) + 0j)
                Sp[ip, ja] = (b_sig / a_sig) / np.sqrt(act.power)
                continue
            Q = 1 if bc.active else 0
            Zm = bc.Zmode(k0)
            const = np.ones(tv.shape[1])

            def inprod1(x, y, z):
                E = fieldf(x, y, z) - Q * bc.port_mode_3d_global(x, y, z, k0)
                m = np.conj(bc.port_mode_3d_global(x, y, z, k0))
                return (E[0]*m[0] + E[1]*m[1] + E[2]*m[2]) / (2 * Zm)

            def inprod2(x, y, z):
                e = bc.port_mode_3d_global(x, y, z, k0)
                m = np.conj(e)
                return (e[0]*m[0] + e[1]*m[1] + e[2]*m[2]) / (2 * Zm)

            fp = surface_integral(mesh.nodes, tv, inprod1, const, 4)
            mp = surface_integral(mesh.nodes, tv, inprod2, const, 4)
            Sp[ip, ja] = (fp / mp) / np.sqrt(act.power)
        act.active = False
    if recip:
        Sp = (Sp + Sp.T) / 2
    if cap_col:
        for j in range(len(ports)):
            Sp[:, j] = Sp[:, j] / max(1.0, np.sum(np.abs(Sp[:, j])**2))
    if cap_single:
        for i in range(len(ports)):
            for j in range(len(ports)):
                Sp[i, j] = Sp[i, j] / max(1.0, np.abs(Sp[i, j]))
    return Sp
