"""Parallel-plate TEM line between two lumped ports (shared by the S-parameter
pin tests and the GPU parity test).

Box L x w x h (x along the line), PEC plates on z = 0 and z = h, the natural
boundary condition (a magnetic wall) on y = 0 and y = w, so the line carries
an exact TEM wave with E along z and characteristic impedance
Z_c = Z0 h / w.  LumpedPorts (field along +z, width w, height h) on the end
faces x = 0 and x = L, or a PEC short on x = L.  The port sheets see a
uniform field, so the S-matrix of the line between two Z_port references has
the transmission-line closed form (ABCD of a line of length L, beta = k0).
"""
from __future__ import annotations

import functools

import numpy as np

from emerge_amd import bc as pbc
from emerge_amd.basis import Nedelec2
from emerge_amd.mesh import Mesh3D, kuhn_box
from oracle import curlcurl as cc
from oracle import fem
from oracle import ports as oports

ETA0 = 376.73031366857


@functools.lru_cache(maxsize=8)
def case(L=20e-3, w=2e-3, h=2e-3, cells=(16, 2, 2), zport=50.0, powers=(1.0, 1.0), short=False):
    nodes, tets = kuhn_box(*cells, size=(L, w, h))
    m = Mesh3D(nodes, tets)
    f = Nedelec2(m)
    one = np.ones(m.n_tets, dtype=np.complex128)
    E, B = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, cc.isotropic(one), cc.isotropic(one))
    pec = np.union1d(m.plane_tris(2, 0.0), m.plane_tris(2, h))
    p1, p2 = m.plane_tris(0, 0.0), m.plane_tris(0, L)
    o_ports = [oports.LumpedPort(p1, 1, w, h, (0, 0, 1), power=powers[0], Z0_port=zport)]
    prod = [pbc.LumpedPort(p1, 1, w, h, (0, 0, 1), power=powers[0], Z0_port=zport)]
    if short:
        pec = np.union1d(pec, p2)
    else:
        o_ports.append(oports.LumpedPort(p2, 2, w, h, (0, 0, 1), power=powers[1], Z0_port=zport))
        prod.append(pbc.LumpedPort(p2, 2, w, h, (0, 0, 1), power=powers[1], Z0_port=zport))
    return dict(mesh=m, basis=f, E=E, B=B, pec=pec, o_ports=o_ports, prod_bcs=[pbc.PEC(pec)] + prod,
                eps_r=one, inv_mu_r=one, L=L, w=w, h=h, zport=zport,
                maps=dict(edge_to_field=f.edge_to_field, tri_to_field=f.tri_to_field))


def oracle_solution(freq, caps=False, **kw):
    c = case(**kw)
    K, b, pv, ids = fem.assemble_freq_matrix(c["mesh"], c["maps"], c["E"], c["B"], c["pec"], c["o_ports"], freq)
    fields = fem.solve_all_ports(K, b, pv, ids)
    k0 = 2 * np.pi * freq / fem.C0
    S = fem.s_parameters(c["mesh"], c["basis"].tet_to_field, fields, c["o_ports"], k0,
                         cap_single=caps, cap_col=caps)
    return dict(K=K, b=b, pv=pv, ids=ids, fields=fields, S=S, k0=k0)


def line_closed_form(k0, L, zc, zref):
    """2x2 S-matrix of a lossless line (impedance zc, beta = k0, length L)
    between two zref terminations, from its ABCD matrix."""
    A = np.cos(k0 * L)
    B = 1j * zc * np.sin(k0 * L)
    C = 1j * np.sin(k0 * L) / zc
    D = A
    den = A + B / zref + C * zref + D
    return np.array([[(A + B / zref - C * zref - D) / den, 2 / den],
                     [2 / den, (-A + B / zref - C * zref + D) / den]])
