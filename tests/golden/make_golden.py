"""Writes tests/golden/golden.npz: small fixtures (inputs and outputs) from
the oracle, for tests that compare against stored values instead of
re-running the oracle.

* element: two tets (the unit corner tet with eps_r = mu_r = 1, and a
  skewed tet with complex eps_r and mu_r), each as a one-tet mesh whose
  vertex ids keep the order `el_vid`: vertices, local edge/face maps
  (ASM/curlcurl.py:36-47), edge lengths and the 20x20 curl-curl / mass
  matrices of oracle.curlcurl.ned2_tet_stiff_mass (ASM/curlcurl.py:180-412)
  in the mesh's local DOF order.
* wr90: a 15 mm WR-90 section (Kuhn 6x3x6, TE10 ports) at 9, 10, 11 GHz:
  the 2x2 S-matrices of the oracle's assembly + SuperLU (reference options)
  + S extraction (MW/microwave_3d.py:1193-1211).
* cavity: the 8x4x4 lumped-port cavity of tests/cavity_case.py at 6 and
  9 GHz: S11 (voltage-line extraction, MW/microwave_3d.py:1271-1292).

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from emerge_amd.workloads import WR90_A, WR90_B, waveguide, waveguide_bcs  # noqa: E402
from oracle import curlcurl as cc  # noqa: E402
from oracle import fem  # noqa: E402
from oracle import ports as oports  # noqa: E402


def single_tet_mesh(verts, vid):
    """Mesh3D of one tet whose vertex ids keep the order of `vid`."""
    from emerge_amd.basis import Nedelec2
    from emerge_amd.mesh import Mesh3D
    rank = np.argsort(np.argsort(vid))
    nodes = np.zeros((3, 4))
    nodes[:, rank] = verts
    m = Mesh3D(nodes, rank.reshape(4, 1))
    return m, Nedelec2(m)


def element_case():
    verts = np.stack([
        np.array([[0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0]]) * 1e-3,
        np.array([[0.1, 1.3, 0.2, 0.4], [0.0, 0.2, 1.1, 0.3], [0.05, 0.1, 0.2, 0.9]]) * 2e-3,
    ], axis=2)                                                   # (3, 4, 2)
    vid = np.array([[0, 1, 2, 3], [7, 3, 11, 5]]).T              # (4, 2) global-id order
    er = cc.isotropic(np.array([1.0 + 0j, 4.2 - 0.08j]))
    ur = cc.isotropic(np.array([1.0 + 0j, 1.3 - 0.05j]))
    lem = np.zeros((2, 6, 2), dtype=np.int64)
    ltm = np.zeros((3, 4, 2), dtype=np.int64)
    L = np.zeros((6, 2))
    X = np.zeros((3, 4, 2))
    for t in range(2):
        m, f = single_tet_mesh(verts[:, :, t], vid[:, t])
        le, lt = cc.local_maps(m.tets, m.edges, m.tris, f.tet_to_field, m.n_edges)
        lem[:, :, t], ltm[:, :, t] = le[:, :, 0], lt[:, :, 0]
        L[:, t] = m.edge_lengths[m.tet_to_edge[:, 0]]
        X[:, :, t] = m.nodes[:, m.tets[:, 0]]
    D, F = cc.ned2_tet_stiff_mass(X, L, lem, ltm, cc.matinv(ur), er)
    return dict(el_verts=verts, el_vid=vid, el_X=X, el_lem=lem, el_ltm=ltm, el_L=L, el_er=er, el_ur=ur,
                el_E=D, el_B=F)


def wr90_case():
    p = waveguide(cells=(6, 3, 6), length=15e-3)
    bcs = waveguide_bcs(p)
    m, f = p.mesh, p.basis
    E, B = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, cc.isotropic(p.eps_r), cc.isotropic(1 / p.inv_mu_r))
    L = p.meta["length"]
    o_ports = [oports.RectangularWaveguide(bcs[1].tri_ids, 1, oports.CS([1, 0, 0], [0, 1, 0], [0, 0, 1],
                                           [WR90_A / 2, WR90_B / 2, 0.0]), (WR90_A, WR90_B)),
               oports.RectangularWaveguide(bcs[2].tri_ids, 2, oports.CS([1, 0, 0], [0, 1, 0], [0, 0, 1],
                                           [WR90_A / 2, WR90_B / 2, L]), (WR90_A, WR90_B))]
    maps = dict(edge_to_field=f.edge_to_field, tri_to_field=f.tri_to_field)
    freqs = np.array([9e9, 10e9, 11e9])
    S = []
    for fr in freqs:
        K, b, pv, ids = fem.assemble_freq_matrix(m, maps, E, B, bcs[0].tri_ids, o_ports, fr)
        S.append(fem.s_parameters(m, f.tet_to_field, fem.solve_all_ports(K, b, pv, ids), o_ports,
                                  2 * np.pi * fr / fem.C0))
    return dict(wr90_cells=np.array([6, 3, 6]), wr90_length=np.array(15e-3), wr90_freqs=freqs, wr90_S=np.array(S))


def cavity_case():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from tests import cavity_case as cvc
    freqs = np.array([6e9, 9e9])
    S = np.array([cvc.oracle_solution(float(fr))["S"] for fr in freqs])
    return dict(cavity_freqs=freqs, cavity_S=S)


if __name__ == "__main__":
    out = {}
    out.update(element_case())
    out.update(wr90_case())
    out.update(cavity_case())
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, {k: v.shape for k, v in out.items()})
