"""Reduced patch-antenna case (the one-million-tet patch workload at coarse
resolution), shared by the GPU test and the fixture script.

Slab 60 x 60 x 1.52 mm, eps_r = 3.38 (1 - 0.0027 j), PEC ground under the
slab and PEC 30 x 40 mm patch on top, inset lumped port (ground to patch,
field along +z, 50 ohm, 1 W) in the plane y = 0, air box 15 mm beyond the
slab closed by a first-order absorbing boundary (gamma = j k0).  Graded
Kuhn grid: 5 mm over the slab, one cell across the substrate, spacing
doubling to 15 mm in the air (13,824 tets, 93,074 DOFs).
"""
from __future__ import annotations

import functools

import numpy as np

from emerge_amd import workloads as w
from oracle import curlcurl as cc
from oracle import fem
from oracle import ports as oports

KW = dict(h_slab=5e-3, n_sub=1, h_air=15e-3, ratio=2.0, air_gap=15e-3, inset=5e-3)
FREQS = (2.2e9, 2.6e9)


@functools.lru_cache(maxsize=1)
def case():
    p = w.patch_antenna(**KW)
    bcs = w.patch_bcs(p)
    m, f = p.mesh, p.basis
    pec, abc, port = bcs[0], bcs[1], bcs[2]
    o_bcs = [oports.AbsorbingBoundary(abc.tri_ids, order=1),
             oports.LumpedPort(port.tri_ids, 1, port.width, port.height, port.direction)]
    return dict(problem=p, mesh=m, basis=f, prod_bcs=bcs, pec=pec.tri_ids, o_bcs=o_bcs,
                maps=dict(edge_to_field=f.edge_to_field, tri_to_field=f.tri_to_field))


@functools.lru_cache(maxsize=1)
def oracle_matrices():
    c = case()
    p = c["problem"]
    return cc.tet_mass_stiffness_matrices(c["mesh"], c["basis"].tet_to_field, cc.isotropic(p.eps_r),
                                          cc.isotropic(1.0 / p.inv_mu_r))


def oracle_system(freq: float):
    c = case()
    E, B = oracle_matrices()
    return fem.assemble_freq_matrix(c["mesh"], c["maps"], E, B, c["pec"], c["o_bcs"], freq)


def oracle_S(freq: float):
    c = case()
    K, b, pv, ids = oracle_system(freq)
    fields = fem.solve_all_ports(K, b, pv, ids)
    k0 = 2 * np.pi * freq / fem.C0
    return fem.s_parameters(c["mesh"], c["basis"].tet_to_field, fields, [c["o_bcs"][1]], k0)
