"""Tensor and frequency-dependent materials in the sweep, against the oracle.

WR-90 guide (11 x 5 x 18 Kuhn cells) with a block of material between
z = 15 and 25 mm.  Oracle side: E and B from the oracle's assembly with the
same (3, 3, T) tensors (re-assembled per frequency when the material depends
on it, ASM/assembler.py:222-252), then SuperLU with the reference's options
and the oracle's S extraction.

Bar: FP64 relative residual <= 1e-8 on the oracle's matrix and |dS| <= 1e-6.
"""
import functools

import numpy as np
import pytest

from emerge_amd import bc as pbc
from emerge_amd._lib import Context, EmgError
from emerge_amd.materials import pml_layer_tensors, tet_centers
from emerge_amd.sweep import FrequencySweep, lossy_permittivity
from emerge_amd.workloads import WR90_A, WR90_B, waveguide
from oracle import curlcurl as cc
from oracle import fem
from oracle import ports as oports
from tests import wg_case

pytestmark = pytest.mark.gpu

FREQS = np.array([8.5e9, 10e9, 11.5e9])


def _block(m):
    zc = m.nodes[2, m.tets].mean(axis=0)
    return (zc > 15e-3) & (zc < 25e-3)


def _tensor_materials(T, blk):
    er = np.zeros((3, 3, T), dtype=np.complex128)
    ur = np.zeros((3, 3, T), dtype=np.complex128)
    for k in range(3):
        er[k, k] = 1.0
        ur[k, k] = 1.0
    er[0, 0, blk] = 2.2 * (1 - 0.01j)
    er[1, 1, blk] = 3.0 * (1 - 0.02j)
    er[2, 2, blk] = 2.6
    er[0, 1, blk] = er[1, 0, blk] = 0.3
    ur[1, 1, blk] = 1.2
    ur[2, 2, blk] = 1.1 - 0.01j
    return er, ur


def _lossy_eps(T, blk, f):
    eps0 = np.where(blk, 2.5, 1.0).astype(np.complex128)
    return lossy_permittivity(eps0, np.where(blk, 0.005, 0.0), np.where(blk, 0.2, 0.0), f)


@functools.lru_cache(maxsize=2)
def _oracle_EB(kind: str, freq: float):
    c = wg_case.case()
    m = c["mesh"]
    T = m.n_tets
    if kind == "tensor":
        er, ur = _tensor_materials(T, _block(m))
    else:
        er, ur = cc.isotropic(_lossy_eps(T, _block(m), freq)), cc.isotropic(np.ones(T, dtype=np.complex128))
    return cc.tet_mass_stiffness_matrices(m, c["basis"].tet_to_field, er, ur)


@functools.lru_cache(maxsize=8)
def _oracle(kind: str, freq: float):
    """(K, b, port vectors, ids, S) of the oracle at `freq`; the tensor
    material does not depend on the frequency (one assembly)."""
    c = wg_case.case()
    E, B = _oracle_EB(kind, 0.0 if kind == "tensor" else freq)
    K, b, pv, ids = fem.assemble_freq_matrix(c["mesh"], c["maps"], E, B, c["pec"], c["o_ports"], freq)
    fields = fem.solve_all_ports(K, b, pv, ids)
    S = fem.s_parameters(c["mesh"], c["basis"].tet_to_field, fields, c["o_ports"], 2 * np.pi * freq / fem.C0)
    return K.tocsr(), b, pv, ids, S


def _check(sw, res, kind):
    assert np.all(res.relres <= 1e-8), res.relres
    for i, f in enumerate(FREQS):
        K, b, pv, ids, S = _oracle(kind, float(f))
        for p in sw.ports:
            x = res.fields[(i, p.port_number)]
            bb = b + pv[p.port_number]
            rr = np.linalg.norm((bb - K @ x)[ids]) / np.linalg.norm(bb[ids])
            assert rr <= 1e-8, (f, p.port_number, rr)
        dS = np.abs(res.S[i] - S).max()
        print(f"f={f:.3e} dS={dS:.2e} iters={res.iters[i].tolist()}", flush=True)
        assert dS <= 1e-6, (f, dS)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_anisotropic_tensor_block_matches_superlu(precision):
    print(f"tensor block precision={precision}", flush=True)
    c = wg_case.case()
    m = c["mesh"]
    er, ur = _tensor_materials(m.n_tets, _block(m))
    sw = FrequencySweep(m, c["basis"], er, None, c["prod_bcs"], mu_r=ur)
    assert not sw.frequency_dependent
    res = sw.solve(FREQS, precision=precision, tol=1e-8, fblock=8)
    _check(sw, res, "tensor")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_conductive_block_reassembled_per_frequency(precision):
    print(f"conductive block precision={precision}", flush=True)
    c = wg_case.case()
    m = c["mesh"]
    T = m.n_tets
    blk = _block(m)
    eps_of = lambda f: _lossy_eps(T, blk, f)  # noqa: E731
    sw = FrequencySweep(m, c["basis"], eps_of, np.ones(T, dtype=np.complex128), c["prod_bcs"])
    assert sw.frequency_dependent
    res = sw.solve(FREQS, precision=precision, tol=1e-8)
    _check(sw, res, "lossy")
    # a static sweep on the materials of the last frequency gives the same S
    st = FrequencySweep(m, c["basis"], eps_of(FREQS[-1]), np.ones(T, dtype=np.complex128), c["prod_bcs"])
    r2 = st.solve(FREQS[-1:], precision="fp64", tol=1e-8)
    assert np.abs(r2.S[0] - res.S[-1]).max() <= 1e-6


@functools.lru_cache(maxsize=1)
def _pml_case():
    """WR-90, 40 mm on 11 x 5 x 16 Kuhn cells (layer boundary on a grid
    plane), TE10 port on z = 0, a 15 mm PML (pmlbox defaults: exponent 1.5,
    deltamax 8) beyond z = 25 mm, PEC on every other exterior triangle."""
    L, Lm, th = 40e-3, 25e-3, 15e-3
    p = waveguide(cells=(11, 5, 16), length=L)
    m, f = p.mesh, p.basis
    T = m.n_tets
    pml = np.nonzero(tet_centers(m)[2] > Lm)[0]
    er = np.zeros((3, 3, T), dtype=np.complex128)
    ur = np.zeros((3, 3, T), dtype=np.complex128)
    for k in range(3):
        er[k, k] = 1.0
        ur[k, k] = 1.0
    e_p, u_p = pml_layer_tensors(m, pml, (WR90_A / 2, WR90_B / 2, Lm / 2), (WR90_A, WR90_B, Lm), (0, 0, 1), th)
    er[:, :, pml] = e_p
    ur[:, :, pml] = u_p
    t1 = m.plane_tris(2, 0.0)
    pec = np.setdiff1d(m.boundary_tris(), t1)
    cs = ([1, 0, 0], [0, 1, 0], [0, 0, 1], [WR90_A / 2, WR90_B / 2, 0.0])
    prod_bcs = [pbc.PEC(pec), pbc.RectangularWaveguide(t1, 1, pbc.CoordinateSystem(*cs), (WR90_A, WR90_B))]
    o_ports = [oports.RectangularWaveguide(t1, 1, oports.CS(*cs), (WR90_A, WR90_B))]
    E, B = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, er, ur)
    maps = dict(edge_to_field=f.edge_to_field, tri_to_field=f.tri_to_field)
    S = {}
    for fq in FREQS:
        K, b, pv, ids = fem.assemble_freq_matrix(m, maps, E, B, pec, o_ports, float(fq))
        fields = fem.solve_all_ports(K, b, pv, ids)
        S[float(fq)] = fem.s_parameters(m, f.tet_to_field, fields, o_ports, 2 * np.pi * fq / fem.C0)
    return dict(mesh=m, basis=f, er=er, ur=ur, prod_bcs=prod_bcs, S=S)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pml_terminated_guide_matches_superlu(precision):
    """The guide ends in a PML instead of a port: S11 equals the oracle's
    (tensors through the oracle's assembly, SuperLU) and stays below 0.02
    (the layer absorbs; an inverted stretch tensor reflects |S11| = 0.27 on
    this mesh)."""
    print(f"pml guide precision={precision}", flush=True)
    c = _pml_case()
    sw = FrequencySweep(c["mesh"], c["basis"], c["er"], None, c["prod_bcs"], mu_r=c["ur"])
    res = sw.solve(FREQS, precision=precision, tol=1e-8, fblock=4)
    assert np.all(res.relres <= 1e-8), res.relres
    for i, fq in enumerate(FREQS):
        S = c["S"][float(fq)]
        dS = np.abs(res.S[i] - S).max()
        print(f"f={fq:.3e} |S11|={abs(res.S[i][0, 0]):.2e} dS={dS:.2e} iters={res.iters[i].tolist()}", flush=True)
        assert dS <= 1e-6, (fq, dS)
        assert abs(res.S[i][0, 0]) < 0.02


def test_refresh_needs_current_reduction():
    """emg_system_refresh after a new pattern is a state error, not a gather
    through the old reduced map."""
    c = wg_case.case()
    m, f = c["mesh"], c["basis"]
    ctx = Context(0)
    ctx.mesh_set(m.nodes, m.tets, f.tet_to_field, f.n_field)
    ctx.pattern_build()
    ctx.material_set(np.ones(m.n_tets, dtype=np.complex128), np.ones(m.n_tets, dtype=np.complex128))
    ctx.assemble()
    ids = np.ascontiguousarray(np.arange(f.n_field), dtype=np.int32)
    from emerge_amd._lib import ptr
    ctx.call("emg_system_reduce", ids.shape[0], ptr(ids))
    ctx.call("emg_system_refresh")
    ctx.pattern_build()
    with pytest.raises(EmgError):
        ctx.call("emg_system_refresh")


def _metal_case():
    """Copper post (cond 5.8e7 S/m > 1e7: a perfect conductor) across the
    guide's narrow side at x in (7.5, 10.4) mm, z in (10, 14.5) mm, and the
    lossy block (cond 0.2 S/m) of the other tests behind it."""
    c = wg_case.case()
    m = c["mesh"]
    T = m.n_tets
    cen = tet_centers(m)
    post = (cen[0] > 7.5e-3) & (cen[0] < 10.4e-3) & (cen[2] > 10e-3) & (cen[2] < 14.5e-3)
    blk = _block(m)
    cond = np.where(post, 5.8e7, np.where(blk, 0.2, 0.0))
    eps = np.where(blk, 2.5 * (1 - 0.005j), 1.0 + 0j).astype(np.complex128)
    return c, m, T, post, cond, eps


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_perfect_conductor_tets_and_lossy_block(precision):
    """Tets above the conductivity limit lose all 20 DOFs, the lossy block
    makes the sweep frequency-dependent (ASM/assembler.py:240-288), against
    the oracle assembled per frequency with eps_r - j cond / (w eps0) on
    every tet and the conductor elimination, + SuperLU."""
    print(f"metal post precision={precision}", flush=True)
    c, m, T, post, cond, eps = _metal_case()
    assert post.sum() >= 6
    sw = FrequencySweep(m, c["basis"], eps, np.ones(T, dtype=np.complex128), c["prod_bcs"], cond=cond)
    assert sw.frequency_dependent and np.array_equal(sw.conductor_tets, np.nonzero(post)[0])
    res = sw.solve(FREQS, precision=precision, tol=1e-8)
    assert np.all(res.relres <= 1e-8), res.relres
    cond3 = cc.isotropic(cond.astype(np.complex128))
    for i, f in enumerate(FREQS):
        er = lossy_permittivity(eps, 0.0, cond, float(f))
        E, B = cc.tet_mass_stiffness_matrices(m, c["basis"].tet_to_field, cc.isotropic(er),
                                              cc.isotropic(np.ones(T, dtype=np.complex128)))
        K, b, pv, ids = fem.assemble_freq_matrix(m, c["maps"], E, B, c["pec"], c["o_ports"], float(f), cond=cond3,
                                                 tet_to_field=c["basis"].tet_to_field)
        assert np.array_equal(np.sort(sw.solve_ids), ids)
        fields = fem.solve_all_ports(K, b, pv, ids)
        S = fem.s_parameters(m, c["basis"].tet_to_field, fields, c["o_ports"], 2 * np.pi * f / fem.C0)
        for p in sw.ports:
            x = res.fields[(i, p.port_number)]
            bb = b + pv[p.port_number]
            rr = np.linalg.norm((bb - K.tocsr() @ x)[ids]) / np.linalg.norm(bb[ids])
            assert rr <= 1e-8, (f, p.port_number, rr)
        dS = np.abs(res.S[i] - S).max()
        print(f"f={f:.3e} dS={dS:.2e} |S11|={abs(S[0, 0]):.3f} iters={res.iters[i].tolist()}", flush=True)
        assert dS <= 1e-6, (f, dS)
