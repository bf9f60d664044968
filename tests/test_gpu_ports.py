"""Port excitation and S extraction on the device against the oracle, for the
cases whose oracle values are pinned by closed forms
(tests/test_oracle_sparam.py): lumped ports on a parallel-plate TEM line,
a shorted line, and port powers other than 1 W for both port kinds.

Bar: FP64 relative residual <= 1e-8 and |dS| <= 1e-6 against the oracle
(SuperLU, reference options).
"""
import numpy as np
import pytest

from emerge_amd.sweep import FrequencySweep
from oracle import fem
from tests import tem_case

pytestmark = pytest.mark.gpu


def _sweep_S(c, freqs, precision):
    sw = FrequencySweep(c["mesh"], c["basis"], c["eps_r"], c["inv_mu_r"], c["prod_bcs"])
    res = sw.solve(freqs, precision=precision, tol=1e-8, fblock=4, keep_fields=False)
    assert np.all(res.relres <= 1e-8), res.relres
    return res


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("powers", [(1.0, 1.0), (4.0, 0.25)])
def test_tem_line_lumped_ports(precision, powers):
    print(f"tem line precision={precision} powers={powers}", flush=True)
    c = tem_case.case(powers=powers)
    freqs = np.array([5e9, 10e9])
    res = _sweep_S(c, freqs, precision)
    for i, fr in enumerate(freqs):
        o = tem_case.oracle_solution(float(fr), caps=True, powers=powers)
        assert np.abs(res.S[i] - o["S"]).max() <= 1e-6, (fr, np.abs(res.S[i] - o["S"]).max())


def test_shorted_line():
    c = tem_case.case(w=7.534e-3, h=1e-3, short=True)
    res = _sweep_S(c, np.array([10e9]), "fp64")
    o = tem_case.oracle_solution(10e9, caps=True, w=7.534e-3, h=1e-3, short=True)
    assert np.abs(res.S[0] - o["S"]).max() <= 1e-6


def test_waveguide_port_power():
    """TE10 ports with powers (4, 0.25): device S against the oracle."""
    from emerge_amd import bc as pbc
    from emerge_amd.workloads import WR90_A, WR90_B, waveguide, waveguide_bcs
    from oracle import curlcurl as cc
    from oracle import ports as oports
    p = waveguide(cells=(6, 3, 6), length=15e-3)
    bcs = waveguide_bcs(p)
    powers = (4.0, 0.25)
    zs = (0.0, 15e-3)
    prod = [bcs[0]] + [pbc.RectangularWaveguide(bcs[k + 1].tri_ids, k + 1,
                                                pbc.CoordinateSystem([1, 0, 0], [0, 1, 0], [0, 0, 1],
                                                                     [WR90_A / 2, WR90_B / 2, zs[k]]),
                                                (WR90_A, WR90_B), power=powers[k]) for k in range(2)]
    o_ports = [oports.RectangularWaveguide(bcs[k + 1].tri_ids, k + 1,
                                           oports.CS([1, 0, 0], [0, 1, 0], [0, 0, 1], [WR90_A / 2, WR90_B / 2, zs[k]]),
                                           (WR90_A, WR90_B), power=powers[k]) for k in range(2)]
    m, f = p.mesh, p.basis
    sw = FrequencySweep(m, f, p.eps_r, p.inv_mu_r, prod)
    fr = 10e9
    res = sw.solve([fr], precision="fp64", tol=1e-8, fblock=2, keep_fields=False)
    assert np.all(res.relres <= 1e-8)
    E, B = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, cc.isotropic(p.eps_r), cc.isotropic(1 / p.inv_mu_r))
    maps = dict(edge_to_field=f.edge_to_field, tri_to_field=f.tri_to_field)
    K, b, pv, ids = fem.assemble_freq_matrix(m, maps, E, B, bcs[0].tri_ids, o_ports, fr)
    So = fem.s_parameters(m, f.tet_to_field, fem.solve_all_ports(K, b, pv, ids), o_ports, 2 * np.pi * fr / fem.C0)
    assert np.abs(res.S[0] - So).max() <= 1e-6, np.abs(res.S[0] - So).max()
