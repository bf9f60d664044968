"""E and H at points from a solved WR-90 field (emg_field_points /
FrequencySweep.interpolate) against the oracle's interpolation with the
reference's owner rule (last tet containing the point).

Bar: E and H within 1e-11 of the oracle, relative to their largest value
(FP64; the two sum the 20 basis terms in different orders, measured 6e-13);
NaN outside the mesh."""
import numpy as np
import pytest

from emerge_amd.sweep import EPS0, C0, FrequencySweep
from emerge_amd.workloads import WR90_A, WR90_B
from oracle import fem
from tests import wg_case

pytestmark = pytest.mark.gpu


def test_fields_at_points_match_oracle():
    c = wg_case.case()
    m, f, p = c["mesh"], c["basis"], c["problem"]
    sw = FrequencySweep(m, f, p.eps_r, p.inv_mu_r, c["prod_bcs"])
    freq = 10e9
    res = sw.solve([freq], precision="fp64", tol=1e-10)
    x = res.fields[(0, 1)]
    rng = np.random.default_rng(5)
    inside = rng.random((3, 300)) * np.array([[WR90_A], [WR90_B], [wg_case.L]])
    shared = (m.nodes[:, m.tris[0]] + m.nodes[:, m.tris[1]] + m.nodes[:, m.tris[2]])[:, ::97] / 3  # faces, edges of tets
    outside = np.array([[-1e-3, 0.5 * WR90_A], [0.5 * WR90_B, 0.5 * WR90_B], [0.01, wg_case.L + 1e-3]])
    P = np.hstack([inside, shared, outside])
    print(f"points={P.shape[1]}", flush=True)
    E, H = sw.interpolate(x, *P, freq=freq)
    allt = np.arange(m.n_tets)
    Eo = np.array(fem.ned2_tet_interp(P, x, m, f.tet_to_field, allt))
    mu0 = 1 / (C0 * C0 * EPS0)
    cst = 1 / (-1j * 2 * np.pi * freq * mu0 * (1 / p.inv_mu_r))
    Ho = fem.ned2_tet_interp_curl(P, x, m, f.tet_to_field, allt, cst)
    n_in = P.shape[1] - 2
    for a, b in ((E, Eo), (H, Ho)):
        assert np.all(np.isnan(a[:, n_in:])) and np.all(np.isnan(b[:, n_in:]))
        scale = np.abs(b[:, :n_in]).max()
        err = np.abs(a[:, :n_in] - b[:, :n_in]).max() / scale
        print(f"rel err {err:.2e}", flush=True)
        assert err <= 1e-11


def test_point_grid_follows_a_new_mesh():
    """The point grid is built per mesh: after emg_mesh_set with a shifted
    copy of the mesh on the same context, owners and fields follow it."""
    from emerge_amd._lib import Context, ptr
    c = wg_case.case()
    m, f = c["mesh"], c["basis"]
    ctx = Context(0)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(f.n_field) + 1j * rng.standard_normal(f.n_field)
    pts = np.ascontiguousarray(rng.random((3, 50)) * np.array([[WR90_A], [WR90_B], [wg_case.L]]))
    out = []
    for shift in (0.0, 0.05):
        ctx.mesh_set(m.nodes + np.array([[0.0], [0.0], [shift]]), m.tets, f.tet_to_field, f.n_field)
        E = np.empty((3, 50), dtype=np.complex128)
        C = np.empty((3, 50), dtype=np.complex128)
        own = np.empty(50, dtype=np.int32)
        q = np.ascontiguousarray(pts + np.array([[0.0], [0.0], [shift]]))
        ctx.call("emg_field_points", ptr(x), 50, ptr(q), ptr(E), ptr(C), ptr(own))
        out.append((E, C, own))
    np.testing.assert_array_equal(out[0][2], out[1][2])
    assert np.all(out[0][2] >= 0)
    for a, b in ((out[0][0], out[1][0]), (out[0][1], out[1][1])):
        assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max()
