"""Pin of the oracle's curl interpolation (oracle.fem.ned2_tet_interp_curl):
central differences of its E interpolation (exact for the quadratic
Nedelec-2 field up to rounding) on random DOFs in a jittered mesh; points
outside every tet give NaN; the constant scales per owner tet."""
import numpy as np

from emerge_amd.basis import Nedelec2
from emerge_amd.mesh import Mesh3D, kuhn_box
from oracle import fem


def test_curl_matches_central_differences():
    nodes, tets = kuhn_box(2, 2, 2, size=(1.0, 0.8, 1.2), jitter=0.2, seed=3)
    m = Mesh3D(nodes, tets)
    f = Nedelec2(m)
    rng = np.random.default_rng(0)
    sol = rng.standard_normal(f.n_field) + 1j * rng.standard_normal(f.n_field)
    ones = np.ones(m.n_tets, dtype=np.complex128)
    for it in (0, 17, m.n_tets - 1):
        P = m.nodes[:, m.tets[:, it]] @ rng.dirichlet(np.ones(4), 5).T
        cu = fem.ned2_tet_interp_curl(P, sol, m, f.tet_to_field, np.array([it]), ones)
        h = 1e-5
        J = []
        for k in range(3):
            d = np.zeros((3, 1))
            d[k] = h
            J.append((np.array(fem.ned2_tet_interp(P + d, sol, m, f.tet_to_field, np.array([it])))
                      - np.array(fem.ned2_tet_interp(P - d, sol, m, f.tet_to_field, np.array([it])))) / (2 * h))
        fd = np.array([J[1][2] - J[2][1], J[2][0] - J[0][2], J[0][1] - J[1][0]])
        assert np.abs(cu - fd).max() <= 1e-9 * np.abs(fd).max()
        c = rng.standard_normal(m.n_tets) + 1j
        cu2 = fem.ned2_tet_interp_curl(P, sol, m, f.tet_to_field, np.array([it]), c)
        np.testing.assert_allclose(cu2, cu * c[it], rtol=1e-13)
    out = fem.ned2_tet_interp_curl(np.array([[5.0], [5.0], [5.0]]), sol, m, f.tet_to_field,
                                   np.arange(m.n_tets), ones)
    assert np.all(np.isnan(out))
