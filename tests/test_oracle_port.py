"""The C++/OpenMP port of the per-tet kernel (the timed CPU baseline)
against the NumPy oracle, entry by entry, including anisotropic tensors."""
import numpy as np
import pytest

from emerge_amd.basis import Nedelec2
from emerge_amd.mesh import Mesh3D, kuhn_box
from oracle import curlcurl as cc
from oracle import port


@pytest.fixture(scope="module", autouse=True)
def built():
    port.build()


@pytest.mark.parametrize("aniso", [False, True])
def test_port_matches_numpy_oracle(aniso):
    m = Mesh3D(*kuhn_box(4, 3, 3, size=(0.02, 0.015, 0.01), jitter=0.15, seed=2))
    f = Nedelec2(m)
    T = m.n_tets
    rng = np.random.default_rng(4)
    er = cc.isotropic(rng.uniform(1, 10, T) * (1 - 0.01j))
    ur = cc.isotropic(np.full(T, 1.2 - 0.1j))
    if aniso:
        er = er + 0.2 * (rng.standard_normal((3, 3, T)) + 0.1j * rng.standard_normal((3, 3, T)))
        ur = ur + 0.1 * rng.standard_normal((3, 3, T))
    args = (m.nodes, m.tets, m.tris, m.edges, m.edge_lengths, f.tet_to_field, m.tet_to_edge, ur, er)
    dE, dB, r, c = cc.matrix_builder(*args)
    pE, pB, pr, pc, nth = port.matrix_builder(*args)
    assert nth >= 1
    assert np.array_equal(r, pr) and np.array_equal(c, pc)
    assert np.abs(pE - dE).max() <= 1e-12 * np.abs(dE).max()
    assert np.abs(pB - dB).max() <= 1e-12 * np.abs(dB).max()
