"""Device results against the committed golden fixtures
(tests/golden/golden.npz): element matrices of two tets through the
assembly, S-matrices of a short WR-90 section and S11 of the lumped-port
cavity through the batched solve."""
import os

import numpy as np
import pytest

from emerge_amd.assembler import MeshOnDevice
from emerge_amd.basis import Nedelec2
from emerge_amd.mesh import Mesh3D
from emerge_amd.sweep import FrequencySweep
from emerge_amd.workloads import waveguide, waveguide_bcs

pytestmark = pytest.mark.gpu

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"), allow_pickle=False)


@pytest.mark.parametrize("t", [0, 1])
def test_single_tet_assembly_matches_fixture(hip_ctx, t):
    verts, vid = G["el_verts"][:, :, t], G["el_vid"][:, t]
    rank = np.argsort(np.argsort(vid))              # keep the fixture's global-id order
    nodes = np.zeros((3, 4))
    nodes[:, rank] = verts
    m = Mesh3D(nodes, rank.reshape(4, 1))
    f = Nedelec2(m)
    dm = MeshOnDevice(hip_ctx, m.nodes, m.tets, f.tet_to_field, f.n_field)
    dm.assemble(np.array([G["el_er"][0, 0, t]]), np.array([1.0 / G["el_ur"][0, 0, t]]))
    E, B = dm.csr()
    idx = f.tet_to_field[:, 0]
    Ed = E.toarray()[np.ix_(idx, idx)]
    Bd = B.toarray()[np.ix_(idx, idx)]
    assert np.abs(Ed - G["el_E"][:, :, t]).max() <= 1e-12 * np.abs(G["el_E"][:, :, t]).max()
    assert np.abs(Bd - G["el_B"][:, :, t]).max() <= 1e-12 * np.abs(G["el_B"][:, :, t]).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_waveguide_sweep_matches_fixture(precision):
    p = waveguide(cells=tuple(int(x) for x in G["wr90_cells"]), length=float(G["wr90_length"]))
    sw = FrequencySweep(p.mesh, p.basis, p.eps_r, p.inv_mu_r, waveguide_bcs(p))
    res = sw.solve(G["wr90_freqs"], precision=precision, method="cocg", fblock=8, restart=20000, max_cycles=30,
                   inner_rel32=1e-2, keep_fields=False)
    assert np.all(res.relres <= 1e-8)
    assert np.abs(res.S - G["wr90_S"]).max() <= 1e-6


def test_cavity_sweep_matches_fixture():
    from tests import cavity_case
    c = cavity_case.case()
    sw = FrequencySweep(c["mesh"], c["basis"], c["eps"], c["inv_mu"], c["prod_bcs"])
    res = sw.solve(G["cavity_freqs"], precision="mixed", method="cocg", fblock=2, restart=20000, max_cycles=40,
                   inner_rel32=1e-2, keep_fields=False)
    assert np.all(res.relres <= 1e-8)
    assert np.abs(res.S - G["cavity_S"]).max() <= 1e-6
