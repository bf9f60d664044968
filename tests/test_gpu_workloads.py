"""The named solve workloads on the device against committed oracle
fixtures (tests/golden/make_sweeps.py: oracle assembly + SuperLU with the
reference's options + the oracle's S extraction):

* the WR-90 guide over all 201 points of the 8-12 GHz grid;
* the 16 jittered WR-90 sections of uneven size (20k-120k DOFs), two points
  each;
* the reduced patch antenna (slab, PEC ground and patch, inset lumped port,
  first-order absorbing boundary), FP64 and mixed precision.

Bar: every system reaches the FP64 relative residual 1e-8 and
|dS| <= 1e-6 against the fixture.
"""
import os

import numpy as np
import pytest

from emerge_amd.sweep import FrequencySweep

pytestmark = pytest.mark.gpu

_FIX = os.path.join(os.path.dirname(__file__), "golden", "sweeps.npz")


class _Fixture:
    """tests/golden/sweeps.npz, loaded on first use."""

    def __getitem__(self, key):
        if not hasattr(self, "_d"):
            self._d = np.load(_FIX, allow_pickle=False)
        return self._d[key]


G = _Fixture()


def test_wr90_all_201_points():
    from tests import wg_case
    c = wg_case.case()
    p = c["problem"]
    sw = FrequencySweep(c["mesh"], c["basis"], p.eps_r, p.inv_mu_r, c["prod_bcs"])
    freqs = G["wr90_freqs"]
    res = sw.solve(freqs, precision="mixed", fblock=32, keep_fields=False)
    assert np.all(res.relres <= 1e-8), res.relres.max()
    dS = np.abs(res.S - G["wr90_S"]).max()
    print(f"wr90 201 points: max |dS| {dS:.2e}", flush=True)
    assert dS <= 1e-6


def test_sixteen_uneven_waveguides():
    from emerge_amd import workloads as w
    probs = w.waveguide_family()
    freqs = G["family_freqs"]
    worst = 0.0
    for i, p in enumerate(probs):
        print(f"family mesh {i}: N={p.basis.n_field}", flush=True)
        sw = FrequencySweep(p.mesh, p.basis, p.eps_r, p.inv_mu_r, w.waveguide_bcs(p))
        res = sw.solve(freqs, precision="mixed", fblock=4, keep_fields=False)
        assert np.all(res.relres <= 1e-8), (i, res.relres.max())
        dS = np.abs(res.S - G["family_S"][i]).max()
        worst = max(worst, dS)
        assert dS <= 1e-6, (i, dS)
        sw.ctx.close()
    assert 20_000 <= min(q.basis.n_field for q in probs) and max(q.basis.n_field for q in probs) <= 125_000
    print(f"family: max |dS| {worst:.2e}", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_patch_antenna_reduced(precision):
    from tests import patch_case
    c = patch_case.case()
    p = c["problem"]
    sw = FrequencySweep(c["mesh"], c["basis"], p.eps_r, p.inv_mu_r, c["prod_bcs"])
    freqs = np.array(G["patch_freqs"])
    print(f"patch precision={precision} N={c['basis'].n_field}", flush=True)
    res = sw.solve(freqs, precision=precision, fblock=2)
    for i, fr in enumerate(freqs):
        K, b, pv, ids = patch_case.oracle_system(float(fr))
        x = res.fields[(i, 1)]
        rhs = b + pv[1]
        rr = np.linalg.norm((rhs - K @ x)[ids]) / np.linalg.norm(rhs[ids])
        assert rr <= 1e-8, (fr, rr)
    dS = np.abs(res.S - G["patch_S"]).max()
    print(f"patch {precision}: |dS| {dS:.2e} iters {res.iters.ravel().tolist()}", flush=True)
    assert dS <= 1e-6


@pytest.mark.parametrize("segmented", [True, False])
def test_sixteen_uneven_waveguides_one_system(segmented):
    """The same 16 meshes as one block-diagonal device system (one launch per
    Krylov kernel for all of them), with per-mesh Krylov scalars (row
    segments) or scalars global over the union: every mesh's S against the
    fixture, the per-mesh residual bound <= 1e-8, and the union's solution
    restricted to a mesh against that mesh's own system (oracle matrices)."""
    from emerge_amd import workloads as w
    from emerge_amd.family import FamilySweep
    from oracle import curlcurl as cc
    from oracle import fem
    from oracle import ports as oports
    probs = w.waveguide_family()
    fam = FamilySweep(probs, [w.waveguide_bcs(p) for p in probs], segmented=segmented)
    if segmented:
        rows = fam.sweep.segment_rows
        assert rows is not None and rows.shape == (len(probs) + 1,) and rows[-1] == fam.sweep.n_red
    freqs = G["family_freqs"]
    print(f"family union (segmented={segmented}): N={fam.basis.n_field}", flush=True)
    S, bound, res = fam.solve(freqs, precision="mixed", fblock=4)
    print(f"family union: iters={res.iters.ravel().tolist()} bound={bound.max():.2e}", flush=True)
    assert np.all(bound <= 1e-8), bound.max()
    worst = max(float(np.abs(S[m] - G["family_S"][m]).max()) for m in range(len(probs)))
    print(f"family union: max |dS| {worst:.2e}", flush=True)
    assert worst <= 1e-6
    # mesh 3 directly: its own oracle system and the union's solution on it
    m = 3
    p = probs[m]
    bc = w.waveguide_bcs(p)
    E, B = cc.tet_mass_stiffness_matrices(p.mesh, p.basis.tet_to_field, cc.isotropic(p.eps_r),
                                          cc.isotropic(1.0 / p.inv_mu_r))
    a, b, L = p.meta["a"], p.meta["b"], p.meta["length"]
    o_ports = [oports.RectangularWaveguide(bc[k + 1].tri_ids, k + 1,
                                           oports.CS([1, 0, 0], [0, 1, 0], [0, 0, 1], [a / 2, b / 2, z]), (a, b))
               for k, z in enumerate((0.0, L))]
    maps = dict(edge_to_field=p.basis.edge_to_field, tri_to_field=p.basis.tri_to_field)
    K, rhs0, pv, ids = fem.assemble_freq_matrix(p.mesh, maps, E, B, bc[0].tri_ids, o_ports, float(freqs[0]))
    for g in range(2):
        x = fam.mesh_field(res.fields[(0, g)], m)
        rhs = rhs0 + pv[g + 1]
        rr = np.linalg.norm((rhs - K @ x)[ids]) / np.linalg.norm(rhs[ids])
        assert rr <= 1e-8, (g, rr)
