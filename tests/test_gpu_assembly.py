"""GPU parity of the volume assembly against the oracle (scipy COO->CSR).

Bar: sparsity pattern (indptr, indices) and scatter map bit-exact; E and B
values within 1e-12 relative to the largest entry of their row; bitwise
reproducible across runs.
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix

from emerge_amd.assembler import HipAssembler, MeshOnDevice
from emerge_amd.basis import Nedelec2
from emerge_amd.mesh import Mesh3D, kuhn_box
from emerge_amd.workloads import cube_assembly
from oracle import curlcurl as cc

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _problem(cells, jitter, seed, size=(1.0, 1.0, 1.0), shuffle=False):
    nodes, tets = kuhn_box(*cells, size=size, jitter=jitter, seed=seed)
    if shuffle:
        rng = np.random.default_rng(seed)
        tets = np.stack([rng.permutation(tets[:, t]) for t in range(tets.shape[1])], axis=1)
    m = Mesh3D(nodes, tets)
    f = Nedelec2(m)
    rng = np.random.default_rng(seed + 11)
    T = m.n_tets
    eps = rng.uniform(1, 10, T) * (1 - 1j * rng.uniform(0, 0.02, T))
    mu = rng.uniform(0.5, 2.0, T) * (1 - 1j * rng.uniform(0, 0.05, T))
    return m, f, eps, mu


def _rowwise_rel_err(A, Bm):
    """max over rows of max|A-B| / max|B| in that row (CSR, same pattern)."""
    d = np.abs(A.data - Bm.data)
    ref = np.abs(Bm.data)
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    rmax = np.zeros(A.shape[0])
    np.maximum.at(rmax, rows, ref)
    return float(np.max(d / np.maximum(rmax[rows], 1e-300)))


@pytest.mark.parametrize("cells,jitter,seed,shuffle", [
    ((2, 2, 1), 0.0, 0, False),
    ((4, 3, 3), 0.15, 1, True),
    ((10, 10, 10), 0.15, 2, False),
])
def test_assembly_matches_oracle(hip_ctx, cells, jitter, seed, shuffle):
    print(f"case cells={cells} jitter={jitter} seed={seed} shuffle={shuffle}", flush=True)
    m, f, eps, mu = _problem(cells, jitter, seed, shuffle=shuffle)
    dm = MeshOnDevice(hip_ctx, m.nodes, m.tets, f.tet_to_field, f.n_field)
    dm.assemble(eps, 1.0 / mu)
    E, B = dm.csr()
    Eo, Bo = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, cc.isotropic(eps), cc.isotropic(mu))
    assert Eo.indices.dtype == np.int32 and Eo.indptr.dtype == np.int32
    np.testing.assert_array_equal(E.indptr, Eo.indptr)
    np.testing.assert_array_equal(E.indices, Eo.indices)
    np.testing.assert_array_equal(B.indptr, Bo.indptr)
    np.testing.assert_array_equal(B.indices, Bo.indices)
    assert _rowwise_rel_err(E, Eo) <= TOL
    assert _rowwise_rel_err(B, Bo) <= TOL


def test_scatter_map_is_coo_to_csr(hip_ctx):
    m, f, eps, mu = _problem((3, 2, 2), 0.1, 4)
    dm = MeshOnDevice(hip_ctx, m.nodes, m.tets, f.tet_to_field, f.n_field)
    cmap = hip_ctx.scatter_map_get()
    indptr, indices = dm.pattern
    t2f = f.tet_to_field
    rows = np.repeat(t2f.T, 20, axis=1).reshape(-1)
    cols = np.tile(t2f.T, (1, 20)).reshape(-1)
    # slot s of COO entry k must hold (rows[k], cols[k])
    srow = np.repeat(np.arange(f.n_field), np.diff(indptr))
    np.testing.assert_array_equal(srow[cmap], rows)
    np.testing.assert_array_equal(indices[cmap], cols)


def test_assembly_bitwise_reproducible(hip_ctx):
    m, f, eps, mu = _problem((6, 5, 4), 0.15, 9)
    dm = MeshOnDevice(hip_ctx, m.nodes, m.tets, f.tet_to_field, f.n_field)
    dm.assemble(eps, 1.0 / mu)
    E1, B1 = hip_ctx.values_get()
    dm.assemble(eps, 1.0 / mu)
    E2, B2 = hip_ctx.values_get()
    assert np.array_equal(E1.view(np.uint64), E2.view(np.uint64))
    assert np.array_equal(B1.view(np.uint64), B2.view(np.uint64))


def test_hip_assembler_dropin_interface():
    """tet_mass_stiffness_matrices(field, er, ur) with (3,3,T) tensors."""
    m, f, eps, mu = _problem((3, 3, 2), 0.1, 5)
    asm = HipAssembler(None)
    E, B = asm.tet_mass_stiffness_matrices(f, cc.isotropic(eps), cc.isotropic(mu))
    Eo, Bo = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, cc.isotropic(eps), cc.isotropic(mu))
    assert _rowwise_rel_err(E, Eo) <= TOL and _rowwise_rel_err(B, Bo) <= TOL
    aniso = cc.isotropic(eps)
    aniso[0, 1] = 0.1
    E, B = asm.tet_mass_stiffness_matrices(f, aniso, cc.isotropic(mu))
    Eo, Bo = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, aniso, cc.isotropic(mu))
    assert _rowwise_rel_err(E, Eo) <= TOL and _rowwise_rel_err(B, Bo) <= TOL


def _random_tensors(T, seed, symmetric=False):
    rng = np.random.default_rng(seed)
    def one(scale):
        A = rng.uniform(-0.3, 0.3, (3, 3, T)) + 1j * rng.uniform(-0.05, 0.05, (3, 3, T))
        if symmetric:
            A = (A + A.transpose(1, 0, 2)) / 2
        for k in range(3):
            A[k, k] += rng.uniform(1.0, scale, T) * (1 - 1j * rng.uniform(0, 0.02, T))
        return A.astype(np.complex128)
    return one(10.0), one(2.0)


@pytest.mark.parametrize("symmetric,seed", [(True, 3), (False, 4)])
def test_tensor_assembly_matches_oracle(hip_ctx, symmetric, seed):
    """General 3x3 complex eps_r and mu_r per tet (PML-like diagonal blocks
    plus off-diagonal coupling; non-symmetric when symmetric=False)."""
    m, f, _, _ = _problem((4, 3, 3), 0.15, seed, shuffle=True)
    er, ur = _random_tensors(m.n_tets, seed, symmetric)
    dm = MeshOnDevice(hip_ctx, m.nodes, m.tets, f.tet_to_field, f.n_field)
    hip_ctx.material_set_tensor(er, ur)
    hip_ctx.assemble()
    E, B = dm.csr()
    Eo, Bo = cc.tet_mass_stiffness_matrices(m, f.tet_to_field, er, ur)
    np.testing.assert_array_equal(E.indices, Eo.indices)
    assert _rowwise_rel_err(E, Eo) <= TOL
    assert _rowwise_rel_err(B, Bo) <= TOL


def test_tensor_path_equals_scalar_path_on_isotropic_media(hip_ctx):
    m, f, eps, mu = _problem((5, 4, 3), 0.15, 8)
    dm = MeshOnDevice(hip_ctx, m.nodes, m.tets, f.tet_to_field, f.n_field)
    dm.assemble(eps, 1.0 / mu)
    E1, B1 = dm.csr()
    hip_ctx.material_set_tensor(cc.isotropic(eps), cc.isotropic(mu))
    hip_ctx.assemble()
    E2, B2 = dm.csr()
    assert _rowwise_rel_err(E2, E1) <= 1e-13 and _rowwise_rel_err(B2, B1) <= 1e-13


def sampled_rows_oracle(m, f, eps, mu, rows):
    """Oracle values of the given global rows, summed from the oracle element
    matrices of only the tets incident to them: {(row, col): (E, B)}."""
    t2f = f.tet_to_field
    rowset = np.zeros(f.n_field, dtype=bool)
    rowset[rows] = True
    tets = np.nonzero(rowset[t2f].any(axis=0))[0]
    lem, ltm = cc.local_maps(m.tets[:, tets], m.edges, m.tris, t2f[:, tets], m.n_edges)
    D, F = cc.ned2_tet_stiff_mass(m.nodes[:, m.tets[:, tets]], m.edge_lengths[m.tet_to_edge[:, tets]],
                                  lem, ltm, cc.matinv(cc.isotropic(mu[tets])),
                                  cc.isotropic(eps[tets]))
    out = {}
    for k, t in enumerate(tets):
        for i in range(20):
            r = t2f[i, t]
            if not rowset[r]:
                continue
            for j in range(20):
                key = (r, t2f[j, t])
                e, b = out.get(key, (0j, 0j))
                out[key] = (e + D[i, j, k], b + F[i, j, k])
    return out


@pytest.fixture(scope="module")
def cube60(hip_ctx):
    """The 60^3 jittered cube (1.296M tets, 8.3M DOFs, seed 0) assembled once
    on the device for the full-size tests."""
    prob = cube_assembly(60, seed=0)
    m, f = prob.mesh, prob.basis
    print(f"full-size cube T={m.n_tets} N={f.n_field}", flush=True)
    dm = MeshOnDevice(hip_ctx, m.nodes, m.tets, f.tet_to_field, f.n_field)
    dm.assemble(prob.eps_r, prob.inv_mu_r)
    E, B = dm.csr()
    yield prob, dm, E, B
    del E, B


def test_full_size_cube_sampled_rows(cube60):
    """60^3 jittered cube: pattern and values of 300 random rows (all four
    DOF blocks) against the oracle, and global pattern invariants (symmetric
    pattern, diagonal present, nnz per tet)."""
    prob, dm, E, B = cube60
    m, f = prob.mesh, prob.basis
    assert E.nnz == dm.nnz and 270 < E.nnz / m.n_tets < 285
    rng = np.random.default_rng(3)
    rows = rng.choice(f.n_field, size=300, replace=False)
    ref = sampled_rows_oracle(m, f, prob.eps_r, 1.0 / prob.inv_mu_r, rows)
    for r in rows:
        cols = E.indices[E.indptr[r]:E.indptr[r + 1]]
        oc = sorted(c for (rr, c) in ref if rr == r)
        np.testing.assert_array_equal(cols, oc)
        ev = E.data[E.indptr[r]:E.indptr[r + 1]]
        bv = B.data[B.indptr[r]:B.indptr[r + 1]]
        eo = np.array([ref[(r, c)][0] for c in oc])
        bo = np.array([ref[(r, c)][1] for c in oc])
        assert np.abs(ev - eo).max() <= TOL * np.abs(eo).max()
        assert np.abs(bv - bo).max() <= TOL * np.abs(bo).max()
    # pattern symmetric, diagonal present
    P = E.copy()
    P.data = np.ones_like(P.data, dtype=np.float64)
    assert (abs(P - P.T)).nnz == 0
    assert np.all(E.diagonal() != 0)


def test_full_size_cube_every_row_against_fixtures(cube60):
    """All 8.3M rows of the 60^3 cube against committed oracle fixtures
    (tests/golden/make_cube60.py): SHA-256 of scipy's indptr / indices
    (bit-exact pattern), and block checksums sum_i u_i (A v)_i over blocks of
    4096 rows for seeded v, u, each within 1e-12 of the block's absolute sum
    sum_i sum_j |A_ij| |v_j| (E and B)."""
    import hashlib
    import os
    prob, dm, E, B = cube60
    G = np.load(os.path.join(os.path.dirname(__file__), "golden", "cube60.npz"), allow_pickle=False)
    assert int(G["n_field"]) == prob.basis.n_field and int(G["nnz"]) == E.nnz
    assert hashlib.sha256(E.indptr.astype(np.int32).tobytes()).hexdigest() == str(G["sha_indptr"])
    assert hashlib.sha256(E.indices.astype(np.int32).tobytes()).hexdigest() == str(G["sha_indices"])
    N = prob.basis.n_field
    rng = np.random.default_rng(int(G["seed_v"]))
    v = rng.normal(size=N) + 1j * rng.normal(size=N)
    u = np.where(np.random.default_rng(int(G["seed_u"])).random(N) < 0.5, -1.0, 1.0)
    blk = int(G["block"])
    nb = (N + blk - 1) // blk
    pad = nb * blk - N
    for name, A in (("E", E), ("B", B)):
        y = A @ v
        c = np.concatenate([y * u, np.zeros(pad, dtype=y.dtype)]).reshape(nb, blk).sum(axis=1)
        err = np.abs(c - G["c" + name]) / G["a" + name]
        print(f"{name}: max block error {err.max():.2e}", flush=True)
        assert err.max() <= 1e-12, (name, int(err.argmax()), err.max())


@pytest.mark.parametrize("use_mfma", [False, True])
def test_batched_element_matrices_match_oracle(hip_ctx, use_mfma):
    """emg_element_matrices (sparse compile-time tables on the vector units,
    or the dense tables on FP64 matrix cores) against the oracle's
    ned2_tet_stiff_mass per tet, in the reference's local DOF order, on a
    shuffled jittered mesh with complex eps_r and mu_r."""
    print(f"element matrices use_mfma={use_mfma}", flush=True)
    m, f, eps, mu = _problem((4, 3, 3), 0.15, 5, shuffle=True)
    dm = MeshOnDevice(hip_ctx, m.nodes, m.tets, f.tet_to_field, f.n_field)
    K, M = dm.element_matrices(eps, 1.0 / mu, use_mfma=use_mfma)
    lem, ltm = cc.local_maps(m.tets, m.edges, m.tris, f.tet_to_field, m.n_edges)
    D, F = cc.ned2_tet_stiff_mass(m.nodes[:, m.tets], m.edge_lengths[m.tet_to_edge], lem, ltm,
                                  cc.matinv(cc.isotropic(mu)), cc.isotropic(eps))
    D, F = D.transpose(2, 0, 1), F.transpose(2, 0, 1)
    eK = np.abs(K - D).max(axis=(1, 2)) / np.abs(D).max(axis=(1, 2))
    eM = np.abs(M - F).max(axis=(1, 2)) / np.abs(F).max(axis=(1, 2))
    print(f"max per-tet relative error K {eK.max():.2e} M {eM.max():.2e}", flush=True)
    assert eK.max() <= TOL and eM.max() <= TOL
