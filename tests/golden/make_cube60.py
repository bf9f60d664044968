"""Writes tests/golden/cube60.npz: full-size fixtures of the 60^3 jittered
Kuhn cube assembly (1,296,000 tets, 8,316,360 DOFs, seed 0, the bench
workload), for a GPU test that covers every row instead of a sample.

* pattern: SHA-256 of scipy's CSR indptr and indices (int32 bytes) of the
  reference's COO layout (ASM/curlcurl.py:145-173, 459-465), computed by
  scipy 1.15 coo -> csr on row blocks (each block's rows hold every COO entry
  of those rows, so its canonical CSR equals those rows of the full one).
* values: for a seeded complex vector v (PCG64 seed 11) the products E v and
  B v, summed tet by tet from the oracle's element matrices
  (oracle.curlcurl.ned2_tet_stiff_mass, ASM/curlcurl.py:180-412), reduced
  to block checksums c_b = sum_{i in block b} u_i y_i over blocks of 4096
  rows with seeded signs u (seed 12), together with the absolute bounds
  a_b = sum_i sum_j |A_ij| |v_j| that scale the test's tolerance.

Run from the repository root: python tests/golden/make_cube60.py
(about 2 minutes and 20 GB of memory on 8 cores).
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from scipy.sparse import coo_matrix  # noqa: E402

from emerge_amd.workloads import cube_assembly  # noqa: E402
from oracle import curlcurl as cc  # noqa: E402

BLOCK = 4096
SEED_V, SEED_U = 11, 12


def seeded_vectors(n):
    rng = np.random.default_rng(SEED_V)
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    u = np.where(np.random.default_rng(SEED_U).random(n) < 0.5, -1.0, 1.0)
    return v, u


def block_checksums(y, a, u):
    nb = (y.shape[0] + BLOCK - 1) // BLOCK
    pad = nb * BLOCK - y.shape[0]
    yu = np.concatenate([y * u, np.zeros(pad, dtype=y.dtype)]).reshape(nb, BLOCK).sum(axis=1)
    ab = np.concatenate([a, np.zeros(pad)]).reshape(nb, BLOCK).sum(axis=1)
    return yu, ab


def pattern_hashes(t2f, N, rows_per_block=400_000):
    """SHA-256 of indptr / indices of coo -> csr of the 400 COO entries per tet."""
    T = t2f.shape[1]
    flat = t2f.T.reshape(-1)                      # incidence k = 20 t + i -> row
    order = np.argsort(flat, kind="stable")
    srows = flat[order]
    h_ptr, h_idx = hashlib.sha256(), hashlib.sha256()
    h_ptr.update(np.zeros(1, dtype=np.int32).tobytes())
    offset = 0
    for r0 in range(0, N, rows_per_block):
        r1 = min(N, r0 + rows_per_block)
        lo, hi = np.searchsorted(srows, [r0, r1])
        inc = order[lo:hi]
        tets = inc // 20
        rows = np.repeat(flat[inc] - r0, 20)
        cols = t2f[:, tets].T.reshape(-1)
        m = coo_matrix((np.ones(rows.shape[0]), (rows, cols)), shape=(r1 - r0, N)).tocsr()
        h_ptr.update((m.indptr[1:].astype(np.int64) + offset).astype(np.int32).tobytes())
        h_idx.update(m.indices.astype(np.int32).tobytes())
        offset += m.nnz
    return h_ptr.hexdigest(), h_idx.hexdigest(), offset


def products(prob, v, chunk=20000):
    m, f = prob.mesh, prob.basis
    N, T = f.n_field, m.n_tets
    yE = np.zeros(N, dtype=np.complex128)
    yB = np.zeros(N, dtype=np.complex128)
    aE = np.zeros(N)
    aB = np.zeros(N)
    t2f = f.tet_to_field
    for s in range(0, T, chunk):
        e = min(T, s + chunk)
        sl = slice(s, e)
        lem, ltm = cc.local_maps(m.tets[:, sl], m.edges, m.tris, t2f[:, sl], m.n_edges)
        D, F = cc.ned2_tet_stiff_mass(m.nodes[:, m.tets[:, sl]], m.edge_lengths[m.tet_to_edge[:, sl]], lem, ltm,
                                      cc.matinv(cc.isotropic(1.0 / prob.inv_mu_r[sl])),
                                      cc.isotropic(prob.eps_r[sl]))
        vl = v[t2f[:, sl]]                                   # (20, t)
        idx = t2f[:, sl].reshape(-1)
        for A, y, a in ((D, yE, aE), (F, yB, aB)):
            loc = np.einsum("ijt,jt->it", A, vl)
            absl = np.einsum("ijt,jt->it", np.abs(A), np.abs(vl))
            y += np.bincount(idx, weights=loc.real.reshape(-1), minlength=N) \
                + 1j * np.bincount(idx, weights=loc.imag.reshape(-1), minlength=N)
            a += np.bincount(idx, weights=absl.reshape(-1), minlength=N)
    return yE, yB, aE, aB


def main():
    t0 = time.time()
    prob = cube_assembly(60, seed=0)
    f = prob.basis
    N = f.n_field
    print(f"mesh T={prob.mesh.n_tets} N={N} ({time.time() - t0:.1f} s)", flush=True)
    hp, hi, nnz = pattern_hashes(f.tet_to_field, N)
    print(f"pattern nnz={nnz} ({time.time() - t0:.1f} s)", flush=True)
    v, u = seeded_vectors(N)
    yE, yB, aE, aB = products(prob, v)
    cE, bE = block_checksums(yE, aE, u)
    cB, bB = block_checksums(yB, aB, u)
    print(f"products ({time.time() - t0:.1f} s)", flush=True)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cube60.npz")
    np.savez_compressed(out, n=np.array(60), seed=np.array(0), n_field=np.array(N), nnz=np.array(nnz),
                        sha_indptr=np.array(hp), sha_indices=np.array(hi), block=np.array(BLOCK),
                        seed_v=np.array(SEED_V), seed_u=np.array(SEED_U), cE=cE, aE=bE, cB=cB, aB=bB)
    print("wrote", out, flush=True)


if __name__ == "__main__":
    main()
