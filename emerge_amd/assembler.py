"""Drop-in assembler backed by the HIP kernels.

``HipAssembler`` mirrors the reference ``Assembler`` plug-in
(ASM/assembler.py:100-108, ``ASM/`` = src/emerge/_emerge/physics/microwave/
assembly/): it is installed as ``model.mw.assembler = HipAssembler(settings)``
(attribute set at physics/microwave/microwave_3d.py:128) and keeps the
``cached_matrices`` semantics.  Its narrow entry point
``tet_mass_stiffness_matrices(field, er, ur) -> (E, B)`` replaces
ASM/curlcurl.py:145-173 with the same inputs (``(3,3,T)`` complex tensors)
and outputs (scipy CSR, int32 indices, identical sparsity pattern).
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix

from ._lib import Context


def isotropic_scalars(tensor: np.ndarray, name: str) -> np.ndarray | None:
    """(3,3,T) tensor -> (T,) scalars when every tet is isotropic (one
    complex scalar times the identity), else None."""
    t = np.asarray(tensor)
    if t.ndim != 3 or t.shape[:2] != (3, 3):
        raise ValueError(f"{name} must have shape (3, 3, T)")
    d = t[0, 0]
    off = np.ones((3, 3), dtype=bool)
    np.fill_diagonal(off, False)
    if np.any(t[off] != 0) or np.any(t[1, 1] != d) or np.any(t[2, 2] != d):
        return None
    return np.ascontiguousarray(d, dtype=np.complex128)


# canonical DOF order (tet vertices ranked by global id): edges (01, 02, 03,
# 12, 13, 23), faces (012, 013, 023, 123); the reference's local pairs and the
# vertex each local face misses (E/mesh3d.py:427, 431)
_CANON_EDGE = {(0, 1): 0, (0, 2): 1, (0, 3): 2, (1, 2): 3, (1, 3): 4, (2, 3): 5}
_LOCAL_EDGE = ((0, 1), (0, 2), (0, 3), (1, 2), (3, 1), (2, 3))
_LOCAL_FACE_MISSING = (3, 1, 2, 0)


def canonical_dof_map(tets: np.ndarray) -> np.ndarray:
    """(20, T): the canonical DOF of each local DOF of every tet (the
    reference's local order is that of tet_to_field)."""
    w = np.asarray(tets)
    rk = (w[None, :, :] < w[:, None, :]).sum(axis=1)        # rk[k, t] = rank of local vertex k
    umap = np.empty((20, w.shape[1]), dtype=np.int64)
    for e, (a, b) in enumerate(_LOCAL_EDGE):
        lo, hi = np.minimum(rk[a], rk[b]), np.maximum(rk[a], rk[b])
        ce = np.zeros(w.shape[1], dtype=np.int64)
        for (p, q), k in _CANON_EDGE.items():
            ce[(lo == p) & (hi == q)] = k
        umap[e], umap[e + 10] = ce, ce + 10
    for f, miss in enumerate(_LOCAL_FACE_MISSING):
        cf = 3 - rk[miss]
        umap[f + 6], umap[f + 16] = cf + 6, cf + 16
    return umap


class MeshOnDevice:
    """A mesh uploaded to one context with its CSR pattern built."""

    def __init__(self, ctx: Context, nodes: np.ndarray, tets: np.ndarray, tet_to_field: np.ndarray,
                 n_field: int):
        self.ctx = ctx
        self.tets = np.asarray(tets)
        ctx.mesh_set(nodes, tets, tet_to_field, n_field)
        self.nnz = ctx.pattern_build()
        self.n_field = int(n_field)
        self.n_tets = tets.shape[1]
        self._pattern = None

    @property
    def pattern(self) -> tuple[np.ndarray, np.ndarray]:
        if self._pattern is None:
            self._pattern = self.ctx.pattern_get()
        return self._pattern

    def assemble(self, eps_r: np.ndarray, inv_mu_r: np.ndarray) -> None:
        self.ctx.material_set(eps_r, inv_mu_r)
        self.ctx.assemble()

    def element_matrices(self, eps_r: np.ndarray, inv_mu_r: np.ndarray, use_mfma: bool = False):
        """Dense element matrices of every tet (ned2_tet_stiff_mass,
        ASM/curlcurl.py:180-412, batched on the device): K, M as
        (T, 20, 20) complex in the reference's local DOF order."""
        from ._lib import ptr
        self.ctx.material_set(eps_r, inv_mu_r)
        T = self.n_tets
        Kc = np.empty((400, T), dtype=np.complex128)
        Mc = np.empty((400, T), dtype=np.complex128)
        self.ctx.call("emg_element_matrices", int(bool(use_mfma)), ptr(Kc), ptr(Mc))
        umap = canonical_dof_map(self.tets)                    # (20, T)
        idx = umap[:, None, :] * 20 + umap[None, :, :]         # (20, 20, T)
        t = np.arange(T)[None, None, :]
        K = Kc[idx, t].transpose(2, 0, 1)
        M = Mc[idx, t].transpose(2, 0, 1)
        return K, M

    def csr(self) -> tuple[csr_matrix, csr_matrix]:
        indptr, indices = self.pattern
        Ev, Bv = self.ctx.values_get()
        N = self.n_field
        E = csr_matrix((Ev, indices, indptr), shape=(N, N))
        B = csr_matrix((Bv, indices, indptr), shape=(N, N))
        E.has_sorted_indices = True
        B.has_sorted_indices = True
        return E, B


class HipAssembler:
    """Assembler plug-in: volume matrices on the GPU."""

    def __init__(self, settings=None, device: int = 0):
        self.settings = settings
        self.cached_matrices = None
        self.device = device
        self._ctx: Context | None = None
        self._mesh_key = None
        self._dev_mesh: MeshOnDevice | None = None

    def _context(self) -> Context:
        if self._ctx is None:
            self._ctx = Context(self.device)
        return self._ctx

    @staticmethod
    def _fingerprint(*arrays) -> tuple:
        """Content key of the arrays the device mesh is built from, so that a
        mutated mesh (or a new one at a recycled address) is rebuilt."""
        try:
            import xxhash
            h = xxhash.xxh3_128()
        except ImportError:  # pragma: no cover
            import hashlib
            h = hashlib.blake2b(digest_size=16)
        shapes = []
        for a in arrays:
            a = np.ascontiguousarray(a)
            shapes.append((a.shape, a.dtype.str))
            h.update(a.view(np.uint8).reshape(-1))
        return tuple(shapes), h.hexdigest()

    def device_mesh(self, field) -> MeshOnDevice:
        mesh = field.mesh
        key = self._fingerprint(mesh.nodes, mesh.tets, field.tet_to_field)
        if self._dev_mesh is None or self._mesh_key != key:
            self._dev_mesh = MeshOnDevice(self._context(), mesh.nodes, mesh.tets, field.tet_to_field,
                                          field.n_field)
            self._mesh_key = key
        return self._dev_mesh

    def tet_mass_stiffness_matrices(self, field, er: np.ndarray, ur: np.ndarray):
        """ASM/curlcurl.py:145-173: (E curl-curl, B mass) as scipy CSR.
        Isotropic media take the scalar kernel; any anisotropic tet (or PML
        tensor) sends the whole mesh through the tensor kernel."""
        eps = isotropic_scalars(er, "er")
        mu = isotropic_scalars(ur, "ur")
        dm = self.device_mesh(field)
        if eps is not None and mu is not None:
            dm.assemble(eps, 1.0 / mu)  # matinv's diagonal path (E/mth/optimized.py:354-357)
        else:
            dm.ctx.material_set_tensor(er, ur)
            dm.ctx.assemble()
        return dm.csr()
