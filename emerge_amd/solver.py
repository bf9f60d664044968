"""Drop-in linear solver plug-in backed by the HIP Krylov kernels.

``SolverHIP`` has the shape of the reference ``Solver`` base class
(E/solver.py:333-371): class attributes ``real_only``, ``req_sorter`` and
``released_gil``; ``initialize / duplicate / set_options / reset``; and
``solve(A, b, precon, reuse_factorization, id) -> (x, SolveReport)``.  It is
registered like any other solver, ``model.set_solver(SolverHIP())``
(E/simmodel.py:726-737, E/solver.py:1079-1090), and sits next to SuperLU,
UMFPACK, PARDISO and cuDSS (E/solver.py:535-773).  ctypes calls release the
GIL, so the solver is legal in the multi-threaded sweep mode
(E/solver.py:1031-1033).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable

import numpy as np
from scipy.sparse import csr_matrix

from ._lib import Context, EmgError, ptr


class SimulationError(Exception):
    """Raised on solver failure (as E/solver.py:295-296)."""


@dataclass
class SolveReport:
    """Per-solve record with the fields of E/solver.py:92-113."""
    simtime: float = -1.0
    jobid: int = -1
    ndof: int = -1
    nnz: int = -1
    ndof_solve: int = -1
    nnz_solve: int = -1
    exit_code: int = 0
    solver: str = "None"
    sorter: str = "None"
    precon: str = "None"
    aux: dict = field(default_factory=dict)
    worker_name: str = "Unknown Worker"

    def add(self, **kwargs):
        for k, v in kwargs.items():
            self.aux[k] = str(v)

    @property
    def mdof(self) -> float:
        return (self.ndof ** 2) / ((self.simtime + 1e-6) * 1e6)


class SolverHIP:
    """Krylov solve on the GPU: GMRES(restart), IDR(s) or COCG on the
    symmetric part inside FP64 refinement (method 'gmres' / 'idr' / 'cocg');
    FP64, or complex64 inner iterations inside FP64 iterative refinement."""

    real_only: bool = False
    req_sorter: bool = False
    released_gil: bool = True

    _METHODS = {"gmres": 0, "idr": 1, "cocg": 2}

    def __init__(self, pre: str = "", precision: str = "fp64", tol: float = 1e-10, restart: int = 150,
                 max_cycles: int = 400, device: int = 0, method: str = "gmres", inner_rel: float | None = None):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if method not in self._METHODS:
            raise ValueError(f"method must be one of {sorted(self._METHODS)}")
        self.method = method
        self.inner_rel = inner_rel if inner_rel is not None else (1e-2 if method == "cocg" else 1e-5)
        self.pre = pre
        self.own_preconditioner = True
        self.initialized = False
        self.precision = precision
        self.tol = tol
        self.restart = restart
        self.max_cycles = max_cycles
        self.device = device
        self._ctx: Context | None = None
        self._n = -1
        self._aux = None          # (G, G^T) of the attached mesh, rows = the reduced DOFs
        self._aux_on_device = False

    def __str__(self) -> str:
        return "SolverHIP"

    def initialize(self) -> None:
        if self.initialized:
            return
        self._ctx = Context(self.device)
        self.initialized = True

    def duplicate(self) -> "SolverHIP":
        out = self.__class__(self.pre, self.precision, self.tol, self.restart, self.max_cycles, self.device,
                             self.method, self.inner_rel)
        out._aux = self._aux
        return out

    def attach(self, mesh, basis, solve_ids=None, pec_tris=None, vector: bool = True) -> "SolverHIP":
        """Give the solver the mesh and DOF maps of the systems it will see,
        so that COCG adds the auxiliary-space correction ``A D_a^-1 A^T``
        with A = [G | P] (discrete gradient of the second-order nodal
        functions, and with ``vector`` the vector-nodal functions lam_a e_d
        interpolated into the Nedelec-2 space) and D_a = diag(A^T A_s A)
        formed on the device per matrix.  SolveRoutine hands
        A[ix_(solve_ids, solve_ids)] (E/solver.py:1266-1268) with solve_ids
        ascending, the DOFs left by the PEC elimination
        (ASM/assembler.py:274-307, 406-407): pass them, or the PEC triangles
        that define them.  Auxiliary functions that touch an eliminated DOF
        are left out."""
        import scipy.sparse as sp
        from .gradient import discrete_gradient, vector_interpolation
        N = basis.n_field
        if solve_ids is None:
            t = np.asarray(pec_tris if pec_tris is not None else [], dtype=np.int64)
            removed = np.concatenate([basis.edge_to_field[:, mesh.tri_to_edge[:, t].reshape(-1)].reshape(-1),
                                      basis.tri_to_field[:, t].reshape(-1)])
            keep = np.ones(N, dtype=bool)
            keep[removed] = False
            solve_ids = np.nonzero(keep)[0]
        solve_ids = np.asarray(solve_ids, dtype=np.int64)
        if np.any(np.diff(solve_ids) <= 0):
            raise ValueError("solve_ids must be ascending (the order SolveRoutine uses)")
        G = discrete_gradient(mesh, basis).tocsr()
        if vector:
            G = sp.hstack([G, vector_interpolation(mesh, basis)], format="csr")
        dropped = np.ones(N, dtype=bool)
        dropped[solve_ids] = False
        touches = np.asarray(abs(G[dropped]).sum(axis=0)).ravel() > 0
        cols = np.nonzero(~touches)[0]
        Gr = G[solve_ids][:, cols].tocsr()
        Gr.sort_indices()
        Gt = Gr.T.tocsr()
        Gt.sort_indices()
        self._aux = (Gr, Gt)
        self._aux_on_device = False
        return self

    def set_options(self, pivoting_threshold: float | None = None) -> None:
        """Pivoting thresholds do not apply to a Krylov solver."""

    def reset(self) -> None:
        self._n = -1

    def solve(self, A, b: np.ndarray, precon=None, reuse_factorization: bool = False, id: int = -1):
        self.initialize()
        t0 = time.time()
        Ac = csr_matrix(A)  # SolveRoutine hands a CSC slice (E/solver.py:1266-1268)
        Ac.sort_indices()
        n = Ac.shape[0]
        if not reuse_factorization or self._n != n:
            ip = np.ascontiguousarray(Ac.indptr, dtype=np.int32)
            ix = np.ascontiguousarray(Ac.indices, dtype=np.int32)
            v = np.ascontiguousarray(Ac.data, dtype=np.complex128)
            self._ctx.call("emg_csr_set", n, ptr(ip), ptr(ix), ptr(v))
            self._n = n
            if self._aux is not None and not self._aux_on_device:
                Gr, Gt = self._aux
                if Gr.shape[0] != n:
                    raise SimulationError(f"SolverHIP: attached mesh has {Gr.shape[0]} reduced DOFs, "
                                          f"the system {n}")
                gp, gi = np.ascontiguousarray(Gr.indptr, np.int32), np.ascontiguousarray(Gr.indices, np.int32)
                gv = np.ascontiguousarray(Gr.data, np.float64)
                tp, ti = np.ascontiguousarray(Gt.indptr, np.int32), np.ascontiguousarray(Gt.indices, np.int32)
                tv = np.ascontiguousarray(Gt.data, np.float64)
                self._ctx.call("emg_csr_aux_set", Gr.shape[1], ptr(gp), ptr(gi), ptr(gv), ptr(tp), ptr(ti),
                               ptr(tv))
                self._aux_on_device = True
        bb = np.ascontiguousarray(np.asarray(b, dtype=np.complex128).reshape(-1))
        x = np.empty(n, dtype=np.complex128)
        rel = np.empty(1)
        it = np.empty(1, dtype=np.int64)
        prec = 1 if self.precision == "mixed" else 0
        try:
            self._ctx.call("emg_solver_config", self._METHODS[self.method], 4, float(self.inner_rel))
            self._ctx.call("emg_csr_solve", 1, ptr(bb), prec, float(self.tol), int(self.restart),
                           int(self.max_cycles), ptr(x), ptr(rel), ptr(it))
        except EmgError as e:
            raise SimulationError(str(e)) from e
        aux = {"precision": self.precision, "method": self.method, "iterations": str(int(it[0])), "relres": f"{rel[0]:.3e}",
               "tolerance": str(self.tol),
               "gradient_space": str(self._aux is not None and self.method == "cocg")}
        return x, SolveReport(solver=str(self), exit_code=0, aux=aux, simtime=time.time() - t0)
