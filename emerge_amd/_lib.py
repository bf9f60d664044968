"""ctypes binding of ``libemerge_hip.so`` (declarations in include/emerge_hip.h).

The library is built in-tree by ``emerge_amd.build``.  There is no CPU
fallback: if the shared library is missing or no HIP device is present, the
calls raise.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libemerge_hip.so"

_c_int = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pf = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes); every exported entry point of emerge_hip.h
SIGNATURES: dict[str, tuple] = {
    "emg_version": (ctypes.c_char_p, []),
    "emg_create": (_c_int, [_c_int, ctypes.POINTER(_p)]),
    "emg_destroy": (_c_int, [_p]),
    "emg_last_error": (_c_int, [_p, ctypes.c_char_p, ctypes.c_size_t]),
    "emg_timer_start": (_c_int, [_p]),
    "emg_timer_stop": (_c_int, [_p, _pf]),
    "emg_synchronize": (_c_int, [_p]),
    "emg_mesh_set": (_c_int, [_p, _i64, _p, _i64, _p, _p, _i64]),
    "emg_pattern_build": (_c_int, [_p, _pi64]),
    "emg_pattern_get": (_c_int, [_p, _p, _p]),
    "emg_scatter_map_get": (_c_int, [_p, _p]),
    "emg_pattern_stats": (_c_int, [_p, _pi64, _pi64, _pi64]),
    "emg_material_set": (_c_int, [_p, _p, _p]),
    "emg_assemble": (_c_int, [_p]),
    "emg_values_get": (_c_int, [_p, _p, _p]),
    "emg_system_reduce": (_c_int, [_p, _i64, _p]),
    "emg_system_refresh": (_c_int, [_p]),
    "emg_dof_maps": (_c_int, [_p, _p, _p, _p]),
    "emg_pec_keep": (_c_int, [_p, _i64, _p, _p, _p]),
    "emg_field_points": (_c_int, [_p, _p, _i64, _p, _p, _p, _p]),
    "emg_farfield": (_c_int, [_p, _i64, _p, _p, _p, _p, _i64, _p, _p, ctypes.c_double, _p, _p]),
    "emg_port_mode_elements": (_c_int, [_p, _i64, _p, _p, _p, ctypes.c_double, _p, _p]),
    "emg_farfield_noalloc": (_c_int, [_p, _i64, _p, _p, _p, _p, _i64, _p, _p, ctypes.c_double, _p, _p]),
    "emg_system_pattern_get": (_c_int, [_p, _p, _p, _pi64]),
    "emg_bc_add": (_c_int, [_p, _i64, _p, _p, ctypes.POINTER(ctypes.c_int32), _pi64]),
    "emg_tri_terms_get": (_c_int, [_p, _p, _p]),
    "emg_surface_set": (_c_int, [_p, _i64, _p, _p, _p, _p, _i64, _p]),
    "emg_surface_get": (_c_int, [_p, _p]),
    "emg_forcing_set": (_c_int, [_p, ctypes.c_int32, _i64, _p, _p, _i64, _p]),
    "emg_forcing_get": (_c_int, [_p, _p]),
    "emg_sweep_solve": (_c_int, [_p, _i64, _p, _p, _p, _c_int, _c_int, ctypes.c_double, _c_int, _c_int,
                                 _p, _p, _p, _p]),
    "emg_port_probe_set": (_c_int, [_p, ctypes.c_int32, _p, _p, _p, _p, _p]),
    "emg_port_norms_get": (_c_int, [_p, _p]),
    "emg_solver_config": (_c_int, [_p, _c_int, _c_int, ctypes.c_double]),
    "emg_aux_set": (_c_int, [_p, _i64, _p, _p, _p, _p, _p, _p, _p, _c_int, _p]),
    "emg_aux_set_terms": (_c_int, [_p, _p, _p]),
    "emg_sweep_escalated": (_c_int, [_p, _p]),
    "emg_solver_set_escalation": (_c_int, [_p, _c_int]),
    "emg_solver_set_safe32": (_c_int, [_p, _c_int]),
    "emg_solver_set_spmv_lanes": (_c_int, [_p, _c_int]),
    "emg_solver_set_segments": (_c_int, [_p, ctypes.c_int32, _p]),
    "emg_topology_build": (_c_int, [_p, _i64, _p, _i64, _p, _p, _p]),
    "emg_topology_get": (_c_int, [_p, _p, _p, _p, _p, _p, _p, _p]),
    "emg_port_points_set": (_c_int, [_p, ctypes.c_int32, _p, _p, _p, _p, _p]),
    "emg_material_set_tensor": (_c_int, [_p, _p, _p]),
    "emg_bc_add_abc2": (_c_int, [_p, _i64, _p, ctypes.POINTER(ctypes.c_int32), _pi64]),
    "emg_csr_set": (_c_int, [_p, _i64, _p, _p, _p]),
    "emg_csr_aux_set": (_c_int, [_p, _i64, _p, _p, _p, _p, _p, _p]),
    "emg_element_matrices": (_c_int, [_p, _c_int, _p, _p]),
    "emg_csr_solve": (_c_int, [_p, ctypes.c_int32, _p, _c_int, ctypes.c_double, _c_int, _c_int, _p, _p, _p]),
}

_lib = None
_lock = threading.Lock()


class EmgError(RuntimeError):
    """Error reported by libemerge_hip (negative status code)."""

    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} failed with status {code}: {msg}")
        self.code = code


def load() -> ctypes.CDLL:
    """Load the shared library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m emerge_amd.build` "
                "(the HIP path has no CPU fallback)")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    if not a.flags.c_contiguous:
        raise ValueError("array must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """One emg_ctx: a device, a HIP stream and the device-resident state."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.emg_create(device, ctypes.byref(h))
        if rc != 0:
            raise EmgError("emg_create", rc, f"no usable HIP device {device}")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.emg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def call(self, name: str, *args) -> None:
        rc = getattr(self.lib, name)(self.h, *args)
        if rc != 0:
            buf = ctypes.create_string_buffer(512)
            self.lib.emg_last_error(self.h, buf, 512)
            raise EmgError(name, rc, buf.value.decode(errors="replace"))

    # ------------------------------------------------------------ wrappers
    def mesh_set(self, nodes: np.ndarray, tets: np.ndarray, tet_to_field: np.ndarray, n_field: int) -> None:
        xyz = np.ascontiguousarray(nodes, dtype=np.float64)
        t = np.ascontiguousarray(tets, dtype=np.int32)
        f = np.ascontiguousarray(tet_to_field, dtype=np.int32)
        self._keep = (xyz, t, f)
        self.call("emg_mesh_set", xyz.shape[1], ptr(xyz), t.shape[1], ptr(t), ptr(f), int(n_field))
        self.n_field = int(n_field)
        self.n_tets = t.shape[1]

    def pattern_build(self) -> int:
        nnz = ctypes.c_int64()
        self.call("emg_pattern_build", ctypes.byref(nnz))
        self.nnz = nnz.value
        return self.nnz

    def pattern_get(self) -> tuple[np.ndarray, np.ndarray]:
        indptr = np.empty(self.n_field + 1, dtype=np.int32)
        indices = np.empty(self.nnz, dtype=np.int32)
        self.call("emg_pattern_get", ptr(indptr), ptr(indices))
        return indptr, indices

    def scatter_map_get(self) -> np.ndarray:
        m = np.empty(400 * self.n_tets, dtype=np.int32)
        self.call("emg_scatter_map_get", ptr(m))
        return m

    def pattern_stats(self) -> dict:
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self.call("emg_pattern_stats", ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"chunks": a.value, "incidences": b.value, "max_rank": c.value}

    def material_set(self, eps_r: np.ndarray, inv_mu_r: np.ndarray) -> None:
        e = np.ascontiguousarray(eps_r, dtype=np.complex128)
        m = np.ascontiguousarray(inv_mu_r, dtype=np.complex128)
        if e.shape != (self.n_tets,) or m.shape != (self.n_tets,):
            raise ValueError("material arrays must have shape (n_tets,)")
        self.call("emg_material_set", ptr(e), ptr(m))

    def material_set_tensor(self, eps_r: np.ndarray, mu_r: np.ndarray) -> None:
        """(3, 3, n_tets) complex tensors eps_r and mu_r (not inverted)."""
        e = np.ascontiguousarray(eps_r, dtype=np.complex128)
        m = np.ascontiguousarray(mu_r, dtype=np.complex128)
        if e.shape != (3, 3, self.n_tets) or m.shape != (3, 3, self.n_tets):
            raise ValueError("material tensors must have shape (3, 3, n_tets)")
        self.call("emg_material_set_tensor", ptr(e), ptr(m))

    def assemble(self) -> None:
        self.call("emg_assemble")

    def values_get(self) -> tuple[np.ndarray, np.ndarray]:
        E = np.empty(self.nnz, dtype=np.complex128)
        B = np.empty(self.nnz, dtype=np.complex128)
        self.call("emg_values_get", ptr(E), ptr(B))
        return E, B

    def timer_start(self) -> None:
        self.call("emg_timer_start")

    def timer_stop(self) -> float:
        ms = ctypes.c_float()
        self.call("emg_timer_stop", ctypes.byref(ms))
        return ms.value

    def synchronize(self) -> None:
        self.call("emg_synchronize")
