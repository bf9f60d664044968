"""Oracle: ctypes wrapper of the C++/OpenMP port of the reference's per-tet
assembly kernel (``oracle/asm_port.cpp``).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the timed CPU baseline
of ``bench.py`` and its own parity test (tests/test_oracle_port.py, which
checks it entry by entry against ``oracle.curlcurl``).  ``build()`` compiles it
with g++ -O3 -fopenmp into ``oracle/_build/``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "asm_port.cpp")
LIB = os.path.join(HERE, "_build", "liboracle_asm.so")


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmd = ["g++", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-std=c++17", SRC, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built (oracle.port.build())")
        lib = ctypes.CDLL(LIB)
        p = ctypes.c_void_p
        i = ctypes.c_int64
        lib.oracle_matrix_builder.argtypes = [i, i, i, i] + [p] * 13
        lib.oracle_matrix_builder.restype = ctypes.c_int
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        lib.oracle_set_threads.restype = None
        _lib = lib
    return _lib


def set_threads(n: int) -> None:
    """OpenMP threads of the port (0 = all host threads)."""
    _load().oracle_set_threads(int(n))


def matrix_builder(nodes, tets, tris, edges, all_edge_lengths, tet_to_field, tet_to_edge, ur, er):
    """Same inputs and outputs as ``oracle.curlcurl.matrix_builder``; also
    returns the number of OpenMP threads used."""
    lib = _load()
    c = lambda a, dt: np.ascontiguousarray(a, dtype=dt)  # noqa: E731
    nodes = c(nodes, np.float64)
    tets = c(tets, np.int64)
    tris = c(tris, np.int64)
    edges = c(edges, np.int64)
    el = c(all_edge_lengths, np.float64)
    t2f = c(tet_to_field, np.int64)
    t2e = c(tet_to_edge, np.int64)
    ur = c(ur, np.complex128)
    er = c(er, np.complex128)
    T = tets.shape[1]
    dE = np.empty(400 * T, dtype=np.complex128)
    dB = np.empty(400 * T, dtype=np.complex128)
    rows = np.empty(400 * T, dtype=np.int64)
    cols = np.empty(400 * T, dtype=np.int64)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    nth = lib.oracle_matrix_builder(nodes.shape[1], T, edges.shape[1], tris.shape[1], P(nodes), P(tets), P(edges),
                                    P(tris), P(el), P(t2f), P(t2e), P(ur), P(er), P(dE), P(dB), P(rows), P(cols))
    return dE, dB, rows, cols, nth
