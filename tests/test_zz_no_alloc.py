"""emg_farfield_noalloc (device buffers, static scratch, no allocation)
against emg_farfield (host buffers, context allocations) through the C-ABI:
bit-identical results on two calls in a row, the second smaller, and a
problem beyond the scratch limit rejected before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(M, N, seed):
    rng = np.random.default_rng(seed)
    E = (rng.normal(size=(3, M)) + 1j * rng.normal(size=(3, M))).astype(np.complex128)
    H = (rng.normal(size=(3, M)) + 1j * rng.normal(size=(3, M))).astype(np.complex128) / 377.0
    E[:, ::7] *= 1e-5        # some samples fall under the 1e-3 level filter
    pos = rng.uniform(-0.02, 0.02, size=(3, M))
    wn = rng.normal(size=(3, M)) * 1e-6
    theta = rng.uniform(0, np.pi, size=N)
    phi = rng.uniform(0, 2 * np.pi, size=N)
    return E, H, pos, wn, theta, phi


def _host_call(ctx, M, N, E, H, pos, wn, theta, phi, k0):
    from emerge_amd._lib import ptr
    Eo = np.empty((3, N), dtype=np.complex128)
    Ho = np.empty((3, N), dtype=np.complex128)
    ctx.call("emg_farfield", M, ptr(E), ptr(H), ptr(pos), ptr(wn), N, ptr(theta), ptr(phi), ctypes.c_double(k0),
             ptr(Eo), ptr(Ho))
    return Eo, Ho


class _Hip:
    """Device buffers for the test through the HIP runtime the library itself
    links (the test allocates; the entry point under test does not)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.bufs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self._ok(self.rt.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self._ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _device_call(ctx, M, N, E, H, pos, wn, theta, phi, k0):
    hip = _Hip()
    try:
        d = [hip.upload(a) for a in (E, H, pos, wn, theta, phi)]
        nan = np.full((3, N), np.nan + 1j * np.nan)
        dEo, dHo = hip.upload(nan), hip.upload(nan)
        ctx.call("emg_farfield_noalloc", M, d[0], d[1], d[2], d[3], N, d[4], d[5], ctypes.c_double(k0), dEo, dHo)
        ctx.synchronize()
        return hip.download(dEo, nan), hip.download(dHo, nan)
    finally:
        hip.free()


def test_noalloc_matches_allocating_entry_point():
    from emerge_amd._lib import Context
    k0 = 2 * np.pi * 10e9 / 299792458
    with Context(0) as ctx:
        for M, N, seed in [(20000, 3000, 1), (3001, 517, 2)]:
            print(f"farfield noalloc M={M} N={N}", flush=True)
            args = _inputs(M, N, seed)
            Eh, Hh = _host_call(ctx, M, N, *args, k0)
            Ed, Hd = _device_call(ctx, M, N, *args, k0)
            assert np.all(np.isfinite(Ed)) and np.all(np.isfinite(Hd))
            assert np.array_equal(Ed, Eh) and np.array_equal(Hd, Hh), \
                (np.abs(Ed - Eh).max() / np.abs(Eh).max(), np.abs(Hd - Hh).max() / np.abs(Hh).max())


def test_noalloc_rejects_beyond_scratch():
    from emerge_amd._lib import Context, EmgError
    M, N = 64, 1_000_000          # per-chunk partial sums: 96 MB > 64 MiB
    with Context(0) as ctx:
        hip = _Hip()
        try:
            a = hip.upload(np.zeros(6 * M))
            d = hip.upload(np.zeros(N))
            o = hip.upload(np.zeros(6 * N))
            with pytest.raises(EmgError, match="scratch"):
                ctx.call("emg_farfield_noalloc", M, a, a, a, a, N, d, d, ctypes.c_double(1.0), o, o)
        finally:
            hip.free()
