"""le_geneig_noalloc (the eigensolve with its workspace in a fixed device array, for callers
that cannot allocate) against le_geneig on the same pencils, through the C ABI: two calls in a
row on one stream, the second smaller (d = 231, then d = 40), eigenvalues and S-normalised
eigenvectors identical to le_geneig's within 1e-12; the size limit and the not-positive-definite
status.  Device buffers come from the HIP runtime the library itself links."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from laughlinedge_amd import lib
    lib()   # loads libamdhip64.so.7 into the process
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


class Dev:
    """one device buffer holding a copy of a float64 / int32 host array"""

    def __init__(self, hip, arr):
        self.hip, self.arr = hip, np.ascontiguousarray(arr)
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(8, self.arr.nbytes)) == 0
        assert hip.hipMemcpy(self.p, self.arr.ctypes.data, self.arr.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.arr)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _pencil(rng, d):
    A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    S = A @ A.conj().T / d + np.eye(d)
    B = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    H = (B + B.conj().T) / 2
    return S, H


def _pack(A):
    return np.ascontiguousarray(np.stack([A.real, A.imag], -1).reshape(-1))


def _noalloc(hip, stream, S, H, vectors=True):
    from laughlinedge_amd import lib
    d = S.shape[0]
    dS, dH = Dev(hip, _pack(S)), Dev(hip, _pack(H))
    dE = Dev(hip, np.full(d, np.nan))
    dV = Dev(hip, np.zeros(2 * d * d)) if vectors else None
    dst = Dev(hip, np.full(1, -7, dtype=np.int32))
    rc = lib().le_geneig_noalloc(ctypes.c_int(d), dS.p, dH.p, dE.p, dV.p if vectors else None, dst.p, stream)
    assert rc == 0, lib().le_last_error()
    assert hip.hipStreamSynchronize(stream) == 0
    ev, st = dE.get(), int(dst.get()[0])
    vec = dV.get() if vectors else None
    for b in (dS, dH, dE, dst) + ((dV,) if vectors else ()):
        b.free()
    return ev, (vec[0::2] + 1j * vec[1::2]).reshape(d, d) if vectors else None, st


def test_noalloc_geneig_matches_geneig():
    from laughlinedge_amd import geneig
    hip = _hip()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    rng = np.random.default_rng(2024)
    try:
        for d in (231, 40):
            print(f"case no-alloc geneig d={d}", flush=True)
            S, H = _pencil(rng, d)
            ev_ref, vec_ref = geneig(S, H, vectors=True)
            ev, vec, st = _noalloc(hip, stream, S, H)
            assert st == 0
            assert np.max(np.abs(ev - ev_ref)) <= 1e-12 * np.max(np.abs(ev_ref))
            assert np.max(np.abs(vec - vec_ref)) <= 1e-10 * np.max(np.abs(vec_ref))
            # the pencil itself: H v = E S v, S-orthonormal
            assert np.allclose(H @ vec, S @ vec * ev, atol=1e-9)
            assert np.allclose(vec.conj().T @ S @ vec, np.eye(d), atol=1e-10)
            ev2, _, st2 = _noalloc(hip, stream, S, H, vectors=False)
            assert st2 == 0 and np.array_equal(ev2, ev)
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_geneig_limits_and_status():
    from laughlinedge_amd import lib
    from laughlinedge_amd.engine import LE_ERR_ARG
    hip = _hip()
    print("case no-alloc geneig limits and status", flush=True)
    big = Dev(hip, np.zeros(2))
    rc = lib().le_geneig_noalloc(ctypes.c_int(257), big.p, big.p, big.p, None, None, None)
    assert rc == LE_ERR_ARG and b"257" in lib().le_last_error()
    big.free()
    S = np.array([[1.0, 2.0], [2.0, 1.0]], dtype=complex)
    H = np.eye(2, dtype=complex)
    _, _, st = _noalloc(hip, None, S, H, vectors=False)
    assert st == 1
