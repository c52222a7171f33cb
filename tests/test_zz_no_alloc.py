"""ehd_poisson_noalloc (the Poisson solve with its workspace in a static
device array, for callers that cannot allocate) against ehd_poisson on the
same charge: bit-identical potentials, on two calls in a row, the second on a
smaller grid (the array keeps the first call's contents).  Grids whose
workspace does not fit and non-FP64 parameters are rejected on the host.

The caller's device buffers come from the HIP runtime the library itself
links (hipMalloc / hipMemcpy through ctypes): the test is the caller that
owns them; the entry point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import ehd_amd as E

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    E.lib()                                   # loads the library and, with it, its HIP runtime
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipDeviceSynchronize.argtypes = []
    for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipDeviceSynchronize):
        f.restype = C.c_int
    return H


def _reference(p, q):
    with E.Box(p) as b:
        b.init()
        b.write_charge(q)
        b.poisson()
        return b.read_fields(("phi",))["phi"]


def _noalloc(H, p, q):
    n = q.size * 8
    dq, dphi = C.c_void_p(), C.c_void_p()
    assert H.hipMalloc(C.byref(dq), n) == 0 and H.hipMalloc(C.byref(dphi), n) == 0
    try:
        qc = np.ascontiguousarray(q)
        assert H.hipMemcpy(dq, qc.ctypes.data, n, H2D) == 0
        rc = E.lib().ehd_poisson_noalloc(C.byref(p), dq, dphi, None)
        assert rc == 0, E.lib().ehd_last_error()
        out = np.empty_like(qc)
        assert H.hipMemcpy(out.ctypes.data, dphi, n, D2H) == 0
        return out
    finally:
        H.hipDeviceSynchronize()
        H.hipFree(dq)
        H.hipFree(dphi)


def _charge(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    return 10.0 * (1.0 + 0.2 * rng.standard_normal((nz, ny, nx)))


def test_poisson_noalloc_matches_ehd_poisson_twice():
    H = _hip()
    for (nx, ny, nz, T, seed) in ((64, 64, 32, 300.0, 1), (48, 40, 20, 200.0, 2)):
        p = E.make_params(nx, ny, nz, T=T, C_=10.0, lam=0.25, precision=E.FP64, init_mode=E.INIT_HYDROSTATIC)
        q = _charge(nx, ny, nz, seed)
        ref = _reference(p, q)
        got = _noalloc(H, p, q)
        print(f"{nx}x{ny}x{nz}: max |noalloc - ehd_poisson| = {np.abs(got - ref).max():.3e}")
        assert np.array_equal(got, ref), f"{nx}x{ny}x{nz} differs"


def test_poisson_noalloc_rejects_on_host():
    L = E.lib()
    big = E.make_params(512, 512, 128, precision=E.FP64)
    rc = L.ehd_poisson_noalloc(C.byref(big), C.c_void_p(16), C.c_void_p(16), None)
    assert rc == -1 and b"EHD_NOALLOC_SCRATCH_BYTES" in L.ehd_last_error()
    f32 = E.make_params(64, 64, 32, precision=E.FP32S)
    rc = L.ehd_poisson_noalloc(C.byref(f32), C.c_void_p(16), C.c_void_p(16), None)
    assert rc == -1 and b"FP64" in L.ehd_last_error()
