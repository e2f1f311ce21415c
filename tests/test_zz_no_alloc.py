"""dqmc_inverse_noalloc (the entry point that allocates nothing) against dqmc_inverse.

Both run the same Gauss-Jordan kernel; the no-allocation form takes caller-owned device
buffers and keeps its scratch in the library's static per-device array. The caller's buffers
are allocated here with the HIP runtime the library itself links (ctypes on libamdhip64, the
same runtime instance as the library's, not a second one). The inverse and log|det| must be
bit-identical to the context-based entry point, on two calls in a row with the second one
smaller (the scratch keeps the first call's contents).
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    """The HIP runtime libdqmc_hip.so is linked against (already loaded in the process)."""
    from dqmc_amd import _capi
    _capi.load()
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _dev_alloc(hip, nbytes):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p


def _say(msg):
    print("[gpu-test] " + msg, flush=True)
    sys.stdout.flush()


def _rand(rng, n):
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) + 3.0 * np.eye(n)


def test_inverse_noalloc_matches_inverse():
    import dqmc_amd
    from dqmc_amd import DQMC, _capi
    dqmc_amd.load()
    lib = _capi.load()
    hip = _hip()
    rng = np.random.default_rng(7007)
    nmax = 400
    dA = _dev_alloc(hip, nmax * nmax * 16)
    dX = _dev_alloc(hip, nmax * nmax * 16)
    dld = _dev_alloc(hip, 8)
    try:
        for L in (10, 6):  # n = 400, then the smaller n = 144
            p = dqmc_amd.Params(L=L, slices=10, hoppings=(1.0, 0.5, 0.5, 1.0))
            n = p.n
            A = np.asfortranarray(_rand(rng, n))
            _say("inverse_noalloc n=%d" % n)
            c = DQMC(p)
            ref, ref_ld = c.inv(A)
            c.close()
            assert hip.hipMemcpy(dA, A.ctypes.data, n * n * 16, H2D) == 0
            _capi.check(lib.dqmc_inverse_noalloc(n, dA, dX, dld, None))
            assert hip.hipDeviceSynchronize() == 0
            got = np.zeros((n, n), dtype=np.complex128, order="F")
            ld = np.zeros(1)
            assert hip.hipMemcpy(got.ctypes.data, dX, n * n * 16, D2H) == 0
            assert hip.hipMemcpy(ld.ctypes.data, dld, 8, D2H) == 0
            assert np.array_equal(got, ref), n
            assert ld[0] == ref_ld, (n, ld[0], ref_ld)
            assert np.abs(got @ A - np.eye(n)).max() < 1e-10
        # sizes beyond the scratch's maximum are rejected before any device work
        rc = lib.dqmc_inverse_noalloc(1025, dA, dX, None, None)
        assert rc == -1  # DQMC_E_ARG
        assert b"1..1024" in lib.dqmc_last_error()
    finally:
        for ptr in (dA, dX, dld):
            hip.hipFree(ptr)
