"""sam_vel_max_noalloc (no counterpart in the reference): the allocation-free
variant of sam_vel_max for callers that cannot allocate device memory.
Inputs and outputs are device buffers (allocated here through the HIP
runtime); the results must equal sam_vel_max's (vel_max.f:11-49) bit for
bit, on two calls in a row, the second on a smaller grid."""
import ctypes

import numpy as np
import pytest

from sam_amd import samhip as S

pytestmark = pytest.mark.gpu


class Hip:
    """the few HIP runtime calls the test needs to own device buffers"""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.L.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0      # hipMemcpyHostToDevice

    def d2h(self, a, src):
        assert self.L.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0      # hipMemcpyDeviceToHost

    def free(self, p):
        self.L.hipFree(p)


def _case(hip, n, strat, i_cfl, seed, dt_in):
    s = S.Sam(S.make_params(n, i_strat=strat, i_cfl=i_cfl, cfl0=0.7, dt0=0.01, grav=9.0, lapse=1.0))
    rng = np.random.default_rng(seed)
    phys = np.asfortranarray(rng.standard_normal((n, n, n, s.lu)) * 1.7 + 0.3)
    want = s.vel_max(phys, dt_in)
    flat = np.ascontiguousarray(phys.ravel(order="F"))
    dp, ds = hip.malloc(flat.nbytes), hip.malloc(17 * 8)
    try:
        hip.h2d(dp, flat)
        hip.h2d(ds, np.full(17, np.nan))
        S._chk(s.L.sam_vel_max_noalloc(s.h, dp, dt_in, ds))
        s.sync()
        got = np.zeros(17)
        hip.d2h(got, ds)
    finally:
        hip.free(dp)
        hip.free(ds)
    s.close()
    return got, want


def test_vel_max_noalloc_matches_vel_max():
    hip = Hip()
    got1, want1 = _case(hip, 64, 1, 1, 3, 0.02)       # 4 fields, CFL dt
    np.testing.assert_array_equal(got1, want1)
    assert want1[16] != 0.02 and np.all(np.isfinite(got1))
    got2, want2 = _case(hip, 32, 0, 0, 4, 0.015)      # smaller grid, 3 fields, fixed dt
    np.testing.assert_array_equal(got2, want2)
    assert want2[16] == 0.015 and got2[3] == 0.0 and got2[11] == 0.0


def test_vel_max_noalloc_rejects_grids_above_its_workspace():
    s = S.Sam(S.make_params(1024, 1024, 256, precision=S.PREC_SPLIT))
    hip = Hip()
    ds = hip.malloc(17 * 8)
    try:
        rc = s.L.sam_vel_max_noalloc(s.h, ds, 0.01, ds)     # refused before any device work
        assert rc == S.SAM_ERR_SIZE
        assert b"partial rows" in s.L.sam_last_error()
    finally:
        hip.free(ds)
        s.close()
