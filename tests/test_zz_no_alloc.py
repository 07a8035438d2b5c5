"""am_allpairs_forces_noalloc (the all-pairs entry point that allocates nothing) against
am_compute_allpairs_coulomb_forces on the same inputs.

The device buffers are the caller's: this test allocates them with the HIP runtime through
ctypes, calls the entry point through the C ABI, and compares with the store entry point:
bit-identical below 4,096 particles (the same unsorted tile kernel and reduction), and within
the all-pairs FP32 tolerance (global relative error <= 1e-5 against the FP64 path, which equals
the oracle bit for bit) at the maximum size, where the store path switches to its sorted kernel.
Two calls in a row, the second smaller, check that the shared scratch array keeps no state that
matters between calls.
"""
import ctypes

import numpy as np
import pytest

import antimatter_amd as am
import oracle

pytestmark = pytest.mark.gpu

E = oracle.E_CHARGE
_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [ctypes.c_int]
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.nbytes = max(self.host.nbytes, 8)
        self.ptr = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        assert hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def noalloc_forces(pos, q, ann):
    n = len(q)
    assert hip().hipSetDevice(0) == 0
    bufs = [DeviceArray(pos.astype(np.float64)), DeviceArray(q.astype(np.float64)),
            DeviceArray(ann.astype(np.uint8)), DeviceArray(np.full((n, 3), np.nan))]
    try:
        rc = am.lib().am_allpairs_forces_noalloc(n, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, None)
        assert rc == 0, am.lib().am_last_error().decode()
        assert hip().hipDeviceSynchronize() == 0
        return bufs[3].read()
    finally:
        for b in bufs:
            b.free()


def store_forces(ctx, pos, q, ann, precision):
    n = len(q)
    s = am.Store(ctx, n)
    try:
        sp = np.where(q > 0, 1, np.where(q < 0, 3, 2)).astype(np.uint8)
        s.upload(pos, np.zeros_like(pos), sp, charge=q, annihilated=ann)
        s.compute_allpairs_coulomb_forces(precision)
        return s.forces()
    finally:
        s.close()


def cloud(n, seed, dead_frac=0.05):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1e-3, 1e-3, (n, 3)) * np.array([1.0, 1.0, 4.0])
    q = np.where(rng.random(n) < 0.5, E, -E)
    q[::97] = 0.0  # a few neutral particles
    q[5] = 2 * E   # and one non-unit charge
    ann = (rng.random(n) < dead_frac).astype(np.uint8)
    return pos, q, ann


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_noalloc_matches_store_entry_point_twice(ctx):
    for n, seed in ((3000, 11), (1000, 12)):  # the second call smaller than the first
        print(f"noalloc: n={n} seed={seed}", flush=True)
        pos, q, ann = cloud(n, seed)
        f = noalloc_forces(pos, q, ann)
        g = store_forces(ctx, pos, q, ann, am.FP32)
        np.testing.assert_array_equal(f, g)
        assert np.all(f[ann == 1] == 0.0) and np.all(f[q == 0.0] == 0.0)
        ref = oracle.allpairs_forces(pos, q, ann)
        assert rel(f, ref) < 1e-5


def test_noalloc_at_maximum_size(ctx):
    n = 65536
    print(f"noalloc: n={n} (maximum)", flush=True)
    pos, q, ann = cloud(n, 13, dead_frac=0.01)
    f = noalloc_forces(pos, q, ann)
    ref = store_forces(ctx, pos, q, ann, am.FP64)  # bit-identical to the oracle's FP64 sum
    assert rel(f, ref) < 1e-5
    g = store_forces(ctx, pos, q, ann, am.FP32)
    assert rel(f, g) < 1e-5
    # a smaller call after the largest one
    pos2, q2, ann2 = cloud(2000, 14)
    np.testing.assert_array_equal(noalloc_forces(pos2, q2, ann2), store_forces(ctx, pos2, q2, ann2, am.FP32))


def test_noalloc_rejects_sizes_beyond_the_maximum():
    L = am.lib()
    assert L.am_allpairs_forces_noalloc(65537, None, None, None, None, None) == -3  # AM_ERR_CAPACITY
    assert "65536" in L.am_last_error().decode()
    assert L.am_allpairs_forces_noalloc(-1, None, None, None, None, None) == -3
    assert L.am_allpairs_forces_noalloc(0, None, None, None, None, None) == 0
