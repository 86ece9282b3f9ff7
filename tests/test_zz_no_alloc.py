"""rsd_per_grpery_inv_fri_noalloc (no counterpart in the reference; for
callers that cannot allocate): its output equals rsd_per_grpery_inv_fri's
bit for bit, on two calls in a row, the second with fewer particles (and a
different NI), and it rejects an NI beyond its scratch.  Device buffers come
from the HIP runtime through ctypes; the library allocates nothing."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


class _Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        return p

    def h2d(self, dst, a):
        a = np.ascontiguousarray(a)
        assert self.lib.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2) == 0

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0


def _noalloc(rsd, hip, conf, radii, lat, ld):
    L = rsd.load_library()
    L.rsd_per_grpery_inv_fri_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.POINTER(rsd.Lattice), C.c_int,
                                                 C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    n = len(radii)
    dpos, drad = hip.malloc(conf.nbytes), hip.malloc(radii.nbytes)
    dG, dcnt = hip.malloc(ld * 11 * n * 8), hip.malloc(4)
    try:
        hip.h2d(dpos, conf)
        hip.h2d(drad, radii)
        rc = L.rsd_per_grpery_inv_fri_noalloc(n, dpos, drad, C.byref(lat), 2, 1.0 / np.pi, dG, ld, dcnt, None)
        assert rc == 0, rsd.load_library().rsd_last_error()
        hip.sync()
        G = np.zeros(ld * 11 * n)
        hip.d2h(G, dG)
        cnt = np.zeros(1, dtype=np.int32)
        hip.d2h(cnt, dcnt)
    finally:
        for p in (dpos, drad, dG, dcnt):
            hip.lib.hipFree(p)
    G = G.reshape(11 * n, ld).T  # column-major (ld x 11N)
    return G[:11 * n, :], int(cnt[0])


def test_noalloc_matches_the_allocating_entry_point(rsd):
    from rsd_amd.configs import random_suspension
    hip = _Hip()
    cases = [(64, 0.3, 3, 1e-4, 0.41), (16, 0.2, 4, 1e-9, 0.0)]  # the second call is smaller, with another NI
    for n, phi, seed, ews, T in cases:
        conf, radii, L = random_suspension(n, phi, seed)
        conf[1] = conf[0] + np.array([1.97, 0.1, 0.0])   # one overlapping pair: the counter moves
        lat = rsd.set_lattice_skew(T, np.full(3, L), 1.0)
        APP, APQ, AQQ, nov = rsd.per_grpery_inv_fri(conf, radii, lat, ews)  # also fills lat's Ewald constants
        print(f"noalloc case N={n} NR={lat.nr} NI={lat.ni} overlaps={nov}", flush=True)
        G, cnt = _noalloc(rsd, hip, conf, radii, lat, 11 * n + 5)
        n6 = 6 * n
        assert np.array_equal(G[:n6, :n6], APP)
        assert np.array_equal(G[:n6, n6:], APQ)
        assert np.array_equal(G[n6:, n6:], AQQ)
        assert np.array_equal(G[n6:, :n6], APQ.T)
        assert cnt == nov >= 1


def test_noalloc_rejects_ni_beyond_its_scratch(rsd):
    from rsd_amd.configs import random_suspension
    hip = _Hip()
    conf, radii, L = random_suspension(8, 0.2, 5)
    lat = rsd.set_lattice_skew(0.0, np.full(3, L), 1.0)
    rsd.calc_lattice_inverse(lat, 1e-4, radii)
    lat.ni = 8
    L_ = rsd.load_library()
    L_.rsd_per_grpery_inv_fri_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.POINTER(rsd.Lattice), C.c_int,
                                                  C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    dG = hip.malloc(88 * 88 * 8)
    try:
        rc = L_.rsd_per_grpery_inv_fri_noalloc(8, dG, dG, C.byref(lat), 2, 1.0 / np.pi, dG, 88, None, None)
    finally:
        hip.lib.hipFree(dG)
    assert rc == -5 and b"NI = 8" in L_.rsd_last_error()
