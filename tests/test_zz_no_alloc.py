"""csg_pair_forces_noalloc, the device-pointer variant of csg_pair_forces that allocates nothing (its scratch is a
static device array), against csg_pair_forces on the same inputs: two calls in a row, the second with fewer sources
and particles, and a call beyond its source limit refused with CSG_ERR_CAPACITY. Tolerances of the parity tests:
distances 1e-9 relative, forces 1e-9 of the largest force (the two share kernels; only the spill capacity, and so
which sources the HBM-table tier takes, differs), energies 1e-12 relative, flags identical."""
import ctypes as C

import numpy as np
import pytest

import curvedspacesim_amd as csa
import workloads as wl
from oracle import pyoracle

pytestmark = pytest.mark.gpu


class Hip:
    """Device buffers of the caller, through the HIP runtime the library itself uses (the same loaded libamdhip64): the
    caller allocates its inputs and outputs; the entry point under test allocates nothing."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _bind(L):
    vp = C.c_void_p
    L.csg_pair_forces_noalloc.restype = C.c_int
    L.csg_pair_forces_noalloc.argtypes = [vp, C.c_int32, vp, vp, C.c_int32, vp, vp, vp, vp, C.c_double, vp, C.c_int32, C.c_double,
                                          vp, vp, vp, vp, vp, vp]
    return L.csg_pair_forces_noalloc


def _case(X, T, n, rc, seed):
    rng = np.random.default_rng(seed)
    f, b = wl.area_uniform(X, T, n, rng)
    off, idx, md = wl.candidates(X, T, f, b, rc)
    return f, b, off, idx, np.minimum(rc, md)


def test_noalloc_matches_pair_forces():
    ctx = csa.Context(0)
    hip = Hip()
    try:
        fn = _bind(ctx.L)
        X, T = csa.bumpy_sphere(5, 20, 0.15, 0.3, 13377)
        ctx.upload_mesh(X, T)
        rc = 0.2
        pot = csa._native.csg_potential(csa.HARMONIC, 1.0 / rc ** 2, rc, 0.0)
        for n in (2000, 700):  # the second call is smaller and reuses the scratch the first left
            print(f"case noalloc pair forces N={n}", flush=True)
            f, b, off, idx, rad = _case(X, T, n, rc, 100 + n)
            ref = ctx.pair_forces(f, b, off, idx, patch_radius=rad, potential=(csa.HARMONIC, 1.0 / rc ** 2, rc), use_sentinel=True,
                                  sentinel=2 * rc)
            P = int(off[-1])
            d0, st0, fl0 = np.zeros(P), np.zeros(3 * P, np.float32), np.zeros(P, np.uint32)
            F0, E0 = np.zeros(3 * n), np.zeros(n)
            ptrs = [hip.put(a) for a in (f.astype(np.int32), b, off.astype(np.int32), idx.astype(np.int32), rad, d0, st0, fl0, F0, E0)]
            df, db, doff, didx, drad, dd, dst, dfl, dF, dE = ptrs
            rc_ = fn(ctx.h, n, df, db, n, None, doff, didx, drad, 1e300, C.byref(pot), 1, 2 * rc, dd, dst, None, dfl, dF, dE)
            assert rc_ == 0, ctx.L.csg_last_error(ctx.h).decode()
            d, fl = hip.get(dd, d0), hip.get(dfl, fl0)
            F, E = hip.get(dF, F0).reshape(n, 3), hip.get(dE, E0)
            hip.free()
            np.testing.assert_array_equal(fl, ref["flags"])
            np.testing.assert_allclose(d, ref["dist"], rtol=1e-9, atol=0)
            assert np.max(np.abs(F - ref["forces"])) <= 1e-9 * np.abs(ref["forces"]).max()
            np.testing.assert_allclose(E, ref["energy"], rtol=1e-12, atol=1e-300)
            # spot check against the oracle
            om = pyoracle.OracleMesh(X, T)
            o = om.compute_forces(f, b, off, idx, rad, csa.HARMONIC, np.array([1.0 / rc ** 2, rc, 0.0]), True, rc, threads=16)
            within = o["dist"] < rc
            np.testing.assert_allclose(d[within], o["dist"][within], rtol=1e-9)
        print("case noalloc beyond its source limit", flush=True)
        big = 4096
        r = fn(ctx.h, big, C.c_void_p(8), C.c_void_p(8), big, None, C.c_void_p(8), C.c_void_p(8), C.c_void_p(8), 1e300, C.byref(pot), 1,
               2 * rc, None, None, None, None, None, None)
        assert r == csa.ERR_CAPACITY and "at most" in ctx.L.csg_last_error(ctx.h).decode()
    finally:
        hip.free()
        ctx.close()
