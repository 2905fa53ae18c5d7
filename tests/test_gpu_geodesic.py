"""Parity of the HIP geodesic + force kernel (through the C-ABI) with the FP64 oracle.

Tolerances: distances 1e-9 relative (both sides compute in FP64; the reference target is 1e-6); start/end tangents
2e-5 absolute (the kernel carries frame rotations in FP32); forces 1e-4 relative to the largest force. Pairs the oracle
flags as ties (another path within 1e-8 leaving in a different direction) are excluded from tangent and force
comparisons: the tangent there depends on which of the equal paths is kept (parity unpinned, as CGAL's choice).
"""
import numpy as np
import pytest

import curvedspacesim_amd as csa
import small_meshes as sm
import workloads as wl
from oracle.pyoracle import OracleMesh

pytestmark = pytest.mark.gpu

DIST_RTOL = 1e-9
TAN_ATOL = 2e-5


@pytest.fixture(scope="module")
def ctx():
    c = csa.Context(0)
    assert c.device_name.startswith("gfx950")
    yield c
    c.close()


def _check_pairs(d, st, et, fl, od, ost, oet, ofl):
    reach = np.isfinite(od)
    assert np.array_equal(np.isfinite(d) & ((fl & 1) == 0), reach)
    np.testing.assert_allclose(d[reach], od[reach], rtol=DIST_RTOL, atol=1e-14)
    ok = reach & ((ofl & 2) == 0) & (od > 1e-9)
    np.testing.assert_allclose(st[ok], ost[ok], atol=TAN_ATOL)
    np.testing.assert_allclose(et[ok], oet[ok], atol=TAN_ATOL)


@pytest.mark.parametrize("name", ["tetrahedron", "cube", "icosahedron", "torus_with_saddles", "l_shape", "bumpy_open_grid"])
def test_whole_mesh_small(ctx, name, rng):
    mk = {"tetrahedron": sm.tetrahedron, "cube": sm.cube, "icosahedron": sm.icosahedron,
          "torus_with_saddles": lambda: sm.torus(6, 4), "l_shape": lambda: sm.l_shape(2),
          "bumpy_open_grid": lambda: sm.grid(5, 4, z=lambda x, y: 0.3 * np.sin(2 * x) * np.cos(1.5 * y))}[name]
    X, T = mk()
    om = OracleMesh(X, T)
    ctx.upload_mesh(X, T)
    ctx.set_submeshing(False)
    ns, nt = 5, 12
    sf, sb = sm.random_points(rng, len(T), ns)
    tf, tb = sm.random_points(rng, len(T), ns * nt)
    off = np.arange(0, ns * nt + 1, nt, dtype=np.int32)
    print(f"case whole_mesh {name}: {ns} sources x {nt} targets", flush=True)
    d, st, et, fl = ctx.distance_batch(sf, sb, off, tf, tb)
    for s in range(ns):
        q = slice(off[s], off[s + 1])
        od, ost, oet, ofl = om.geodesic(sf[s], sb[s], tf[q], tb[q])
        _check_pairs(d[q], st[q], et[q], fl[q], od, ost, oet, ofl)


def _patch_case(ctx, X, T, N, rc, rng, n_src=None, potential=(csa.TRUNC_GAUSSIAN, 1.0, 0.05, 0.15), kind=2, params=(1.0, 0.05, 0.15)):
    om = OracleMesh(X, T)
    ctx.upload_mesh(X, T)
    f, b = wl.area_uniform(X, T, N, rng)
    off, idx, md = wl.candidates(X, T, f, b, rc)
    rad = np.minimum(rc, md)
    n_src = n_src or N
    off = off[: n_src + 1]
    idx = idx[: off[-1]]
    rad = rad[:n_src]
    print(f"case patch nF={len(T)} N={N} rc={rc} sources={n_src} pairs={off[-1]}", flush=True)
    g = ctx.pair_forces(f, b, off, idx, patch_radius=rad, rmax=np.inf, potential=potential, use_sentinel=True,
                        sentinel=2 * rc)
    # the oracle driver takes every particle as a source: the unchecked ones get empty candidate lists
    full_off = np.append(off, np.full(N - n_src, off[-1])).astype(np.int32)
    full_md = np.append(md[:n_src], np.zeros(N - n_src))
    o = om.compute_forces(f, b, full_off, idx, full_md, kind, np.array(params, float), True, rc, threads=8)
    od, ost, oet, ofl = o["dist"], o["tan"], o["end"], o["flags"]
    d = g["dist"]
    np.testing.assert_allclose(d, od, rtol=DIST_RTOL, atol=1e-14)
    ok = ((ofl & 3) == 0) & (od > 1e-9)
    np.testing.assert_allclose(g["start"][ok], ost[ok], atol=TAN_ATOL)
    np.testing.assert_allclose(g["end"][ok], oet[ok], atol=TAN_ATOL)
    # forces of sources without tie pairs
    tie_src = np.zeros(n_src, bool)
    for s in range(n_src):
        tie_src[s] = np.any(ofl[off[s]:off[s + 1]] & 2)
    F, oF = g["forces"], o["forces"][:n_src]
    scale = max(np.abs(oF).max(), 1e-30)
    np.testing.assert_allclose(F[~tie_src], oF[~tie_src], atol=1e-4 * scale)
    return g, o


def test_patch_torus_harmonic(ctx, rng):
    X, T = csa.jittered_torus(3.0, 1.0, 60, 30, 0.02, 7)
    rc = 0.6
    _patch_case(ctx, X, T, 300, rc, rng, potential=(csa.HARMONIC, 1.0 / rc ** 2, rc), kind=0, params=(1.0 / rc ** 2, rc, 0))


def test_patch_bumpy_sphere_truncated_gaussian(ctx, rng):
    X, T = csa.bumpy_sphere(4, 20, 0.15, 0.3, 13377)
    _patch_case(ctx, X, T, 400, 0.3, rng, potential=(csa.TRUNC_GAUSSIAN, 1.0, 0.1, 0.3), kind=2, params=(1.0, 0.1, 0.3))


def test_patch_bumpy_sphere_l6_config_scale(ctx, rng):
    """Bumpy sphere L6 (81,920 faces) at the r_c = 0.15 of the N = 20,000 workload, 200 sources checked."""
    X, T = csa.bumpy_sphere(6, 20, 0.15, 0.3, 13377)
    _patch_case(ctx, X, T, 20000, 0.15, rng, n_src=200)


def test_overflow_sources_fall_back(ctx, rng):
    """Patches larger than the LDS tables (here 0.6 on an L5 sphere: ~2,000 faces) go through the HBM variant."""
    X, T = csa.bumpy_sphere(5, 20, 0.15, 0.3, 5)
    _patch_case(ctx, X, T, 600, 0.6, rng, n_src=40, potential=(csa.HARMONIC, 1.0, 0.6), kind=0, params=(1.0, 0.6, 0))


def test_edge_cases(ctx):
    X, T = sm.grid(8, 8)
    om = OracleMesh(X, T)
    ctx.upload_mesh(X, T)
    ctx.set_submeshing(True, 0.5)
    sb = np.array([[1 / 3, 1 / 3, 1 / 3]] * 3)
    sf = np.array([70, 70, 70], np.int32)
    # source 0: no targets; source 1: targets in its own face; source 2: one near and one disconnected target
    tf = np.array([70, 70, 71, 0], np.int32)
    tb = np.array([[0.2, 0.3, 0.5], [0.6, 0.2, 0.2], [0.3, 0.3, 0.4], [0.2, 0.3, 0.5]])
    off = np.array([0, 0, 2, 4], np.int32)
    thr = np.array([0.1, 0.1, 0.1])
    d, st, et, fl = ctx.distance_batch(sf, sb, off, tf, tb, thr)
    for s in (1, 2):
        q = slice(off[s], off[s + 1])
        od, ost, oet, ofl = om.distance(sf[s], sb[s], tf[q], tb[q], True, 0.5, 0.1)
        np.testing.assert_allclose(d[q], od, rtol=DIST_RTOL)
        np.testing.assert_allclose(st[q], ost, atol=TAN_ATOL)
    assert d[3] == 1.0 and np.all(st[3] == [0, 0, 1]) and (fl[3] & csa.PAIR_SENTINEL)


def test_truncation_radius(ctx, rng):
    X, T = csa.jittered_torus(3.0, 1.0, 80, 40, 0.02, 3)
    ctx.upload_mesh(X, T)
    om = OracleMesh(X, T)
    f, b = wl.area_uniform(X, T, 200, rng)
    off, idx, md = wl.candidates(X, T, f, b, 0.8)
    g = ctx.pair_forces(f, b, off, idx, patch_radius=np.minimum(0.8, md), rmax=0.4)
    o = om.compute_forces(f, b, off, idx, md, 3, np.zeros(3), True, 0.8, threads=4)
    inside = o["dist"] <= 0.4
    np.testing.assert_allclose(g["dist"][inside], o["dist"][inside], rtol=DIST_RTOL)
    assert np.all((g["flags"][~inside] & (csa.PAIR_UNREACHED | csa.PAIR_BEYOND)) != 0)
