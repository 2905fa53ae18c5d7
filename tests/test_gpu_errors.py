"""Error paths and the reference Gaussian on the device, through the C-ABI.

* csg_mesh_upload refuses what triangulatedMeshSpace::loadMeshFromFile would not accept (src/models/
  triangulatedMeshSpace.cpp:51-63) with CSG_ERR_MESH;
* failed walks keep their particle and report the reference's throw site (triangulatedMeshSpace.cpp:522-548,
  meshUtilities.cpp:184-195) per particle, and as CSG_ERR_WALK from the system's move;
* a source beyond the step's fixed geodesic tables is solved by its last tier (forces equal the oracle's), and a
  candidate-capacity overflow is reported as CSG_ERR_CAPACITY by csg_system_compute_forces and by the next
  csg_system_sync after an asynchronous step;
* gaussianRepulsion (src/forces/gaussianRepulsion.cpp:3-12) has no cutoff (its range stays the default 1,
  src/forces/baseForce.h:57), so a candidate the submesh does not connect to the source still pushes with the
  sentinel distance 2 maximumDistance along (0, 0, 1) (triangulatedMeshSpace.cpp:197-203): a flattened ellipsoid puts
  such pairs across its two faces.

Force tolerance: 1e-4 of the largest force, sources with cut-locus pairs excepted (FP32 start tangents; the same bar as
the system tests)."""
import numpy as np
import pytest

import curvedspacesim_amd as csa
import small_meshes as sm
import workloads as wl
from oracle.pyoracle import OracleMesh, OracleSystem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csa.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("case", ["out_of_range", "degenerate", "non_manifold", "flipped", "zero_area", "empty"])
def test_mesh_upload_rejects(ctx, case):
    X, T = sm.tetrahedron()
    T = T.copy()
    code = csa.ERR_MESH
    if case == "out_of_range":
        T[0, 0] = 9
    elif case == "degenerate":
        T[0, 1] = T[0, 0]
    elif case == "non_manifold":
        X = np.vstack([X, [[2.0, 0.0, 0.0]]])
        T = np.vstack([T, [[1, 0, 4]]])
    elif case == "flipped":
        T[0] = T[0][[0, 2, 1]]
    elif case == "zero_area":  # a flat fan whose first face has three collinear corners
        X = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [1, 1, 0]], float)
        T = np.array([[0, 1, 2], [0, 2, 3]])
    else:
        X, T = np.zeros((0, 3)), np.zeros((0, 3), np.int32)
        code = csa.ERR_ARGUMENT
    print(f"case mesh upload {case}", flush=True)
    with pytest.raises(csa.CsgError) as e:
        ctx.upload_mesh(X, T)
    assert e.value.code == code


def test_walk_failures_reported_per_particle(ctx):
    X, T = sm.grid(3, 3)
    ctx.upload_mesh(X, T)
    f0, b0 = np.array([8, 8, 8, 8], np.int32), np.array([[0.5, 0.25, 0.25]] * 4)
    disp = np.array([[5.0, 0.1, 0.0], [2.5, 3.5, 0.0], [np.nan, 0.0, 0.0], [0.2, 0.1, 0.0]])
    print("case walk failure statuses", flush=True)
    f, b, d, v, st = ctx.transport_batch(f0, b0, disp)
    assert st[0] == -9 and st[1] == -9 and st[2] == -4 and st[3] >= 0
    assert np.array_equal(f[:3], f0[:3]) and np.allclose(b[:3], b0[:3])


def test_system_move_failure_raises(ctx):
    X, T = sm.grid(3, 3)
    ctx.upload_mesh(X, T)
    ctx.set_submeshing(True, 0.5)
    s = csa.System(ctx, [8, 1], [[0.5, 0.25, 0.25], [0.3, 0.3, 0.4]])
    s.configure((csa.HARMONIC, 1.0, 0.5), 0.5, True, 0.5)
    print("case system move over the border", flush=True)
    with pytest.raises(csa.CsgError) as e:
        s.move([[5.0, 0.1, 0.0], [0.01, 0.0, 0.0]])
    assert e.value.code == csa.ERR_WALK
    f, b, _, _ = s.get()
    assert f[0] == 8 and np.allclose(b[0], [0.5, 0.25, 0.25])
    assert s.status()["walk_failures"] == 1
    s.move([[0.01, 0.0, 0.0], [0.01, 0.0, 0.0]])  # later calls report only new failures


def test_sources_beyond_the_fixed_tables(ctx, rng):
    """Patches of 2,000 to 5,000 faces (bumpy sphere L6, range 0.5) exceed the LDS tables and the step's fixed HBM
    tables (1,536 faces): the step's last tier, sized from a bound on the patch size of the configured radius, solves
    them, in csg_system_compute_forces and in an asynchronous step alike, and the forces match the oracle."""
    X, T = csa.bumpy_sphere(6, 20, 0.15, 0.3, 13377)
    f, b = wl.area_uniform(X, T, 300, rng)
    pot = (csa.HARMONIC, 4.0, 0.5)
    ctx.upload_mesh(X, T)
    ctx.set_submeshing(True, 0.5)
    s = csa.System(ctx, f, b)
    s.configure(pot, 0.5, True, 0.5)
    print("case sources beyond the fixed tables: compute_forces", flush=True)
    F, E = s.compute_forces()
    st = s.status()
    assert st["unsolved"] == 0 and st["cand_overflow"] == 0
    om = OracleMesh(X, T)
    o = OracleSystem(om, f, b, np.zeros((len(f), 3)), pot, 0.5, True, True, 0.5, threads=16)
    o.compute_forces()
    off, idx = s.pairs()
    assert np.array_equal(off, o.last["off"]) and np.array_equal(idx, o.last["idx"])
    fl = o.last["flags"]
    tie = np.array([np.any(fl[off[i]:off[i + 1]] & 2) for i in range(len(f))])
    scale = np.abs(o.F).max()
    assert np.all(np.abs(F - o.F).max(1)[~tie] <= 1e-4 * scale)
    assert abs(E - o.last["energy"]) <= 1e-6 * abs(o.last["energy"])
    print("case sources beyond the fixed tables: asynchronous steps", flush=True)
    s.nve(1e-4, 2)
    s.sync()
    assert s.status()["unsolved"] == 0


def test_candidate_overflow_reported(ctx, rng):
    """400 particles packed into a cap of radius 0.05 with range 0.2: every particle has 399 candidates, beyond the
    capacity sized from the mean density (8 times the expected count plus 64). csg_system_compute_forces and the sync
    after an asynchronous step report CSG_ERR_CAPACITY instead of returning truncated forces."""
    X, T = csa.icosphere(5)
    C = X[T].mean(1)
    near = np.where(np.linalg.norm(C - np.array([0.0, 0.0, 1.0]), axis=1) < 0.05)[0]
    f = rng.choice(near, 400).astype(np.int32)
    b = rng.dirichlet([2.0, 2.0, 2.0], 400)
    ctx.upload_mesh(X, T)
    ctx.set_submeshing(True, 0.2)
    s = csa.System(ctx, f, b)
    s.configure((csa.HARMONIC, 1.0, 0.2), 0.2, True, 0.2)
    print("case candidate overflow", flush=True)
    with pytest.raises(csa.CsgError) as e:
        s.compute_forces()
    assert e.value.code == csa.ERR_CAPACITY and "candidate" in str(e.value)
    s.nve(1e-4, 1)
    with pytest.raises(csa.CsgError) as e:
        s.sync()
    assert e.value.code == csa.ERR_CAPACITY
    assert s.status()["cand_overflow"] > 0


def _flat_ellipsoid(rng, n_side=60):
    """Icosphere L4 flattened to 3 x 3 x 0.1: its two faces are 0.2 apart near the middle and joined only at the rim,
    3 away. Particles near the middle of both faces."""
    X, T = csa.icosphere(4)
    X = X * np.array([3.0, 3.0, 0.1])
    C = X[T].mean(1)
    top = np.where((C[:, 2] > 0) & (np.abs(C[:, 0]) < 1.2) & (np.abs(C[:, 1]) < 1.2))[0]
    bot = np.where((C[:, 2] < 0) & (np.abs(C[:, 0]) < 1.2) & (np.abs(C[:, 1]) < 1.2))[0]
    f = np.concatenate([rng.choice(top, n_side), rng.choice(bot, n_side)]).astype(np.int32)
    b = rng.dirichlet([2.0, 2.0, 2.0], len(f))
    return X, T, f, b


def test_reference_gaussian_with_sentinel_pairs(ctx, rng):
    X, T, f, b = _flat_ellipsoid(rng)
    alpha, sigma, rng_, maxd = 1.0, 0.4, 1.0, 1.0
    pot = (csa.GAUSSIAN, alpha, sigma)
    om = OracleMesh(X, T)
    o = OracleSystem(om, f, b, np.zeros((len(f), 3)), pot, rng_, True, True, maxd, threads=16)
    r = o.compute_forces()
    sent = (r["flags"] & 1) != 0  # the oracle marks unreached pairs with bit 0; they carry the sentinel distance
    assert sent.sum() > 50  # pairs across the two faces
    np.testing.assert_allclose(r["dist"][sent], 2 * maxd)
    print("case reference Gaussian: system path", flush=True)
    ctx.upload_mesh(X, T)
    ctx.set_submeshing(True, maxd)
    s = csa.System(ctx, f, b)
    s.configure(pot, rng_, True, rng_)
    F, E = s.compute_forces()
    fl = r["flags"]
    off = o.last["off"]
    tie = np.array([np.any(fl[off[i]:off[i + 1]] & 2) for i in range(len(f))])
    scale = np.abs(o.F).max()
    assert np.all(np.abs(F - o.F).max(1)[~tie] <= 1e-4 * scale)
    assert abs(E - r["energy"]) <= 1e-9 * abs(r["energy"])
    print("case reference Gaussian: pair_forces path", flush=True)
    off2, idx2, md = o.candidates()
    g = ctx.pair_forces(f, b, off2, idx2, patch_radius=np.minimum(maxd, md), potential=pot, use_sentinel=True,
                        sentinel=2 * maxd)
    gs = (g["flags"] & csa.PAIR_SENTINEL) != 0
    assert np.array_equal(gs, sent)
    np.testing.assert_allclose(g["dist"], r["dist"], rtol=1e-9)
    np.testing.assert_allclose(g["start"][sent], np.tile([0.0, 0.0, 1.0], (int(sent.sum()), 1)))
    assert np.all(np.abs(g["forces"] - o.F).max(1)[~tie] <= 1e-4 * scale)
