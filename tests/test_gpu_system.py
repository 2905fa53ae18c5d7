"""Parity of the device-resident step (cell list, fused forces, integrators) with the oracle's restatement of the
reference control flow, from identical states.

Tolerances: candidate lists identical (order included); forces 1e-4 of the largest force (the kernel's start tangents
carry FP32 frame rotations); one integrator step: positions 1e-8 and velocities 1e-6 relative to their scale."""
import numpy as np
import pytest

import curvedspacesim_amd as csa
from curvedspacesim_amd import workloads
from oracle.pyoracle import OracleMesh, OracleSystem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csa.Context(0)
    yield c
    c.close()


def _oracle(w, threads=16):
    om = OracleMesh(w.X, w.T)
    return om, OracleSystem(om, w.face, w.bary, w.vel, w.potential, w.range, w.cell_list, w.submeshing,
                            w.maximum_distance, threads=threads)


def _tie_sources(osys):
    fl = osys.last["flags"]
    off = osys.last["off"]
    return np.array([np.any(fl[off[i]:off[i + 1]] & 2) for i in range(osys.N)])


# N = 8192 fills exactly two 4,096-entry tiles of the candidate-offset scan; 20000 ends in a partial tile. Config 1
# (whole mesh, truncated) is solved over patches of the truncation radius; at N = 256 its 255 targets per source
# exceed the first LDS tables and take the larger tiers
@pytest.mark.parametrize("cfg,n", [(2, 2000), (3, 4000), (1, 64), (1, 256), (3, 20000), (2, 8192)])
def test_candidates_and_forces(ctx, cfg, n):
    w = workloads.config(cfg, n=n)
    print(f"case system forces config {cfg} N={n}", flush=True)
    sysm = workloads.make_system(ctx, w)
    F, E = sysm.compute_forces()
    off, idx = sysm.pairs()
    om, osys = _oracle(w)
    r = osys.compute_forces()
    assert np.array_equal(off, osys.last["off"]) and np.array_equal(idx, osys.last["idx"])
    tie = _tie_sources(osys)
    scale = np.abs(osys.F).max()
    bad = np.where(~tie & (np.abs(F - osys.F).max(1) > 1e-4 * scale))[0]
    for i in bad[:5]:
        print(f"source {i}: gpu {F[i]} oracle {osys.F[i]}", flush=True)
    assert len(bad) == 0
    st = sysm.status()
    assert st["unsolved"] == 0 and st["cand_overflow"] == 0 and st["pairs"] == off[-1]


def test_whole_mesh_distances_over_patches(ctx):
    """Whole-mesh mode with a finite truncation radius is solved over patches of that radius plus the longest edge:
    every pair within the radius gets the distance of the oracle's whole-mesh solve (no patch, no truncation), to
    1e-9 relative; the pairs beyond it carry no force either way."""
    w = workloads.config(1, n=256)
    print("case whole-mesh distances over patches config 1 N=256", flush=True)
    sysm = workloads.make_system(ctx, w)
    sysm.record_pairs(True)
    F, E = sysm.compute_forces()
    off, idx = sysm.pairs()
    d = sysm.pair_distances()
    sysm.record_pairs(False)
    om, osys = _oracle(w)
    r = osys.compute_forces()
    assert np.array_equal(off, osys.last["off"]) and np.array_equal(idx, osys.last["idx"])
    within = r["dist"] < w.range
    assert within.sum() > 500
    np.testing.assert_allclose(d[within], r["dist"][within], rtol=1e-9, atol=0)
    assert np.all(d[~within] >= w.range)  # beyond the radius: unreached, or a path no shorter than the whole mesh's
    assert sysm.status()["unsolved"] == 0


@pytest.mark.parametrize("cfg,n", [(2, 2000), (3, 4000)])
def test_one_step_matches_oracle(ctx, cfg, n):
    w = workloads.config(cfg, n=n)
    print(f"case system step config {cfg} N={n} integrator {w.integrator}", flush=True)
    sysm = workloads.make_system(ctx, w)
    om, osys = _oracle(w)
    if w.integrator == "nve":
        # the reference starts velocity Verlet with zero forces; take one step, then compare a second one
        sysm.nve(w.dt, 1)
        osys.nve_step(w.dt)
        sysm.nve(w.dt, 1)
        osys.nve_step(w.dt)
    else:
        sysm.nvt(w.dt, w.temperature, w.tau, w.chain, 1)
        osys.nvt_step(w.dt, w.temperature, w.tau, w.chain)
    f, b, v, F = sysm.get()
    P = om.positions(f, b)
    Po = om.positions(osys.face, osys.bary)
    np.testing.assert_allclose(P, Po, atol=1e-8)
    vs = np.abs(osys.vel).max()
    np.testing.assert_allclose(v, osys.vel, atol=1e-6 * vs)
    st = sysm.status()
    assert st["walk_failures"] == 0 and st["unsolved"] == 0
    if w.integrator == "nvt":
        assert abs(st["kinetic"] - osys.KE) < 1e-6 * osys.KE


@pytest.mark.parametrize("cfg,n", [(2, 2000), (3, 4000)])
def test_forces_at_an_evolved_state(ctx, cfg, n):
    """Forces after MD steps (particles moved off their initial faces, the step buffers reused from step to step): the
    patches hold as many faces and vertices, summed over the sources, as those of a system set up in a new context
    from the same state, and candidates and forces match the oracle's at that state."""
    w = workloads.config(cfg, n=n)
    print(f"case forces at an evolved state config {cfg} N={n}", flush=True)
    sysm = workloads.make_system(ctx, w)
    sysm.nvt(w.dt, w.temperature, w.tau, w.chain, 6)
    f, b, v, _ = sysm.get()
    F, E = sysm.compute_forces()
    off, idx = sysm.pairs()
    reused = sysm.patch_stats()
    c2 = csa.Context(0)
    try:
        fresh = workloads.make_system(c2, w)
        fresh.set_state(f, b, v)
        F2, E2 = fresh.compute_forces()
        ps = fresh.patch_stats()
    finally:
        c2.close()
    assert reused["faces"] == ps["faces"] and reused["verts"] == ps["verts"] and reused["max_faces"] == ps["max_faces"]
    om = OracleMesh(w.X, w.T)
    osys = OracleSystem(om, f, b, v, w.potential, w.range, w.cell_list, w.submeshing, w.maximum_distance, threads=16)
    osys.compute_forces()
    assert np.array_equal(off, osys.last["off"]) and np.array_equal(idx, osys.last["idx"])
    tie = _tie_sources(osys)
    scale = np.abs(osys.F).max()
    for got in (F, F2):
        bad = np.where(~tie & (np.abs(got - osys.F).max(1) > 1e-4 * scale))[0]
        assert len(bad) == 0
    assert sysm.status()["unsolved"] == 0


def test_sources_beyond_every_lds_tier(ctx):
    """Patches of about 2,000 faces (harmonic sigma = 0.6 on a bumpy L5 sphere) exceed the first LDS tables and both
    larger tiers (1,024 and 1,536 faces); the HBM-table kernel solves them, each of its workgroups testing chunks of 64
    statuses for the sources left unsolved. Candidates and forces match the oracle's and nothing is left unsolved."""
    rng = np.random.default_rng(21)
    X, T = csa.bumpy_sphere(5, 20, 0.15, 0.3, 5)
    n, sigma = 300, 0.6
    f, b = workloads.area_uniform(X, T, n, rng)
    w = workloads.Workload("bumpy sphere L5, N=300, harmonic sigma=0.6", X, T, f, b, np.zeros((n, 3)),
                           (csa.HARMONIC, 1.0 / sigma ** 2, sigma), sigma, True, True, sigma, sigma, "nve", 0.005, 0.1)
    print("case sources beyond every LDS tier, bumpy sphere L5 N=300 sigma=0.6", flush=True)
    sysm = workloads.make_system(ctx, w)
    F, E = sysm.compute_forces()
    off, idx = sysm.pairs()
    assert sysm.patch_stats()["max_faces"] > 1536  # only the HBM tables hold such a patch
    om, osys = _oracle(w)
    osys.compute_forces()
    assert np.array_equal(off, osys.last["off"]) and np.array_equal(idx, osys.last["idx"])
    tie = _tie_sources(osys)
    bad = np.where(~tie & (np.abs(F - osys.F).max(1) > 1e-4 * np.abs(osys.F).max()))[0]
    assert len(bad) == 0
    assert sysm.status()["unsolved"] == 0


def test_nvt_full_size_runs(ctx):
    """The N = 20,000 workload for 20 steps: no errors, and the chain's kinetic energy equals that of the final
    velocities (the last chain half-step scales both by the same factor, noseHooverNVT.cpp:46-58)."""
    w = workloads.config(3)
    sysm = workloads.make_system(ctx, w)
    sysm.nvt(w.dt, w.temperature, w.tau, w.chain, 20)
    st = sysm.status()
    assert st["walk_failures"] == 0 and st["unsolved"] == 0 and st["cand_overflow"] == 0
    f, b, v, F = sysm.get()
    ke = 0.5 * (v * v).sum()
    assert np.isfinite(ke) and abs(st["kinetic"] - ke) < 1e-9 * ke


@pytest.mark.parametrize("cfg,n", [(3, 4000), (2, 2000)])
def test_lds_variants_and_autotune(ctx, cfg, n):
    """Every variant of the LDS kernel gives the same forces (to the sum order of a source's pairs and FP32 tangents of
    tied windows), and the autotuner keeps one of them: on the sphere's patches of about 380 faces and on the torus's
    of about 70."""
    w = workloads.config(cfg, n=n)
    print(f"case lds variants config {cfg} N={n}", flush=True)
    sysm = workloads.make_system(ctx, w)
    out = []
    for v in range(csa.LDS_VARIANTS):
        sysm.set_lds_variant(v)
        out.append(sysm.compute_forces())
    scale = np.abs(out[0][0]).max()
    for F, E in out[1:]:
        assert np.max(np.abs(F - out[0][0])) <= 1e-6 * scale
        assert abs(E - out[0][1]) <= 1e-12 * abs(out[0][1])
    chosen, ms = sysm.autotune()
    assert 0 <= chosen < csa.LDS_VARIANTS and len(ms) == csa.LDS_VARIANTS and np.all(ms > 0)
    assert sysm.status()["unsolved"] == 0


def test_autotune_cache(ctx, tmp_path):
    """The autotuner's JSON cache: a first system is timed and recorded; a second system of the same workload takes the
    recorded choice without timing (the same times come back, and the same forces); another particle count is a
    different key."""
    cache = str(tmp_path / "tuning.json")
    w = workloads.config(3, n=3000)
    print("case autotune cache config 3 N=3000", flush=True)
    sysm = workloads.make_system(ctx, w)
    chosen, ms = sysm.autotune(cache=cache)
    key = sysm.tuning_key()
    assert csa._native.lookup_tuning(cache, key)[0] == chosen
    F0, _ = sysm.compute_forces()
    again = workloads.make_system(ctx, w)
    assert again.tuning_key() == key
    c2, ms2 = again.autotune(cache=cache)
    assert c2 == chosen and np.array_equal(ms2, ms)
    F1, _ = again.compute_forces()
    assert np.max(np.abs(F1 - F0)) <= 1e-6 * np.abs(F0).max()  # tied windows' FP32 tangents may differ between runs
    other = workloads.make_system(ctx, workloads.config(3, n=2000))
    assert other.tuning_key() != key and csa._native.lookup_tuning(cache, other.tuning_key()) is None


@pytest.mark.parametrize("integrator", ["nve", "nvt"])
def test_graph_replays_match_eager(integrator):
    """Graph mode: one captured step replayed several times, with host reads and synchronisation between the replays,
    against the same steps launched directly on an identical system (separate contexts). Tolerances: positions 1e-10,
    velocities 1e-9 of their scale (the two runs may order equal-length window work differently)."""
    print(f"case graph replays {integrator}", flush=True)
    w = workloads.config(2, n=2000)
    ca, cb = csa.Context(0), csa.Context(0)
    try:
        a = workloads.make_system(ca, w)
        b = workloads.make_system(cb, w)
        b.use_graph(True)
        om = OracleMesh(w.X, w.T)
        for rep in range(3):
            if integrator == "nve":
                a.nve(w.dt, 2)
                b.nve(w.dt, 2)
            else:
                a.nvt(w.dt, 0.5, 1.0, 2, 2)
                b.nvt(w.dt, 0.5, 1.0, 2, 2)
            fa, ba, va, Fa = a.get()
            fb, bb, vb, Fb = b.get()
            np.testing.assert_allclose(om.positions(fb, bb), om.positions(fa, ba), atol=1e-10)
            np.testing.assert_allclose(vb, va, atol=1e-9 * np.abs(va).max())
        assert b.status()["unsolved"] == 0 and b.status()["walk_failures"] == 0
    finally:
        ca.close()
        cb.close()
