"""Many uneven independent systems stepped by the same launches (particle groups on a union mesh): each replica's
candidate sets, forces and two velocity-Verlet steps against the oracle run on that replica alone.

Tolerances as in the single-system tests: forces 1e-4 of the replica's largest force (sources with a cut-locus pair
excluded, where the tangent is ambiguous), positions 1e-8, velocities 1e-6 of the largest speed. The union's
coordinates are the replicas' translated along x, so the candidate test and the solver see coordinates that differ by
the rounding of the translation; the candidate order differs from a lone run (one cell grid for all replicas), the
candidate sets do not."""
import numpy as np
import pytest

import curvedspacesim_amd as csa
from curvedspacesim_amd import screen, workloads
from oracle.pyoracle import OracleMesh, OracleSystem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csa.Context(0)
    yield c
    c.close()


def _replica_oracle(rep, w, threads=16):
    p0, n = rep["particle_offset"], rep["n"]
    face = w.face[p0:p0 + n] - rep["face_offset"]
    om = OracleMesh(rep["X"], rep["T"])
    s = rep["sigma"]
    osys = OracleSystem(om, face, w.bary[p0:p0 + n], w.vel[p0:p0 + n], (csa.HARMONIC, 1.0 / s ** 2, s), s, True, True, s,
                        threads=threads)
    return om, osys


def test_replicas_forces_and_steps(ctx):
    w = workloads.replicas(seed=4242, count=6, n_max=1024, kinds=(0, 1, 2))
    print(f"case replicas: {w.name}", flush=True)
    sysm = workloads.make_system(ctx, w)
    F, E = sysm.compute_forces()
    off, idx = sysm.pairs()
    st = sysm.status()
    assert st["unsolved"] == 0 and st["cand_overflow"] == 0
    oracles = []
    for r, rep in enumerate(w.meta["replicas"]):
        p0, n = rep["particle_offset"], rep["n"]
        om, osys = _replica_oracle(rep, w)
        osys.compute_forces()
        ooff, oidx = osys.last["off"], osys.last["idx"]
        for i in range(n):
            mine = np.sort(idx[off[p0 + i]:off[p0 + i + 1]] - p0)
            assert np.array_equal(mine, np.sort(oidx[ooff[i]:ooff[i + 1]])), f"replica {r} source {i} candidates"
        fl = osys.last["flags"]
        tie = np.array([np.any(fl[ooff[i]:ooff[i + 1]] & 2) for i in range(n)])
        scale = np.abs(osys.F).max()
        err = np.abs(F[p0:p0 + n] - osys.F).max(1)
        assert np.all(err[~tie] <= 1e-4 * scale), f"replica {r} ({rep['mesh']}): force error {err[~tie].max() / scale}"
        oracles.append((om, osys))
    sysm.nve(w.dt, 2)
    f, b, v, _ = sysm.get()
    for r, rep in enumerate(w.meta["replicas"]):
        om, osys = oracles[r]  # both start from the forces computed above
        osys.nve_step(w.dt)
        osys.nve_step(w.dt)
        p0, n = rep["particle_offset"], rep["n"]
        P = om.positions(f[p0:p0 + n] - rep["face_offset"], b[p0:p0 + n])
        np.testing.assert_allclose(P, om.positions(osys.face, osys.bary), atol=1e-8)
        np.testing.assert_allclose(v[p0:p0 + n], osys.vel, atol=1e-6 * np.abs(osys.vel).max())
    assert sysm.status()["walk_failures"] == 0


def test_groups_refused_beyond_the_cell_range_and_all_to_all(ctx):
    w = workloads.replicas(seed=7, count=2, n_max=300, kinds=(0,))
    sysm = workloads.make_system(ctx, w)
    with pytest.raises(csa.CsgError):  # a group range beyond the configured one (the cell size)
        sysm.set_groups(w.meta["groups"], w.meta["pots"], [2 * w.range] * 2, w.meta["maxd"])
    sysm.configure(w.potential, w.range, False, w.rmax)
    with pytest.raises(csa.CsgError):
        sysm.set_groups(w.meta["groups"], w.meta["pots"], w.meta["ranges"], w.meta["maxd"])


def _sampled_oracle_forces(rep, w, sample, threads=16):
    """The oracle's forces of the sampled sources of one replica (the other sources get empty lists; each sampled
    source keeps its candidates in order, so its force sums in the same order), and its full candidate lists."""
    om, osys = _replica_oracle(rep, w, threads)
    off, idx, md = osys.candidates()
    keep = np.zeros(len(off) - 1, bool)
    keep[sample] = True
    counts = np.where(keep, np.diff(off), 0)
    soff = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    sidx = np.concatenate([idx[off[i]:off[i + 1]] for i in np.where(keep)[0]]).astype(np.int32)
    r = om.compute_forces(osys.face, osys.bary, soff, sidx, md, osys.kind, osys.params, True, rep["sigma"], threads=threads)
    tie = np.array([np.any(r["flags"][soff[i]:soff[i + 1]] & 2) for i in range(len(keep))])
    return off, idx, r["forces"], tie


def test_32_replicas_candidates_and_forces(ctx):
    """The 32-replica workload (N log-uniform 256-8192 per replica, 50,183 particles): every replica's candidate sets,
    and the forces of 64 seeded sources per replica, against the oracle run on that replica alone (each replica is one
    curvedSpaceSimulation.cpp:72-147 run)."""
    w = workloads.config(5)
    print(f"case 32 replicas: {w.name}", flush=True)
    sysm = workloads.make_system(ctx, w)
    F, E = sysm.compute_forces()
    off, idx = sysm.pairs()
    st = sysm.status()
    assert st["unsolved"] == 0 and st["cand_overflow"] == 0
    pick = np.random.default_rng(5)
    for r, rep in enumerate(w.meta["replicas"]):
        p0, n = rep["particle_offset"], rep["n"]
        sample = np.sort(pick.choice(n, min(n, 64), replace=False))
        ooff, oidx, oF, tie = _sampled_oracle_forces(rep, w, sample)
        for i in range(n):
            mine = np.sort(idx[off[p0 + i]:off[p0 + i + 1]] - p0)
            assert np.array_equal(mine, np.sort(oidx[ooff[i]:ooff[i + 1]])), f"replica {r} source {i} candidates"
        scale = max(np.abs(oF[sample]).max(), 1e-300)
        err = np.abs(F[p0 + sample] - oF[sample]).max(1)
        ok = ~tie[sample]
        assert np.all(err[ok] <= 1e-4 * scale), f"replica {r} ({rep['mesh']}, N={n}): force error {err[ok].max() / scale}"


def test_32_replicas_screened_forces_match_plain(ctx):
    """The learned screen on the 32-replica workload (particle groups, each pair screened against its group's
    range): trained on MD snapshots of two other 32-replica sets, calibrated on two more; after 25 and 50
    velocity-Verlet steps no pair within its cutoff is dropped and the screened forces equal the plain ones."""
    print("case 32 replicas: build the screen", flush=True)
    built = screen.build(ctx, config=5, train_md_steps=(0, 20, 40), holdout_md_steps=(25, 50))
    w = workloads.config(5)
    sysm = workloads.make_system(ctx, w)
    done = 0
    for n in (25, 50):
        workloads.step(sysm, w, n - done)
        done = n
        print(f"case 32 replicas screened vs plain after {n} steps", flush=True)
        sysm.record_pairs(True)
        F0, E0 = sysm.compute_forces()
        d = sysm.pair_distances()
        X = sysm.screen_features()
        rc = screen.pair_ranges(sysm, w)
        sysm.record_pairs(False)
        drop = screen.dropped(ctx, built["params"], built["threshold"], X)
        assert not np.any(drop & (d < rc)), f"{int(np.sum(drop & (d < rc)))} within-cutoff pairs would be dropped"
        sysm.set_screen(built["params"], built["threshold"])
        F1, E1 = sysm.compute_forces()
        st = sysm.screen_stats()
        sysm.set_screen(None)
        assert st["candidates"] == len(d) and st["kept"] == int(np.sum(~drop))
        assert np.max(np.abs(F1 - F0)) <= 1e-6 * np.abs(F0).max()
        assert abs(E1 - E0) <= 1e-12 * abs(E0)
        print(f"32 replicas after {n} steps: {int(drop.sum())} of {len(d)} pairs dropped, beyond cutoff {np.mean(d >= rc):.4%}",
              flush=True)
