"""Checkpoint and resume of the device-resident system: a run interrupted by a checkpoint (written to disk, read back
and restored into a fresh context) continues the same trajectory as an uninterrupted run. The reference restores
positions, velocities and forces (simpleModelDatabase::readState, src/databases/simpleModelDatabase.cpp:122-158);
the Nose-Hoover chain is restored as well. Tolerances as for graph replays against direct launches: positions 1e-10,
velocities 1e-9 of their scale (equal-length windows may be processed in another order)."""
import numpy as np
import pytest

import curvedspacesim_amd as csa
from curvedspacesim_amd import trajectory, workloads
from oracle.pyoracle import OracleMesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", ["npz", "h5"])
@pytest.mark.parametrize("integrator", ["nve", "nvt"])
def test_resume_continues_the_trajectory(tmp_path, integrator, fmt):
    """The checkpoint goes through a file of either format: an .npz archive, or an HDF5 file in the reference's
    simpleModelDatabase layout (plus the Nose-Hoover chain datasets)."""
    if fmt == "h5" and not trajectory.available():
        pytest.skip("libcsg_h5.so not built (no HDF5 C library)")
    w = workloads.config(2, n=2000)
    print(f"case checkpoint resume {integrator} through {fmt}", flush=True)

    def run(sysm, steps):
        if integrator == "nve":
            sysm.nve(w.dt, steps)
        else:
            sysm.nvt(w.dt, 0.5, 1.0, 2, steps)

    ca, cb, cc = csa.Context(0), csa.Context(0), csa.Context(0)
    try:
        a = workloads.make_system(ca, w)
        run(a, 6)
        fa, ba, va, Fa = a.get()

        b = workloads.make_system(cb, w)
        run(b, 3)
        ck = b.checkpoint(time=3 * w.dt)
        csa.save_checkpoint(tmp_path / f"ck.{fmt}", ck)
        cb.close()
        back = csa.load_checkpoint(tmp_path / f"ck.{fmt}")
        for k in csa._native.CHECKPOINT_FIELDS:
            np.testing.assert_array_equal(back[k], ck[k])
        assert (back["nhParams"][0] == 2) == (integrator == "nvt")

        # a fresh context, set up from the same workload, then the checkpoint's state
        c = workloads.make_system(cc, w)
        c.restore(back)
        run(c, 3)
        fc, bc, vc, Fc = c.get()
        om = OracleMesh(w.X, w.T)
        np.testing.assert_allclose(om.positions(fc, bc), om.positions(fa, ba), atol=1e-10)
        np.testing.assert_allclose(vc, va, atol=1e-9 * np.abs(va).max())
        assert c.status()["unsolved"] == 0 and c.status()["walk_failures"] == 0
    finally:
        ca.close()
        cc.close()


def test_chainless_checkpoint_clears_the_thermostat():
    """A checkpoint taken before any NVT step, restored into a system whose chain already ran, leaves no chain: the
    next NVT steps start a new one, as they do in the run the checkpoint came from."""
    w = workloads.config(2, n=2000)
    print("case checkpoint without chain", flush=True)
    ca, cb = csa.Context(0), csa.Context(0)
    try:
        a = workloads.make_system(ca, w)
        a.nve(w.dt, 2)
        ck = a.checkpoint()
        assert ck["nhParams"][0] == 0 and ck["nhChain"].shape == (0, 4)
        a.nvt(w.dt, 0.5, 1.0, 2, 3)
        fa, ba, va, _ = a.get()

        b = workloads.make_system(cb, w)
        b.nvt(w.dt, 0.5, 1.0, 2, 2)  # a chain that the restore must drop
        b.restore(ck)
        b.nvt(w.dt, 0.5, 1.0, 2, 3)
        fb, bb, vb, _ = b.get()
        om = OracleMesh(w.X, w.T)
        np.testing.assert_allclose(om.positions(fb, bb), om.positions(fa, ba), atol=1e-10)
        np.testing.assert_allclose(vb, va, atol=1e-9 * np.abs(va).max())
    finally:
        ca.close()
        cb.close()


def test_snapshot_overlaps_later_steps(tmp_path):
    """csg_system_snapshot queues the state's copy behind the steps queued so far: steps queued after it do not change
    what csg_system_snapshot_read returns, which equals the state at that point bit for bit (positions: the oracle's
    barycentric interpolation, to rounding). A later snapshot replaces it; reading before any snapshot is an error."""
    w = workloads.config(2, n=2000)
    print("case asynchronous snapshot", flush=True)
    c = csa.Context(0)
    try:
        s = workloads.make_system(c, w)
        with pytest.raises(csa.CsgError):
            s.snapshot_read()
        s.nvt(w.dt, 0.5, 1.0, 2, 3)
        fa, ba, va, Fa = s.get()
        s.snapshot()
        s.nvt(w.dt, 0.5, 1.0, 2, 4)
        snap = s.snapshot_read()
        np.testing.assert_array_equal(snap["faceIndex"], fa)
        np.testing.assert_array_equal(snap["barycentricPosition"], ba)
        np.testing.assert_array_equal(snap["velocity"], va)
        np.testing.assert_array_equal(snap["force"], Fa)
        om = OracleMesh(w.X, w.T)
        np.testing.assert_allclose(snap["R3position"], om.positions(fa, ba), atol=1e-12)
        fb, bb, vb, Fb = s.get()
        assert not np.array_equal(vb, va)
        s.snapshot()
        again = s.snapshot_read()
        np.testing.assert_array_equal(again["faceIndex"], fb)
        np.testing.assert_array_equal(again["barycentricPosition"], bb)
        np.testing.assert_array_equal(again["velocity"], vb)
        if trajectory.available():
            with trajectory.TrajectoryDatabase(tmp_path / "traj.h5", len(fa), trajectory.REPLACE, True, False, True) as db:
                db.write_snapshot(3 * w.dt, snap)
                db.write_snapshot(7 * w.dt, again)
                r0, r1 = db.read(0), db.read(1)
            np.testing.assert_array_equal(r0["faceIndex"], fa)
            np.testing.assert_array_equal(r0["R3position"], snap["R3position"])
            np.testing.assert_array_equal(r1["velocity"], vb)
            assert r1["time"][0] == 7 * w.dt
    finally:
        c.close()
