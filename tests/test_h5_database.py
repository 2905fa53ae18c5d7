"""Trajectory files in the layout of the reference's simpleModelDatabase (src/databases/simpleModelDatabase.cpp:8-158)
through libcsg_h5.so, on the CPU: records written and read back bit for bit, appended in readwrite mode, the
Nose-Hoover chain records, the reference's error cases, and the file's layout read by a separate program that uses the
HDF5 C API alone (tests/cpp/h5_layout_check.cpp): two-dimensional datasets, rows unlimited, chunks of one row, native
int and double, as the reference's baseHDF5Database reads them. Skipped where the HDF5 library is not built."""
import os
import subprocess

import numpy as np
import pytest

import curvedspacesim_amd as csa
from curvedspacesim_amd import build, trajectory

pytestmark = pytest.mark.skipif(not trajectory.available(), reason="libcsg_h5.so not built (no HDF5 C library)")

HERE = os.path.dirname(os.path.abspath(__file__))


def _state(rng, n):
    return dict(face=rng.integers(0, 1000, n).astype(np.int32), bary=rng.dirichlet([1, 1, 1], n), r3=rng.normal(size=(n, 3)),
                vel=rng.normal(size=(n, 3)), force=rng.normal(size=(n, 3)), types=rng.integers(0, 3, n).astype(np.int32))


def test_write_snapshot_saves_what_the_file_keeps(tmp_path, rng):
    """write_snapshot takes the dict of System.snapshot_read(): velocities and forces go in when the file saves them
    and are left out when it does not."""
    n = 23
    s = _state(rng, n)
    snap = dict(faceIndex=s["face"], barycentricPosition=s["bary"], R3position=s["r3"], velocity=s["vel"], force=s["force"])
    with trajectory.TrajectoryDatabase(tmp_path / "all.h5", n, trajectory.REPLACE, True, False, True) as db:
        db.write_snapshot(0.25, snap)
        r = db.read(0)
    assert r["time"][0] == 0.25
    for k, v in (("faceIndex", s["face"]), ("barycentricPosition", s["bary"]), ("R3position", s["r3"]), ("velocity", s["vel"]),
                 ("force", s["force"])):
        np.testing.assert_array_equal(r[k], v)
    with trajectory.TrajectoryDatabase(tmp_path / "positions.h5", n, trajectory.REPLACE, False, False, False) as db:
        db.write_snapshot(0.5, snap)
        r = db.read(0)
    assert "velocity" not in r and "force" not in r
    np.testing.assert_array_equal(r["R3position"], s["r3"])


def test_round_trip_and_append(tmp_path, rng):
    path = tmp_path / "traj.h5"
    n = 37
    states = [_state(rng, n) for _ in range(4)]
    with trajectory.TrajectoryDatabase(path, n, trajectory.REPLACE, True, True, True) as db:
        for k, s in enumerate(states[:3]):
            db.write(0.5 * k, s["face"], s["bary"], s["r3"], s["vel"], s["force"], s["types"])
        assert db.records == 3
    with trajectory.TrajectoryDatabase(path, n, trajectory.READWRITE, True, True, True) as db:
        db.write(1.5, states[3]["face"], states[3]["bary"], states[3]["r3"], states[3]["vel"], states[3]["force"], states[3]["types"])
        assert db.records == 4
    with trajectory.TrajectoryDatabase(path, n, trajectory.READONLY, True, True, True) as db:
        for k, s in enumerate(states):
            r = db.read(k)
            assert r["time"][0] == 0.5 * k and r["record"] == k
            np.testing.assert_array_equal(r["faceIndex"], s["face"])
            np.testing.assert_array_equal(r["barycentricPosition"], s["bary"])
            np.testing.assert_array_equal(r["R3position"], s["r3"])
            np.testing.assert_array_equal(r["velocity"], s["vel"])
            np.testing.assert_array_equal(r["force"], s["force"])
            np.testing.assert_array_equal(r["type"], s["types"])
        assert db.read()["record"] == 3  # the last record by default (readDataset's record = -1)


def test_thermostat_records(tmp_path, rng):
    path = tmp_path / "nh.h5"
    s = _state(rng, 5)
    chain = rng.normal(size=(3, 4))
    with trajectory.TrajectoryDatabase(path, 5, trajectory.REPLACE, True, False, True) as db:
        db.write(0.0, s["face"], s["bary"], s["r3"], s["vel"], s["force"])
        c0, p0 = db.read_thermostat()
        assert len(c0) == 0 and p0[0] == 0  # no thermostat record yet
        db.write_thermostat(2, chain, 0.5, 1.0, 7.25)
        c, p = db.read_thermostat()
    np.testing.assert_array_equal(c, chain)
    np.testing.assert_array_equal(p, [2, 0.5, 1.0, 7.25])


def test_checkpoint_round_trip_through_hdf5(tmp_path, rng):
    n, M = 11, 2
    ck = dict(time=np.array([0.125]), R3position=rng.normal(size=(n, 3)), barycentricPosition=rng.dirichlet([1, 1, 1], n),
              faceIndex=rng.integers(0, 50, n).astype(np.int32), velocity=rng.normal(size=(n, 3)), force=rng.normal(size=(n, 3)),
              nhChain=rng.normal(size=(M + 1, 4)), nhParams=np.array([M, 0.5, 1.0, 3.0]))
    csa.save_checkpoint(tmp_path / "ck.h5", ck)
    back = csa.load_checkpoint(tmp_path / "ck.h5")
    for k in ck:
        np.testing.assert_array_equal(back[k], ck[k])


def test_reference_error_cases(tmp_path, rng):
    with pytest.raises(csa.CsgError, match="does not exist"):
        trajectory.TrajectoryDatabase(tmp_path / "absent.h5", 4)
    path = tmp_path / "e.h5"
    s = _state(rng, 4)
    with trajectory.TrajectoryDatabase(path, 4, trajectory.REPLACE) as db:
        with pytest.raises(csa.CsgError, match="past the end"):
            db.read(0)
        with pytest.raises(csa.CsgError, match="every saved quantity"):
            db.write(0.0, s["face"], s["bary"], s["r3"])  # velocities are saved by default
        db.write(0.0, s["face"], s["bary"], s["r3"], s["vel"])
        with pytest.raises(csa.CsgError, match="past the end"):
            db.read(1)
    with trajectory.TrajectoryDatabase(path, 4, trajectory.READONLY) as db:
        with pytest.raises(csa.CsgError, match="readonly"):
            db.write(0.0, s["face"], s["bary"], s["r3"], s["vel"])
    with pytest.raises(csa.CsgError, match="another number of particles"):
        trajectory.TrajectoryDatabase(path, 5, trajectory.READONLY)


def test_layout_read_with_the_hdf5_c_api(tmp_path, rng):
    path = tmp_path / "layout.h5"
    n = 9
    s = _state(rng, n)
    with trajectory.TrajectoryDatabase(path, n, trajectory.REPLACE, True, True, True) as db:
        for k in range(2):
            db.write(float(k), s["face"], s["bary"], s["r3"], s["vel"], s["force"], s["types"])
    exe = str(tmp_path / "h5_layout_check")
    inc, lib = os.path.join(build.HDF5_DIR, "include"), os.path.join(build.HDF5_DIR, "lib")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + inc, os.path.join(HERE, "cpp", "h5_layout_check.cpp"), "-L" + lib, "-lhdf5",
                    "-Wl,-rpath," + lib, "-o", exe], check=True)
    names = ["time", "R3position", "barycentricPosition", "faceIndex", "velocity", "force", "type"]
    out = subprocess.run([exe, str(path)] + names, capture_output=True, text=True, check=True).stdout.split("\n")
    got = {line.split()[0]: line.split()[1:] for line in out if line}
    cols = dict(time=1, R3position=3 * n, barycentricPosition=3 * n, faceIndex=n, velocity=3 * n, force=3 * n, type=n)
    for name in names:
        kind = "int" if name in ("faceIndex", "type") else "double"
        assert got[name] == ["2", "2", str(cols[name]), "1", "1", str(cols[name]), kind], (name, got[name])
