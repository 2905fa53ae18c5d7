"""HDF5 trajectory and checkpoint files in the layout of the reference's simpleModelDatabase, through libcsg_h5.so
(include/csg_h5.h). The library is built only where the HDF5 C library is found; available() says whether it was.

A record holds time, R3position, barycentricPosition and faceIndex, plus velocity, force and type when the file saves
them (src/databases/simpleModelDatabase.cpp:37-50); this build can add a record of the Nose-Hoover chain (nhChain,
nhParams), which the reference does not save."""
import ctypes as C
import os

import numpy as np

from ._native import CsgError, _ptr

PKG = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(PKG, "libcsg_h5.so")
READONLY, READWRITE, REPLACE = 0, 1, 2
_L = None


def available():
    return os.path.exists(LIB)


def _lib():
    global _L
    if _L is None:
        if not available():
            raise CsgError("HDF5 support is not built (libcsg_h5.so needs the HDF5 C library; see curvedspacesim_amd/build.py)")
        L = C.CDLL(LIB)
        L.csgh5_last_error.restype = C.c_char_p
        L.csgh5_open.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
        L.csgh5_close.argtypes = [C.c_void_p]
        L.csgh5_records.restype = C.c_int64
        L.csgh5_records.argtypes = [C.c_void_p]
        L.csgh5_write.argtypes = [C.c_void_p, C.c_double] + [C.c_void_p] * 6
        L.csgh5_read.restype = C.c_int64
        L.csgh5_read.argtypes = [C.c_void_p, C.c_int64, C.c_void_p] + [C.c_void_p] * 6
        L.csgh5_write_thermostat.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.csgh5_read_thermostat.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.csgh5_columns.restype = C.c_int64
        L.csgh5_columns.argtypes = [C.c_char_p, C.c_char_p]
        _L = L
    return _L


class TrajectoryDatabase:
    """simpleModelDatabase(numberOfParticles, fn, mode, saveVelocities, saveTypes, saveForces)."""

    def __init__(self, path, n, mode=READONLY, save_velocities=True, save_types=False, save_forces=False):
        self.L = _lib()
        self.N = int(n)
        self.velocity, self.type, self.force = bool(save_velocities), bool(save_types), bool(save_forces)
        h = C.c_void_p()
        if self.L.csgh5_open(os.fspath(path).encode(), self.N, int(mode), int(self.velocity), int(self.type), int(self.force),
                             C.byref(h)):
            raise CsgError(f"opening {path} failed: {self.L.csgh5_last_error().decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.csgh5_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _fail(self, what):
        raise CsgError(f"{what} failed: {self.L.csgh5_last_error().decode()}")

    @property
    def records(self):
        n = self.L.csgh5_records(self.h)
        if n < 0:
            self._fail("counting records")
        return int(n)

    def write(self, time, face, bary, r3, vel=None, force=None, types=None):
        """Appends a record (writeState)."""
        N = self.N
        f = np.ascontiguousarray(face, np.int32).reshape(N)
        b = np.ascontiguousarray(bary, np.float64).reshape(N, 3)
        p = np.ascontiguousarray(r3, np.float64).reshape(N, 3)
        v = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(N, 3)
        F = None if force is None else np.ascontiguousarray(force, np.float64).reshape(N, 3)
        t = None if types is None else np.ascontiguousarray(types, np.int32).reshape(N)
        if self.L.csgh5_write(self.h, float(time), _ptr(f), _ptr(b), _ptr(p), _ptr(v), _ptr(F), _ptr(t)):
            self._fail("writing a record")

    def write_snapshot(self, time, snap, types=None):
        """Appends a record from System.snapshot_read() (or System.checkpoint()): velocities and forces as the file
        saves them."""
        self.write(time, snap["faceIndex"], snap["barycentricPosition"], snap["R3position"],
                   snap["velocity"] if self.velocity else None, snap["force"] if self.force else None, types)

    def read(self, record=-1):
        """Record `record` (the last for -1) as a dict of the quantities the file saves (readState, plus R3position)."""
        N = self.N
        out = dict(time=np.zeros(1), faceIndex=np.zeros(N, np.int32), barycentricPosition=np.zeros((N, 3)),
                   R3position=np.zeros((N, 3)))
        if self.velocity:
            out["velocity"] = np.zeros((N, 3))
        if self.force:
            out["force"] = np.zeros((N, 3))
        if self.type:
            out["type"] = np.zeros(N, np.int32)
        r = self.L.csgh5_read(self.h, int(record), _ptr(out["time"]), _ptr(out["faceIndex"]), _ptr(out["barycentricPosition"]),
                              _ptr(out["R3position"]), _ptr(out.get("velocity")), _ptr(out.get("force")), _ptr(out.get("type")))
        if r < 0:
            self._fail("reading a record")
        out["record"] = int(r)
        return out

    def write_thermostat(self, M, chain, T, tau, kinetic):
        c = np.ascontiguousarray(chain, np.float64).reshape(-1) if int(M) > 0 else None
        if c is not None and c.size != 4 * (int(M) + 1):
            raise ValueError("the chain must hold 4 (M + 1) values")
        if self.L.csgh5_write_thermostat(self.h, int(M), _ptr(c), float(T), float(tau), float(kinetic)):
            self._fail("writing the thermostat")

    def read_thermostat(self, record=-1):
        """(chain (M + 1) x 4, params [M, T, tau, kinetic energy]); an empty chain when the file holds none."""
        params = np.zeros(4)
        M = self.L.csgh5_read_thermostat(self.h, int(record), None, _ptr(params))
        if M < 0:
            self._fail("reading the thermostat")
        chain = np.zeros((M + 1, 4)) if M > 0 else np.zeros((0, 4))
        if M > 0 and self.L.csgh5_read_thermostat(self.h, int(record), _ptr(chain), _ptr(params)) < 0:
            self._fail("reading the thermostat")
        return chain, params


def save_checkpoint(path, ckpt, append=False):
    """Writes a System.checkpoint() as a record of an HDF5 file (with velocities and forces, and the chain)."""
    n = len(ckpt["faceIndex"])
    with TrajectoryDatabase(path, n, READWRITE if append else REPLACE, True, False, True) as db:
        db.write(float(np.asarray(ckpt["time"]).ravel()[0]), ckpt["faceIndex"], ckpt["barycentricPosition"], ckpt["R3position"],
                 ckpt["velocity"], ckpt["force"])
        M, T, tau, ke = np.asarray(ckpt["nhParams"], float)
        db.write_thermostat(int(M), ckpt["nhChain"], T, tau, ke)


def load_checkpoint(path, record=-1):
    """A System.restore() argument from a record of an HDF5 file written by save_checkpoint (the file must save
    velocities and forces); the chain is that of the same record."""
    with TrajectoryDatabase(path, _particles(path), READONLY, True, False, True) as db:
        r = db.read(record)
        chain, params = db.read_thermostat(r["record"])
    return dict(time=r["time"], R3position=r["R3position"], barycentricPosition=r["barycentricPosition"], faceIndex=r["faceIndex"],
                velocity=r["velocity"], force=r["force"], nhChain=chain, nhParams=params)


def _particles(path):
    """Number of particles of a file: the columns of its faceIndex dataset."""
    L = _lib()
    n = L.csgh5_columns(os.fspath(path).encode(), b"faceIndex")
    if n < 0:
        raise CsgError(f"reading {path} failed: {L.csgh5_last_error().decode()}")
    return int(n)
