"""ORACLE — test infrastructure only.

ctypes bindings of liboracle.so, the FP64 CPU restatement of the reference's geodesic force path. Imported only by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, always as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build(quiet=True):
    """Compile the oracle libraries with make (g++)."""
    r = subprocess.run(["make", "-C", HERE, "-j4"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib(fast=False, path=None):
    name = path or ("liboracle_fast.so" if fast else "liboracle.so")
    if name in _LIBS:
        return _LIBS[name]
    if path is None:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
    L = C.CDLL(path)
    L.orc_last_error.restype = C.c_char_p
    L.orc_mesh_create.restype = C.c_void_p
    L.orc_mesh_create.argtypes = [C.c_int, _f64p, C.c_int, _i32p]
    L.orc_mesh_destroy.argtypes = [C.c_void_p]
    L.orc_mesh_counts.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 4
    L.orc_mesh_adjacency.argtypes = [C.c_void_p, _i32p, _i32p]
    L.orc_mesh_angle_sums.argtypes = [C.c_void_p, _f64p]
    L.orc_patch.argtypes = [C.c_void_p, C.c_int, _f64p, C.c_int, _i32p, C.c_double, _i32p, C.c_int]
    L.orc_geodesic.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _f64p, C.c_int, _i32p, _f64p, C.c_double,
                               _f64p, _f64p, _f64p, _i32p, C.c_void_p]
    L.orc_distance.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, _f64p, C.c_int, _i32p, _f64p, C.c_double,
                               _f64p, _f64p, _f64p, _i32p]
    L.orc_mesh_set_boundary.argtypes = [C.c_void_p, C.c_int]
    L.orc_transport.argtypes = [C.c_void_p, C.POINTER(C.c_int), _f64p, _f64p, C.c_int, C.c_void_p]
    L.orc_transport_batch.argtypes = [C.c_void_p, C.c_int, _i32p, _f64p, _f64p, C.c_int, C.c_void_p, _i32p]
    L.orc_positions.argtypes = [C.c_void_p, C.c_int, _i32p, _f64p, _f64p]
    L.orc_cell_candidates.restype = C.c_longlong
    L.orc_cell_candidates.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_double, C.c_double, _i32p, _i32p, _f64p,
                                      C.c_longlong]
    L.orc_pair_force.restype = C.c_double
    L.orc_pair_force.argtypes = [C.c_int, _f64p, C.c_double, C.POINTER(C.c_double)]
    L.orc_compute_forces.argtypes = [C.c_void_p, C.c_int, _i32p, _f64p, C.c_int, C.c_double, _i32p, _i32p, C.c_void_p,
                                     C.c_int, _f64p, _f64p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_double), C.c_int]
    L.orc_move.argtypes = [C.c_void_p, C.c_int, _i32p, _f64p, _f64p, C.c_void_p, C.c_int]
    _LIBS[name] = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleMesh:
    """Mesh with adjacency, owned by the oracle library."""

    def __init__(self, xyz, tri, fast=False, libpath=None):
        self.L = lib(fast, libpath)
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
        h = self.L.orc_mesh_create(len(self.xyz), self.xyz.ravel(), len(self.tri), self.tri.ravel())
        if not h:
            raise ValueError(self.L.orc_last_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_mesh_destroy(self.h)
            self.h = None

    @property
    def nF(self):
        return len(self.tri)

    @property
    def nV(self):
        return len(self.xyz)

    def counts(self):
        v = [C.c_int() for _ in range(4)]
        self.L.orc_mesh_counts(self.h, *[C.byref(x) for x in v])
        return dict(nV=v[0].value, nF=v[1].value, border=v[2].value, saddle=v[3].value)

    def adjacency(self):
        a = np.zeros(3 * self.nF, np.int32)
        e = np.zeros(3 * self.nF, np.int32)
        self.L.orc_mesh_adjacency(self.h, a, e)
        return a.reshape(-1, 3), e.reshape(-1, 3)

    def angle_sums(self):
        out = np.zeros(self.nV, np.float64)
        self.L.orc_mesh_angle_sums(self.h, out)
        return out

    def positions(self, face, bary):
        face = np.ascontiguousarray(face, np.int32)
        bary = np.ascontiguousarray(bary, np.float64).reshape(-1, 3)
        out = np.zeros((len(face), 3))
        self.L.orc_positions(self.h, len(face), face, bary.ravel(), out.ravel())
        return out

    def patch(self, src_face, src_bary, tgt_faces, radius):
        tgt = np.ascontiguousarray(tgt_faces, np.int32)
        out = np.zeros(self.nF, np.int32)
        n = self.L.orc_patch(self.h, int(src_face), np.ascontiguousarray(src_bary, np.float64), len(tgt), tgt,
                             float(radius), out, self.nF)
        if n < 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return out[:n].copy()

    def geodesic(self, src_face, src_bary, tgt_face, tgt_bary, patch=None, rmax=np.inf, stats=False):
        tgt_face = np.ascontiguousarray(tgt_face, np.int32)
        tgt_bary = np.ascontiguousarray(tgt_bary, np.float64).reshape(-1, 3)
        n = len(tgt_face)
        d = np.zeros(n)
        st = np.zeros((n, 3))
        et = np.zeros((n, 3))
        fl = np.zeros(n, np.int32)
        st_ = np.zeros(4, np.int64)
        if patch is None:
            pp, npatch = None, -1
        else:
            patch = np.ascontiguousarray(patch, np.int32)
            pp, npatch = _ptr(patch), len(patch)
        rc = self.L.orc_geodesic(self.h, pp, npatch, int(src_face), np.ascontiguousarray(src_bary, np.float64), n,
                                 tgt_face, tgt_bary.ravel(), float(rmax), d, st.ravel(), et.ravel(), fl, _ptr(st_))
        if rc:
            raise RuntimeError(f"orc_geodesic failed ({rc}): {self.L.orc_last_error().decode()}")
        if stats:
            return d, st, et, fl, dict(created=int(st_[0]), processed=int(st_[1]), vertex_events=int(st_[2]))
        return d, st, et, fl

    def distance(self, src_face, src_bary, tgt_face, tgt_bary, submeshing, maximum_distance, threshold=1e20):
        tgt_face = np.ascontiguousarray(tgt_face, np.int32)
        tgt_bary = np.ascontiguousarray(tgt_bary, np.float64).reshape(-1, 3)
        n = len(tgt_face)
        d = np.zeros(n)
        st = np.zeros((n, 3))
        et = np.zeros((n, 3))
        fl = np.zeros(n, np.int32)
        rc = self.L.orc_distance(self.h, int(bool(submeshing)), float(maximum_distance), int(src_face),
                                 np.ascontiguousarray(src_bary, np.float64), n, tgt_face, tgt_bary.ravel(),
                                 float(threshold), d, st.ravel(), et.ravel(), fl)
        if rc:
            raise RuntimeError(f"orc_distance failed ({rc}): {self.L.orc_last_error().decode()}")
        return d, st, et, fl

    BOUNDARY_MODES = {"closed": 0, "absorbing": 1, "tangential": 2}

    def set_boundary(self, mode):
        """Walk behaviour at the mesh border: "closed" (an error), "absorbing" or "tangential"."""
        if self.L.orc_mesh_set_boundary(self.h, self.BOUNDARY_MODES[mode]):
            raise ValueError(mode)

    def transport(self, face, bary, disp, vecs=None):
        """Walk many particles; returns (face, bary, disp, vecs, status)."""
        face = np.array(face, np.int32).ravel().copy()
        n = len(face)
        bary = np.array(bary, np.float64).reshape(n, 3).copy()
        disp = np.array(disp, np.float64).reshape(n, 3).copy()
        nvec = 0
        if vecs is not None:
            vecs = np.array(vecs, np.float64).reshape(n, -1, 3).copy()
            nvec = vecs.shape[1]
        status = np.zeros(n, np.int32)
        self.L.orc_transport_batch(self.h, n, face, bary.ravel(), disp.ravel(), nvec,
                                   _ptr(vecs) if vecs is not None else None, status)
        return face, bary, disp, vecs, status

    def compute_forces(self, face, bary, offsets, idx, maxdist, kind, params, submeshing, maximum_distance,
                       threads=1, want_pairs=True):
        face = np.ascontiguousarray(face, np.int32)
        bary = np.ascontiguousarray(bary, np.float64).reshape(-1, 3)
        N = len(face)
        offsets = np.ascontiguousarray(offsets, np.int32)
        idx = np.ascontiguousarray(idx, np.int32)
        P = len(idx)
        forces = np.zeros((N, 3))
        pd = np.zeros(P) if want_pairs else None
        pt = np.zeros((P, 3)) if want_pairs else None
        pe = np.zeros((P, 3)) if want_pairs else None
        pf = np.zeros(P, np.int32) if want_pairs else None
        md = None if maxdist is None else np.ascontiguousarray(maxdist, np.float64)
        e = C.c_double()
        rc = self.L.orc_compute_forces(self.h, N, face, bary.ravel(), int(bool(submeshing)), float(maximum_distance),
                                       offsets, idx, _ptr(md), int(kind),
                                       np.ascontiguousarray(params, np.float64), forces.ravel(), _ptr(pd), _ptr(pt),
                                       _ptr(pe), _ptr(pf), C.byref(e), int(threads))
        if rc:
            raise RuntimeError(f"orc_compute_forces failed ({rc}): {self.L.orc_last_error().decode()}")
        return dict(forces=forces, dist=pd, tan=pt, end=pe, flags=pf, energy=e.value)

    def move(self, face, bary, disp, vel=None, threads=1):
        face = np.ascontiguousarray(face, np.int32).copy()
        bary = np.ascontiguousarray(bary, np.float64).reshape(-1, 3).copy()
        v = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(-1, 3).copy()
        bad = self.L.orc_move(self.h, len(face), face, bary.ravel(), np.ascontiguousarray(disp, np.float64).ravel(),
                              _ptr(v), int(threads))
        return face, bary, v, bad


def cell_candidates(pos, lo, hi, grid, rng, cap=None):
    """Reference cell-list candidates: returns (offsets, idx, maxdist)."""
    L = lib()
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    N = len(pos)
    cap = cap or max(16, N * 64)
    while True:
        off = np.zeros(N + 1, np.int32)
        idx = np.zeros(cap, np.int32)
        md = np.zeros(N)
        tot = L.orc_cell_candidates(N, pos.ravel(), np.asarray(lo, np.float64), np.asarray(hi, np.float64),
                                    float(grid), float(rng), off, idx, md, cap)
        if tot >= 0:
            return off, idx[:tot].copy(), md
        cap = -tot


def pair_force(kind, params, d):
    L = lib()
    sc = C.c_double()
    e = L.orc_pair_force(int(kind), np.ascontiguousarray(params, np.float64), float(d), C.byref(sc))
    return e, sc.value


class OracleSystem:
    """Reference control flow of the integrators over the oracle's forces and walk (test infrastructure).

    velocityVerletNVE::performUpdate (src/updaters/velocityVerletNVE.cpp:3-29) and noseHooverNVT::performUpdate with
    propagateChain / propagatePositionsVelocities (src/updaters/noseHooverNVT.cpp:3-150), including the chain
    initialisation of setBathVariables (kinetic energy set to Q0).
    """

    def __init__(self, mesh, face, bary, vel, potential, rng_range, cell_list, submeshing, maximum_distance,
                 mass=None, threads=8):
        self.m = mesh
        self.face = np.array(face, np.int32).copy()
        self.bary = np.array(bary, np.float64).reshape(-1, 3).copy()
        self.vel = np.array(vel, np.float64).reshape(-1, 3).copy()
        self.N = len(self.face)
        self.F = np.zeros((self.N, 3))
        self.mass = np.ones(self.N) if mass is None else np.asarray(mass, float)
        self.kind = int(potential[0])
        self.params = np.array(list(potential[1:]) + [0.0] * (3 - len(potential[1:])), float)
        self.range = rng_range
        self.cell_list = cell_list
        self.submeshing = submeshing
        self.maximum_distance = maximum_distance
        self.threads = threads
        X = mesh.xyz
        self.lo, self.hi = np.minimum(X.min(0), 0), np.maximum(X.max(0), 0)
        self.chain = None

    def candidates(self):
        P = self.m.positions(self.face, self.bary)
        if self.cell_list:
            return cell_candidates(P, self.lo, self.hi, self.range, self.range)
        N = self.N
        off = np.arange(0, N * (N - 1) + 1, N - 1, dtype=np.int32)
        idx = np.array([j for i in range(N) for j in range(N) if j != i], np.int32)
        return off, idx, np.full(N, 1e20)

    def compute_forces(self):
        off, idx, md = self.candidates()
        r = self.m.compute_forces(self.face, self.bary, off, idx, md, self.kind, self.params, self.submeshing,
                                  self.maximum_distance, threads=self.threads)
        self.F = r["forces"]
        self.last = dict(off=off, idx=idx, **r)
        return r

    def move(self, disp):
        self.face, self.bary, self.vel, bad = self.m.move(self.face, self.bary, disp, self.vel, threads=self.threads)
        if bad:
            raise RuntimeError(f"{bad} walks failed")

    def nve_step(self, dt):
        disp = dt * self.vel + 0.5 * dt * dt * self.F
        self.vel = self.vel + 0.5 * dt * self.F
        self.move(disp)
        self.compute_forces()
        self.vel = self.vel + 0.5 * dt * self.F

    def _chain_half(self):
        b, M, T, dt = self.chain, self.M, self.T, self.dt
        dt2, dt4, dt8 = 0.5 * dt, 0.25 * dt, 0.125 * dt
        for ii in range(M - 1, 0, -1):
            b[ii, 2] = (b[ii - 1, 3] * b[ii - 1, 1] ** 2 - T) / b[ii, 3]
            ef = np.exp(-dt8 * b[ii + 1, 1])
            b[ii, 1] = (b[ii, 1] * ef + b[ii, 2] * dt4) * ef
        b[0, 2] = 2.0 * self.KE / b[0, 3] - 1.0
        ef = np.exp(-dt8 * b[1, 1])
        b[0, 1] = (b[0, 1] * ef + b[0, 2] * dt4) * ef
        for ii in range(M):
            b[ii, 0] += dt2 * b[ii, 1]
        s = np.exp(-dt2 * b[0, 1])
        self.KE = s * s * self.KE
        b[0, 2] = 2.0 * self.KE / b[0, 3] - 1.0
        ef = np.exp(-dt8 * b[1, 1])
        b[0, 1] = (b[0, 1] * ef + b[0, 2] * dt4) * ef
        for ii in range(1, M):
            b[ii, 2] = (b[ii - 1, 3] * b[ii - 1, 1] ** 2 - T) / b[ii, 3]
            ef = np.exp(-dt8 * b[ii + 1, 1])
            b[ii, 1] = (b[ii, 1] * ef + b[ii, 2] * dt4) * ef
        return s

    def nvt_step(self, dt, T, tau=1.0, M=2):
        if self.chain is None:
            self.M, self.T, self.dt = M, T, dt
            self.chain = np.zeros((M + 1, 4))
            self.chain[0, 3] = 2.0 * (self.N - 1) * T * tau * tau
            self.chain[1:, 3] = T * tau * tau
            self.KE = self.chain[0, 3]
        dt2 = 0.5 * dt
        self.vel = self.vel * self._chain_half()
        self.move(dt2 * self.vel)
        self.compute_forces()
        self.vel = self.vel + (dt / self.mass)[:, None] * self.F
        self.KE = float((0.5 * self.mass * (self.vel * self.vel).sum(1)).sum())
        self.move(dt2 * self.vel)
        self.vel = self.vel * self._chain_half()
