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
