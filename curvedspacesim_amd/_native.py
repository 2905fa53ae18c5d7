"""ctypes bindings of libcsg.so (C-ABI of the HIP geodesic force path) and libcsg_host.so (C++ host layer).

Both libraries are built in-tree by curvedspacesim_amd/build.py. Loading fails loudly when they are missing: there is
no Python or CPU fallback for anything in this package.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
_csg = None
_host = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
VP = C.c_void_p


class CsgError(RuntimeError):
    """A C-ABI call failed; `code` is its csg_status (CSG_ERR_* of include/csg.h)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class csg_potential(C.Structure):
    _fields_ = [("kind", C.c_int32), ("p0", C.c_double), ("p1", C.c_double), ("p2", C.c_double)]


HARMONIC, GAUSSIAN, TRUNC_GAUSSIAN, NO_POTENTIAL = 0, 1, 2, 3
# csg_status codes (include/csg.h)
ERR_ARGUMENT, ERR_DEVICE, ERR_MESH, ERR_STATE, ERR_WALK, ERR_CAPACITY = -1, -2, -3, -4, -5, -6
PAIR_UNREACHED, PAIR_BEYOND, PAIR_SENTINEL = 1, 4, 16
SCREEN_FEATURES, SCREEN_PARAMS = 16, 5313
# variants of the LDS geodesic kernel (CSG_LDS_VARIANTS), in the library's order
LDS_VARIANT_NAMES = ["256_lanes_2_per_cu", "512_lanes_2_per_cu", "256_lanes_3_per_cu", "256_lanes_4_per_cu"]
LDS_VARIANTS = len(LDS_VARIANT_NAMES)


def _path(name):
    p = os.path.join(PKG, name)
    if not os.path.exists(p):
        raise CsgError(f"{name} is not built; run curvedspacesim_amd/build.py (or __graft_entry__.build())")
    return p


def csg():
    """libcsg.so with argument types declared."""
    global _csg
    if _csg is not None:
        return _csg
    L = C.CDLL(_path("libcsg.so"))
    L.csg_ctx_create.argtypes = [C.POINTER(VP), C.c_int]
    L.csg_ctx_destroy.argtypes = [VP]
    L.csg_last_error.argtypes = [VP]
    L.csg_last_error.restype = C.c_char_p
    L.csg_device_name.argtypes = [VP]
    L.csg_device_name.restype = C.c_char_p
    L.csg_mesh_upload.argtypes = [VP, C.c_int32, _f64p, C.c_int32, _i32p]
    L.csg_mesh_info.argtypes = [VP, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double), _f64p, _f64p]
    L.csg_set_submeshing.argtypes = [VP, C.c_int32, C.c_double]
    L.csg_set_boundary.argtypes = [VP, C.c_int32]
    L.csg_distance_batch.argtypes = [VP, C.c_int32, _i32p, _f64p, _i32p, _i32p, _f64p, VP, VP, VP, VP, VP]
    L.csg_pair_forces.argtypes = [VP, C.c_int32, _i32p, _f64p, C.c_int32, VP, _i32p, _i32p, VP, C.c_double,
                                  C.POINTER(csg_potential), C.c_int32, C.c_double, VP, VP, VP, VP, VP, VP]
    L.csg_transport_batch.argtypes = [VP, C.c_int32, _i32p, _f64p, _f64p, C.c_int32, VP, _i32p]
    L.csg_positions.argtypes = [VP, C.c_int32, _i32p, _f64p, _f64p]
    L.csg_system_init.argtypes = [VP, C.c_int32, _i32p, _f64p, VP, VP]
    L.csg_system_configure.argtypes = [VP, C.POINTER(csg_potential), C.c_double, C.c_int32, C.c_double, C.c_int32]
    L.csg_system_compute_forces.argtypes = [VP, VP, VP]
    L.csg_system_move.argtypes = [VP, _f64p, C.c_int32]
    L.csg_system_nve.argtypes = [VP, C.c_double, C.c_int32]
    L.csg_system_nvt.argtypes = [VP, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32]
    L.csg_system_sync.argtypes = [VP]
    L.csg_system_get.argtypes = [VP, VP, VP, VP, VP]
    L.csg_system_set_state.argtypes = [VP, VP, VP, VP, VP]
    L.csg_system_snapshot.argtypes = [VP]
    L.csg_system_snapshot_read.argtypes = [VP, VP, VP, VP, VP, VP]
    L.csg_system_thermostat.argtypes = [VP, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double), VP,
                                        C.POINTER(C.c_double)]
    L.csg_system_set_thermostat.argtypes = [VP, C.c_int32, C.c_double, C.c_double, VP, C.c_double]
    L.csg_system_status.argtypes = [VP, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    L.csg_system_pairs.argtypes = [VP, VP, VP, C.c_int64]
    L.csg_system_profile.argtypes = [VP, C.c_int32]
    L.csg_system_kernel_times.argtypes = [VP, _f64p, _i32p]
    L.csg_system_patch_stats.argtypes = [VP, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"), _i32p]
    L.csg_system_fallback_counts.argtypes = [VP, _i32p]
    _f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.csg_system_autotune.argtypes = [VP, C.POINTER(C.c_int32), _f64p]
    L.csg_system_set_lds_variant.argtypes = [VP, C.c_int32]
    L.csg_system_use_graph.argtypes = [VP, C.c_int32]
    L.csg_system_record_pairs.argtypes = [VP, C.c_int32]
    L.csg_system_set_groups.argtypes = [VP, C.c_int32, VP, VP, VP, VP]
    L.csg_system_pair_distances.argtypes = [VP, _f64p, C.c_int64]
    L.csg_system_screen_features.argtypes = [VP, _f32p, C.c_int64]
    L.csg_system_set_screen.argtypes = [VP, VP, C.c_float]
    L.csg_system_screen_stats.argtypes = [VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.csg_screen_forward.argtypes = [VP, _f32p, C.c_int32, _f32p, _f32p]
    L.csg_screen_gradient.argtypes = [VP, _f32p, C.c_int32, _f32p, _f32p, _f32p, C.POINTER(C.c_float)]
    L.csg_screen_train.argtypes = [VP, _f32p, C.c_int32, _f32p, _f32p, C.c_int32, C.c_int32, C.c_float, C.c_uint32, VP]
    _csg = L
    return L


def host():
    """libcsg_host.so with argument types declared."""
    global _host
    if _host is not None:
        return _host
    csg()  # dependency
    L = C.CDLL(_path("libcsg_host.so"))
    L.csgh_last_error.restype = C.c_char_p
    L.csgh_mesh_icosphere.argtypes = [C.c_int, C.c_double, C.POINTER(VP)]
    L.csgh_mesh_torus.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_uint32, C.POINTER(VP)]
    L.csgh_mesh_bumpy_sphere.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint32, C.POINTER(VP)]
    L.csgh_mesh_read_off.argtypes = [C.c_char_p, C.POINTER(VP)]
    L.csgh_mesh_from_arrays.argtypes = [C.c_int, _f64p, C.c_int, _i32p, C.POINTER(VP)]
    L.csgh_mesh_size.argtypes = [VP, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.csgh_mesh_copy.argtypes = [VP, _f64p, _i32p]
    L.csgh_mesh_free.argtypes = [VP]
    _host = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(VP)


def _hmesh(fn, *args):
    L = host()
    h = VP()
    rc = getattr(L, fn)(*args, C.byref(h))
    if rc != 0:
        raise CsgError(L.csgh_last_error().decode())
    nv, nf = C.c_int(), C.c_int()
    L.csgh_mesh_size(h, C.byref(nv), C.byref(nf))
    xyz = np.zeros(3 * nv.value)
    tri = np.zeros(3 * nf.value, np.int32)
    L.csgh_mesh_copy(h, xyz, tri)
    L.csgh_mesh_free(h)
    return xyz.reshape(-1, 3), tri.reshape(-1, 3)


def icosphere(level, radius=1.0):
    return _hmesh("csgh_mesh_icosphere", int(level), float(radius))


def jittered_torus(R=3.0, r=1.0, nu=200, nv=100, jitter=0.02, seed=13377):
    return _hmesh("csgh_mesh_torus", float(R), float(r), int(nu), int(nv), float(jitter), int(seed))


def bumpy_sphere(level=6, n_bumps=20, amplitude=0.15, width=0.3, seed=13377):
    return _hmesh("csgh_mesh_bumpy_sphere", int(level), int(n_bumps), float(amplitude), float(width), int(seed))


def read_off(path):
    return _hmesh("csgh_mesh_read_off", path.encode())


class Context:
    """One device context of libcsg.so with a mesh."""

    def __init__(self, device=0):
        self.L = csg()
        h = VP()
        rc = self.L.csg_ctx_create(C.byref(h), int(device))
        if rc != 0:
            raise CsgError(f"csg_ctx_create failed ({rc}): no gfx950 device {device} or HIP runtime error")
        self.h = h

    # --- learned pair screen (csg_screen_*) ---
    def screen_forward(self, params, X):
        p = np.ascontiguousarray(params, np.float32)
        X = np.ascontiguousarray(X, np.float32).reshape(-1, SCREEN_FEATURES)
        out = np.zeros(len(X), np.float32)
        self._check(self.L.csg_screen_forward(self.h, p, len(X), X, out), "csg_screen_forward")
        return out

    def screen_gradient(self, params, X, y):
        p = np.ascontiguousarray(params, np.float32)
        X = np.ascontiguousarray(X, np.float32).reshape(-1, SCREEN_FEATURES)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(SCREEN_PARAMS, np.float32)
        loss = C.c_float()
        self._check(self.L.csg_screen_gradient(self.h, p, len(X), X, y, g, C.byref(loss)), "csg_screen_gradient")
        return g, loss.value

    def screen_train(self, params, X, y, steps, batch=4096, lr=1e-3, seed=1):
        p = np.array(params, np.float32).copy()
        X = np.ascontiguousarray(X, np.float32).reshape(-1, SCREEN_FEATURES)
        y = np.ascontiguousarray(y, np.float32)
        losses = np.zeros(max(int(steps), 1), np.float32)
        self._check(self.L.csg_screen_train(self.h, p, len(X), X, y, int(steps), int(batch), float(lr), int(seed),
                                            losses.ctypes.data), "csg_screen_train")
        return p, losses[:steps]

    def close(self):
        if getattr(self, "h", None):
            self.L.csg_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc, what):
        if rc != 0:
            raise CsgError(f"{what} failed ({rc}): {self.L.csg_last_error(self.h).decode()}", rc)

    @property
    def device_name(self):
        return self.L.csg_device_name(self.h).decode()

    def upload_mesh(self, xyz, tri):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        tri = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
        self._check(self.L.csg_mesh_upload(self.h, len(xyz), xyz.ravel(), len(tri), tri.ravel()), "csg_mesh_upload")
        self.nV, self.nF = len(xyz), len(tri)
        self._mesh_digest = hashlib.sha1(xyz.tobytes() + tri.tobytes()).hexdigest()[:16]

    def mesh_info(self):
        nv, nf, area = C.c_int32(), C.c_int32(), C.c_double()
        lo, hi = np.zeros(3), np.zeros(3)
        self._check(self.L.csg_mesh_info(self.h, C.byref(nv), C.byref(nf), C.byref(area), lo, hi), "csg_mesh_info")
        return dict(nV=nv.value, nF=nf.value, area=area.value, lo=lo, hi=hi)

    BOUNDARY_MODES = {"closed": 0, "absorbing": 1, "tangential": 2}

    def set_boundary(self, mode):
        """What every walk of this context does at the border of an open mesh: "closed" (an error, the closed
        triangulatedMeshSpace), "absorbing" (absorbingOpenMeshSpace: stop on the border) or "tangential"
        (tangentialOpenMeshSpace: slide along it)."""
        if mode not in self.BOUNDARY_MODES:
            raise ValueError(f"unknown boundary mode {mode!r}")
        self._check(self.L.csg_set_boundary(self.h, self.BOUNDARY_MODES[mode]), "csg_set_boundary")

    def set_submeshing(self, enabled, maximum_distance=1.0):
        self._check(self.L.csg_set_submeshing(self.h, int(bool(enabled)), float(maximum_distance)), "csg_set_submeshing")
        self._submeshing = (bool(enabled), float(maximum_distance))

    def distance_batch(self, src_face, src_bary, tgt_offsets, tgt_face, tgt_bary, threshold=None):
        src_face = np.ascontiguousarray(src_face, np.int32)
        n = len(src_face)
        off = np.ascontiguousarray(tgt_offsets, np.int32)
        P = int(off[-1]) if n else 0
        d = np.zeros(P)
        st = np.zeros((P, 3))
        et = np.zeros((P, 3))
        fl = np.zeros(P, np.uint32)
        thr = None if threshold is None else np.ascontiguousarray(threshold, np.float64)
        self._check(self.L.csg_distance_batch(self.h, n, src_face, np.ascontiguousarray(src_bary, np.float64).ravel(), off,
                                              np.ascontiguousarray(tgt_face, np.int32),
                                              np.ascontiguousarray(tgt_bary, np.float64).ravel(), _ptr(thr), _ptr(d),
                                              _ptr(st), _ptr(et), _ptr(fl)), "csg_distance_batch")
        return d, st, et, fl

    def pair_forces(self, face, bary, tgt_offsets, tgt_idx, src_idx=None, patch_radius=None, rmax=np.inf,
                    potential=None, use_sentinel=False, sentinel=0.0):
        face = np.ascontiguousarray(face, np.int32)
        bary = np.ascontiguousarray(bary, np.float64).reshape(-1, 3)
        off = np.ascontiguousarray(tgt_offsets, np.int32)
        nsrc = len(off) - 1
        idx = np.ascontiguousarray(tgt_idx, np.int32)
        P = int(off[-1])
        sidx = None if src_idx is None else np.ascontiguousarray(src_idx, np.int32)
        rad = None if patch_radius is None else np.ascontiguousarray(patch_radius, np.float64)
        pot = None
        if potential is not None:
            pot = csg_potential(int(potential[0]), *[float(x) for x in (list(potential[1:]) + [0, 0, 0])[:3]])
        d = np.zeros(P)
        st = np.zeros((P, 3), np.float32)
        et = np.zeros((P, 3), np.float32)
        fl = np.zeros(P, np.uint32)
        F = np.zeros((nsrc, 3))
        E = np.zeros(nsrc)
        rm = float(rmax) if np.isfinite(rmax) else 1e300
        self._check(self.L.csg_pair_forces(self.h, len(face), face, bary.ravel(), nsrc, _ptr(sidx), off, idx, _ptr(rad), rm,
                                           C.byref(pot) if pot is not None else None, int(bool(use_sentinel)),
                                           float(sentinel), _ptr(d), _ptr(st), _ptr(et), _ptr(fl), _ptr(F), _ptr(E)),
                    "csg_pair_forces")
        return dict(dist=d, start=st, end=et, flags=fl, forces=F, energy=E)

    def transport_batch(self, face, bary, disp, vecs=None):
        face = np.array(face, np.int32).ravel().copy()
        n = len(face)
        bary = np.array(bary, np.float64).reshape(n, 3).copy()
        disp = np.array(disp, np.float64).reshape(n, 3).copy()
        nvec = 0
        if vecs is not None:
            vecs = np.array(vecs, np.float64).reshape(n, -1, 3).copy()
            nvec = vecs.shape[1]
        status = np.zeros(n, np.int32)
        rc = self.L.csg_transport_batch(self.h, n, face, bary.ravel(), disp.ravel(), nvec,
                                        _ptr(vecs) if vecs is not None else None, status)
        if rc not in (0, -5):
            self._check(rc, "csg_transport_batch")
        return face, bary, disp, vecs, status

    def positions(self, face, bary):
        face = np.ascontiguousarray(face, np.int32)
        out = np.zeros((len(face), 3))
        self._check(self.L.csg_positions(self.h, len(face), face, np.ascontiguousarray(bary, np.float64).ravel(),
                                         out.ravel()), "csg_positions")
        return out


class System:
    """Device-resident particle system of a Context (csg_system_* entry points). A context holds one system: creating
    another System on the same context replaces it, and the older handle then refuses every call."""

    def __getattribute__(self, name):
        if not name.startswith("_") and name not in ("ctx", "L", "N"):
            if object.__getattribute__(self, "_gen") != getattr(object.__getattribute__(self, "ctx"), "_sys_gen", 0):
                raise CsgError("this System was replaced by a newer System on the same Context")
        return object.__getattribute__(self, name)

    def __init__(self, ctx, face, bary, vel=None, mass=None):
        ctx._sys_gen = getattr(ctx, "_sys_gen", 0) + 1
        self._gen = ctx._sys_gen
        self.ctx = ctx
        self.L = ctx.L
        face = np.ascontiguousarray(face, np.int32)
        self.N = len(face)
        v = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        m = None if mass is None else np.ascontiguousarray(mass, np.float64)
        ctx._check(self.L.csg_system_init(ctx.h, self.N, face, np.ascontiguousarray(bary, np.float64).ravel(), _ptr(v),
                                          _ptr(m)), "csg_system_init")

    def configure(self, potential, range_, cell_list=True, rmax=np.inf, max_patch_faces=0):
        pot = csg_potential(int(potential[0]), *[float(x) for x in (list(potential[1:]) + [0, 0, 0])[:3]])
        rm = float(rmax) if np.isfinite(rmax) else 1e300
        self.ctx._check(self.L.csg_system_configure(self.ctx.h, C.byref(pot), float(range_), int(bool(cell_list)), rm,
                                                    int(max_patch_faces)), "csg_system_configure")
        self._config = [[float(x) for x in potential], float(range_), bool(cell_list), rm, int(max_patch_faces)]

    def compute_forces(self):
        F = np.zeros((self.N, 3))
        e = C.c_double()
        self.ctx._check(self.L.csg_system_compute_forces(self.ctx.h, _ptr(F), C.byref(e)), "csg_system_compute_forces")
        return F, e.value

    def move(self, disp, transport_velocity=True):
        self.ctx._check(self.L.csg_system_move(self.ctx.h, np.ascontiguousarray(disp, np.float64).ravel(),
                                               int(bool(transport_velocity))), "csg_system_move")

    def nve(self, dt, steps):
        self.ctx._check(self.L.csg_system_nve(self.ctx.h, float(dt), int(steps)), "csg_system_nve")

    def nvt(self, dt, T, tau=1.0, M=2, steps=1):
        self.ctx._check(self.L.csg_system_nvt(self.ctx.h, float(dt), float(T), float(tau), int(M), int(steps)),
                        "csg_system_nvt")

    def sync(self):
        self.ctx._check(self.L.csg_system_sync(self.ctx.h), "csg_system_sync")

    def get(self):
        f = np.zeros(self.N, np.int32)
        b = np.zeros((self.N, 3))
        v = np.zeros((self.N, 3))
        F = np.zeros((self.N, 3))
        self.ctx._check(self.L.csg_system_get(self.ctx.h, _ptr(f), _ptr(b), _ptr(v), _ptr(F)), "csg_system_get")
        return f, b, v, F

    def snapshot(self):
        """Queues a copy of the state behind the steps queued so far and returns at once; snapshot_read() waits for it
        only. For trajectory writers (simpleModelDatabase::writeState, src/databases/simpleModelDatabase.cpp:63-120)
        that keep the device busy: queue steps, snapshot, queue more steps, then write the snapshot while they run."""
        self.ctx._check(self.L.csg_system_snapshot(self.ctx.h), "csg_system_snapshot")

    def snapshot_read(self):
        """The last snapshot as a dict: faceIndex, barycentricPosition, R3position, velocity, force."""
        out = dict(faceIndex=np.zeros(self.N, np.int32), barycentricPosition=np.zeros((self.N, 3)),
                   R3position=np.zeros((self.N, 3)), velocity=np.zeros((self.N, 3)), force=np.zeros((self.N, 3)))
        self.ctx._check(self.L.csg_system_snapshot_read(self.ctx.h, *(_ptr(out[k]) for k in
                                                                      ("faceIndex", "barycentricPosition", "R3position",
                                                                       "velocity", "force"))), "csg_system_snapshot_read")
        return out

    def set_state(self, face=None, bary=None, vel=None, forces=None):
        """simpleModelDatabase::readState counterpart: replaces the given fields of the device state."""
        f = None if face is None else np.ascontiguousarray(face, np.int32)
        b, v, F = (None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1, 3) for a in (bary, vel, forces))
        for a in (f, b, v, F):
            if a is not None and len(a) != self.N:
                raise ValueError("state arrays must hold one entry per particle")
        self.ctx._check(self.L.csg_system_set_state(self.ctx.h, _ptr(f), _ptr(b), _ptr(v), _ptr(F)), "csg_system_set_state")

    def checkpoint(self, time=0.0):
        """Everything a resumed run needs, as arrays: the record of the reference's simpleModelDatabase
        (src/databases/simpleModelDatabase.cpp:37-50: time, R3position, barycentricPosition, faceIndex, velocity, force)
        and the Nose-Hoover chain, which the reference does not save (nhChain, (M + 1) x 4; nhParams = M, T, tau and the
        chain's kinetic energy; M = 0 before any NVT step)."""
        f, b, v, F = self.get()
        M, T, tau, ke = C.c_int32(), C.c_double(), C.c_double(), C.c_double()
        h = self.ctx.h
        self.ctx._check(self.L.csg_system_thermostat(h, C.byref(M), C.byref(T), C.byref(tau), None, C.byref(ke)),
                        "csg_system_thermostat")
        chain = np.zeros((M.value + 1 if M.value else 0, 4))
        if M.value:
            self.ctx._check(self.L.csg_system_thermostat(h, C.byref(M), C.byref(T), C.byref(tau), _ptr(chain),
                                                         C.byref(ke)), "csg_system_thermostat")
        return dict(time=np.array([float(time)]), R3position=self.ctx.positions(f, b), barycentricPosition=b,
                    faceIndex=f, velocity=v, force=F, nhChain=chain,
                    nhParams=np.array([M.value, T.value, tau.value, ke.value]))

    def restore(self, ckpt):
        """Sets the state of a checkpoint: positions, velocities, forces and the Nose-Hoover chain (a checkpoint taken
        before any NVT step clears the chain, so that the next NVT step starts a new one)."""
        M, T, tau, ke = ckpt["nhParams"]
        chain = np.ascontiguousarray(ckpt["nhChain"], np.float64)
        if int(M) > 0 and chain.shape != (int(M) + 1, 4):
            raise ValueError("Nose-Hoover chain of the wrong shape")
        self.set_state(ckpt["faceIndex"], ckpt["barycentricPosition"], ckpt["velocity"], ckpt["force"])
        self.ctx._check(self.L.csg_system_set_thermostat(self.ctx.h, int(M), float(T), float(tau),
                                                         _ptr(chain) if int(M) > 0 else None, float(ke)),
                        "csg_system_set_thermostat")

    def status(self):
        p, o, u, w, k = C.c_int64(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        self.ctx._check(self.L.csg_system_status(self.ctx.h, C.byref(p), C.byref(o), C.byref(u), C.byref(w), C.byref(k)),
                        "csg_system_status")
        return dict(pairs=p.value, cand_overflow=o.value, unsolved=u.value, walk_failures=w.value, kinetic=k.value)

    def pairs(self):
        off = np.zeros(self.N + 1, np.int32)
        self.ctx._check(self.L.csg_system_pairs(self.ctx.h, _ptr(off), None, 0), "csg_system_pairs")
        idx = np.zeros(max(int(off[-1]), 1), np.int32)
        self.ctx._check(self.L.csg_system_pairs(self.ctx.h, _ptr(off), _ptr(idx), len(idx)), "csg_system_pairs")
        return off, idx[: off[-1]]

    def profile(self, enable=True):
        self.ctx._check(self.L.csg_system_profile(self.ctx.h, int(bool(enable))), "csg_system_profile")

    def kernel_times(self):
        ms = np.zeros(6)
        n = np.zeros(6, np.int32)
        self.ctx._check(self.L.csg_system_kernel_times(self.ctx.h, ms, n), "csg_system_kernel_times")
        names = ["neighbors", "geodesic_lds", "geodesic_hbm", "transport", "patch", "screen"]
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(names)}

    def tuning_key(self):
        """What the fastest LDS variant depends on: the device, the mesh (a digest of its arrays), the particle count,
        the potential, range, cell list, truncation and patch bound, submeshing, and the particle groups."""
        sub = getattr(self.ctx, "_submeshing", None)
        groups = getattr(self, "_groups", None)
        return json.dumps([self.ctx.device_name, getattr(self.ctx, "_mesh_digest", None), self.N, getattr(self, "_config", None),
                           sub, groups], separators=(",", ":"))

    def autotune(self, cache=None):
        """Times each workgroup size of the LDS geodesic kernel (the variants of CSG_LDS_VARIANTS) and keeps the
        fastest: (chosen, ms per force computation for each variant). With cache, the path of a JSON file: a system
        whose tuning_key() the file holds takes the recorded choice without timing (and returns the recorded times);
        otherwise the result of the timing is added to the file."""
        key = self.tuning_key() if cache else None
        if cache:
            hit = lookup_tuning(cache, key)
            if hit is not None:
                self.set_lds_variant(hit[0])
                return hit
        v = C.c_int32()
        ms = np.zeros(LDS_VARIANTS)
        self.ctx._check(self.L.csg_system_autotune(self.ctx.h, C.byref(v), ms), "csg_system_autotune")
        if cache:
            record_tuning(cache, key, v.value, ms)
        return v.value, ms

    def use_graph(self, enable=True):
        self.ctx._check(self.L.csg_system_use_graph(self.ctx.h, int(bool(enable))), "csg_system_use_graph")

    def set_lds_variant(self, variant):
        self.ctx._check(self.L.csg_system_set_lds_variant(self.ctx.h, int(variant)), "csg_system_set_lds_variant")

    def set_groups(self, group_of, potentials, ranges, maximum_distances):
        """Independent particle groups (see csg_system_set_groups): group index per particle, and per group a potential
        tuple, an interaction range and a maximum distance. group_of=None clears."""
        if group_of is None:
            self.ctx._check(self.L.csg_system_set_groups(self.ctx.h, 0, None, None, None, None), "csg_system_set_groups")
            self._groups = None
            return
        g = np.ascontiguousarray(group_of, np.int32)
        n = len(potentials)
        self._groups = [hashlib.sha1(g.tobytes()).hexdigest()[:16], [[float(x) for x in p] for p in potentials],
                        [float(x) for x in ranges], [float(x) for x in maximum_distances]]
        pots = (csg_potential * n)(*[csg_potential(int(p[0]), *[float(x) for x in (list(p[1:]) + [0, 0, 0])[:3]])
                                     for p in potentials])
        r = np.ascontiguousarray(ranges, np.float64)
        m = np.ascontiguousarray(maximum_distances, np.float64)
        self.ctx._check(self.L.csg_system_set_groups(self.ctx.h, n, _ptr(g), pots, _ptr(r), _ptr(m)),
                        "csg_system_set_groups")

    def record_pairs(self, enable=True):
        self.ctx._check(self.L.csg_system_record_pairs(self.ctx.h, int(bool(enable))), "csg_system_record_pairs")

    def pair_distances(self):
        off, _ = self.pairs()
        d = np.zeros(max(int(off[-1]), 1))
        self.ctx._check(self.L.csg_system_pair_distances(self.ctx.h, d, len(d)), "csg_system_pair_distances")
        return d[: off[-1]]

    def screen_features(self):
        off, _ = self.pairs()
        X = np.zeros((max(int(off[-1]), 1), SCREEN_FEATURES), np.float32)
        self.ctx._check(self.L.csg_system_screen_features(self.ctx.h, X, len(X)), "csg_system_screen_features")
        return X[: off[-1]]

    def set_screen(self, params=None, threshold=0.0):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        if p is not None and p.size != SCREEN_PARAMS:
            raise CsgError(f"screen parameters must hold {SCREEN_PARAMS} floats")
        self.ctx._check(self.L.csg_system_set_screen(self.ctx.h, None if p is None else p.ctypes.data, float(threshold)),
                        "csg_system_set_screen")

    def screen_stats(self):
        a, b = C.c_int64(), C.c_int64()
        self.ctx._check(self.L.csg_system_screen_stats(self.ctx.h, C.byref(a), C.byref(b)), "csg_system_screen_stats")
        return dict(candidates=a.value, kept=b.value)

    def fallback_counts(self):
        r = np.zeros(6, np.int32)
        self.ctx._check(self.L.csg_system_fallback_counts(self.ctx.h, r), "csg_system_fallback_counts")
        return dict(zip(["faces", "verts", "queue", "targets", "steps", "frontier"], r.tolist()))

    def patch_stats(self):
        s = np.zeros(7, np.int64)
        m = np.zeros(7, np.int32)
        self.ctx._check(self.L.csg_system_patch_stats(self.ctx.h, s, m), "csg_system_patch_stats")
        return dict(faces=int(s[0]), verts=int(s[1]), steps=int(s[2]), items=int(s[3]), max_faces=int(m[0]),
                    max_steps=int(m[2]), max_items=int(m[3]), us_patch=s[4] * 0.01, us_propagate=s[5] * 0.01,
                    us_epilogue=s[6] * 0.01)


def _read_tuning(path):
    try:
        with open(path) as fh:
            table = json.load(fh)
        return table if isinstance(table, dict) else {}
    except (OSError, ValueError):
        return {}


def lookup_tuning(path, key):
    """The recorded (variant, ms per variant) for key in the autotuner's JSON file, or None (no file, no entry, or an
    entry that does not fit this library's variants)."""
    e = _read_tuning(path).get(key)
    try:
        v, ms = int(e["variant"]), np.array(e["ms"], np.float64)
    except (TypeError, KeyError, ValueError):
        return None
    if not 0 <= v < LDS_VARIANTS or len(ms) != LDS_VARIANTS or e.get("variants") != list(LDS_VARIANT_NAMES):
        return None
    return v, ms


def record_tuning(path, key, variant, ms):
    """Adds or replaces key's entry in the autotuner's JSON file (written to a temporary file, then renamed over it)."""
    table = _read_tuning(path)
    table[key] = {"variant": int(variant), "ms": [float(x) for x in ms], "variants": list(LDS_VARIANT_NAMES)}
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as fh:
        json.dump(table, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


CHECKPOINT_FIELDS = ("time", "R3position", "barycentricPosition", "faceIndex", "velocity", "force", "nhChain", "nhParams")


def _is_hdf5(path):
    return str(path).endswith((".h5", ".hdf5"))


def save_checkpoint(path, ckpt):
    """Writes a System.checkpoint(): to an HDF5 file in the reference's simpleModelDatabase layout when the name ends
    in .h5 or .hdf5 (curvedspacesim_amd.trajectory), else to an uncompressed .npz archive of plain arrays."""
    if _is_hdf5(path):
        from . import trajectory

        return trajectory.save_checkpoint(path, ckpt)
    np.savez(path, **{k: np.asarray(ckpt[k]) for k in CHECKPOINT_FIELDS})


def load_checkpoint(path):
    """Reads a checkpoint written by save_checkpoint (the last record of an HDF5 file; from an .npz, plain arrays only,
    object arrays are refused)."""
    if _is_hdf5(path):
        from . import trajectory

        return trajectory.load_checkpoint(path)
    with np.load(path, allow_pickle=False) as z:
        missing = [k for k in CHECKPOINT_FIELDS if k not in z.files]
        if missing:
            raise ValueError(f"not a checkpoint: missing {missing}")
        return {k: z[k].copy() for k in CHECKPOINT_FIELDS}
