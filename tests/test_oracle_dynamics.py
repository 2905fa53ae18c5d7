"""Pins the oracle's integrators (OracleSystem, restating velocityVerletNVE::performUpdate,
src/updaters/velocityVerletNVE.cpp:3-29, and noseHooverNVT::performUpdate, src/updaters/noseHooverNVT.cpp:42-139) by
what the mathematics fixes, independently of the device code:

* free flight on a plane is a straight line at constant velocity;
* a head-on collision of two particles under harmonic repulsion follows its closed form, and the velocity-Verlet
  error at a fixed time falls as dt^2;
* a polyhedral cylinder and its exact unrolling (a flat grid with the same faces) are isometric, so the same
  (face, barycentric) state gives the same trajectory on both: this pins the geodesic tangents, the parallel
  transport of velocities across bent edges and the integrator together;
* the velocity-Verlet energy error falls as dt^2;
* the Nose-Hoover chain conserves its extended energy.

The order tests keep the force smooth along the measured stretch: a kink of the force (the harmonic potential's edge
at d = sigma) inside a step gives that step an O(dt^2) local error whose constant depends on where in the step the
kink falls, which would blur the dt^2 ratio. The collision run therefore ends inside the contact (whose onset falls
on a step boundary, where the force is zero), and the energy tests use the truncated Gaussian with a far cutoff (6 s,
where the force is e^-18 of its scale).

Harmonic repulsion is the reference's convention, U = (k/2)(sigma - d)^2 per unordered pair with F = -k(sigma - d) t
(src/forces/harmonicRepulsion.cpp:3-34); the reference's energy sums over ordered pairs (src/forces/baseForce.cpp:
33-44), so the potential energy of the system is half of it. Masses are 1.
"""
import numpy as np
import pytest

import small_meshes as sm
from oracle.pyoracle import OracleMesh, OracleSystem

HARMONIC, TRUNC_GAUSSIAN = 0, 2
SMOOTH = (TRUNC_GAUSSIAN, 1.0, 0.3, 1.8)  # epsilon, s, cutoff 6 s
NX, NY, R, HY = 24, 16, 1.5, 0.4
CHORD = sm.cylinder_chord(NX, R)


def locate_flat(x, y, nx=NX, hx=CHORD, hy=HY):
    """(face, barycentrics) of the point (x, y) on grid(nx, ., hx, hy)."""
    i, j = int(np.floor(x / hx)), int(np.floor(y / hy))
    u, v = x / hx - i, y / hy - j
    if v <= u:  # corners (i,j), (i+1,j), (i+1,j+1)
        return 2 * (j * nx + i), np.array([1 - u, u - v, v])
    return 2 * (j * nx + i) + 1, np.array([1 - v, u, v - u])  # corners (i,j), (i+1,j+1), (i,j+1)


def column_chord_dir(face):
    """Unit direction of the cylinder's faces of this column along increasing angle (the unrolled x axis)."""
    i = (np.asarray(face) // 2) % NX
    t0, t1 = 2 * np.pi * i / NX, 2 * np.pi * (i + 1) / NX
    c = np.stack([np.cos(t1) - np.cos(t0), np.sin(t1) - np.sin(t0), np.zeros_like(t0)], -1)
    return c / np.linalg.norm(c, axis=-1, keepdims=True)


def flat_to_cylinder_vel(face, v):
    return v[:, :1] * column_chord_dir(face) + v[:, 1:2] * np.array([0.0, 0.0, 1.0])


def cylinder_to_flat_vel(face, v):
    return np.stack([(v * column_chord_dir(face)).sum(1), v[:, 2], np.zeros(len(v))], 1)


def strip_system(rng, n, speed):
    """n particles in the middle of the strip (away from its border and the cylinder's seam)."""
    xs = rng.uniform(3.0, 6.0, n)
    ys = rng.uniform(2.2, 4.2, n)
    fb = [locate_flat(x, y) for x, y in zip(xs, ys)]
    face = np.array([f for f, _ in fb], np.int32)
    bary = np.array([b for _, b in fb])
    ang = rng.uniform(0, 2 * np.pi, n)
    vel = speed * np.stack([np.cos(ang), np.sin(ang), np.zeros(n)], 1)
    return face, bary, vel


def make(mesh, face, bary, vel, k, sigma):
    return OracleSystem(mesh, face, bary, vel, (HARMONIC, k, sigma), sigma, True, True, sigma, threads=4)


def make_smooth(mesh, face, bary, vel):
    rc = SMOOTH[3]
    return OracleSystem(mesh, face, bary, vel, SMOOTH, rc, True, True, rc, threads=4)


def energy(osys):
    return 0.5 * float((osys.vel * osys.vel).sum()) + 0.5 * osys.last["energy"]


def test_free_flight_is_a_straight_line():
    """Particles out of each other's range move by n dt v exactly, whatever edges they cross."""
    X, T = sm.grid(NX, NY, CHORD, HY)
    m = OracleMesh(X, T)
    pts = [(3.1, 2.3), (6.7, 4.45)]
    face = np.array([locate_flat(*p)[0] for p in pts], np.int32)
    bary = np.array([locate_flat(*p)[1] for p in pts])
    vel = np.array([[0.83, 0.31, 0.0], [-0.52, 0.47, 0.0]])
    osys = make(m, face, bary, vel, 10.0, 0.5)
    osys.compute_forces()
    dt, steps = 0.01, 150
    for _ in range(steps):
        osys.nve_step(dt)
    P = m.positions(osys.face, osys.bary)
    np.testing.assert_allclose(P[:, :2], np.array(pts) + steps * dt * vel[:, :2], atol=1e-11)
    np.testing.assert_allclose(osys.vel, vel, atol=1e-14)


def _collision(dt, t_end=1.0, k=10.0, sigma=0.5, v0=0.5, d0=1.0, xc=4.5, y=2.3):
    X, T = sm.grid(NX, NY, CHORD, HY)
    m = OracleMesh(X, T)
    pts = [(xc - d0 / 2, y), (xc + d0 / 2, y)]
    face = np.array([locate_flat(*p)[0] for p in pts], np.int32)
    bary = np.array([locate_flat(*p)[1] for p in pts])
    vel = np.array([[v0, 0.0, 0.0], [-v0, 0.0, 0.0]])
    osys = make(m, face, bary, vel, k, sigma)
    osys.compute_forces()
    for _ in range(int(round(t_end / dt))):
        osys.nve_step(dt)
    return m.positions(osys.face, osys.bary)[:, 0], osys.vel[:, 0]


def _collision_exact(t, k=10.0, sigma=0.5, v0=0.5, d0=1.0, xc=4.5):
    """Separation d(t): free approach at relative speed 2 v0, then d'' = 2 k (sigma - d) while d < sigma (each
    particle feels k (sigma - d) away from the other), so d = sigma - A sin(w (t - tc)) with w = sqrt(2k), A = 2 v0 / w,
    then free separation."""
    w = np.sqrt(2 * k)
    A = 2 * v0 / w
    tc = (d0 - sigma) / (2 * v0)
    te = tc + np.pi / w
    if t <= tc:
        d, dd = d0 - 2 * v0 * t, -2 * v0
    elif t <= te:
        d, dd = sigma - A * np.sin(w * (t - tc)), -A * w * np.cos(w * (t - tc))
    else:
        d, dd = sigma + 2 * v0 * (t - te), 2 * v0
    return np.array([xc - d / 2, xc + d / 2]), np.array([-dd / 2, dd / 2])


def test_head_on_collision_closed_form_second_order():
    x_ex, v_ex = _collision_exact(1.0)
    errs = []
    for dt in (0.02, 0.01, 0.005):
        x, v = _collision(dt)
        errs.append(max(np.abs(x - x_ex).max(), np.abs(v - v_ex).max()))
    assert errs[-1] < 2e-4, errs
    # a first-order slip (a dropped dt^2/2 F term, a full kick instead of two halves) would give ratios near 2
    for a, b in zip(errs, errs[1:]):
        assert 3.2 < a / b < 4.8, errs


def test_cylinder_trajectory_equals_its_unrolling(rng):
    """Velocity Verlet with harmonic repulsion on a bent prism cylinder and on the flat grid it unrolls to."""
    face, bary, vel = strip_system(rng, 14, 0.6)
    Xf, Tf = sm.grid(NX, NY, CHORD, HY)
    Xc, Tc = sm.cylinder(NX, NY, R, HY)
    assert np.array_equal(Tc % 10 ** 9, Tc) and len(Tc) == len(Tf)
    mf, mc = OracleMesh(Xf, Tf), OracleMesh(Xc, Tc)
    sf = make(mf, face, bary, vel, 10.0, 0.5)
    scy = make(mc, face, bary, flat_to_cylinder_vel(face, vel), 10.0, 0.5)
    sf.compute_forces()
    scy.compute_forces()
    assert np.abs(sf.F).max() > 1.0  # particles interact
    for _ in range(60):
        sf.nve_step(0.01)
        scy.nve_step(0.01)
    # the cylinder's (face, bary) read on the flat grid is the flat run's position
    np.testing.assert_allclose(mf.positions(scy.face, scy.bary), mf.positions(sf.face, sf.bary), atol=1e-9)
    np.testing.assert_allclose(cylinder_to_flat_vel(scy.face, scy.vel), sf.vel, atol=1e-9)
    # velocities stay in the tangent planes of the bent faces
    nrm = np.cross(Xc[Tc[scy.face, 1]] - Xc[Tc[scy.face, 0]], Xc[Tc[scy.face, 2]] - Xc[Tc[scy.face, 0]])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    assert np.abs((scy.vel * nrm).sum(1)).max() < 1e-9


def _nve_energy_error(dt, t_end, seed):
    face, bary, vel = strip_system(np.random.default_rng(seed), 14, 0.6)
    Xc, Tc = sm.cylinder(NX, NY, R, HY)
    s = make_smooth(OracleMesh(Xc, Tc), face, bary, flat_to_cylinder_vel(face, vel))
    s.compute_forces()
    e0 = energy(s)
    worst = 0.0
    for _ in range(int(round(t_end / dt))):
        s.nve_step(dt)
        worst = max(worst, abs(energy(s) - e0))
    return worst, e0


def test_nve_energy_error_is_second_order():
    e1, E = _nve_energy_error(0.01, 0.6, 5)
    e2, _ = _nve_energy_error(0.005, 0.6, 5)
    assert e1 < 2e-3 * E, (e1, E)
    assert 3.0 < e1 / e2 < 5.0, (e1, e2)


def _potential_now(s):
    """Potential energy of the current positions. An NVT step evaluates forces at its half-step positions
    (noseHooverNVT.cpp:121-123), so the energy of that evaluation is not the end-of-step one."""
    F, last = s.F, s.last
    s.compute_forces()
    u = 0.5 * s.last["energy"]
    s.F, s.last = F, last
    return u


def _nh_extended(s, T):
    """Conserved quantity of the reference's chain: with v0' = (2 K - Q0) / Q0 - v0 v1 (noseHooverNVT.cpp:82-86, the
    first mass Q0 doubling as the target 2 K) and v_i' = (Q_{i-1} v_{i-1}^2 - T) / Q_i - v_i v_{i+1}, the time
    derivative of K + U + sum_i Q_i v_i^2 / 2 + Q0 xi_0 + T sum_{i>=1} xi_i telescopes to zero."""
    b = s.chain
    M = s.M
    ext = 0.5 * float((b[:M, 3] * b[:M, 1] ** 2).sum()) + b[0, 3] * b[0, 0] + T * float(b[1:M, 0].sum())
    return 0.5 * float((s.vel * s.vel).sum()) + _potential_now(s) + ext


def _nh_drift(dt, t_end, T=0.5, seed=9):
    face, bary, vel = strip_system(np.random.default_rng(seed), 14, 0.9)
    Xc, Tc = sm.cylinder(NX, NY, R, HY)
    s = make_smooth(OracleMesh(Xc, Tc), face, bary, flat_to_cylinder_vel(face, vel))
    s.compute_forces()
    # the first half step of the chain uses the kinetic energy set to Q0 by setBathVariables (noseHooverNVT.cpp:28-36),
    # not the particles' own: the extended energy is conserved from the end of the first step on
    s.nvt_step(dt, T)
    h0 = _nh_extended(s, T)
    worst, kes = 0.0, []
    for _ in range(int(round(t_end / dt)) - 1):
        s.nvt_step(dt, T)
        worst = max(worst, abs(_nh_extended(s, T) - h0))
        kes.append(0.5 * float((s.vel * s.vel).sum()))
    return worst, max(kes) - min(kes), s


def test_nose_hoover_extended_energy_is_conserved():
    d1, swing, s = _nh_drift(0.01, 1.0)
    d2, _, _ = _nh_drift(0.005, 1.0)
    assert abs(s.chain[0, 1]) > 1e-3  # the thermostat acts
    assert d1 < 0.02 * swing, (d1, swing)
    assert 3.0 < d1 / d2 < 5.0, (d1, d2)


@pytest.mark.parametrize("M", [2, 3])
def test_nose_hoover_extended_energy_longer_chains(M):
    face, bary, vel = strip_system(np.random.default_rng(21), 10, 0.9)
    Xc, Tc = sm.cylinder(NX, NY, R, HY)
    s = make_smooth(OracleMesh(Xc, Tc), face, bary, flat_to_cylinder_vel(face, vel))
    s.compute_forces()
    s.nvt_step(0.005, 0.5, 1.0, M)
    h0 = _nh_extended(s, 0.5)
    worst = 0.0
    for _ in range(100):
        s.nvt_step(0.005, 0.5, 1.0, M)
        worst = max(worst, abs(_nh_extended(s, 0.5) - h0))
    assert worst < 1e-4 * abs(h0), (worst, h0)
