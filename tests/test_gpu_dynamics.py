"""The device-resident integrators (kVVFirst ... kNHChain behind csg_system_nve / csg_system_nvt, replacing
velocityVerletNVE::performUpdate, src/updaters/velocityVerletNVE.cpp:3-29, and noseHooverNVT::performUpdate,
src/updaters/noseHooverNVT.cpp:42-139) held to what the mathematics fixes, as tests/test_oracle_dynamics.py holds
the oracle: the bent prism cylinder against the oracle's run on its flat unrolling, the dt^2 order of the
velocity-Verlet energy error, and the conservation of the Nose-Hoover extended energy.

Tolerances: the kernel keeps start tangents in FP32, so forces carry relative errors of about 1e-7; over 60 steps
positions and velocities then agree with the FP64 oracle to 1e-6."""
import numpy as np
import pytest

import curvedspacesim_amd as csa
import small_meshes as sm
from oracle.pyoracle import OracleMesh
from test_oracle_dynamics import (CHORD, HY, NX, NY, R, SMOOTH, cylinder_to_flat_vel, flat_to_cylinder_vel, make,
                                  strip_system)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csa.Context(0)
    yield c
    c.close()


def gpu_system(ctx, X, T, face, bary, vel, pot, rng_):
    ctx.upload_mesh(X, T)
    ctx.set_submeshing(True, rng_)
    s = csa.System(ctx, face, bary, vel)
    s.configure(pot, rng_, True, rng_)
    return s


def test_cylinder_matches_oracle_unrolled(ctx, rng):
    face, bary, vel = strip_system(rng, 14, 0.6)
    Xc, Tc = sm.cylinder(NX, NY, R, HY)
    Xf, Tf = sm.grid(NX, NY, CHORD, HY)
    print("case gpu cylinder against oracle flat strip, 60 NVE steps", flush=True)
    g = gpu_system(ctx, Xc, Tc, face, bary, flat_to_cylinder_vel(face, vel), (csa.HARMONIC, 10.0, 0.5), 0.5)
    g.compute_forces()
    mf = OracleMesh(Xf, Tf)
    o = make(mf, face, bary, vel, 10.0, 0.5)
    o.compute_forces()
    for _ in range(60):
        o.nve_step(0.01)
    g.nve(0.01, 60)
    f, b, v, _ = g.get()
    st = g.status()
    assert st["walk_failures"] == 0 and st["unsolved"] == 0
    np.testing.assert_allclose(mf.positions(f, b), mf.positions(o.face, o.bary), atol=1e-6)
    np.testing.assert_allclose(cylinder_to_flat_vel(f, v), o.vel, atol=1e-6)


def _energy(g):
    _, _, v, _ = g.get()
    _, e = g.compute_forces()
    return 0.5 * float((v * v).sum()) + 0.5 * e, v


def _nve_error(ctx, dt, t_end=0.6, seed=5):
    face, bary, vel = strip_system(np.random.default_rng(seed), 14, 0.6)
    Xc, Tc = sm.cylinder(NX, NY, R, HY)
    g = gpu_system(ctx, Xc, Tc, face, bary, flat_to_cylinder_vel(face, vel), SMOOTH, SMOOTH[3])
    e0, _ = _energy(g)
    worst = 0.0
    for _ in range(int(round(t_end / dt))):
        g.nve(dt, 1)
        worst = max(worst, abs(_energy(g)[0] - e0))
    return worst, e0


def test_nve_energy_error_is_second_order(ctx):
    print("case gpu NVE energy order", flush=True)
    e1, E = _nve_error(ctx, 0.01)
    e2, _ = _nve_error(ctx, 0.005)
    assert e1 < 2e-3 * E, (e1, E)
    assert 3.0 < e1 / e2 < 5.0, (e1, e2)


def _nh_run(ctx, dt, t_end=1.0, T=0.5, seed=9):
    face, bary, vel = strip_system(np.random.default_rng(seed), 14, 0.9)
    Xc, Tc = sm.cylinder(NX, NY, R, HY)
    g = gpu_system(ctx, Xc, Tc, face, bary, flat_to_cylinder_vel(face, vel), SMOOTH, SMOOTH[3])
    g.compute_forces()

    def extended():
        c = g.checkpoint()
        M = int(c["nhParams"][0])
        b = c["nhChain"]
        e, v = _energy(g)
        return e + 0.5 * float((b[:M, 3] * b[:M, 1] ** 2).sum()) + b[0, 3] * b[0, 0] + T * float(b[1:M, 0].sum()), v

    # the extended energy is conserved from the end of the first step (the first chain half step uses the kinetic
    # energy set to Q0 by setBathVariables, noseHooverNVT.cpp:28-36)
    g.nvt(dt, T, 1.0, 2, 1)
    h0, _ = extended()
    worst, kes = 0.0, []
    for _ in range(int(round(t_end / dt)) - 1):
        g.nvt(dt, T, 1.0, 2, 1)
        h, v = extended()
        worst = max(worst, abs(h - h0))
        kes.append(0.5 * float((v * v).sum()))
    return worst, max(kes) - min(kes)


def test_nose_hoover_extended_energy_is_conserved(ctx):
    print("case gpu Nose-Hoover extended energy", flush=True)
    d1, swing = _nh_run(ctx, 0.01)
    d2, _ = _nh_run(ctx, 0.005)
    assert d1 < 0.02 * swing, (d1, swing)
    assert 3.0 < d1 / d2 < 5.0, (d1, d2)
