"""Pins of the oracle's particle dynamics and flow forcing against published
laws and exact solutions (CPU only).

The oracle follows src/particle_kernel.cu:2290-3484 (springs, pair
collisions, motion, rotation) and src/cuda_bluebottle.cu:1753-1982 (time
step, forcing) line by line.  These tests fix those formulas from outside
the sources:

- pair lubrication: the singular parts of Jeffrey & Onishi's (1984)
  resistance functions for two unequal spheres, X^A_11 (normal
  translation), Y^A_11 (tangential translation) and Y^C_11 (rotation),
  as the difference between the gap h and the cut-off hN at which the
  reference switches lubrication on, for equal (beta = 1) and unequal
  (beta = 1.5, 0.5) radii; the translation-rotation coupling obeys the
  reciprocal theorem (force from rotation and torque from translation
  carry one coefficient);
- pair contact: the Hertz law F = (4/3) E* sqrt(R*) delta^(3/2) with
  1/E* = (1-s1^2)/E1 + (1-s2^2)/E2 and 1/R* = 1/a1 + 1/a2, on the line of
  centres, no torque, and the contact list entry;
- spring: Hooke's law -k (|d| - l) d/|d| about the anchor, also for a
  centre beyond a periodic face; a sphere on a spring converges to the
  oscillator at sqrt(k/m), at first order in dt (the reference's velocity
  update lags half a step for a position-dependent force);
- free fall: with no fluid force the acceleration is (1 - rho_f/rho_p) g;
- rotation of the body axes: Rodrigues' formula for the angle |omega| dt;
- forcing: a laminar Poiseuille profile under the constant pressure
  gradient is a steady solution of the discrete equations (exact for a
  quadratic), so it stays unchanged over several steps;
- time step: the velocity branch of the CFL formula on a known maximum.
"""
import math

import numpy as np
import pytest

import bbconfig
import orc

L = 48.0


def pair_box(tmp, a1=4.0, a2=4.0, **p2):
    """Two spheres in a periodic box, no gravity; sphere 0 of radius a1,
    sphere 1 of radius a2 (extra keys for sphere 1 in p2)."""
    kw, parts = bbconfig.sediment(n=(48, 48, 48), nparts=2, a=max(a1, a2),
                                  seed=1)
    kw["g"] = ((0.0, 0.0), (0.0, 0.0), (0.0, 0.0))
    parts[0].update(r=a1, x=10.0, y=10.0, z=10.0, E=2e7, sigma=0.3)
    parts[1].update(r=a2, x=30.0, y=30.0, z=30.0, E=1e7, sigma=0.45, **p2)
    d = str(tmp)
    bbconfig.write_case(d, **kw)
    bbconfig.write_parts(d, parts)
    o = orc.Oracle(d)
    o.setup()
    return o, kw["nu"] * kw["rho_f"]


def place_pair(o, a1, a2, h, n, U1=(0, 0, 0), W1=(0, 0, 0)):
    """Sphere 0 at the box centre, sphere 1 at gap h along -n (so that n
    points from sphere 1 to sphere 0), sphere 1 at rest."""
    n = np.asarray(n, float) / np.linalg.norm(n)
    x0 = np.array([L / 2, L / 2, L / 2])
    x1 = x0 - (a1 + a2 + h) * n
    z = np.zeros(3)
    o.set_motion(0, x0, np.asarray(U1, float), np.asarray(W1, float))
    o.set_motion(1, x1, z, z)
    o.collide()
    s = o.part_state(0)
    return (np.array([s["iFx"], s["iFy"], s["iFz"]]),
            np.array([s["iLx"], s["iLy"], s["iLz"]]), n)


def jeffrey_onishi(beta):
    """Singular coefficients (g1 of 1/xi, g2 of log(1/xi)) of X^A_11 and
    g2 of Y^A_11, Y^B_11, Y^C_11 (Jeffrey & Onishi 1984, table 1)."""
    b = beta
    return {"XA_g1": 2 * b * b / (1 + b) ** 3,
            "XA_g2": b * (1 + 7 * b + b * b) / (5 * (1 + b) ** 3),
            "YA_g2": 4 * b * (2 + b + 2 * b * b) / (15 * (1 + b) ** 3),
            "YB_g2": -b * (4 + b) / (5 * (1 + b) ** 2),
            "YC_g2": 2 * b / (5 * (1 + b))}


@pytest.mark.parametrize("a1,a2", [(4.0, 4.0), (4.0, 6.0), (6.0, 3.0)])
@pytest.mark.parametrize("h", [0.3, 1.7])
def test_pair_lubrication_jeffrey_onishi(tmp_path, a1, a2, h):
    o, mu = pair_box(tmp_path, a1, a2)
    try:
        beta = a2 / a1
        hN = 1.0 * a1                       # (l/a) = 1
        jo = jeffrey_onishi(beta)
        xi, xiN = 2 * h / (a1 + a2), 2 * hN / (a1 + a2)
        ln = math.log(xiN / xi)
        n = np.array([0.3, -0.5, 0.8])
        # normal approach
        Un = 0.01
        F, T, n = place_pair(o, a1, a2, h, n, U1=-Un * n / np.linalg.norm(n))
        XA = jo["XA_g1"] * (1 / xi - 1 / xiN) + jo["XA_g2"] * ln
        F0 = -6 * math.pi * mu * a1 * XA * (-Un) * n
        print(f"beta {beta} h {h}: normal F {F} JO {F0}")
        assert np.linalg.norm(F - F0) <= 1e-12 * np.linalg.norm(F0)
        assert np.linalg.norm(T) <= 1e-12 * np.linalg.norm(F0) * a1
        # tangential sliding
        t = np.cross(n, [0.0, 0.0, 1.0])
        t /= np.linalg.norm(t)
        F, T, _ = place_pair(o, a1, a2, h, n, U1=0.01 * t)
        F0 = -6 * math.pi * mu * a1 * jo["YA_g2"] * ln * 0.01 * t
        assert np.linalg.norm(F - F0) <= 1e-12 * np.linalg.norm(F0)
        # torque from sliding; reciprocity fixes its coefficient to the one
        # of the force from rotation (below)
        c_lt = np.cross(0.01 * t, n)
        kLT = -np.dot(T, c_lt) / np.dot(c_lt, c_lt)
        # rotation about an axis normal to n
        W = 0.002 * t
        F, T, _ = place_pair(o, a1, a2, h, n, W1=W)
        c_fr = np.cross(W, n)
        kFR = np.dot(F, c_fr) / np.dot(c_fr, c_fr)
        # the coupling's magnitude is Jeffrey & Onishi's Y^B_11 (force
        # normalised by 4 pi mu a^2)
        assert kFR == pytest.approx(4 * math.pi * mu * a1 ** 2 *
                                    abs(jo["YB_g2"]) * ln, rel=1e-12)
        assert kLT == pytest.approx(kFR, rel=1e-12)
        # rotation about n: torque only, Y^C_11 (normalised by 8 pi mu a^3)
        F, T, _ = place_pair(o, a1, a2, h, n, W1=0.002 * n)
        T0 = -8 * math.pi * mu * a1 ** 3 * jo["YC_g2"] * ln * 0.002 * n
        assert np.linalg.norm(T - T0) <= 1e-12 * np.linalg.norm(T0)
        assert np.linalg.norm(F) * a1 <= 1e-12 * np.linalg.norm(T0)
        # beyond the cut-off: nothing
        F, T, _ = place_pair(o, a1, a2, hN + 0.1, n, U1=0.01 * n)
        assert np.all(F == 0) and np.all(T == 0)
    finally:
        o.teardown()


@pytest.mark.parametrize("a1,a2,delta", [(4.0, 4.0, 0.02), (4.0, 6.0, 0.05),
                                         (6.0, 3.0, 0.01)])
def test_pair_contact_hertz(tmp_path, a1, a2, delta):
    o, mu = pair_box(tmp_path, a1, a2)
    try:
        n = np.array([-0.6, 0.0, 0.8])
        F, T, n = place_pair(o, a1, a2, -delta, n)
        E1, s1, E2, s2 = 2e7, 0.3, 1e7, 0.45
        Es = 1.0 / ((1 - s1 ** 2) / E1 + (1 - s2 ** 2) / E2)
        Rs = 1.0 / (1.0 / a1 + 1.0 / a2)
        F0 = 4.0 / 3.0 * Es * math.sqrt(Rs) * delta ** 1.5 * n
        print(f"Hertz {np.linalg.norm(F0):.6e}, oracle {np.linalg.norm(F):.6e}")
        assert np.linalg.norm(F - F0) <= 1e-12 * np.linalg.norm(F0)
        assert np.linalg.norm(T) <= 1e-12 * np.linalg.norm(F0) * a1
        iSt, St, ncp, ncw = o.contacts(0)
        assert iSt.count(1) == 1 and ncp == 1
    finally:
        o.teardown()


def spring_box(tmp, k=2.0, anchor=(20.0, 24.0, 26.0), l=3.0, pos=None,
               rho=2.0):
    kw, parts = bbconfig.sediment(n=(48, 48, 48), nparts=1, a=4.0, seed=1,
                                  rho_ratio=rho)
    kw["g"] = ((0.0, 0.0), (0.0, 0.0), (0.0, 0.0))
    p = (22.0, 25.0, 27.5) if pos is None else pos
    parts[0].update(x=p[0], y=p[1], z=p[2], spring_k=k, spring_x=anchor[0],
                    spring_y=anchor[1], spring_z=anchor[2], spring_l=l,
                    rotating=0)
    d = str(tmp)
    bbconfig.write_case(d, **kw)
    bbconfig.write_parts(d, parts)
    o = orc.Oracle(d)
    o.setup()
    return o


def test_spring_hooke(tmp_path):
    k, l = 2.0, 3.0
    anchor = np.array([20.0, 24.0, 26.0])
    o = spring_box(tmp_path, k=k, anchor=anchor, l=l)
    try:
        z = np.zeros(3)
        vol = 4.0 / 3.0 * math.pi * 4.0 ** 3
        m = 2.0 * vol
        o.set_scalar("dt", 0.01)
        for x in (np.array([22.0, 25.0, 27.5]), np.array([14.0, 30.0, 21.0]),
                  np.array([20.5, 24.0, 26.0])):
            o.set_motion(0, x, z, z)
            o.move_sub()                    # collide, move_a, collide
            s = o.part_state(0)
            d = x - anchor
            Fk = -k * (np.linalg.norm(d) - l) * d / np.linalg.norm(d)
            acc = np.array([s["udot"], s["vdot"], s["wdot"]])
            assert np.allclose(acc * m, Fk, rtol=1e-13, atol=1e-15)
    finally:
        o.teardown()


def test_spring_across_periodic_face(tmp_path):
    """A centre stored beyond the x face is measured from its wrapped
    position (src/particle_kernel.cu:2318-2320)."""
    k, l = 1.5, 0.0
    anchor = np.array([2.0, 24.0, 24.0])
    o = spring_box(tmp_path, k=k, anchor=anchor, l=l)
    try:
        z = np.zeros(3)
        vol = 4.0 / 3.0 * math.pi * 4.0 ** 3
        m = 2.0 * vol
        o.set_scalar("dt", 0.01)
        x = np.array([-1.0, 24.0, 25.0])     # wraps to 47.0
        o.set_motion(0, x, z, z)
        o.move_sub()
        s = o.part_state(0)
        xw = x + np.array([L, 0.0, 0.0])
        Fk = -k * (xw - anchor)
        acc = np.array([s["udot"], s["vdot"], s["wdot"]])
        assert np.allclose(acc * m, Fk, rtol=1e-13, atol=1e-15)
    finally:
        o.teardown()


def oscillate(o, x0, anchor, dt, tend):
    z = np.zeros(3)
    o.set_scalar("dt", dt)
    o.set_motion(0, np.asarray(x0, float), z, z)
    nsteps = int(round(tend / dt))
    for _ in range(nsteps):
        o.move_sub()
        o.move_b()
    return np.array([o.part_state(0)[c] for c in "xyz"])


def test_spring_oscillator_converges(tmp_path):
    """x'' = -(k/m) (x - anchor) from rest at amplitude A: after one period
    the sphere is back near its start, and the error falls in proportion to
    dt.  The update is first order for a force that depends on the
    position alone: the velocity takes the trapezoid of the acceleration at
    the step's start and the stored one of the step before
    (src/particle_kernel.cu:3368-3410), half a step behind the midpoint; a
    wrong spring sign, mass or factor would not converge to the
    oscillator."""
    k = 1.0
    anchor = np.array([24.0, 24.0, 24.0])
    o = spring_box(tmp_path, k=k, anchor=anchor, l=0.0,
                   pos=(25.0, 24.0, 24.0))
    try:
        m = 2.0 * 4.0 / 3.0 * math.pi * 4.0 ** 3
        period = 2 * math.pi * math.sqrt(m / k)
        x0 = anchor + np.array([1.0, 0.0, 0.0])
        errs = []
        for nper in (200, 400, 800):
            x = oscillate(o, x0, anchor, period / nper, period)
            errs.append(np.linalg.norm(x - x0))
        print("oscillator errors after one period", errs)
        assert errs[0] < 0.15
        assert 1.8 < errs[0] / errs[1] < 2.2
        assert 1.8 < errs[1] / errs[2] < 2.2
    finally:
        o.teardown()


def test_free_fall_acceleration(tmp_path):
    kw, parts = bbconfig.sediment(n=(48, 48, 48), nparts=1, a=4.0, seed=1,
                                  rho_ratio=3.0)
    kw["g"] = ((0.2, 0.0), (0.0, 0.0), (-9.81, 0.0))
    parts[0].update(x=24.0, y=24.0, z=24.0)
    d = str(tmp_path)
    bbconfig.write_case(d, **kw)
    bbconfig.write_parts(d, parts)
    with orc.Oracle(d) as o:
        o.set_scalar("dt", 0.01)
        o.lib.orc_forcing()                 # sets g for this time
        o.move_sub()
        s = o.part_state(0)
        acc = np.array([s["udot"], s["vdot"], s["wdot"]])
        want = (1.0 - 1.0 / 3.0) * np.array([0.2, 0.0, -9.81])
        assert np.allclose(acc, want, rtol=1e-15, atol=0)
        # trapezoidal velocity from rest: u = dt/2 (udot + udot0), udot0 = 0
        u = np.array([s["u"], s["v"], s["w"]])
        assert np.allclose(u, 0.005 * want, rtol=1e-15, atol=0)


def rodrigues(axis, theta):
    k = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(theta) * K + (1 - math.cos(theta)) * K @ K


@pytest.mark.parametrize("om", [(0.0, 0.0, 0.3), (0.2, -0.1, 0.25)])
def test_body_axes_rotation(tmp_path, om):
    kw, parts = bbconfig.sediment(n=(48, 48, 48), nparts=1, a=4.0, seed=1)
    kw["g"] = ((0.0, 0.0), (0.0, 0.0), (0.0, 0.0))
    parts[0].update(x=24.0, y=24.0, z=24.0, translating=0)
    d = str(tmp_path)
    bbconfig.write_case(d, **kw)
    bbconfig.write_parts(d, parts)
    with orc.Oracle(d) as o:
        dt = 0.05
        o.set_scalar("dt", dt)
        om = np.array(om)
        o.set_motion(0, np.full(3, 24.0), np.zeros(3), om)
        s0 = o.part_state(0)
        ax0 = np.array([[s0["axx"], s0["axy"], s0["axz"]],
                        [s0["ayx"], s0["ayy"], s0["ayz"]],
                        [s0["azx"], s0["azy"], s0["azz"]]])
        assert np.allclose(ax0, np.eye(3))
        nstep = 3
        for _ in range(nstep):
            o.move_b()
        s = o.part_state(0)
        ax = np.array([[s["axx"], s["axy"], s["axz"]],
                       [s["ayx"], s["ayy"], s["ayz"]],
                       [s["azx"], s["azy"], s["azz"]]])
        R = rodrigues(om, np.linalg.norm(om) * dt * nstep)
        # each body axis (a row) turns with the body: a' = R a
        assert np.allclose(ax, (R @ ax0.T).T, rtol=0, atol=1e-14)
        assert np.allclose(ax @ ax.T, np.eye(3), atol=1e-14)


def test_poiseuille_profile_is_steady(tmp_path):
    """Periodic in x and y, no-slip walls at z = -1 and 1, constant
    pressure gradient dp/dx = -G: u = G/(2 mu) (1 - z^2) balances the
    gradient exactly in the discrete equations (the 7-point Laplacian and
    the wall ghost formula are exact on a quadratic), so it does not
    change; a wrong sign or factor of the forcing would accelerate it."""
    G = 1.0
    kw = bbconfig.channel(n=(8, 4, 16), amp=0.0, init_cond="QUIESCENT")
    kw["extensions"] = {"fixed_dt": 1e-3}
    d = str(tmp_path)
    bbconfig.write_case(d, **kw)
    with orc.Oracle(d) as o:
        mu = kw["nu"] * kw["rho_f"]
        nz = 16
        dz = 2.0 / nz
        zc = -1.0 + (np.arange(nz + 2) - 0.5) * dz
        prof = G / (2 * mu) * (1 - zc ** 2)
        u = o.field("u")
        u[...] = prof[None, None, :]
        o.field("u0")[...] = prof[None, None, :]
        ref = np.array(u)[1:-1, 1:-1, 1:-1].copy()
        o.step(3)
        got = np.array(o.field("u"))[1:-1, 1:-1, 1:-1]
        err = np.abs(got - ref).max() / ref.max()
        print("Poiseuille drift after 3 steps", err)
        assert err < 1e-12
        for c in "vw":
            assert np.abs(np.array(o.field(c))[1:-1, 1:-1, 1:-1]).max() < 1e-12


def test_time_step_velocity_branch(tmp_path):
    kw = bbconfig.channel(n=(16, 8, 12), amp=0.0, init_cond="QUIESCENT")
    kw["extensions"] = {}
    d = str(tmp_path)
    bbconfig.write_case(d, **kw)
    with orc.Oracle(d) as o:
        rng = np.random.default_rng(4)
        for c, vmax in (("u", 1.7), ("v", -0.9), ("w", 0.4)):
            f = o.field(c)
            f[...] = rng.uniform(-0.5, 0.5, f.shape) * abs(vmax)
            f[3, 2, 5] = vmax
        o.lib.orc_find_dt()
        dt = o.scalar("dt")
        dx, dy, dz = 4 * math.pi / 16, 2 * math.pi / 8, 2.0 / 12
        nu = kw["nu"]
        want = kw["CFL"] / (1.7 / dx + 2 * nu / dx ** 2 + 0.9 / dy
                            + 2 * nu / dy ** 2 + 0.4 / dz + 2 * nu / dz ** 2)
        assert dt == pytest.approx(want, rel=1e-14)
