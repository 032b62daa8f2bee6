"""Particle-wall forces of the oracle (src/particle_kernel.cu:2347-3073).

Pinned by closed forms and by the symmetries of a box, not by restating the
formulas:
- lubrication normal to each face: 6 pi mu a^2 U (1/h - 1/hN), repulsive
  (the sphere-wall lubrication asymptote the reference uses);
- contact normal to each face with no relative motion: the Hertz force
  (4/3) E* sqrt(a) delta^(3/2), E* = E / (2 (1 - sigma^2));
- lubrication with arbitrary translation and rotation: a rotation of the
  box taking W to S to B (x -> y -> z) and the mirrors W <-> E, S <-> N,
  B <-> T carry forces and torques with them (vectors and pseudo-vectors);
- contact with tangential slip (walls at rest): S and B are images under
  the rotation; the upper faces N, T, E apply their lower partner's law to
  the reversed slip, and the W and E faces differ from S as the reference
  writes them (W: tangential spring with the normal stiffness k; E: no z
  torque), each checked as a relation between faces;
- the contact list: a contact enters code -10 - face once, lubrication
  range clears it; a periodic face adds no force but keeps the list.
"""
import math

import numpy as np
import pytest

import bbconfig
import orc

FACES = "WESNBT"
AX = {"W": 0, "E": 0, "S": 1, "N": 1, "B": 2, "T": 2}
SG = {"W": 1.0, "E": -1.0, "S": 1.0, "N": -1.0, "B": 1.0, "T": -1.0}
L = 24.0
A = 4.0


def box(tmp, walls="xyz"):
    kw, parts = bbconfig.sediment(n=(24, 24, 24), nparts=1, a=A, seed=1,
                                  walls=walls)
    kw["g"] = ((0.0, 0.0), (0.0, 0.0), (0.0, 0.0))
    parts[0].update(x=L / 2, y=L / 2, z=L / 2)
    d = str(tmp)
    bbconfig.write_case(d, **kw)
    bbconfig.write_parts(d, parts)
    o = orc.Oracle(d)
    o.setup()
    return o, kw


def pos_at(face, h, shift=(0.0, 0.0, 0.0)):
    """Centre at gap h from the face (inside the box), other coordinates
    at the box centre plus `shift`."""
    p = np.array([L / 2, L / 2, L / 2]) + np.asarray(shift)
    a = AX[face]
    p[a] = (A + h) if SG[face] > 0 else L - (A + h)
    return p


def forces(o, pos, vel, om, vel0=None, om0=None):
    o.set_motion(0, np.asarray(pos, float), np.asarray(vel, float),
                 np.asarray(om, float),
                 None if vel0 is None else np.asarray(vel0, float),
                 None if om0 is None else np.asarray(om0, float))
    o.collide()
    s = o.part_state(0)
    F = np.array([s["iFx"], s["iFy"], s["iFz"]])
    T = np.array([s["iLx"], s["iLy"], s["iLz"]])
    return F, T


def R(v):
    """Rotation taking e_x -> e_y -> e_z -> e_x (W -> S -> B, E -> N -> T)."""
    v = np.asarray(v, float)
    return np.array([v[2], v[0], v[1]])


def M(ax):
    m = np.ones(3)
    m[ax] = -1.0
    return m


def close(a, b, tol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) <= tol * scale


@pytest.mark.parametrize("face", FACES)
def test_normal_lubrication_closed_form(tmp_path, face):
    o, kw = box(tmp_path)
    mu = kw["nu"] * kw["rho_f"]
    h, U = 0.3 * A, 0.02
    hN = 1 * A                           # (l/a) = 1
    vel = np.zeros(3)
    vel[AX[face]] = -SG[face] * U        # approaching the face
    F, T = forces(o, pos_at(face, h), vel, np.zeros(3))
    want = np.zeros(3)
    want[AX[face]] = SG[face] * 6 * math.pi * mu * A * A * U * (1 / h - 1 / hN)
    o.teardown()
    assert close(F, want, 1e-13), (F, want)
    assert np.all(T == 0.0)


@pytest.mark.parametrize("face", FACES)
def test_contact_hertz_normal(tmp_path, face):
    o, kw = box(tmp_path)
    delta = 0.01
    F, T = forces(o, pos_at(face, -delta), np.zeros(3), np.zeros(3))
    E, sig = 1e7, 0.5
    Estar = E / (2 * (1 - sig * sig))
    want = np.zeros(3)
    want[AX[face]] = SG[face] * 4.0 / 3.0 * Estar * math.sqrt(A) * delta ** 1.5
    ist, st, npart, nwall = o.contacts(0)
    code = -10 - FACES.index(face)
    # the contact enters the list once
    forces(o, pos_at(face, -delta), np.zeros(3), np.zeros(3))
    ist2, _, _, nwall2 = o.contacts(0)
    # leaving for the lubrication range clears it
    forces(o, pos_at(face, 0.5 * A), np.zeros(3), np.zeros(3))
    ist3, _, _, _ = o.contacts(0)
    o.teardown()
    assert close(F, want, 1e-12), (F, want)
    assert np.all(T == 0.0)
    assert ist.count(code) == 1 and nwall == 1
    assert ist2.count(code) == 1 and nwall2 == 1
    assert code not in ist3


def test_lubrication_symmetries(tmp_path):
    o, _ = box(tmp_path)
    rng = np.random.default_rng(11)
    vel = rng.uniform(-0.03, 0.03, 3)
    om = rng.uniform(-0.01, 0.01, 3)
    shift = np.array([0.3, -0.4, 0.2])
    h = 0.4 * A
    got = {}
    for face in FACES:
        k = "WSB".find(face)
        if k < 0:
            continue
        v, w, s = vel, om, shift
        for _ in range(k):
            v, w, s = R(v), R(w), R(s)
        got[face] = forces(o, pos_at(face, h, s), v, w)
        # the mirror image at the opposite face
        ax = AX[face]
        m = M(ax)
        got[FACES[FACES.index(face) + 1]] = (
            forces(o, pos_at(FACES[FACES.index(face) + 1], h, m * s), m * v,
                   -m * w),
            m)
    o.teardown()
    FW, TW = got["W"]
    assert np.any(FW != 0) and np.any(TW != 0)
    FS, TS = got["S"]
    FB, TB = got["B"]
    assert close(FS, R(FW)) and close(TS, R(TW))
    assert close(FB, R(R(FW))) and close(TB, R(R(TW)))
    for lo, hi in (("W", "E"), ("S", "N"), ("B", "T")):
        (Fh, Th), m = got[hi]
        Fl, Tl = got[lo]
        assert close(Fh, m * Fl), (hi, Fh, m * Fl)
        assert close(Th, -m * Tl), (hi, Th, -m * Tl)


def contact(o, face, vel, om, vel0, om0, shift, delta=0.02):
    return forces(o, pos_at(face, -delta, shift), vel, om, vel0, om0)


def test_contact_face_forms(tmp_path):
    """Contact with slip and rotation (walls at rest).  The S and B faces
    are images of each other under the box rotation.  Each upper face (N,
    T) applies its lower partner's law to the slip with the opposite sign,
    as the reference writes them (V = -(U_t +- (a + h/2) w)): with the same
    tangential motion and the normal velocity reversed, its force is the
    negative of the lower face's and its torque the same."""
    o, _ = box(tmp_path)
    rng = np.random.default_rng(5)
    vel = rng.uniform(-0.02, 0.02, 3)
    vel0 = rng.uniform(-0.02, 0.02, 3)
    om = rng.uniform(-0.01, 0.01, 3)
    om0 = rng.uniform(-0.01, 0.01, 3)
    shift = np.array([0.2, 0.1, -0.3])
    FS, TS = contact(o, "S", vel, om, vel0, om0, shift)
    FB, TB = contact(o, "B", R(vel), R(om), R(vel0), R(om0), R(shift))
    assert np.any(FS != 0) and np.any(TS != 0)
    assert close(FB, R(FS)) and close(TB, R(TS))
    for lo, hi, ax, v, w, v0, w0 in (
            ("S", "N", 1, vel, om, vel0, om0),
            ("B", "T", 2, R(vel), R(om), R(vel0), R(om0))):
        Fl, Tl = contact(o, lo, v, w, v0, w0, np.zeros(3))
        m = M(ax)
        Fh, Th = contact(o, hi, m * v, w, m * v0, w0, np.zeros(3))
        assert close(Fh, -Fl), (hi, Fh, Fl)
        assert close(Th, Tl), (hi, Th, Tl)
    o.teardown()


def test_contact_x_faces(tmp_path):
    """The W and E faces with slip but no rotation, below the Coulomb
    limit, against the S face: equal normal forces; the W face's tangential
    spring takes the normal stiffness k where the others take kt; the E
    face is the W law with kt and the slip reversed, and adds no z torque."""
    o, _ = box(tmp_path)
    zero = np.zeros(3)
    vS = np.array([1e-4, 0.0, -2e-4])        # tangential at S: x and z
    FS, TS = contact(o, "S", vS, zero, vS, zero, zero)
    # R^-1 takes the S configuration to W (y -> x, z -> y, x -> z)
    Rinv = lambda v: np.array([v[1], v[2], v[0]])
    vW = Rinv(vS)
    FW, TW = contact(o, "W", vW, zero, vW, zero, zero)
    m = M(0)
    FE, TE = contact(o, "E", m * vW, zero, m * vW, zero, zero)
    E, sig, delta = 1e7, 0.5, 0.02
    Hi = 0.5 * E / (1 + sig)
    kt = 8.0 / (2 * (1 - sig * sig) / Hi) * math.sqrt(A) * math.sqrt(delta)
    k = 4.0 / 3.0 / (2 * (1 - sig * sig) / E) * math.sqrt(A)
    o.teardown()
    assert abs(FS[0]) > 0 and abs(FS[2]) > 0
    assert close(FW[0], Rinv(FS)[0])
    assert close(FW[1:], Rinv(FS)[1:] * (k / kt), 1e-10)
    assert close(TW, Rinv(TS) * (k / kt), 1e-10)
    assert close(FE[0], -FW[0])
    assert close(FE[1:], -FW[1:] * (kt / k), 1e-10)
    assert close(TE[1], TW[1] * (kt / k), 1e-10)
    assert TE[0] == 0.0 and TE[2] == 0.0


def test_periodic_face_keeps_list_without_force(tmp_path):
    o, _ = box(tmp_path, walls="yz")         # x periodic
    pos = np.array([1.0, L / 2, L / 2])      # overlaps the x = 0 plane
    F, T = forces(o, pos, np.array([-0.01, 0.0, 0.0]), np.zeros(3))
    ist, _, _, nwall = o.contacts(0)
    o.teardown()
    assert np.all(F == 0) and np.all(T == 0)
    assert -10 in ist and nwall == 1
