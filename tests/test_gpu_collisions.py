"""GPU parity of the particle interaction and motion branches against the
FP64 host oracle: pair contact (Hertz spring, dashpot, Mindlin friction and
the contact lists), unequal-sphere lubrication, contact through a periodic
face, the anchor spring, body-axis rotation, and a box with enough spheres
that the collision bins hold several of them
(src/particle_kernel.cu:578-905, 2303-2345, 3075-3484; src/particle.c:744-834).

Bar: 1e-10 relative to the oracle (the device's sqrt, log, pow, sin and cos
are not the host libm, and the force bodies group their terms differently
from the oracle's restatement), contact lists and collision counts equal.
"""
import ctypes

import numpy as np
import pytest

import bbconfig
import bbhip
import orc
from helpers import interior, rel_l2

pytestmark = pytest.mark.gpu

AXES = ("axx", "axy", "axz", "ayx", "ayy", "ayz", "azx", "azy", "azz")
MOTION = ("x", "y", "z", "u", "v", "w", "ox", "oy", "oz")


def _set_device_motion(sim, state):
    sim.lib.cuda_part_pull()
    P = ctypes.POINTER(bbhip.PartStruct).in_dll(sim.lib, "parts")
    for k, (x, v, w) in enumerate(state):
        for name, val in zip(("x", "y", "z", "u", "v", "w", "u0", "v0",
                              "w0", "ox", "oy", "oz", "ox0", "oy0", "oz0"),
                             tuple(x) + tuple(v) + tuple(v) + tuple(w)
                             + tuple(w)):
            setattr(P[k], name, val)
    sim.lib.cuda_part_push()


def _substep_case(case_dir, n, parts, state):
    """One particle sub-step (collisions, velocity update, collisions) from
    the same state on the device and in the oracle."""
    kw, _ = bbconfig.sediment(n=n, nparts=len(parts), a=4.0, seed=5)
    for p, (x, _, _) in zip(parts, state):
        p.update(x=x[0], y=x[1], z=x[2])
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    with orc.Oracle(case_dir) as o:
        for k, (x, v, w) in enumerate(state):
            o.set_motion(k, np.array(x), np.array(v), np.array(w))
        o.move_sub()
        want = [o.part_state(k) for k in range(len(state))]
        lists = [o.contacts(k) for k in range(len(state))]
    with bbhip.Simulation(case_dir, kw["n"]) as s:
        assert ctypes.c_int.in_dll(s.lib, "nparts").value == len(state)
        _set_device_motion(s, state)
        s.lib.cuda_move_parts_sub()
        got = s.parts()
    return want, lists, got


def _compare_substep(want, lists, got, tol=1e-10):
    worst = 0.
    for k in range(len(want)):
        for q in ("iFx", "iFy", "iFz", "iLx", "iLy", "iLz", "u", "v", "w",
                  "ox", "oy", "oz"):
            a, b = got[k][q], want[k][q]
            e = abs(a - b) / max(abs(b), 1e-300)
            worst = max(worst, e if b != 0 else abs(a))
            assert (a == b) or e <= tol, (k, q, a, b)
        ist, st, npart, nwall = lists[k]
        assert list(got[k]["iSt"]) == ist, (k, got[k]["iSt"], ist)
        assert np.allclose(got[k]["St"], st, rtol=1e-12, atol=0), k
        assert got[k]["ncoll_part"] == npart, k
        assert got[k]["ncoll_wall"] == nwall, k
    print("worst relative difference", f"{worst:.2e}")


def test_pair_contact_one_substep(case_dir):
    """Two overlapping pairs (overlaps 0.03 and 0.05, one of them through
    the periodic x face) with slip and rotation, and an unequal pair
    (radii 4 and 3) inside the lubrication range."""
    parts = [dict(r=4.0, x=0, y=0, z=0, rho=2.0) for _ in range(6)]
    parts[4]["r"] = 3.0
    state = [  # position, velocity, rotation
        ((10.0, 10.0, 10.0), (0.02, 0.003, -0.001), (0.001, -0.002, 0.003)),
        ((17.97, 10.2, 10.1), (-0.01, 0.004, 0.002), (-0.002, 0.001, 0.001)),
        ((1.5, 24.0, 20.0), (-0.015, 0.002, 0.001), (0.002, 0.001, -0.002)),
        ((25.55, 24.3, 19.8), (0.01, -0.003, 0.002), (0.001, -0.001, 0.002)),
        ((16.0, 22.0, 4.0), (0.001, -0.02, 0.002), (-0.003, 0.002, 0.001)),
        ((16.4, 30.3, 4.5), (0.002, 0.01, -0.001), (0.002, -0.001, 0.003)),
    ]
    want, lists, got = _substep_case(case_dir, (32, 32, 32), parts, state)
    _compare_substep(want, lists, got)
    # both contacts registered on both sides, the lubrication pair not
    for k, partner in ((0, 1), (1, 0), (2, 3), (3, 2)):
        assert partner in lists[k][0], (k, lists[k])
        assert lists[k][2] == 1
    for k in (4, 5):
        assert lists[k][2] == 0
        assert abs(want[k]["iFy"]) > 0.


def test_contact_lists_persist_and_clear(case_dir):
    """Four sub-steps (dt 1e-4) of a pair that starts in contact (overlap
    1e-4) and separates: the contact keeps its first Stokes number while it
    lasts and leaves the list once the gap opens (lubrication branch)."""
    parts = [dict(r=4.0, x=0, y=0, z=0, rho=2.0) for _ in range(2)]
    state = [((12.0, 16.0, 16.0), (-0.4, 0.0, 0.01), (0.0, 0.0, 0.01)),
             ((19.9999, 16.0, 16.0), (0.4, 0.0, -0.01), (0.0, 0.01, 0.0))]
    kw, _ = bbconfig.sediment(n=(32, 32, 32), nparts=2, a=4.0, seed=5,
                              extensions=dict(fixed_dt=1e-4))
    for p, (x, _, _) in zip(parts, state):
        p.update(x=x[0], y=x[1], z=x[2])
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    want = []
    with orc.Oracle(case_dir) as o:
        for k, (x, v, w) in enumerate(state):
            o.set_motion(k, np.array(x), np.array(v), np.array(w))
        for _ in range(4):
            o.move_sub()
            o.move_b()
            want.append(([o.part_state(k) for k in range(2)],
                         [o.contacts(k) for k in range(2)]))
    got = []
    with bbhip.Simulation(case_dir, kw["n"]) as s:
        _set_device_motion(s, state)
        for _ in range(4):
            s.lib.cuda_move_parts_sub()
            s.lib.cuda_update_part_position()
            got.append(s.parts())
    for step, ((wst, wl), g) in enumerate(zip(want, got)):
        for k in range(2):
            for q in MOTION + ("iFx", "iLy"):
                a, b = g[k][q], wst[k][q]
                assert a == b or abs(a - b) <= 1e-10 * max(abs(b), 1e-300), \
                    (step, k, q, a, b)
            assert list(g[k]["iSt"]) == wl[k][0], (step, k)
            assert np.allclose(g[k]["St"], wl[k][1], rtol=1e-12, atol=0)
            assert g[k]["ncoll_part"] == wl[k][2]
    assert want[0][1][0][0][0] == 1             # in contact at first
    assert want[-1][1][0][0][0] == -1           # and cleared at the end


def _run_steps(case_dir, kw, parts, nsteps, state=None):
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    out = {}
    with orc.Oracle(case_dir) as o:
        if state:
            for k, (x, v, w) in enumerate(state):
                o.set_motion(k, np.array(x), np.array(v), np.array(w))
        lamb = []
        for _ in range(nsteps):
            o.step()
            lamb.append(int(o.scalar("lamb_iters")))
        out["orc_parts"] = [o.part_state(k) for k in range(o.nparts())]
        out["orc"] = {f: np.array(o.field(f)) for f in ("u", "v", "w", "p")}
        out["lamb_orc"] = lamb
    with bbhip.Simulation(case_dir, kw["n"]) as s:
        if state:
            _set_device_motion(s, state)
        lamb = []
        for _ in range(nsteps):
            s.step()
            lamb.append(s.lamb_stats()[0])
        s.pull()
        out["gpu_parts"] = s.parts()
        out["gpu"] = {f: np.array(s.field(f)) for f in ("u", "v", "w", "p")}
        out["lamb_gpu"] = lamb
    return out


def _check_steps(out, keys, tol=1e-10):
    for f in ("u", "v", "w", "p"):
        e = rel_l2(interior(out["gpu"][f]), interior(out["orc"][f]))
        print(f, f"rel L2 {e:.2e}")
        assert e <= tol, f
    assert out["lamb_gpu"] == out["lamb_orc"]
    worst = 0.
    for k in keys:
        a = np.array([p[k] for p in out["gpu_parts"]])
        b = np.array([p[k] for p in out["orc_parts"]])
        e = rel_l2(a, b)
        worst = max(worst, e)
        assert e <= tol, (k, a, b)
    print("worst particle rel L2", f"{worst:.2e}")


def test_spring_and_rotating_sphere(case_dir):
    """A sphere on a spring (anchor 1.5 dx away, rest length 0.5) that
    starts rotating about a tilted axis, and a free rotating sphere: three
    full timesteps; positions, velocities, rotation rates and body axes."""
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=2, a=4.0, seed=9)
    parts[0].update(x=10.0, y=12.0, z=14.0, spring_k=50.0, spring_x=11.0,
                    spring_y=12.5, spring_z=15.0, spring_l=0.5)
    parts[1].update(x=23.0, y=22.0, z=20.0)
    state = [((10.0, 12.0, 14.0), (0.0, 0.0, 0.0), (0.02, -0.01, 0.03)),
             ((23.0, 22.0, 20.0), (0.01, 0.0, 0.0), (-0.01, 0.04, 0.02))]
    out = _run_steps(case_dir, kw, parts, 3, state)
    _check_steps(out, MOTION + AXES + ("Fx", "Fz", "Lx", "Lz"))
    g = out["gpu_parts"]
    assert abs(g[0]["kFx"]) > 0. and abs(g[0]["kFz"]) > 0.
    a = np.array([g[1][k] for k in AXES]).reshape(3, 3)
    assert np.allclose(a @ a.T, np.eye(3), atol=1e-12)   # still orthonormal
    assert abs(a[0, 0] - 1.) > 1e-8                      # and turned


def test_many_spheres_64(case_dir):
    """A 64^3 periodic box with 48 spheres of radius 4 (gaps from 0.4 dx):
    bins holding several spheres, pairs inside the lubrication range and
    images across every periodic face; two timesteps."""
    kw, parts = bbconfig.sediment(n=(64, 64, 64), nparts=48, a=4.0,
                                  seed=20260104)
    bins = {}
    for p in parts:
        key = tuple(int(p[c] // 9.0) for c in "xyz")
        bins[key] = bins.get(key, 0) + 1
    assert max(bins.values()) >= 2
    out = _run_steps(case_dir, kw, parts, 2)
    _check_steps(out, MOTION + ("Fx", "Fy", "Fz", "iFx", "iFy", "iFz"))
    assert sum(abs(p["iFx"]) > 0 for p in out["orc_parts"]) >= 4
