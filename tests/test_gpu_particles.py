"""GPU parity of the particle / Physalis path against the FP64 host oracle.

Cases (boxes with dx = 1, spheres of radius a at seeded positions, Lamb
order 4, gravity from Ga = 100):
  - a few spheres, two timesteps with the full Lamb iteration;
  - a sphere cut by the periodic faces (cage, nodes and BCs through the
    images);
  - two spheres closing in (lubrication inside the interaction length);
  - spheres at no-slip walls (cages cut by a wall, nodes beyond it, wall
    lubrication and contact), with walls in z only and on all faces;
  - the mixed-precision pressure solve with particles.
Bar: integer fields (phase, phase_shell, face flags) identical; u, v, w, p
and every particle quantity (position, velocity, forces, Lamb
coefficients) within relative L2 1e-10 of the oracle (the device's sin,
cos, acos, pow and log are not the host libm, so a last-bit difference in
a node angle can propagate through the solve); Lamb iteration counts equal.
"""
import numpy as np
import pytest

import bbconfig
import bbhip
import orc
from helpers import describe, interior, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p")
IFIELDS = ("phase", "phase_shell", "flag_u", "flag_v", "flag_w")
PSCAL = ("x", "y", "z", "u", "v", "w", "ox", "oy", "oz", "Fx", "Fy", "Fz",
         "Lx", "Ly", "Lz", "iFx", "iFy", "iFz")
PCOEF = ("pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re", "chinm_im")


def run(case_dir, kw, parts, nsteps, precision=None):
    kw = dict(kw)
    if precision:
        kw["extensions"] = dict(kw.get("extensions") or {},
                                pp_precision=precision)
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    out = {"orc": {}, "gpu": {}, "lamb_orc": [], "lamb_gpu": []}
    with orc.Oracle(case_dir) as o:
        out["orc_ifields0"] = {f: np.array(o.field(f)) for f in IFIELDS}
        for _ in range(nsteps):
            o.step()
            out["lamb_orc"].append(int(o.scalar("lamb_iters")))
        for f in FIELDS + IFIELDS:
            out["orc"][f] = np.array(o.field(f))
        out["orc_parts"] = [o.part_state(n) for n in range(o.nparts())]
    with bbhip.Simulation(case_dir, kw["n"]) as s:
        s.pull()
        out["gpu_ifields0"] = {f: np.array(s.field(f)) for f in IFIELDS}
        for _ in range(nsteps):
            s.step()
            out["lamb_gpu"].append(s.lamb_stats()[0])
        s.pull()
        for f in FIELDS + IFIELDS:
            out["gpu"][f] = np.array(s.field(f))
        out["gpu_parts"] = s.parts()
    return out


def check(out, tol=1e-10, fields=FIELDS):
    for f in IFIELDS:
        a, b = out["gpu_ifields0"][f], out["orc_ifields0"][f]
        assert np.array_equal(interior(a), interior(b)), f"initial {f}"
        a, b = out["gpu"][f], out["orc"][f]
        assert np.array_equal(interior(a), interior(b)), f
    errs = {f: rel_l2(interior(out["gpu"][f]), interior(out["orc"][f]))
            for f in fields}
    print("fields rel L2:", {f: f"{e:.2e}" for f, e in errs.items()})
    print("lamb iterations gpu", out["lamb_gpu"], "oracle", out["lamb_orc"])
    for f in fields:
        assert errs[f] <= tol, describe(f, interior(out["gpu"][f]),
                                        interior(out["orc"][f]))
    assert out["lamb_gpu"] == out["lamb_orc"]
    g = out["gpu_parts"]
    o = out["orc_parts"]
    assert len(g) == len(o)
    worst = 0.
    for k in PSCAL:
        a = np.array([p[k] for p in g])
        b = np.array([p[k] for p in o])
        e = rel_l2(a, b)
        worst = max(worst, e)
        assert e <= tol, (k, a, b)
    for k in PCOEF:
        a = np.array([p[k] for p in g])
        b = np.array([p[k] for p in o])
        e = rel_l2(a, b)
        worst = max(worst, e)
        assert e <= tol, (k, a, b)
    print("worst particle rel L2", f"{worst:.2e}")


def test_sediment_three_spheres(case_dir):
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=3, a=4.0, seed=7)
    check(run(case_dir, kw, parts, 2))


def test_sphere_across_periodic_faces(case_dir):
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=1, a=4.0, seed=7)
    parts[0].update(x=1.25, y=31.5, z=0.4)
    check(run(case_dir, kw, parts, 2))


def test_lubrication_pair(case_dir):
    """Two spheres 0.6 dx apart (inside (l/a) a) with opposite velocities."""
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=2, a=4.0, seed=3)
    parts[0].update(x=11.0, y=16.0, z=16.0, u=0.05)
    parts[1].update(x=19.6, y=16.3, z=16.2, u=-0.05)
    out = run(case_dir, kw, parts, 2)
    check(out)
    assert abs(out["orc_parts"][0]["iFx"]) > 0.


def test_walls_bottom_top(case_dir):
    """Periodic in x and y, no-slip walls at z = 0 and z = 40: spheres
    inside the lubrication range of the bottom and the top wall (cages cut
    by a wall, quadrature nodes beyond it) and one between."""
    kw, parts = bbconfig.sediment(n=(32, 32, 40), nparts=3, a=4.0, seed=7,
                                  walls="z")
    parts[0].update(x=9.3, y=10.1, z=4.4, w=-0.02, u=0.01)
    parts[1].update(x=22.4, y=20.7, z=35.5, v=0.01, w=0.01)
    parts[2].update(x=16.2, y=15.8, z=20.3)
    out = run(case_dir, kw, parts, 2)
    check(out)
    assert abs(out["orc_parts"][0]["iFz"]) > 0.
    assert abs(out["orc_parts"][1]["iFz"]) > 0.


def test_walls_all_faces(case_dir):
    """A closed box (no-slip walls on all six faces): spheres inside the
    lubrication range of the W, E, S and N faces, moving along and toward
    them."""
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=4, a=4.0, seed=7,
                                  walls="xyz")
    parts[0].update(x=4.5, y=8.2, z=10.1, u=-0.01, v=0.01, w=-0.01)
    parts[1].update(x=27.3, y=24.3, z=20.2, u=0.02, w=0.01)
    parts[2].update(x=20.1, y=4.4, z=8.3, v=-0.02, u=0.01)
    parts[3].update(x=12.2, y=27.6, z=24.1, u=-0.01, v=0.01)
    out = run(case_dir, kw, parts, 2)
    check(out)
    for n, k in ((0, "iFx"), (1, "iFx"), (2, "iFy"), (3, "iFy")):
        assert abs(out["orc_parts"][n][k]) > 0., (n, k)


def test_wall_contacts_one_substep(case_dir):
    """Wall contact (overlaps of 0.02 to 0.05) at the W, N and T faces and
    lubrication at the E, S and B faces, with slip and rotation: one
    particle sub-step (collisions, velocity update, collisions) from the
    same state on the device and in the oracle; forces, torques,
    velocities and the contact lists (codes, Stokes numbers, counts)."""
    import ctypes
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=6, a=4.0, seed=7,
                                  walls="xyz")
    state = [  # position, velocity, rotation
        ((3.98, 8.0, 9.0), (-0.01, 0.004, 0.002), (0.001, -0.002, 0.003)),
        ((27.2, 24.0, 20.0), (0.02, 0.003, -0.001), (0.002, 0.001, -0.001)),
        ((20.0, 4.6, 8.0), (0.001, -0.02, 0.002), (-0.003, 0.002, 0.001)),
        ((12.0, 28.05, 24.0), (0.003, 0.01, -0.002), (0.001, 0.003, -0.002)),
        ((8.0, 20.0, 4.7), (0.002, -0.001, -0.02), (0.002, -0.001, 0.003)),
        ((24.0, 12.0, 28.04), (-0.002, 0.003, 0.01), (-0.001, 0.002, 0.002)),
    ]
    for p, (x, _, _) in zip(parts, state):
        p.update(x=x[0], y=x[1], z=x[2])
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    with orc.Oracle(case_dir) as o:
        for n, (x, v, w) in enumerate(state):
            o.set_motion(n, np.array(x), np.array(v), np.array(w))
        o.move_sub()
        want = [o.part_state(n) for n in range(len(state))]
        lists = [o.contacts(n) for n in range(len(state))]
    with bbhip.Simulation(case_dir, kw["n"]) as s:
        n = ctypes.c_int.in_dll(s.lib, "nparts").value
        assert n == len(state)
        s.lib.cuda_part_pull()
        P = ctypes.POINTER(bbhip.PartStruct).in_dll(s.lib, "parts")
        for k, (x, v, w) in enumerate(state):
            for name, val in zip(("x", "y", "z", "u", "v", "w", "u0", "v0",
                                  "w0", "ox", "oy", "oz", "ox0", "oy0",
                                  "oz0"), x + v + v + w + w):
                setattr(P[k], name, val)
        s.lib.cuda_part_push()
        s.lib.cuda_move_parts_sub()
        got = s.parts()
    worst = 0.
    for k in range(len(state)):
        for q in ("iFx", "iFy", "iFz", "iLx", "iLy", "iLz", "u", "v", "w",
                  "ox", "oy", "oz"):
            a, b = got[k][q], want[k][q]
            e = abs(a - b) / max(abs(b), 1e-300)
            worst = max(worst, e if b != 0 else abs(a))
            assert (a == b) or e <= 1e-10, (k, q, a, b)
        ist, st, _, nwall = lists[k]
        assert list(got[k]["iSt"]) == ist, (k, got[k]["iSt"], ist)
        assert np.allclose(got[k]["St"], st, rtol=1e-12, atol=0), k
        assert got[k]["ncoll_wall"] == nwall, k
    print("worst relative difference", f"{worst:.2e}")
    codes = [lists[k][0] for k in range(len(state))]
    assert -10 in codes[0] and -13 in codes[3] and -15 in codes[5]
    for k in (1, 2, 4):
        assert abs(want[k]["iF" + "xyz"[(k == 2) + 2 * (k == 4)]]) > 0.


def test_mixed_precision_with_particles(case_dir):
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=3, a=4.0, seed=7)
    out = run(case_dir, kw, parts, 2, precision="MIXED")
    errs = {f: rel_l2(interior(out["gpu"][f]), interior(out["orc"][f]))
            for f in FIELDS}
    print("mixed vs FP64 oracle, fields:",
          {f: f"{e:.2e}" for f, e in errs.items()})
    for f in FIELDS:
        assert errs[f] <= 1e-6, f
    for k in ("w", "Fz"):
        a = np.array([p[k] for p in out["gpu_parts"]])
        b = np.array([p[k] for p in out["orc_parts"]])
        print(k, f"{rel_l2(a, b):.2e}")
        assert rel_l2(a, b) <= 1e-6
