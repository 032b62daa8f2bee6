"""GPU parity of the particle / Physalis path against the FP64 host oracle.

Cases (triply periodic boxes, dx = 1, spheres of radius a at seeded
positions, Lamb order 4, gravity from Ga = 100):
  - a few spheres, two timesteps with the full Lamb iteration;
  - a sphere cut by the periodic faces (cage, nodes and BCs through the
    images);
  - two spheres closing in (lubrication inside the interaction length);
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
