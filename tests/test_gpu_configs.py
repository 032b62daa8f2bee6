"""The workload shapes the benchmark names, at test sizes, against the FP64
host oracle:
  - Rayleigh-Benard convection with particles: hot bottom and cold top
    walls, periodic sides, Boussinesq coupling of the fluid and of each
    sphere, five conducting spheres of radius 4 dx with particle heat
    transfer (src/bluebottle.c:168-407, src/scalar_kernel.cu:4-1360);
    plain FP64, and the mixed-precision solve with the learned initial
    guess;
  - sedimentation with the learned initial guess (network trained online
    on the first solves, then the guess of every solve), with the FP64 and
    with the mixed-precision PCG.
The plain form is held to 1e-10 relative L2, Lamb and scalar iteration
counts equal.  A run that starts its solves elsewhere (learned guess) or
solves in FP32 (mixed) stops at another point of the same FP64 tolerance
ball (r, M^-1 r) <= 1e-16 (b, b): every one of its solves is checked to
meet that test, and its fields, the pressure correction phi included,
are held to the difference two runs of that many solves can show
(solve_difference_bounds, with the extreme eigenvalues of the case's
masked operator from tests/golden/mixed_bounds.json).  (A guess that kept
values in cells particles have entered since would act as boundary data
of the fluid rows and move the solution off the reference's by far more:
1e-2 in u in the sedimentation case.)
"""
import json
import os
import ctypes

import numpy as np
import pytest

import bbconfig
import bbhip
import orc
from helpers import (describe, field_bounds, interior, rel_l2,
                     solve_difference_bounds)

with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "mixed_bounds.json")) as _f:
    EIGEN = json.load(_f)

pytestmark = pytest.mark.gpu

FLUID = ("u", "v", "w", "p")
PKEYS = ("x", "y", "z", "u", "v", "w", "ox", "oy", "oz", "Fx", "Fy", "Fz",
         "Lx", "Ly", "Lz")

WARM_KNOBS = {"bb_warm_train_solves": 8, "bb_warm_steps": 20,
              "bb_warm_crop": 16, "bb_warm_net_all_passes": 1}


def rb_particles():
    """Rayleigh-Benard 64 x 64 x 32 on [0, 2) x [0, 2) x [0, 1], Ra 1e4,
    Pr 1, with five spheres of radius 4 dx = 0.125 (density ratio 2, Lamb
    order 4), two of them inside each other's lubrication range, each with
    its own temperature and heat transfer (scalar orders 1 and 2)."""
    kw = bbconfig.rayleigh_benard(n=(64, 64, 32), Ra=1e4)
    kw.update(lamb_max_iter=50, lamb_residual=1e-2, lamb_relax=1.0,
              lamb_cut=1e-4)
    centres = [(0.4, 0.5, 0.4), (1.2, 0.4, 0.6), (0.6, 1.35, 0.5),
               (1.6, 1.5, 0.45), (1.2, 0.66, 0.6)]
    parts = []
    for k, (x, y, z) in enumerate(centres):
        parts.append(dict(r=0.125, x=x, y=y, z=z, rho=2.0, order=4,
                          s=0.3 - 0.15 * k, update=1, cp=1.5, srs_r=1.2,
                          s_order=2 if k % 2 == 0 else 1))
    return kw, parts


def run_case(case_dir, kw, parts, nsteps, fields, precision=None,
             warm=None, scalar=False):
    kw = dict(kw)
    ext = dict(kw.get("extensions") or {})
    if precision:
        ext["pp_precision"] = precision
    if warm:
        ext["pp_warmstart"] = warm
    kw["extensions"] = ext
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    out = {"lamb_orc": [], "lamb_gpu": [], "stats": []}
    with orc.Oracle(case_dir) as o:
        for _ in range(nsteps):
            o.step()
            out["lamb_orc"].append(int(o.scalar("lamb_iters")))
        out["orc"] = {f: np.array(o.field(f)) for f in fields}
        out["orc_parts"] = []
        for k in range(o.nparts()):
            st = o.part_state(k)
            if scalar:
                s, q, _, _ = o.part_scalar(k)
                st.update(s=s, q=q)
            out["orc_parts"].append(st)
    L = bbhip.load()
    saved = {k: ctypes.c_int.in_dll(L, k).value for k in WARM_KNOBS}
    if warm:
        for k, v in WARM_KNOBS.items():
            ctypes.c_int.in_dll(L, k).value = v
    try:
        with bbhip.Simulation(case_dir, kw["n"]) as s:
            net0 = ctypes.c_int.in_dll(L, "bb_warm_net_solves").value
            for _ in range(nsteps):
                s.step()
                out["lamb_gpu"].append(s.lamb_stats()[0])
                out["stats"].append(s.pp_stats())
            s.pull()
            out["gpu"] = {f: np.array(s.field(f)) for f in fields}
            out["gpu_parts"] = s.parts()
            out["net_solves"] = \
                ctypes.c_int.in_dll(L, "bb_warm_net_solves").value - net0
    finally:
        for k, v in saved.items():
            ctypes.c_int.in_dll(L, k).value = v
    return out


def compare(out, fields, tol, ptol, pkeys=PKEYS, counts=True):
    """tol, ptol: a number, or a dict of per-field bounds (fields) and one
    bound for the particle quantities."""
    errs = {f: rel_l2(interior(out["gpu"][f]), interior(out["orc"][f]))
            for f in fields}
    print("fields rel L2:", {f: f"{e:.2e}" for f, e in errs.items()})
    print("lamb passes gpu", out["lamb_gpu"], "oracle", out["lamb_orc"])
    for f in fields:
        bound = tol[f] if isinstance(tol, dict) else tol
        assert errs[f] <= bound, describe(f, interior(out["gpu"][f]),
                                          interior(out["orc"][f]))
    if counts:
        assert out["lamb_gpu"] == out["lamb_orc"]
    worst = 0.
    for k in pkeys:
        a = np.array([p[k] for p in out["gpu_parts"]])
        b = np.array([p[k] for p in out["orc_parts"]])
        e = rel_l2(a, b)
        worst = max(worst, e)
        assert e <= ptol, (k, a, b)
    print("worst particle rel L2", f"{worst:.2e}")
    for st in out["stats"]:
        assert st["resid"] <= 1e-8, st


def test_rayleigh_benard_particles_plain(case_dir):
    kw, parts = rb_particles()
    fields = FLUID + ("s",)
    out = run_case(case_dir, kw, parts, 2, fields, scalar=True)
    compare(out, fields, 1e-10, 1e-10, PKEYS + ("s", "q"))
    assert max(abs(p["q"]) for p in out["orc_parts"]) > 0
    assert max(abs(p["iFx"]) for p in out["orc_parts"]) > 0


def derived(out, case, fields):
    """Bounds of the fields and particle quantities of a run whose solves
    started elsewhere, from the stopping test and the masked operator's
    eigenvalues; the particle quantities take the larger field bound."""
    e = EIGEN[case]
    nsolves = sum(out["lamb_gpu"])
    b = solve_difference_bounds(e["lam_min"], e["lam_max"], e["diag_max"],
                                nsolves)
    fb = field_bounds(b, {f: out["orc"][f] for f in FLUID})
    fb.update({f: b["phi"] for f in fields if f not in fb})  # phi, s
    print("bounds derived from the stopping test:",
          {f: f"{v:.2e}" for f, v in fb.items()})
    return fb, max(b["phi"], max(fb[c] for c in ("u", "v", "w")))


def test_rayleigh_benard_particles_mixed_learned(case_dir):
    """Mixed-precision solve with the learned guess: 8 training solves, then
    the network's guess on every solve; every solve at the FP64 test."""
    kw, parts = rb_particles()
    fields = FLUID + ("s", "phi")
    out = run_case(case_dir, kw, parts, 2, fields, precision="MIXED",
                   warm="LEARNED", scalar=True)
    fb, pb = derived(out, "rb_particles", fields)
    compare(out, fields, fb, pb, PKEYS + ("s", "q"), counts=False)
    assert out["net_solves"] > 0


@pytest.mark.parametrize("precision", ["PLAIN", "MIXED"])
def test_sediment_learned_guess(case_dir, precision):
    """Sedimentation box 32^3 with four spheres of radius 4, three steps,
    with the learned initial guess (8 training solves, then every solve)."""
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=4, a=4.0, seed=11)
    fields = FLUID + ("phi",)
    out = run_case(case_dir, kw, parts, 3, fields, precision=precision,
                   warm="LEARNED")
    fb, pb = derived(out, "sediment32_4", fields)
    compare(out, fields, fb, pb, counts=False)
    assert out["net_solves"] > 0
