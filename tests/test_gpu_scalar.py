"""GPU parity of scalar transport with Boussinesq coupling against the FP64
host oracle (src/scalar_kernel.cu:4-301, src/cuda_scalar.cu:149-257,
949-957, src/bluebottle.c:189-192, 293-341).

Cases: Rayleigh-Benard (Dirichlet walls, seeded initial noise), the same
box with a Neumann flux at the bottom wall, a periodic channel carrying a
scalar between Dirichlet walls with the Boussinesq force across the mean
flow, and a run whose time step follows s_D > nu.  The scalar path is FP64
arithmetic in the reference's order on both sides: u, v, w, p and s agree
bit for bit, the PCG iteration counts and dt are equal.
"""
import numpy as np
import pytest

import bbconfig
import bbhip
import orc
from helpers import bitwise_equal, describe, interior

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "s")


def run_pair(case_dir, kw, nsteps):
    bbconfig.write_case(case_dir, **kw)
    out = {"gpu": {}, "orc": {}, "iters_gpu": [], "iters_orc": []}
    with orc.Oracle(case_dir) as o:
        for _ in range(nsteps):
            out["iters_orc"].append(o.step())
        for f in FIELDS:
            out["orc"][f] = np.array(o.field(f))
        out["orc_dt"] = o.scalar("dt")
    with bbhip.Simulation(case_dir, kw["n"]) as s:
        for _ in range(nsteps):
            s.step()
            out["iters_gpu"].append(s.pp_stats()["niter"])
        s.pull()
        for f in FIELDS:
            out["gpu"][f] = np.array(s.field(f))
        out["gpu_dt"] = s.scalar("dt")
    return out


def check_bitwise(out):
    print("iters gpu", out["iters_gpu"], "oracle", out["iters_orc"])
    for f in FIELDS:
        a, b = interior(out["gpu"][f]), interior(out["orc"][f])
        print(describe(f, a, b))
        assert bitwise_equal(a, b), describe(f, a, b)
    assert out["iters_gpu"] == out["iters_orc"]
    assert out["gpu_dt"] == out["orc_dt"]


def test_rayleigh_benard(case_dir):
    kw = bbconfig.rayleigh_benard(n=(32, 32, 16), Ra=1e5)
    out = run_pair(case_dir, kw, 5)
    check_bitwise(out)
    s = interior(out["gpu"]["s"])
    assert np.ptp(s) > 0.05                  # the walls have diffused in


def test_neumann_bottom_flux(case_dir):
    kw = bbconfig.rayleigh_benard(n=(24, 20, 12), Ra=1e5)
    kw["scalar"]["bc_s"]["B"] = ("NEUMANN", -0.75)
    out = run_pair(case_dir, kw, 4)
    check_bitwise(out)


def test_channel_with_scalar(case_dir):
    """Periodic channel 64x32x32: scalar 1 / 0 on the walls, gravity along
    y (across the mean flow) so the Boussinesq force acts on v."""
    kw = bbconfig.channel(n=(64, 32, 32))
    sbc = {f: ("PERIODIC",) for f in "WESN"}
    sbc.update({"B": ("DIRICHLET", 1.0), "T": ("DIRICHLET", 0.0)})
    kw["scalar"] = dict(s_D=0.004, s_k=0.0, s_beta=0.5, s_ref=0.5,
                        init_s=(0.5, 0.1), bc_s=sbc)
    kw["g"] = ((0.0, 0.0), (-1.0, 0.0), (0.0, 0.0))
    out = run_pair(case_dir, kw, 3)
    check_bitwise(out)


def test_time_step_from_diffusivity(case_dir):
    """Pr = 0.5: s_D > nu, so dt follows s_D."""
    kw = bbconfig.rayleigh_benard(n=(16, 16, 8), Ra=1e4, Pr=0.5)
    out = run_pair(case_dir, kw, 3)
    check_bitwise(out)


# ------------------------------ particle heat -------------------------------
PART_SCALARS = ("s", "q", "x", "y", "z", "u", "v", "w", "Fz")


def heat_case(nparts, walls="", n=(32, 32, 32), seed=7):
    kw, parts = bbconfig.sediment(n=n, nparts=nparts, a=4.0, seed=seed,
                                  walls=walls)
    per = ("PERIODIC",)
    sbc = {f: per for f in "WESNBT"}
    if "z" in walls:
        sbc["B"] = ("DIRICHLET", 1.0)
        sbc["T"] = ("NEUMANN", 0.05)
    kw["scalar"] = dict(s_D=0.05, s_k=1.0, s_beta=0.2, s_ref=0.5,
                        init_s=(0.5, 0.05), bc_s=sbc)
    for k, p in enumerate(parts):
        p.update(s=0.2 + 0.3 * k, update=1, cp=1.5, srs_r=1.2,
                 s_order=2 if k % 2 == 0 else 1)
    return kw, parts


def run_heat(case_dir, kw, parts, nsteps):
    import ctypes
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    out = {"gpu": {}, "orc": {}, "lamb_gpu": [], "lamb_orc": [],
           "siters_gpu": [], "siters_orc": []}
    with orc.Oracle(case_dir) as o:
        for _ in range(nsteps):
            o.step()
            out["lamb_orc"].append(int(o.scalar("lamb_iters")))
            out["siters_orc"].append(int(o.scalar("scalar_iters")))
        for f in FIELDS:
            out["orc"][f] = np.array(o.field(f))
        out["orc_parts"] = []
        for k in range(o.nparts()):
            st = o.part_state(k)
            s, q, are, aim = o.part_scalar(k)
            st.update(s=s, q=q, anm_re=are, anm_im=aim)
            out["orc_parts"].append(st)
    with bbhip.Simulation(case_dir, kw["n"]) as sim:
        for _ in range(nsteps):
            sim.step()
            out["lamb_gpu"].append(sim.lamb_stats()[0])
            out["siters_gpu"].append(
                ctypes.c_int.in_dll(sim.lib, "bb_scalar_iters").value)
        sim.pull()
        for f in FIELDS:
            out["gpu"][f] = np.array(sim.field(f))
        out["gpu_parts"] = sim.parts()
    return out


def check_heat(out, tol=1e-10):
    from helpers import rel_l2
    print("lamb iterations", out["lamb_gpu"], out["lamb_orc"],
          "scalar iterations", out["siters_gpu"], out["siters_orc"])
    for f in FIELDS:
        a, b = interior(out["gpu"][f]), interior(out["orc"][f])
        e = rel_l2(a, b)
        print(f, f"rel L2 {e:.2e}")
        assert e <= tol, describe(f, a, b)
    assert out["lamb_gpu"] == out["lamb_orc"]
    assert out["siters_gpu"] == out["siters_orc"]
    worst = 0.
    for k in PART_SCALARS + ("anm_re", "anm_im"):
        a = np.array([np.asarray(p[k]) for p in out["gpu_parts"]], float)
        b = np.array([np.asarray(p[k]) for p in out["orc_parts"]], float)
        e = rel_l2(a, b)
        worst = max(worst, e)
        assert e <= tol, (k, a, b)
    print("worst particle rel L2", f"{worst:.2e}")


def test_particle_heat_periodic(case_dir):
    """Three spheres at different temperatures in a periodic box with
    Boussinesq coupling: scalar series in particle cells, node quadrature,
    heat flux and particle temperature update."""
    kw, parts = heat_case(3)
    out = run_heat(case_dir, kw, parts, 2)
    check_heat(out)
    assert max(out["siters_orc"]) > 1
    assert all(abs(p["q"]) > 0 for p in out["orc_parts"])


def test_particle_heat_near_walls(case_dir):
    """z walls with a Dirichlet bottom and a Neumann top scalar condition:
    spheres close enough that scalar nodes fall beyond the walls (wall
    codes from the scalar BCs)."""
    kw, parts = heat_case(2, walls="z", n=(32, 32, 40))
    parts[0].update(x=9.3, y=10.1, z=4.4)
    parts[1].update(x=22.4, y=20.7, z=35.5)
    out = run_heat(case_dir, kw, parts, 2)
    check_heat(out)
