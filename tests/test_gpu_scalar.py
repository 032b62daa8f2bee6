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
