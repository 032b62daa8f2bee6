"""GPU parity: libbluebottle.so (HIP, gfx950) against the FP64 host oracle.

Bar: relative L2 <= 1e-10 on u, v, w, p (the flow path is FP64 with the same
operation order and the same canonical reduction order on both sides, so the
fields are expected to be bit-identical; that stronger property is asserted
separately)."""
import ctypes

import numpy as np
import pytest

import bbconfig
import bbhip
import orc
from helpers import bitwise_equal, describe, interior, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "phi", "u_star", "v_star", "w_star",
          "conv_u", "conv_v", "conv_w", "diff_u", "diff_v", "diff_w")


def set_knobs(variant=3, tzb=0, batch=16, defer=0):
    """variant 3: search kernel with one wavefront per x row where the grid
    allows it, else as 2; 2: over z-columns of tzb tiles (0: automatic);
    0: direct neighbour loads (the cross-check implementation); defer = 1:
    phi updated every second iteration with both terms (odd iterations
    leave it)."""
    lib = bbhip.load()
    ctypes.c_int.in_dll(lib, "bb_defer_phi").value = defer
    ctypes.c_int.in_dll(lib, "bb_spmv_variant").value = variant
    ctypes.c_int.in_dll(lib, "bb_spmv_tzb").value = tzb
    ctypes.c_int.in_dll(lib, "bb_pcg_batch").value = batch


@pytest.fixture(autouse=True)
def default_knobs():
    """Each test sets its knobs; the library keeps them, so put the
    defaults back for the tests that follow."""
    yield
    set_knobs(defer=1)


def run_pair(case_dir, kw, nsteps, variant=3, tzb=0, batch=16, defer=0):
    n = kw["n"]
    set_knobs(variant, tzb, batch, defer)
    bbconfig.write_case(case_dir, **kw)
    out = {"gpu": {}, "orc": {}, "iters_gpu": [], "iters_orc": []}
    with orc.Oracle(case_dir) as o:
        for _ in range(nsteps):
            out["iters_orc"].append(o.step())
        for f in FIELDS:
            out["orc"][f] = np.array(o.field(f))
        out["orc_dt"] = o.scalar("dt")
    with bbhip.Simulation(case_dir, n) as s:
        for _ in range(nsteps):
            s.step()
            out["iters_gpu"].append(s.pp_stats()["niter"])
        s.pull()
        for f in FIELDS:
            out["gpu"][f] = np.array(s.field(f))
        out["gpu_dt"] = s.scalar("dt")
    return out


def check(out, fields=("u", "v", "w", "p"), tol=1e-10):
    msgs = [describe(f, interior(out["gpu"][f]), interior(out["orc"][f]))
            for f in fields]
    print("\n".join(msgs))
    print("iters gpu", out["iters_gpu"], "oracle", out["iters_orc"])
    for f in fields:
        assert rel_l2(interior(out["gpu"][f]), interior(out["orc"][f])) <= tol, \
            describe(f, interior(out["gpu"][f]), interior(out["orc"][f]))


def test_cavity_16_five_steps(case_dir):
    out = run_pair(case_dir, bbconfig.cavity(n=16, steps=5), 5)
    check(out)
    assert out["iters_gpu"] == out["iters_orc"]


def test_cavity_32_full_100_steps(case_dir):
    """Lid-driven cavity, 32^3, Re 100, dt 1e-3, 100 steps, PCG to 1e-8."""
    out = run_pair(case_dir, bbconfig.cavity(n=32, steps=100), 100)
    check(out)
    assert out["iters_gpu"] == out["iters_orc"]


@pytest.mark.parametrize("variant,tzb", [(0, 0), (2, 1), (2, 2), (3, 0)])
def test_cavity_bitwise(case_dir, variant, tzb):
    out = run_pair(case_dir, bbconfig.cavity(n=16, steps=5), 5, variant, tzb)
    for f in FIELDS:
        assert bitwise_equal(out["gpu"][f], out["orc"][f]), \
            describe(f, out["gpu"][f], out["orc"][f])


def test_channel_small(case_dir):
    """Periodic channel (x, y periodic, z walls) at 64x32x32, 5 steps."""
    out = run_pair(case_dir, bbconfig.channel(n=(64, 32, 32)), 5)
    check(out)
    assert out["iters_gpu"] == out["iters_orc"]
    assert out["gpu_dt"] == out["orc_dt"]


@pytest.mark.parametrize("variant,tzb,defer",
                         [(0, 0, 0), (2, 1, 0), (2, 4, 0), (2, 4, 1),
                          (2, 1, 1), (2, 8, 1), (3, 0, 0), (3, 0, 1)])
def test_channel_bitwise(case_dir, variant, tzb, defer):
    out = run_pair(case_dir, bbconfig.channel(n=(64, 32, 32)), 3, variant,
                   tzb, defer=defer)
    for f in FIELDS:
        assert bitwise_equal(out["gpu"][f], out["orc"][f]), \
            describe(f, out["gpu"][f], out["orc"][f])


@pytest.mark.parametrize("variant,tzb,batch,defer",
                         [(0, 0, 16, 0), (0, 0, 5, 0), (2, 2, 1, 0),
                          (2, 4, 5, 0), (2, 3, 16, 0), (2, 2, 1, 1),
                          (2, 4, 5, 1), (2, 3, 16, 1), (2, 8, 16, 1),
                          (3, 0, 16, 0), (3, 0, 5, 1)])
def test_ragged_sizes(case_dir, variant, tzb, batch, defer):
    """x not a multiple of the 64-lane wavefront, y not a multiple of the
    row group, z not a multiple of the tile depth (and the tile count in z,
    3, not a multiple of the z-column length), shear walls."""
    kw = bbconfig.cavity(n=16, steps=3)
    kw["n"] = (37, 11, 23)
    kw["hi"] = (1.3, 0.7, 1.1)
    kw["vel_bc"] = bbconfig.walls(lid=("v", "E", 0.5))
    out = run_pair(case_dir, kw, 3, variant, tzb, batch, defer)
    check(out)
    assert out["iters_gpu"] == out["iters_orc"]
    for f in FIELDS:
        assert bitwise_equal(out["gpu"][f], out["orc"][f]), \
            describe(f, out["gpu"][f], out["orc"][f])


def test_periodic_all_and_wide_x(case_dir):
    """Triply periodic box (wrapped neighbours in every direction) with an x
    extent of several 64-cell tiles and a ragged last one."""
    per = ("PERIODIC",)
    vel = {c: {f: per for f in "WESNBT"} for c in "uvw"}
    kw = dict(n=(450, 6, 5), lo=(0.0, 0.0, 0.0), hi=(4.5, 0.6, 0.5),
              rho_f=1.0, nu=0.01, duration=1.0, pp_max_iter=100000,
              pp_residual=1e-8, p_bc={f: per for f in "WESNBT"}, vel_bc=vel,
              init_cond="QUIESCENT",
              extensions={"fixed_dt": 1e-3, "perturb": (5, 0.2)})
    out = run_pair(case_dir, kw, 3, variant=2, tzb=1)
    check(out)
    for f in FIELDS:
        assert bitwise_equal(out["gpu"][f], out["orc"][f]), \
            describe(f, out["gpu"][f], out["orc"][f])
    # rows of 450 cells are wider than one wavefront's 256: the row variant
    # falls back to the z-column kernel
    out = run_pair(case_dir, kw, 3, variant=3)
    assert ctypes.c_int.in_dll(bbhip.load(), "bb_spmv_last").value == 2
    check(out)
    for f in FIELDS:
        assert bitwise_equal(out["gpu"][f], out["orc"][f]), \
            describe(f, out["gpu"][f], out["orc"][f])


@pytest.mark.parametrize("variant,defer", [(2, 0), (0, 0), (2, 1), (3, 0),
                                           (3, 1)])
def test_periodic_all_25d(case_dir, variant, defer):
    per = ("PERIODIC",)
    vel = {c: {f: per for f in "WESNBT"} for c in "uvw"}
    kw = dict(n=(40, 12, 10), lo=(0.0, 0.0, 0.0), hi=(4.0, 1.2, 1.0),
              rho_f=1.0, nu=0.01, duration=1.0, pp_max_iter=100000,
              pp_residual=1e-8, p_bc={f: per for f in "WESNBT"}, vel_bc=vel,
              init_cond="QUIESCENT",
              extensions={"fixed_dt": 1e-3, "perturb": (6, 0.2)})
    out = run_pair(case_dir, kw, 3, variant=variant, tzb=2, defer=defer)
    for f in FIELDS:
        assert bitwise_equal(out["gpu"][f], out["orc"][f]), \
            describe(f, out["gpu"][f], out["orc"][f])
