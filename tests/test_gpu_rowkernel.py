"""The search and update kernels with one wavefront per x row against the
FP64 oracle and against the kernels that take one z-column or one tile per
workgroup.

The row kernel holds 1, 2 or 4 cells per lane (rows of up to 64, 128 or 256
cells), loads the y neighbours from an LDS plane of its rows plus a halo row
on either side and keeps the z neighbours in registers.  Its arithmetic per
cell and the order of its tile partials are those of the z-column kernel, so
the plain form must match the oracle bit for bit and the mixed form must
give the same bits as the z-column kernel, whatever the rows and planes per
workgroup.  The row update streams whole rows of r, A p, p, phi (and the
previous direction, the fluid flags) through a ring of planes in flight,
with the arithmetic and partials of the tile update kernel: the same
bits again, with and without particles."""
import ctypes

import numpy as np
import pytest

import bbconfig
import bbhip
import orc
from helpers import bitwise_equal, describe

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "phi")


def knobs(variant=3, ry=0, zc=0, fp64=1, depth=0, uvar=1, ufp64=0, uzc=0,
          udepth=0, pairs=None):
    """fp64 = 0: the plain form's FP64 search takes the z-column kernel
    (the row kernel by default); ufp64 = 1: its FP64 update takes the row
    kernel too (by default only the mixed form's FP32 one does); uvar 0:
    the tile update kernel."""
    lib = bbhip.load()
    ctypes.c_int.in_dll(lib, "bb_spmv_variant").value = variant
    ctypes.c_int.in_dll(lib, "bb_row_fp64").value = fp64
    ctypes.c_int.in_dll(lib, "bb_row_ry").value = ry
    ctypes.c_int.in_dll(lib, "bb_row_zc").value = zc
    ctypes.c_int.in_dll(lib, "bb_row_depth").value = depth
    ctypes.c_int.in_dll(lib, "bb_update_variant").value = uvar
    ctypes.c_int.in_dll(lib, "bb_urow_fp64").value = ufp64
    ctypes.c_int.in_dll(lib, "bb_urow_zc").value = uzc
    ctypes.c_int.in_dll(lib, "bb_urow_depth").value = udepth
    ctypes.c_int.in_dll(lib, "bb_defer_phi").value = 1
    ctypes.c_int.in_dll(lib, "bb_pcg_batch").value = 16
    if pairs is not None:
        ctypes.c_int.in_dll(lib, "bb_row_pairs").value = pairs


@pytest.fixture(autouse=True)
def default_knobs():
    lib = bbhip.load()
    pairs0 = ctypes.c_int.in_dll(lib, "bb_row_pairs").value
    yield
    knobs(pairs=pairs0)


def last_kernel():
    lib = bbhip.load()
    row = (ctypes.c_int * 4).in_dll(lib, "bb_row_last")
    return ctypes.c_int.in_dll(lib, "bb_spmv_last").value, tuple(row)


def last_update():
    return ctypes.c_int.in_dll(bbhip.load(), "bb_update_last").value


def box(n, per_x, per_y):
    """A box with shear walls where a direction is not periodic, so every
    face kind (wall ghost, wrapped neighbour) meets the stencil."""
    per = ("PERIODIC",)
    wall = ("DIRICHLET", 0.0, 0.0)
    vel = {}
    for c in "uvw":
        vel[c] = {}
        for f in "WESNBT":
            p = (per_x and f in "WE") or (per_y and f in "SN")
            vel[c][f] = per if p else wall
    vel["v"]["T"] = ("DIRICHLET", 0.4, 0.0)      # a moving lid drives it
    pb = {f: (per if (per_x and f in "WE") or (per_y and f in "SN")
              else ("NEUMANN", 0.0)) for f in "WESNBT"}
    return dict(n=n, lo=(0.0, 0.0, 0.0),
                hi=(0.05 * n[0], 0.05 * n[1], 0.05 * n[2]),
                rho_f=1.0, nu=0.01, duration=1.0, pp_max_iter=100000,
                pp_residual=1e-8, p_bc=pb, vel_bc=vel, init_cond="QUIESCENT",
                extensions={"fixed_dt": 1e-3, "perturb": (7, 0.2)})


def run_gpu(case_dir, kw, nsteps, parts=None):
    n = kw["n"]
    bbconfig.write_case(case_dir, **kw)
    if parts is not None:
        bbconfig.write_parts(case_dir, parts)
    out = {"f": {}, "iters": []}
    with bbhip.Simulation(case_dir, n) as s:
        for _ in range(nsteps):
            s.step()
            out["iters"].append(s.pp_stats()["niter"])
        s.pull()
        for f in FIELDS:
            out["f"][f] = np.array(s.field(f))
    out["kernel"] = last_kernel()
    out["update"] = last_update()
    return out


def run_oracle(case_dir, kw, nsteps):
    bbconfig.write_case(case_dir, **kw)
    out = {"f": {}, "iters": []}
    with orc.Oracle(case_dir) as o:
        for _ in range(nsteps):
            out["iters"].append(o.step())
        for f in FIELDS:
            out["f"][f] = np.array(o.field(f))
    return out


# nx: 100 and 128 take two cells per lane (100: the last 64-cell tile and
# the last lanes partly outside the row), 132 and 256 four (132: 31 lanes
# outside the row); ny = 11 leaves the second group of 8 rows ragged (rows
# past the halo hold zeros), nz = 13 a ragged last group of planes.
@pytest.mark.parametrize("nx,per_x,per_y,zc,cpr", [
    (100, 0, 1, 0, 2), (128, 1, 0, 8, 2), (132, 1, 1, 0, 4),
    (256, 0, 0, 16, 4), (256, 1, 1, 8, 4)])
def test_row_kernel_plain_bitwise(case_dir, nx, per_x, per_y, zc, cpr):
    kw = box((nx, 11, 13), per_x, per_y)
    ref = run_oracle(case_dir, kw, 2)
    knobs(variant=3, zc=zc, fp64=1, ufp64=1, uzc=zc)
    out = run_gpu(case_dir, kw, 2)
    kern, (c, ry, z, _) = out["kernel"]
    assert kern == 3 and c == cpr and ry == 8, out["kernel"]
    assert out["update"] == 1
    assert z == (zc if zc else 8)
    print("iters gpu", out["iters"], "oracle", ref["iters"])
    assert out["iters"] == ref["iters"]
    for f in FIELDS:
        assert bitwise_equal(out["f"][f], ref["f"][f]), \
            describe(f, out["f"][f], ref["f"][f])


@pytest.mark.parametrize("pairs", [0, 1])
@pytest.mark.parametrize("nx,per_x,per_y,zc", [
    (132, 1, 1, 0), (256, 0, 0, 16), (256, 1, 1, 8)])
def test_row_kernel_fp64_layouts_bitwise(case_dir, nx, per_x, per_y, zc,
                                         pairs):
    """The FP64 row search with four cells per lane, in its two layouts:
    each lane loading and storing its own run of four (pairs 0), or lane m
    loading r and p_old of cells 2m, 2m + 1 and 128 + 2m, 128 + 2m + 1,
    passing p_new through LDS and storing p_new and Ap in the same pairs
    (pairs 1): the oracle's bits either way.  nx = 132 leaves lanes and
    pairs past the row, ny = 11 a second group of rows past the domain."""
    kw = box((nx, 11, 13), per_x, per_y)
    ref = run_oracle(case_dir, kw, 2)
    knobs(variant=3, zc=zc, fp64=1, pairs=pairs)
    out = run_gpu(case_dir, kw, 2)
    kern, (c, _, _, _) = out["kernel"]
    assert kern == 3 and c == 4, out["kernel"]
    assert out["iters"] == ref["iters"]
    for f in FIELDS:
        assert bitwise_equal(out["f"][f], ref["f"][f]), \
            describe(f, out["f"][f], ref["f"][f])


def test_row_kernel_plane_groups_same_bits(case_dir):
    """Planes per workgroup 8, 24 and 64 (one workgroup spans the whole
    z extent of 40) and the z-column kernel: one result."""
    kw = box((64, 20, 40), 1, 0)
    res = []
    for variant, zc in ((2, 0), (3, 8), (3, 24), (3, 64)):
        knobs(variant=variant, zc=zc, fp64=1)
        res.append(run_gpu(case_dir, kw, 2))
        assert res[-1]["kernel"][0] == variant
    for r in res[1:]:
        assert r["iters"] == res[0]["iters"]
        for f in FIELDS:
            assert bitwise_equal(r["f"][f], res[0]["f"][f]), f


@pytest.mark.parametrize("n", [(128, 20, 40), (256, 12, 24)])
def test_row_kernel_mixed_same_bits(case_dir, n):
    """The mixed form's FP32 search: row kernel with 8 and 16 rows per
    workgroup, several plane groups and 2 to 4 planes of loads in flight
    against the z-column kernel, same iterations and the same bits in every
    field."""
    kw = bbconfig.channel(n=n, precision="MIXED")
    res = []
    for variant, ry, zc, depth in ((2, 0, 0, 0), (3, 16, 0, 2), (3, 8, 8, 3),
                                   (3, 16, 24, 4)):
        knobs(variant=variant, ry=ry, zc=zc, depth=depth)
        res.append(run_gpu(case_dir, kw, 2))
        kern, (c, r_y, z, d) = res[-1]["kernel"]
        assert kern == variant
        if variant == 3:
            assert r_y == ry and c == n[0] // 64 and d == depth
    for r in res[1:]:
        assert r["iters"] == res[0]["iters"]
        for f in FIELDS:
            assert bitwise_equal(r["f"][f], res[0]["f"][f]), f


@pytest.mark.parametrize("n", [(128, 20, 40), (256, 12, 24)])
def test_row_update_mixed_same_bits(case_dir, n):
    """The mixed form's FP32 update: row kernel with 2 and 4 planes in
    flight and several plane groups against the tile kernel."""
    kw = bbconfig.channel(n=n, precision="MIXED")
    res = []
    for uvar, uzc, udepth in ((0, 0, 0), (1, 0, 4), (1, 8, 2), (1, 24, 4)):
        knobs(uvar=uvar, uzc=uzc, udepth=udepth)
        res.append(run_gpu(case_dir, kw, 2))
        assert res[-1]["update"] == uvar
    for r in res[1:]:
        assert r["iters"] == res[0]["iters"]
        for f in FIELDS:
            assert bitwise_equal(r["f"][f], res[0]["f"][f]), f


@pytest.mark.parametrize("precision", ["PLAIN", "MIXED"])
def test_row_kernels_particles_same_bits(case_dir, precision):
    """Sedimentation, 64^3 with 6 spheres: the row search and update (the
    update reading the fluid flags of the particle cells) against the
    z-column search and tile update, in both forms."""
    kw, parts = bbconfig.sediment(n=(64, 64, 64), nparts=6, a=5.0, seed=13)
    kw["extensions"] = dict(kw.get("extensions") or {},
                            pp_precision=precision)
    res = []
    for variant, uvar in ((2, 0), (3, 1)):
        knobs(variant=variant, fp64=1, uvar=uvar, ufp64=1)
        res.append(run_gpu(case_dir, kw, 2, parts))
        assert res[-1]["kernel"][0] == variant
        assert res[-1]["update"] == uvar
    assert res[1]["iters"] == res[0]["iters"]
    for f in FIELDS:
        assert bitwise_equal(res[1]["f"][f], res[0]["f"][f]), f
