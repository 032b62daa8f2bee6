"""Restart files (src/domain.c:3005-3261, src/recorder.c:338-391).

A run of 4 steps and a run of 2 steps, restart file, new process state
from that file and 2 more steps give the same fields and particle states
bit for bit (particles, Lamb coefficients, scalar and particle heat
included).  The file has the reference's layout: its size is the sum of
the reference's blocks, and its header holds ttime, dt0, dt, stepnum.
"""
import ctypes
import os
import struct

import numpy as np
import pytest

import bbconfig
import bbhip
from helpers import bitwise_equal, describe, interior, read_restart

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "s")
PART_KEYS = ("x", "y", "z", "u", "v", "w", "ox", "oy", "oz", "Fx", "Fy",
             "Fz", "Lx", "Ly", "Lz", "iFx", "iFy", "iFz", "s", "q", "pnm_re",
             "pnm_im", "phinm_re", "chinm_im", "anm_re", "anm_im", "St", "iSt")


def case():
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=3, a=4.0, seed=7)
    per = ("PERIODIC",)
    kw["scalar"] = dict(s_D=0.05, s_k=1.0, s_beta=0.2, s_ref=0.5,
                        init_s=(0.5, 0.05), bc_s={f: per for f in "WESNBT"})
    for k, p in enumerate(parts):
        p.update(s=0.2 + 0.3 * k, update=1, cp=1.5, srs_r=1.2, s_order=1)
    return kw, parts


def run(case_dir, n, nsteps, restart=False, write=False):
    lib = bbhip.load()
    ctypes.c_int.in_dll(lib, "use_restart").value = 1 if restart else 0
    try:
        with bbhip.Simulation(case_dir, n) as s:
            for _ in range(nsteps):
                s.step()
            s.pull()
            parts = s.parts()
            if write:
                cwd = os.getcwd()
                os.chdir(case_dir)          # the file goes to ./input
                try:
                    lib.out_restart()
                finally:
                    os.chdir(cwd)
            out = {f: np.array(s.field(f)) for f in FIELDS}
            out["parts"] = parts
            out["ttime"] = s.scalar("ttime")
            out["dt"] = s.scalar("dt")
            out["dt0"] = s.scalar("dt0")
            out["stepnum"] = s.scalar("stepnum", ctypes.c_int)
    finally:
        ctypes.c_int.in_dll(lib, "use_restart").value = 0
    return out


def expected_size(n, nparts, part_bytes, scalar):
    nx, ny, nz = n
    gcc = (nx + 2) * (ny + 2) * (nz + 2)
    gfx = (nx + 3) * (ny + 2) * (nz + 2)
    gfy = (nx + 2) * (ny + 3) * (nz + 2)
    gfz = (nx + 2) * (ny + 2) * (nz + 3)
    size = 3 * 8 + 2 * 4 + 3 * 8                  # times, step counters
    size += 7 * 8 * (gfx + gfy + gfz)             # u..u_star, v.., w..
    size += 3 * 8 * gcc + 2 * 4 * gcc             # p, phi, p0, phase(s)
    size += 4 * (gfx + gfy + gfz)                 # flags
    size += 4 + nparts * part_bytes               # particles
    size += 4 + 3 * 8                             # ncoeffs_max, pid, gradP.z
    if scalar:
        size += 6 * 8 * gcc + 4
    return size


def test_restart_round_trip(case_dir):
    kw, parts = case()
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    n = kw["n"]
    full = run(case_dir, n, 4)
    half = run(case_dir, n, 2, write=True)
    path = os.path.join(case_dir, "input", "restart.config-0")
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        ttime, dt0, dt, stepnum = struct.unpack("<dddi", f.read(28))
    assert size == expected_size(n, 3, ctypes.sizeof(bbhip.PartStruct), True)
    assert (ttime, dt0, dt, stepnum) == (half["ttime"], half["dt0"],
                                         half["dt"], half["stepnum"])
    rest = run(case_dir, n, 2, restart=True)
    assert rest["stepnum"] == full["stepnum"] == 4
    assert rest["ttime"] == full["ttime"] and rest["dt"] == full["dt"]
    for f in FIELDS:
        a, b = interior(rest[f]), interior(full[f])
        assert bitwise_equal(a, b), describe(f, a, b)
    for pa, pb in zip(rest["parts"], full["parts"]):
        for k in PART_KEYS:
            assert np.array_equal(np.asarray(pa[k]), np.asarray(pb[k])), k


# ---------------------------------------------------------------------------
# The file against the FP64 oracle: read by helpers.read_restart (the
# reference's block order and layouts), its fields and particles compared
# with the oracle's state after the same two steps, then the run restarted
# from it compared with the oracle's next two steps.

def test_restart_file_against_oracle(case_dir):
    """After two steps the GPU's restart file, read by the reader above,
    holds the oracle's fields, flags and particles (1e-10; integers equal);
    restarted from it, the GPU's next two steps match the oracle's."""
    import orc
    from helpers import rel_l2
    kw, parts = case()
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    n = kw["n"]
    run(case_dir, n, 2, write=True)
    path = os.path.join(case_dir, "input", "restart.config-0")
    f = read_restart(path, n, len(parts), True)
    fields = ("u", "v", "w", "p", "phi", "u0", "v0", "w0", "p0", "conv0_u",
              "diff0_v", "s", "s0")
    with orc.Oracle(case_dir) as o:
        o.step(2)
        assert int(f["stepnum"]) == 2
        for q in ("dt", "ttime"):         # dt follows the velocities
            assert abs(f[q] - o.scalar(q)) <= 1e-12 * abs(o.scalar(q)), q
        worst = 0.
        for name in fields:
            a, b = interior(f[name]), interior(np.array(o.field(name)))
            e = rel_l2(a, b)
            worst = max(worst, e)
            assert e <= 1e-10, describe(name, a, b)
        for name in ("phase", "phase_shell", "flag_u", "flag_v", "flag_w"):
            assert np.array_equal(interior(f[name]),
                                  interior(np.array(o.field(name)))), name
        for k in range(len(parts)):
            want = o.part_state(k)
            for q in ("x", "y", "z", "u", "v", "w", "ox", "Fz", "Lx"):
                a, b = getattr(f["parts"][k], q), want[q]
                assert abs(a - b) <= 1e-10 * max(abs(b), 1e-300), (k, q)
            assert np.allclose(np.array(f["parts"][k].pnm_re[:]),
                               want["pnm_re"], rtol=1e-9, atol=1e-12)
        print("restart file vs oracle: worst field rel L2", f"{worst:.2e}")
        o.step(2)
        want = {name: np.array(o.field(name)) for name in FIELDS}
        lamb_orc = int(o.scalar("lamb_iters"))
    rest = run(case_dir, n, 2, restart=True)
    for name in FIELDS:
        a, b = interior(rest[name]), interior(want[name])
        e = rel_l2(a, b)
        print(name, f"after restart vs oracle rel L2 {e:.2e}")
        assert e <= 1e-10, describe(name, a, b)
    assert rest["stepnum"] == 4
    assert lamb_orc > 0


def test_rec_logs(case_dir):
    """solver_expd.rec holds one line per pressure solve with its step and
    iteration count as the library counted them and a residual at the
    stopping test; iter_etime.rec holds one line per step with the Lamb
    iteration count of the oracle's same step."""
    import orc
    kw, parts = bbconfig.sediment(n=(32, 32, 32), nparts=3, a=4.0, seed=7)
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    lib = bbhip.load()
    slog = (ctypes.c_int * 8192).in_dll(lib, "bb_solve_log")
    solves = ctypes.c_long.in_dll(lib, "bb_pcg_solves_total")
    with orc.Oracle(case_dir) as o:
        lamb_orc = []
        for _ in range(3):
            o.step()
            lamb_orc.append(int(o.scalar("lamb_iters")))
    expect = []
    with bbhip.Simulation(case_dir, kw["n"], record=True) as s:
        for step in range(1, 4):
            n0 = solves.value
            s.step()
            expect += [(step, slog[m % 8192]) for m in range(n0, solves.value)]
    rec = os.path.join(case_dir, "record")
    rows = [l.split() for l in open(os.path.join(rec, "solver_expd.rec"))
            if l.strip()][1:]
    got = [(int(r[0]), int(r[3])) for r in rows]
    assert got == expect
    assert all(float(r[4]) <= 1e-8 for r in rows if int(r[3]) > 0)
    rows = [l.split() for l in open(os.path.join(rec, "iter_etime.rec"))
            if l.strip()][1:]
    assert [int(r[0]) for r in rows] == [1, 2, 3]
    assert [int(r[2]) for r in rows] == lamb_orc
    print("solves logged", len(got), "Lamb iterations per step", lamb_orc)
