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
from helpers import bitwise_equal, describe, interior

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
