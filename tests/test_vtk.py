"""ParaView output (src/vtk.c:29-405, src/recorder.c:91-155).

CPU: the writer is host code.  A child process fills the domain from a
flow.config, points the host arrays (reference layouts) at seeded random
values and a small particle array, and writes two outputs; the files are
parsed as XML and every value is checked against the arrays, viewed [i, j, k]
through the layout maps the rest of the suite uses: the cell-centre
velocity is the mean of the two faces, phase is min(phase, 0), coordinates
are the cell corners, and out.pvd lists both outputs at their times.
The values are printed as the reference prints them (%lf, six decimals;
particles %e), which sets the tolerance.

GPU: a run with VTK_FLOW_FIELD in record.config writes the initial output
and one output per step; the last output agrees with the pulled fields.
"""
import ctypes
import os
import subprocess
import sys
import textwrap
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import bbconfig
import bbhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

WRITER = """
import ctypes, os, sys
import numpy as np
sys.path.insert(0, {pkg!r})
import bbhip
lib = bbhip.load()
os.chdir({case!r})
lib.domain_read_input()
lib.domain_fill()
nx, ny, nz = {n!r}
rng = np.random.default_rng(5)
keep = {{}}
def host(name, grid, dtype, ct, values):
    ni, nj, nk = bbhip.grid_dims(grid, (nx, ny, nz))
    a = np.ascontiguousarray(values(( ni + 2) * (nj + 2) * (nk + 2)), dtype)
    keep[name] = a
    ctypes.c_void_p.in_dll(lib, name).value = a.ctypes.data
for name, grid in (("p", 0), ("u", 1), ("v", 2), ("w", 3)):
    host(name, grid, np.float64, ctypes.c_double,
         lambda m: rng.uniform(-2, 2, m))
host("phase", 0, np.int32, ctypes.c_int, lambda m: rng.integers(-1, 4, m))
host("phase_shell", 0, np.int32, ctypes.c_int,
     lambda m: rng.integers(0, 2, m))
P = (bbhip.PartStruct * 3)()
for k in range(3):
    P[k].N = 10 + k
    P[k].r = 0.1 + 0.05 * k
    P[k].x, P[k].y, P[k].z = rng.uniform(0, 1, 3)
    P[k].u, P[k].v, P[k].w = rng.uniform(-1, 1, 3)
keep["parts"] = np.array([[P[k].N, P[k].r, P[k].x, P[k].y, P[k].z, P[k].u,
                           P[k].v, P[k].w] for k in range(3)])
ctypes.c_void_p.in_dll(lib, "parts").value = ctypes.addressof(P)
ctypes.c_int.in_dll(lib, "nparts").value = 3
ctypes.c_int.in_dll(lib, "bb_verbose").value = 0
ctypes.c_double.in_dll(lib, "rec_vtk_dt").value = 0.5
ctypes.c_int.in_dll(lib, "stepnum").value = 0
ctypes.c_double.in_dll(lib, "ttime").value = 0.0
lib.vtk_recorder_init()
lib.vtk_recorder_write()                  # stepnum 0 always writes
first = {{k: v.copy() for k, v in keep.items()}}
keep["p"] *= -1.0
ctypes.c_int.in_dll(lib, "stepnum").value = 3
ctypes.c_double.in_dll(lib, "ttime").value = 0.625
ctypes.c_double.in_dll(lib, "rec_vtk_ttime_out").value = 0.625
lib.vtk_recorder_write()
lib.vtk_recorder_write()                  # 0.125 left: not due
np.savez("arrays.npz", **{{"0_" + k: v for k, v in first.items()}},
         **{{"1_" + k: v for k, v in keep.items()}},
         clock=ctypes.c_double.in_dll(lib, "rec_vtk_ttime_out").value,
         count=ctypes.c_int.in_dll(lib, "rec_vtk_stepnum_out").value)
"""


def numbers(elem, dtype=float):
    return np.array(elem.text.split(), dtype=dtype)


def arrays(path):
    root = ET.parse(path).getroot()
    return root, {e.get("Name"): e for e in root.iter("DataArray")}


GRIDS = (("p", 0), ("u", 1), ("v", 2), ("w", 3), ("phase", 0),
         ("phase_shell", 0))


def expected_cells(g, n):
    """Cell values and cell-centre velocities from [i, j, k] views of the
    host arrays (ghosts included), flattened in VTK order (x fastest)."""
    nx, ny, nz = n
    inner = (slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 1))
    order = lambda a: a.transpose(2, 1, 0).ravel()
    vel = np.stack([
        0.5 * (g["u"][1:nx + 1, 1:ny + 1, 1:nz + 1]
               + g["u"][2:nx + 2, 1:ny + 1, 1:nz + 1]),
        0.5 * (g["v"][1:nx + 1, 1:ny + 1, 1:nz + 1]
               + g["v"][1:nx + 1, 2:ny + 2, 1:nz + 1]),
        0.5 * (g["w"][1:nx + 1, 1:ny + 1, 1:nz + 1]
               + g["w"][1:nx + 1, 1:ny + 1, 2:nz + 2])], axis=-1)
    return {"p": order(g["p"][inner]),
            "phase": order(np.minimum(g["phase"][inner], 0)),
            "phase_shell": order(g["phase_shell"][inner]),
            "vel": vel.transpose(2, 1, 0, 3).reshape(-1)}


def check_flow_piece(path, exp, n, lo, hi, tol=5.1e-7):
    nx, ny, nz = n
    root, da = arrays(path)
    assert root.get("type") == "RectilinearGrid"
    piece = root.find("RectilinearGrid/Piece")
    assert piece.get("Extent") == f"0 {nx} 0 {ny} 0 {nz}"
    for name in ("p", "vel"):
        got = numbers(da[name])
        assert got.shape == exp[name].shape, name
        assert np.max(np.abs(got - exp[name])) <= tol, name
    for name in ("phase", "phase_shell"):
        assert np.array_equal(numbers(da[name], np.int64), exp[name]), name
    for a, name in enumerate("xyz"):
        want = lo[a] + np.arange(n[a] + 1) * (hi[a] - lo[a]) / n[a]
        assert np.max(np.abs(numbers(da[name]) - want)) <= tol, name


def test_writer_files_and_values(case_dir):
    n = (6, 5, 4)
    lo, hi = (-0.5, 0.0, 1.0), (0.7, 2.0, 1.5)
    kw = bbconfig.cavity(n=8)
    kw.update(n=n, lo=lo, hi=hi)
    bbconfig.write_case(case_dir, **kw)
    code = WRITER.format(pkg=os.path.join(ROOT, "bluebottle-3.0_amd"),
                         case=case_dir, n=n)
    r = subprocess.run([sys.executable, "-c", textwrap.dedent(code)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    z = np.load(os.path.join(case_dir, "arrays.npz"))
    assert z["count"] == 2 and abs(z["clock"] - 0.125) < 1e-15
    out = os.path.join(case_dir, "output")

    pvd = ET.parse(os.path.join(out, "out.pvd")).getroot()
    sets = [(d.get("timestep"), d.get("part"), d.get("file"))
            for d in pvd.iter("DataSet")]
    assert sets == [("0.000000e+00", "0", "out_0.pvtr"),
                    ("0.000000e+00", "1", "out_0.pvtp"),
                    ("6.250000e-01", "0", "out_1.pvtr"),
                    ("6.250000e-01", "1", "out_1.pvtp")]

    for step in (0, 1):
        g = {k: bbhip.host_to_ijk(z[f"{step}_{k}"], grid, n)
             for k, grid in GRIDS}
        exp = expected_cells(g, n)
        check_flow_piece(os.path.join(out, f"out_{step}_0_of_1.vtr"), exp, n,
                         lo, hi)
        proot = ET.parse(os.path.join(out, f"out_{step}.pvtr")).getroot()
        grid = proot.find("PRectilinearGrid")
        assert grid.get("WholeExtent") == f"0 {n[0]} 0 {n[1]} 0 {n[2]}"
        assert [p.get("Source") for p in grid.iter("Piece")] == \
            [f"out_{step}_0_of_1.vtr"]
        names = [e.get("Name") for e in grid.find("PCellData")]
        assert names == ["p", "phase", "phase_shell", "vel"]

        parts = z[f"{step}_parts"]
        root, da = arrays(os.path.join(out, f"out_{step}_0_of_1.vtp"))
        piece = root.find("PolyData/Piece")
        assert piece.get("NumberOfPoints") == "3"
        pts = numbers(root.find("PolyData/Piece/Points/DataArray"))
        np.testing.assert_allclose(pts, parts[:, 2:5].ravel(), rtol=6e-6)
        assert np.array_equal(numbers(da["n"], np.int64), [10, 11, 12])
        np.testing.assert_allclose(numbers(da["r"]), parts[:, 1], rtol=6e-6)
        np.testing.assert_allclose(numbers(da["pvel"]), parts[:, 5:8].ravel(),
                                   rtol=6e-6)
        pp = ET.parse(os.path.join(out, f"out_{step}.pvtp")).getroot()
        assert [p.get("Source") for p in pp.iter("Piece")] == \
            [f"out_{step}_0_of_1.vtp"]


RECORD = """=== OUTPUT FORMATS ===
CGNS_FLOW_FIELD -1

CGNS_PARTICLES -1

VTK_FLOW_FIELD {dt}

=== RESTART CAPABILITY ===
WRITE_RESTART_MINUTES -1
WRITE_RESTART_COMPLETE 0
RESTART_STOP 0
"""


@pytest.mark.gpu
def test_run_writes_output(case_dir):
    kw, parts = bbconfig.sediment(n=(24, 24, 24), nparts=2, a=3.0, seed=3)
    bbconfig.write_case(case_dir, **kw)
    bbconfig.write_parts(case_dir, parts)
    with open(os.path.join(case_dir, "input", "record.config"), "w") as f:
        f.write(RECORD.format(dt=1e-12))       # every step is due
    n = kw["n"]
    with bbhip.Simulation(case_dir, n) as s:
        s.step(2)
        s.pull()
        s.lib.cuda_dom_pull_phase()
        g = {k: np.array(s.field(k)) for k, _ in GRIDS}
        got_parts = s.parts()
        dt_last = s.scalar("dt0")
        count = s.scalar("rec_vtk_stepnum_out", ctypes.c_int)
    assert count == 3                           # initial + two steps
    out = os.path.join(case_dir, "output")
    sets = [d.get("file") for d in
            ET.parse(os.path.join(out, "out.pvd")).getroot().iter("DataSet")]
    assert sets == [f"out_{i}.{e}" for i in range(3) for e in ("pvtr", "pvtp")]
    exp = expected_cells(g, n)
    scale = max(1.0, float(np.max(np.abs(exp["p"]))))
    lo, hi = kw.get("lo", (0.0,) * 3), kw.get("hi", (1.0,) * 3)
    check_flow_piece(os.path.join(out, "out_2_0_of_1.vtr"), exp, n, lo, hi,
                     tol=5.1e-7 * scale)
    # particles: the output comes before the step's position update
    # (src/bluebottle.c:376-387), so velocities match the pulled state and
    # positions lag it by at most one step's motion
    root, da = arrays(os.path.join(out, "out_2_0_of_1.vtp"))
    vel = np.array([[p["u"], p["v"], p["w"]] for p in got_parts]).ravel()
    np.testing.assert_allclose(numbers(da["pvel"]), vel, rtol=6e-6,
                               atol=1e-12)
    assert np.array_equal(numbers(da["n"], np.int64),
                          [p["N"] for p in got_parts])
    pos = np.array([[p["x"], p["y"], p["z"]] for p in got_parts]).ravel()
    moved = dt_last * np.max(np.abs(vel)) + 6e-6 * np.max(np.abs(pos))
    assert np.max(np.abs(numbers(root.find(
        "PolyData/Piece/Points/DataArray")) - pos)) <= moved
    assert np.any(exp["phase"] == 0) and np.any(exp["phase"] == -1)
