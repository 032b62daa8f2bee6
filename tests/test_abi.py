"""C-ABI boundary and host logic (CPU only, no GPU calls).

- libbluebottle.so loads and exports every function and global that
  include/*.h declares;
- the flow.config parser reads the reference grammar (src/domain.c:72-916),
  the optional EXTENSIONS block, and fails loudly on malformed input;
- domain_fill produces the reference grid bookkeeping (src/domain.c:918-1146).
"""
import ctypes
import glob
import os
import re
import subprocess
import sys
import textwrap

import pytest

import bbconfig
import bbhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    funcs, vars_ = set(), set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?!typedef|extern|#|return)[A-Za-z_][\w\s\*]*?"
                             r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, re.M):
            funcs.add(m.group(1))
        for m in re.finditer(r"^\s*extern\s+([^;(]+);", text, re.M):
            decl = m.group(1)
            decl = re.sub(r"^(unsigned\s+long\s+long|long\s+int|[A-Za-z_]\w*)\s+",
                          "", decl.strip())
            for part in decl.split(","):
                name = part.strip().lstrip("*").strip()
                if re.fullmatch(r"[A-Za-z_]\w*", name):
                    vars_.add(name)
    return funcs, vars_


def test_library_exports_every_declared_symbol():
    lib = bbhip.load()
    funcs, vars_ = declared_symbols()
    assert len(funcs) > 30 and len(vars_) > 60
    missing = [f for f in sorted(funcs) if not hasattr(lib, f)]
    assert not missing, f"functions not exported: {missing}"
    missing_v = []
    for v in sorted(vars_):
        try:
            ctypes.c_char.in_dll(lib, v)
        except ValueError:
            missing_v.append(v)
    assert not missing_v, f"globals not exported: {missing_v}"


def test_library_exports_what_the_reference_host_calls(tmp_path):
    """Every function and global that the reference's host C files
    (src/bluebottle.c, domain.c, particle.c, recorder.c, scalar.c, rng.c,
    vtk.c) take from the files this library replaces (the CUDA sources and
    src/mpi_comm.c) is exported, so those objects link against the library
    in their place.  The list is tests/golden/host_calls.json, extracted
    from the reference sources by tests/golden/extract_host_calls.py.  A C
    program that references every one of them links against the library."""
    import json
    fx = json.load(open(os.path.join(ROOT, "tests", "golden",
                                     "host_calls.json")))
    funcs = {k for k, v in fx["functions"].items()
             if not v["test_build_only"]}
    globs = set(fx["globals"])
    assert len(funcs) >= 65 and "mpi_cuda_exchange_Gcc" in funcs
    lib_path = bbhip.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path],
                         capture_output=True, text=True, check=True).stdout
    syms = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = sorted((funcs | globs) - syms)
    assert not missing, f"not exported: {missing}"
    # link check: a program that takes the address of each symbol
    src = tmp_path / "host_link.c"
    body = ["extern char " + n + "[];" for n in sorted(funcs | globs)]
    body.append("void *tab[] = {" + ", ".join(sorted(funcs | globs)) + "};")
    body.append("int main(void) { return tab[0] == 0; }")
    src.write_text("\n".join(body) + "\n")
    exe = tmp_path / "host_link"
    r = subprocess.run(["gcc", "-o", str(exe), str(src), "-L",
                        os.path.dirname(lib_path), "-lbluebottle",
                        "-Wl,--no-as-needed"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def run_parse(case_dir, extra=""):
    """Run domain_read_input + domain_fill in a child process (the parser
    exits the process on malformed input, like the reference)."""
    code = textwrap.dedent(f"""
        import ctypes, os, sys
        sys.path.insert(0, {os.path.join(ROOT, 'bluebottle-3.0_amd')!r})
        import bbhip
        lib = bbhip.load()
        os.chdir({case_dir!r})
        lib.domain_read_input()
        lib.domain_fill()
        d = lambda n: ctypes.c_double.in_dll(lib, n).value
        i = lambda n: ctypes.c_int.in_dll(lib, n).value
        print("OK", d("rho_f"), d("nu"), d("mu"), i("pp_max_iter"),
              d("pp_residual"), i("init_cond"), i("out_plane"),
              d("fixed_dt"), i("pp_precision"), d("perturb_amp"))
    """) + "\n" + extra
    return subprocess.run([sys.executable, "-c", code], capture_output=True,
                          text=True, timeout=120)


def test_parse_reference_grammar(tmp_path):
    kw = bbconfig.cavity(n=8)
    kw["extensions"] = None            # a plain reference-format file
    bbconfig.write_case(str(tmp_path), **kw)
    r = run_parse(str(tmp_path))
    assert r.returncode == 0, r.stderr
    f = r.stdout.split()
    assert f[0] == "OK"
    assert float(f[1]) == 1.0 and float(f[2]) == 0.01 and float(f[3]) == 0.01
    assert int(f[4]) == 100000 and float(f[5]) == 1e-8
    assert int(f[6]) == 0 and int(f[7]) == 10        # QUIESCENT, HOMOGENEOUS
    assert float(f[8]) == 0.0 and int(f[9]) == 0      # defaults


def test_parse_extensions(tmp_path):
    kw = bbconfig.channel(n=(16, 8, 8), precision="MIXED")
    kw["extensions"]["fixed_dt"] = 2.5e-3
    bbconfig.write_case(str(tmp_path), **kw)
    r = run_parse(str(tmp_path))
    assert r.returncode == 0, r.stderr
    f = r.stdout.split()
    assert int(f[6]) == 2                              # CHANNEL
    assert float(f[8]) == 2.5e-3 and int(f[9]) == 1 and float(f[10]) == 0.1


def test_parse_rejects_misspelled_key(tmp_path):
    """The reference's own sim/input/flow.config writes `s_alpha` where the
    parser expects `s_beta` (src/domain.c:168) and silently desynchronises;
    here the read stops with a message naming the expected text."""
    kw = bbconfig.cavity(n=8)
    bbconfig.write_case(str(tmp_path), **kw)
    path = os.path.join(str(tmp_path), "input", "flow.config")
    text = open(path).read().replace("s_beta 0", "s_alpha 20.")
    open(path, "w").write(text)
    r = run_parse(str(tmp_path))
    assert r.returncode != 0
    assert "s_beta" in r.stderr


def test_parse_rejects_bad_bc_and_decomposition(tmp_path):
    kw = bbconfig.cavity(n=8)
    bbconfig.write_case(str(tmp_path), **kw)
    path = os.path.join(str(tmp_path), "input", "flow.config")
    good = open(path).read()
    open(path, "w").write(good.replace("bc.uT DIRICHLET 1.0 0.0",
                                       "bc.uT SLIP 1.0 0.0"))
    r = run_parse(str(tmp_path))
    assert r.returncode != 0 and "DIRICHLET" in r.stderr
    open(path, "w").write(good.replace("(In, Jn, Kn) 1 1 1",
                                       "(In, Jn, Kn) 2 1 1"))
    r = run_parse(str(tmp_path))
    assert r.returncode != 0 and "1 1 1" in r.stderr
    open(path, "w").write(good + "bogus_key 1\n")   # inside EXTENSIONS
    r = run_parse(str(tmp_path))
    assert r.returncode != 0 and "bogus_key" in r.stderr


def test_domain_fill_grid_bookkeeping(tmp_path):
    """Strides and index ranges of src/domain.c:934-1146 for every grid."""
    kw = bbconfig.channel(n=(10, 7, 5))
    bbconfig.write_case(str(tmp_path), **kw)
    extra = textwrap.dedent("""
        class G(ctypes.Structure):
            _fields_ = [(n, ctypes.c_int) for n in (
                "is ie in isb ieb inb js je jn jsb jeb jnb ks ke kn ksb keb knb "
                "_is _ie _isb _ieb _js _je _jsb _jeb _ks _ke _ksb _keb "
                "s1 s1b s2 s2b s3 s3b s2_i s2_j s2_k s2b_i s2b_j s2b_k").split()]
        class D(ctypes.Structure):
            _fields_ = [("Gcc", G), ("Gfx", G), ("Gfy", G), ("Gfz", G)]
        dd = D.in_dll(lib, "DOM")
        for g in ("Gcc", "Gfx", "Gfy", "Gfz"):
            x = getattr(dd, g)
            print(g, x.in_, x.jn, x.kn, x.ie, x.ieb, x.s1b, x.s2b, x.s3b)
    """).replace("x.in_", "getattr(x, 'in')")
    r = run_parse(str(tmp_path), extra)
    assert r.returncode == 0, r.stderr
    rows = {l.split()[0]: list(map(int, l.split()[1:]))
            for l in r.stdout.splitlines()[1:]}
    nx, ny, nz = 10, 7, 5
    # Gcc: x-fastest, s1b = inb, s2b = inb*jnb
    assert rows["Gcc"] == [nx, ny, nz, nx, nx + 1, nx + 2, (nx + 2) * (ny + 2),
                           (nx + 2) * (ny + 2) * (nz + 2)]
    # Gfx: y-fastest (GFX_LOC), s1b = jnb, s2b = jnb*knb
    assert rows["Gfx"] == [nx + 1, ny, nz, nx + 1, nx + 2, ny + 2,
                           (ny + 2) * (nz + 2), (nx + 3) * (ny + 2) * (nz + 2)]
    # Gfy: z-fastest (GFY_LOC), s1b = knb, s2b = knb*inb
    assert rows["Gfy"] == [nx, ny + 1, nz, nx, nx + 1, nz + 2,
                           (nz + 2) * (nx + 2), (nx + 2) * (ny + 3) * (nz + 2)]
    assert rows["Gfz"] == [nx, ny, nz + 1, nx, nx + 1, nx + 2,
                           (nx + 2) * (ny + 2), (nx + 2) * (ny + 2) * (nz + 3)]


def test_executable_fails_cleanly_on_missing_input(tmp_path):
    exe = os.path.join(ROOT, "bluebottle-3.0_amd", "bin", "bluebottle")
    r = subprocess.run([exe], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0
    assert "flow.config" in r.stderr


def test_rejects_out_of_scope_options(tmp_path):
    kw = bbconfig.cavity(n=8)
    bbconfig.write_case(str(tmp_path), **kw)
    path = os.path.join(str(tmp_path), "input", "flow.config")
    good = open(path).read()
    open(path, "w").write(good.replace("Kp 0", "Kp 1.5"))
    r = run_parse(str(tmp_path))
    assert r.returncode != 0 and "PID" in r.stderr
    # scalar transport, also with particles, is read without complaint
    open(path, "w").write(good.replace("SCALAR 0", "SCALAR 1"))
    r = run_parse(str(tmp_path))
    assert r.returncode == 0, r.stderr
    kw, parts = bbconfig.sediment(n=(24, 24, 24), nparts=1, a=3.0, seed=1)
    bbconfig.write_case(str(tmp_path), **kw)
    bbconfig.write_parts(str(tmp_path), parts)
    good = open(path).read()
    open(path, "w").write(good.replace("SCALAR 0", "SCALAR 1"))
    code = textwrap.dedent(f"""
        import os, sys
        sys.path.insert(0, {os.path.join(ROOT, 'bluebottle-3.0_amd')!r})
        import bbhip
        lib = bbhip.load()
        os.chdir({str(tmp_path)!r})
        lib.domain_read_input()
        lib.domain_fill()
        lib.parts_read_input()
        print("OK")
    """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "OK" in r.stdout, r.stderr


def test_restart_file_layout(tmp_path):
    """out_restart (no device involved: host arrays supplied here) writes
    the reference's blocks in the reference's order and layouts: every
    array is filled with (its number in the file) * 1e6 + its host index,
    and tests/helpers.read_restart, which maps [i, j, k] to the host index
    of src/domain.c (Gfx y-fastest, Gfy z-fastest, Gcc and Gfz
    x-fastest), must find each value where that formula puts it; the
    particle and scalar blocks and the scalars around them included."""
    n = (6, 5, 4)
    kw = bbconfig.channel(n=n)
    kw["scalar"] = dict(s_D=0.1, s_k=1.0, s_beta=0.1, s_ref=0.0,
                        init_s=(0.0, 0.0),
                        bc_s={f: ("PERIODIC",) for f in "WESNBT"})
    bbconfig.write_case(str(tmp_path), **kw)
    extra = textwrap.dedent("""
        import numpy as np
        sys.path.insert(0, {tests!r})
        nx, ny, nz = {n!r}
        sizes = {{"x": (nx + 3) * (ny + 2) * (nz + 2),
                  "y": (nx + 2) * (ny + 3) * (nz + 2),
                  "z": (nx + 2) * (ny + 2) * (nz + 3),
                  "c": (nx + 2) * (ny + 2) * (nz + 2)}}
        order = []
        for comp, kind in (("u", "x"), ("v", "y"), ("w", "z")):
            for nm in (comp, comp + "0", "diff0_" + comp, "conv0_" + comp,
                       "diff_" + comp, "conv_" + comp, comp + "_star"):
                order.append((nm, kind, "d"))
        order += [("p", "c", "d"), ("phi", "c", "d"), ("p0", "c", "d"),
                  ("phase", "c", "i"), ("phase_shell", "c", "i"),
                  ("flag_u", "x", "i"), ("flag_v", "y", "i"),
                  ("flag_w", "z", "i")]
        order += [(nm, "c", "d") for nm in ("s", "s0", "s_conv", "s_conv0",
                                            "s_diff", "s_diff0")]
        keep = []
        for num, (nm, kind, t) in enumerate(order):
            m = sizes[kind]
            if t == "d":
                a = (num * 1e6 + np.arange(m)).astype(np.float64)
            else:
                a = (num * 100000 + np.arange(m)).astype(np.int32)
            keep.append(a)
            ctypes.c_void_p.in_dll(lib, nm).value = a.ctypes.data
        parts = (bbhip.PartStruct * 2)()
        for k in range(2):
            parts[k].N = k
            parts[k].x = 1.5 + k
            parts[k].pnm_re[3] = 7.0 + k
        ctypes.c_void_p.in_dll(lib, "parts").value = ctypes.addressof(parts)
        ctypes.c_int.in_dll(lib, "nparts").value = 2
        ctypes.c_int.in_dll(lib, "ncoeffs_max").value = 15
        ctypes.c_int.in_dll(lib, "s_ncoeffs_max").value = 4
        ctypes.c_double.in_dll(lib, "ttime").value = 0.25
        ctypes.c_double.in_dll(lib, "dt").value = 0.125
        ctypes.c_int.in_dll(lib, "stepnum").value = 7
        lib.out_restart()
        from helpers import read_restart
        f = read_restart(os.path.join("input", "restart.config-0"),
                         (nx, ny, nz), 2, True)
        ib = {{"x": nx + 3, "y": nx + 2, "z": nx + 2, "c": nx + 2}}
        jb = {{"x": ny + 2, "y": ny + 3, "z": ny + 2, "c": ny + 2}}
        kb = {{"x": nz + 2, "y": nz + 2, "z": nz + 3, "c": nz + 2}}
        bad = 0
        for num, (nm, kind, t) in enumerate(order):
            a = f[nm]
            I, J, K = np.meshgrid(np.arange(ib[kind]), np.arange(jb[kind]),
                                  np.arange(kb[kind]), indexing="ij")
            if kind == "x":
                idx = J + K * jb[kind] + I * jb[kind] * kb[kind]
            elif kind == "y":
                idx = K + I * kb[kind] + J * kb[kind] * ib[kind]
            else:
                idx = I + J * ib[kind] + K * ib[kind] * jb[kind]
            want = num * (1e6 if t == "d" else 100000) + idx
            bad += int(not np.array_equal(a, want.astype(a.dtype)))
        print("LAYOUT", bad, f["stepnum"], f["ttime"], f["dt"],
              [p.N for p in f["parts"]], [p.x for p in f["parts"]],
              f["parts"][1].pnm_re[3], f["ncoeffs_max"], f["s_ncoeffs_max"])
    """).format(tests=os.path.join(ROOT, "tests"), n=n)
    r = run_parse(str(tmp_path), extra)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("LAYOUT")][0]
    assert line == ("LAYOUT 0 7 0.25 0.125 [0, 1] [1.5, 2.5] 8.0 15 4"), line
