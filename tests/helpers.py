"""Comparison helpers shared by the parity tests."""
import ctypes

import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    if nb == 0.0:
        return d
    return d / nb


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def bitwise_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def describe(name, a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    diff = np.abs(a - b)
    idx = np.unravel_index(np.argmax(diff), diff.shape) if diff.size else ()
    return (f"{name}: rel_l2={rel_l2(a, b):.3e} max_abs={diff.max():.3e} "
            f"at {idx} gpu={a[idx] if diff.size else None} "
            f"oracle={b[idx] if diff.size else None}")


def grid_eigen_range(kw):
    """Smallest nonzero and largest eigenvalue bound and the largest
    diagonal entry of (-A) (the 7-point Laplacian the PCG solves) on the
    uniform grid of a case without particles: per direction the smallest
    nonzero eigenvalue is 4 sin^2(pi/n)/h^2 (periodic) or
    4 sin^2(pi/(2n))/h^2 (zero-flux walls); the largest is below
    sum 4/h^2."""
    n, lo, hi = kw["n"], kw["lo"], kw["hi"]
    h = [(hi[a] - lo[a]) / n[a] for a in range(3)]
    per = [kw["p_bc"][f][0] == "PERIODIC" for f in "WSB"]
    lmin = min(4.0 * np.sin(np.pi / (n[a] if per[a] else 2 * n[a])) ** 2
               / h[a] ** 2 for a in range(3))
    lmax = sum(4.0 / h[a] ** 2 for a in range(3))
    return lmin, lmax, sum(2.0 / h[a] ** 2 for a in range(3))


def solve_difference_bounds(lam_min, lam_max, diag_max, nsolves, tol=1e-8):
    """Largest relative difference between two runs whose pressure solves
    all pass the PCG's stopping test (r, M^-1 r) <= tol^2 (b, b), with
    M = diag(-A) <= diag_max:
      ||r|| <= sqrt(diag_max) tol ||b|| for each solve;
      the error e = A^-1 r of a solve against the exact discrete solution
      phi* then has ||e|| / ||phi*|| <= kappa sqrt(diag_max) tol
      (||e|| <= ||r|| / lam_min, ||phi*|| >= ||b|| / lam_max), and
      its gradient ||G e|| / ||G phi*|| = ||e||_A / ||phi*||_A
      <= sqrt(kappa diag_max) tol;
      two runs differ by at most twice that per solve, and the differences
      of successive solves add (the step map is taken as non-expansive).
    Returns the bound for phi-like fields (phi, p) and for the projected
    velocity relative to the norm of the whole velocity field."""
    kappa = lam_max / lam_min
    r = np.sqrt(diag_max) * tol
    return {"phi": 2.0 * nsolves * kappa * r,
            "grad": 2.0 * nsolves * np.sqrt(kappa) * r}


def field_bounds(b, fields_ref):
    """Per-field bounds from solve_difference_bounds: p and phi take the
    phi bound, u, v and w the gradient bound scaled by ||U|| / ||u_c||
    (the reference run's fields)."""
    U = np.sqrt(sum(np.linalg.norm(interior(fields_ref[c])) ** 2
                    for c in ("u", "v", "w")))
    out = {}
    for f in fields_ref:
        if f in ("u", "v", "w"):
            nc = np.linalg.norm(interior(fields_ref[f]))
            out[f] = b["grad"] * U / max(nc, 1e-300)
        elif f in ("p", "phi"):
            out[f] = b["phi"]
    return out


# ---------------------------------------------------------------------------
# Reader of the reference's restart file: block order of
# src/domain.c:3025-3086; host arrays indexed as src/domain.c:918-1487 lays
# them out (Gcc and Gfz x-fastest, Gfx y-fastest, Gfy z-fastest, one ghost
# layer, the face grid one longer in its own direction).

def restart_grid(n, kind):
    nx, ny, nz = n
    ib, jb, kb = nx + 2, ny + 2, nz + 2
    if kind == "x":
        ib += 1
    elif kind == "y":
        jb += 1
    elif kind == "z":
        kb += 1
    return ib, jb, kb


def restart_to_ijk(flat, n, kind):
    """[i, j, k] view of a host array in the reference layout."""
    ib, jb, kb = restart_grid(n, kind)
    if kind == "x":                     # j + k jb + i jb kb
        return flat.reshape(ib, kb, jb).transpose(0, 2, 1)
    if kind == "y":                     # k + i kb + j kb ib
        return flat.reshape(jb, ib, kb).transpose(1, 0, 2)
    return flat.reshape(kb, jb, ib).transpose(2, 1, 0)   # x-fastest


def read_restart(path, n, nparts, scalar):
    """Dict of the blocks of a restart file (fields as [i, j, k] arrays,
    particles as part_struct records); asserts the file ends there."""
    import bbhip
    data = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(data, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    out = {}
    out["ttime"], out["dt0"], out["dt"] = take("<f8", 3)
    out["stepnum"], out["vtk_stepnum"] = take("<i4", 2)
    out["clocks"] = take("<f8", 3)
    for comp, kind in (("u", "x"), ("v", "y"), ("w", "z")):
        size = int(np.prod(restart_grid(n, kind)))
        for name in (comp, comp + "0", "diff0_" + comp, "conv0_" + comp,
                     "diff_" + comp, "conv_" + comp, comp + "_star"):
            out[name] = restart_to_ijk(take("<f8", size), n, kind)
    size = int(np.prod(restart_grid(n, "c")))
    for name in ("p", "phi", "p0"):
        out[name] = restart_to_ijk(take("<f8", size), n, "c")
    for name in ("phase", "phase_shell"):
        out[name] = restart_to_ijk(take("<i4", size), n, "c")
    for name, kind in (("flag_u", "x"), ("flag_v", "y"), ("flag_w", "z")):
        out[name] = restart_to_ijk(
            take("<i4", int(np.prod(restart_grid(n, kind)))), n, kind)
    out["nparts"] = int(take("<i4", 1)[0])
    assert out["nparts"] == nparts
    psz = ctypes.sizeof(bbhip.PartStruct)
    out["parts"] = [bbhip.PartStruct.from_buffer_copy(data, off + k * psz)
                    for k in range(nparts)]
    off += nparts * psz
    out["ncoeffs_max"] = int(take("<i4", 1)[0])
    out["pid_int"], out["pid_back"], out["gradPz"] = take("<f8", 3)
    if scalar:
        for name in ("s", "s0", "s_conv", "s_conv0", "s_diff", "s_diff0"):
            out[name] = restart_to_ijk(take("<f8", size), n, "c")
        out["s_ncoeffs_max"] = int(take("<i4", 1)[0])
    assert off == len(data), (off, len(data))
    return out
