"""Writers for Bluebottle input files (flow.config, decomp.config).

The text follows the grammar that src/domain.c:72-916 reads, in its rigid
order, plus the optional trailing EXTENSIONS block understood by this build
(fixed_dt, perturb, pp_precision, pp_inner_tol, pp_warmstart).  A file
without the EXTENSIONS block is a plain reference input file.
"""
import math

import numpy as np
import os

FACES = "WESNBT"


def _fmt(x):
    if isinstance(x, float):
        return repr(x)
    return str(x)


def flow_config(*, n=(32, 32, 32), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                rho_f=1.0, nu=0.01, duration=1.0, CFL=0.5,
                pp_max_iter=10000, pp_residual=1e-8,
                lamb_max_iter=50, lamb_residual=1e-2, lamb_relax=1.0,
                lamb_cut=1e-4, p_bc=None, vel_bc=None,
                init_cond="QUIESCENT", out_plane="HOMOGENEOUS",
                gradP=((0.0, 0.0), (0.0, 0.0), (0.0, 0.0)),
                g=((0.0, 0.0), (0.0, 0.0), (0.0, 0.0)),
                p_bc_tdelay=0.0, g_bc_tdelay=0.0, v_bc_tdelay=0.0,
                osci_f=0.0, scalar=None, extensions=None):
    """Return flow.config text.

    p_bc:   dict face -> ("PERIODIC",) or ("NEUMANN", 0.0)
    vel_bc: dict comp ('u','v','w') -> dict face -> ("PERIODIC",) |
            ("DIRICHLET", Dm, Da) | ("NEUMANN",)
    scalar: None (SCALAR 0) or a dict with s_D, s_k, s_beta, s_ref,
            init_s=(s_init, s_init_rand) and bc_s: face -> ("PERIODIC",) |
            ("DIRICHLET", value) | ("NEUMANN", flux); SCALAR 1.
    extensions: dict of optional keys (fixed_dt, perturb=(seed, amp),
            pp_precision, pp_inner_tol, pp_warmstart).
    """
    sc = scalar or {}
    p_bc = p_bc or {f: ("NEUMANN", 0.0) for f in FACES}
    vel_bc = vel_bc or {c: {f: ("DIRICHLET", 0.0, 0.0) for f in FACES}
                        for c in "uvw"}
    L = []
    L.append("GLOBAL DOMAIN")
    for ax, a in zip("XYZ", range(3)):
        L.append(f"({ax}s, {ax}e, {ax}n) {_fmt(float(lo[a]))} "
                 f"{_fmt(float(hi[a]))} {int(n[a])}")
    L.append("")
    L.append("MPI/GPU SUBDOMAIN DECOMPOSITION")
    L.append("(In, Jn, Kn) 1 1 1")
    L.append("")
    L.append("PHYSICAL PARAMETERS")
    L.append(f"rho_f {_fmt(float(rho_f))}")
    L.append(f"nu {_fmt(float(nu))}")
    for k in ("s_D", "s_k", "s_beta", "s_ref"):
        L.append(f"{k} {_fmt(float(sc.get(k, 0.0)))}")
    L.append("")
    L.append("SIMULATION PARAMETERS")
    L.append(f"duration {_fmt(float(duration))}")
    L.append(f"CFL {_fmt(float(CFL))}")
    L.append(f"SCALAR {1 if scalar else 0}")
    L.append(f"pp_max_iter {int(pp_max_iter)}")
    L.append(f"pp_residual {_fmt(float(pp_residual))}")
    L.append(f"lamb_max_iter {int(lamb_max_iter)}")
    L.append(f"lamb_residual {_fmt(float(lamb_residual))}")
    L.append(f"lamb_relax {_fmt(float(lamb_relax))}")
    L.append(f"lamb_cut {_fmt(float(lamb_cut))}")
    L.append(f"s_lamb_cut {_fmt(float(lamb_cut))}")
    L.append("")
    L.append("BOUNDARY CONDITIONS")
    L.append(f"v_bc_tdelay {_fmt(float(v_bc_tdelay))}")
    L.append("PRESSURE")
    for f in FACES:
        t = p_bc[f]
        L.append(f"bc.p{f} " + " ".join(_fmt(x) for x in t))
    for c, head in zip("uvw", ("X-VELOCITY", "Y-VELOCITY", "Z-VELOCITY")):
        L.append(head)
        for f in FACES:
            t = vel_bc[c][f]
            L.append(f"bc.{c}{f} " + " ".join(_fmt(x) for x in t))
    L.append("SCALAR")
    for f in FACES:
        if "bc_s" in sc:
            t = sc["bc_s"][f]
            L.append(f"bc_s.s{f} " + " ".join(_fmt(x) for x in t))
        elif p_bc[f][0] == "PERIODIC":
            L.append(f"bc_s.s{f} PERIODIC")
        else:
            L.append(f"bc_s.s{f} NEUMANN 0")
    L.append("")
    L.append("INITIAL CONDITION")
    L.append(f"init_cond {init_cond}")
    si = sc.get("init_s", (0.0, 0.0))
    L.append(f"init_s {_fmt(float(si[0]))} {_fmt(float(si[1]))}")
    L.append("")
    L.append("SOLVABILITY ENFORCEMENT PLANE")
    L.append(f"out_plane {out_plane}")
    L.append("")
    L.append("SIMULATION DRIVING CONDITIONS")
    L.append(f"p_bc_tdelay {_fmt(float(p_bc_tdelay))}")
    for ax, (m, a) in zip("xyz", gradP):
        L.append(f"gradP.{ax} {_fmt(float(m))} {_fmt(float(a))}")
    L.append(f"g_bc_tdelay {_fmt(float(g_bc_tdelay))}")
    for ax, (m, a) in zip("xyz", g):
        L.append(f"g.{ax} {_fmt(float(m))} {_fmt(float(a))}")
    L.append("")
    L.append("PID CONTROLLER GAINS")
    L.append("Kp 0")
    L.append("Ki 0")
    L.append("Kd 0")
    L.append("")
    L.append("TURBULENT FORCING PARAMETERS")
    L.append("turbA 0")
    L.append(f"osci_f {_fmt(float(osci_f))}")
    if extensions:
        L.append("")
        L.append("EXTENSIONS")
        for k, v in extensions.items():
            if isinstance(v, (tuple, list)):
                L.append(f"{k} " + " ".join(_fmt(x) for x in v))
            else:
                L.append(f"{k} {_fmt(v)}")
    return "\n".join(L) + "\n"


def decomp_config(n, lo, hi):
    L = ["(I, J, K) 0 0 0"]
    for ax, a in zip("XYZ", range(3)):
        L.append(f"({ax}s, {ax}e, {ax}n) {_fmt(float(lo[a]))} "
                 f"{_fmt(float(hi[a]))} {int(n[a])}")
    return "\n".join(L) + "\n"


def write_case(dirpath, **kw):
    """Write input/flow.config and input/decomp.config under dirpath."""
    os.makedirs(os.path.join(dirpath, "input"), exist_ok=True)
    n = kw.get("n", (32, 32, 32))
    lo = kw.get("lo", (0.0, 0.0, 0.0))
    hi = kw.get("hi", (1.0, 1.0, 1.0))
    with open(os.path.join(dirpath, "input", "flow.config"), "w") as f:
        f.write(flow_config(**kw))
    with open(os.path.join(dirpath, "input", "decomp.config"), "w") as f:
        f.write(decomp_config(n, lo, hi))
    return dirpath


def walls(lid=None):
    """All-wall velocity BCs (no slip); lid=(comp, face, value) moves one."""
    bc = {c: {f: ("DIRICHLET", 0.0, 0.0) for f in FACES} for c in "uvw"}
    if lid is not None:
        c, f, val = lid
        bc[c][f] = ("DIRICHLET", float(val), 0.0)
    return bc


def cavity(n=32, steps=100, dt=1e-3, **over):
    """Lid-driven cavity: unit cube, no-slip walls, top lid u = 1, Re = 100,
    fixed dt, Jacobi-PCG to 1e-8, zero initial fields."""
    kw = dict(n=(n, n, n), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
              rho_f=1.0, nu=0.01, duration=steps * dt - 0.5 * dt,
              pp_max_iter=100000, pp_residual=1e-8,
              p_bc={f: ("NEUMANN", 0.0) for f in FACES},
              vel_bc=walls(lid=("u", "T", 1.0)),
              init_cond="QUIESCENT", out_plane="HOMOGENEOUS",
              extensions={"fixed_dt": dt})
    kw.update(over)
    return kw


def channel(n=(256, 128, 128), seed=20260101, amp=0.1, precision="PLAIN",
            **over):
    """Periodic channel at Re_tau = 180: periodic in x and y, no-slip walls
    in z at -1 and 1, constant mean pressure gradient -1, laminar profile plus
    a seeded uniform(-amp, amp) perturbation."""
    per = ("PERIODIC",)
    vel = {c: {"W": per, "E": per, "S": per, "N": per,
               "B": ("DIRICHLET", 0.0, 0.0), "T": ("DIRICHLET", 0.0, 0.0)}
           for c in "uvw"}
    pb = {"W": per, "E": per, "S": per, "N": per,
          "B": ("NEUMANN", 0.0), "T": ("NEUMANN", 0.0)}
    ext = {"perturb": (seed, amp), "pp_precision": precision}
    kw = dict(n=n, lo=(0.0, 0.0, -1.0), hi=(4 * math.pi, 2 * math.pi, 1.0),
              rho_f=1.0, nu=1.0 / 180.0, duration=1e9, CFL=0.5,
              pp_max_iter=100000, pp_residual=1e-8, p_bc=pb, vel_bc=vel,
              init_cond="CHANNEL", out_plane="HOMOGENEOUS",
              gradP=((-1.0, 0.0), (0.0, 0.0), (0.0, 0.0)), extensions=ext)
    kw.update(over)
    return kw


# ------------------------------- particles --------------------------------
def part_config(parts, l_over_a=1.0):
    """Return part.config text in the grammar of src/particle.c:66-391.

    parts: list of dicts with keys r, x, y, z and optional u v w, rho, E,
    sigma, e_dry, coeff_fric, order, rs_r, spring_k, spring_x, spring_y,
    spring_z, spring_l, translating, rotating, and the scalar keys s,
    update, cp, srs_r (scalar node radius / r), s_order.
    """
    L = [f"n {len(parts)}", f"(l/a) {_fmt(float(l_over_a))}"]
    for p in parts:
        g = lambda k, d: float(p.get(k, d))
        L.append("")
        L.append(f"r {_fmt(g('r', 1.0))}")
        L.append(f"(x, y, z) {_fmt(g('x', 0))} {_fmt(g('y', 0))} "
                 f"{_fmt(g('z', 0))}")
        L.append(f"(u, v, w) {_fmt(g('u', 0))} {_fmt(g('v', 0))} "
                 f"{_fmt(g('w', 0))}")
        L.append("(aFx, aFy, aFz) 0 0 0")
        L.append("(aLx, aLy, aLz) 0 0 0")
        L.append(f"rho {_fmt(g('rho', 2.0))}")
        L.append(f"E {_fmt(g('E', 1e7))}")
        L.append(f"sigma {_fmt(g('sigma', 0.5))}")
        L.append(f"e_dry {_fmt(g('e_dry', 0.98))}")
        L.append(f"coeff_fric {_fmt(g('coeff_fric', 0.5))}")
        L.append(f"order {int(p.get('order', 4))}")
        L.append(f"rs/r {_fmt(g('rs_r', 1.2))}")
        L.append(f"spring_k {_fmt(g('spring_k', 0))}")
        L.append(f"spring (x, y, z) {_fmt(g('spring_x', 0))} "
                 f"{_fmt(g('spring_y', 0))} {_fmt(g('spring_z', 0))}")
        L.append(f"spring_l {_fmt(g('spring_l', 0))}")
        L.append(f"translating {int(p.get('translating', 1))}")
        L.append(f"rotating {int(p.get('rotating', 1))}")
        L.append(f"s {_fmt(g('s', 0))}")
        L.append(f"update {int(p.get('update', 0))}")
        L.append(f"cp {_fmt(g('cp', 0))}")
        L.append(f"rs {_fmt(g('srs_r', 0))}")
        L.append(f"s_order {int(p.get('s_order', 0))}")
    return "\n".join(L) + "\n"


def write_parts(dirpath, parts, l_over_a=1.0):
    os.makedirs(os.path.join(dirpath, "input"), exist_ok=True)
    with open(os.path.join(dirpath, "input", "part.config"), "w") as f:
        f.write(part_config(parts, l_over_a))


def random_spheres(n, box, a, seed, gap=0.5, max_tries=2000000):
    """Seeded rejection sampler of n non-overlapping spheres of radius a in
    a periodic box (surface gap >= `gap`, minimum-image distances)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = np.asarray(box, dtype=np.float64)
    pos = np.zeros((0, 3))
    tries = 0
    while len(pos) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could not place {n} spheres")
        c = rng.uniform(0.0, 1.0, 3) * L
        if len(pos):
            d = pos - c
            d -= L * np.round(d / L)
            if np.min(np.sum(d * d, axis=1)) < (2 * a + gap) ** 2:
                continue
        pos = np.vstack([pos, c])
    return pos


def sediment(n=(64, 64, 64), nparts=8, a=6.0, seed=20260102, rho_ratio=2.0,
             order=4, nu=0.1, galileo=100.0, lamb_residual=1e-2,
             lamb_max_iter=50, walls="", **over):
    """Sedimentation box [0, nx) x [0, ny) x [0, nz), dx = 1, spheres of
    radius a at seeded random positions, gravity in -z set from the Galileo
    number Ga = sqrt((rho_p/rho_f - 1) g d^3) / nu.  Periodic in every
    direction, except that the face pairs named in `walls` ("x", "y", "z")
    are no-slip walls with a Neumann pressure condition."""
    per = ("PERIODIC",)
    vel = {c: {f: per for f in FACES} for c in "uvw"}
    pbc = {f: per for f in FACES}
    for ax, fs in (("x", "WE"), ("y", "SN"), ("z", "BT")):
        if ax in walls:
            for f in fs:
                pbc[f] = ("NEUMANN", 0.0)
                for c in "uvw":
                    vel[c][f] = ("DIRICHLET", 0.0, 0.0)
    d = 2.0 * a
    gz = (galileo * nu) ** 2 / ((rho_ratio - 1.0) * d ** 3)
    kw = dict(n=n, lo=(0.0, 0.0, 0.0), hi=tuple(float(x) for x in n),
              rho_f=1.0, nu=nu, duration=1e9, CFL=0.5, pp_max_iter=100000,
              pp_residual=1e-8, p_bc=pbc, vel_bc=vel,
              init_cond="QUIESCENT", out_plane="HOMOGENEOUS",
              lamb_max_iter=lamb_max_iter, lamb_residual=lamb_residual,
              lamb_relax=1.0, lamb_cut=1e-4,
              g=((0.0, 0.0), (0.0, 0.0), (-gz, 0.0)), extensions={})
    kw.update(over)
    pos = random_spheres(nparts, n, a, seed)
    parts = [dict(r=a, x=p[0], y=p[1], z=p[2], rho=rho_ratio, order=order)
             for p in pos]
    return kw, parts


def rayleigh_benard(n=(256, 256, 128), Ra=1e6, Pr=1.0, amp=0.01, **over):
    """Rayleigh-Benard convection: [0, 2) x [0, 2) x [0, 1], periodic in x
    and y, no-slip walls at z = 0 and z = 1 with s = +0.5 and -0.5 there,
    Boussinesq coupling with s_beta g = 1 (g = -z), nu = sqrt(Pr / Ra),
    s_D = 1 / sqrt(Ra Pr); s starts at 0 plus the reference's seeded
    uniform(-amp, amp) noise per cell."""
    per = ("PERIODIC",)
    wall = ("DIRICHLET", 0.0, 0.0)
    vel = {c: {"W": per, "E": per, "S": per, "N": per, "B": wall, "T": wall}
           for c in "uvw"}
    pb = {"W": per, "E": per, "S": per, "N": per,
          "B": ("NEUMANN", 0.0), "T": ("NEUMANN", 0.0)}
    sbc = {"W": per, "E": per, "S": per, "N": per,
           "B": ("DIRICHLET", 0.5), "T": ("DIRICHLET", -0.5)}
    nu = math.sqrt(Pr / Ra)
    scalar = dict(s_D=1.0 / math.sqrt(Ra * Pr), s_k=0.0, s_beta=1.0,
                  s_ref=0.0, init_s=(0.0, amp), bc_s=sbc)
    kw = dict(n=n, lo=(0.0, 0.0, 0.0), hi=(2.0, 2.0, 1.0), rho_f=1.0, nu=nu,
              duration=1e9, CFL=0.5, pp_max_iter=100000, pp_residual=1e-8,
              p_bc=pb, vel_bc=vel, init_cond="QUIESCENT",
              out_plane="HOMOGENEOUS",
              g=((0.0, 0.0), (0.0, 0.0), (-1.0, 0.0)), scalar=scalar,
              extensions={})
    kw.update(over)
    return kw
