"""Scalar transport of the oracle (src/scalar_kernel.cu:4-301,
src/cuda_scalar.cu:149-257, 949-957), pinned without restating it:

- a Fourier mode in a periodic box decays by the discrete Laplacian's
  eigenvalue (2 cos(k h) - 2) / h^2, with Euler on the first step and
  Adams-Bashforth 2 (1.5 / -0.5 at constant dt) after;
- the flux form conserves the sum of s in a periodic box for any velocity
  field, divergence-free or not;
- a linear profile between Dirichlet walls, and a profile with the imposed
  slope between Neumann walls, are steady (the ghost-cell forms);
- a uniform s drives w = dt (-g s_beta (s - s_ref)) in one step from rest
  (the Boussinesq sign and face average);
- with s_D > nu the time step follows s_D (CFL / sum 2 s_D / h^2).
"""
import math

import numpy as np
import pytest

import bbconfig
import orc

L = 16


def periodic_case(tmp, D=0.05, g=0.0, beta=0.0, sref=0.0, fixed_dt=0.02,
                  bc=None, n=(L, L, L), walls_z=False):
    per = ("PERIODIC",)
    wall = ("DIRICHLET", 0.0, 0.0)
    vel = {c: {f: per for f in "WESNBT"} for c in "uvw"}
    pb = {f: per for f in "WESNBT"}
    sbc = {f: per for f in "WESNBT"}
    if walls_z:
        for f in "BT":
            pb[f] = ("NEUMANN", 0.0)
            for c in "uvw":
                vel[c][f] = wall
    if bc:
        sbc.update(bc)
    scalar = dict(s_D=D, s_k=0.0, s_beta=beta, s_ref=sref, init_s=(0.0, 0.0),
                  bc_s=sbc)
    ext = {"fixed_dt": fixed_dt} if fixed_dt else {}
    kw = dict(n=n, lo=(0.0, 0.0, 0.0), hi=tuple(float(x) for x in n),
              rho_f=1.0, nu=0.01, duration=1e9, CFL=0.5, pp_max_iter=100000,
              pp_residual=1e-10, p_bc=pb, vel_bc=vel, init_cond="QUIESCENT",
              out_plane="HOMOGENEOUS", g=((0.0, 0.0), (0.0, 0.0), (g, 0.0)),
              scalar=scalar, extensions=ext)
    d = str(tmp)
    bbconfig.write_case(d, **kw)
    o = orc.Oracle(d)
    o.setup()
    return o, kw


def centres(n):
    return (np.arange(n + 2) - 0.5)


def set_s(o, f):
    for name in ("s", "s0"):
        a = o.field(name)
        a[...] = f


def test_fourier_mode_decay(tmp_path):
    D, dt = 0.05, 0.02
    o, _ = periodic_case(tmp_path, D=D, fixed_dt=dt)
    x = centres(L)
    k = 2 * math.pi / L
    mode = np.cos(k * x)[:, None, None] * np.ones((1, L + 2, L + 2))
    set_s(o, mode)
    lam = D * (2 * math.cos(k) - 2)          # h = 1
    amps = [1.0]
    for step in range(3):
        o.step()
        s = np.array(o.field("s"))[1:-1, 1:-1, 1:-1]
        m = mode[1:-1, 1:-1, 1:-1]
        amps.append(float(np.sum(s * m) / np.sum(m * m)))
        resid = s - amps[-1] * m
        assert np.max(np.abs(resid)) <= 1e-14
    o.teardown()
    want = [1.0, 1.0 + dt * lam]
    for _ in range(2):
        want.append(want[-1] + dt * (1.5 * lam * want[-1] - 0.5 * lam * want[-2]))
    assert np.allclose(amps, want, rtol=0, atol=1e-14), (amps, want)


def test_flux_form_conserves_sum(tmp_path):
    o, _ = periodic_case(tmp_path, D=0.03, fixed_dt=0.05)
    rng = np.random.default_rng(3)
    set_s(o, rng.uniform(0.0, 1.0, (L + 2, L + 2, L + 2)))
    for ax, name in enumerate(("u0", "v0", "w0")):
        a = o.field(name)
        a[...] = rng.uniform(-0.3, 0.3, a.shape)
        # the face at x = L is the face at x = 0 (periodic)
        idx_hi = [slice(None)] * 3
        idx_lo = [slice(None)] * 3
        idx_hi[ax], idx_lo[ax] = L + 1, 1
        a[tuple(idx_hi)] = a[tuple(idx_lo)]
    before = float(np.sum(np.array(o.field("s"))[1:-1, 1:-1, 1:-1]))
    o.step()
    after = float(np.sum(np.array(o.field("s"))[1:-1, 1:-1, 1:-1]))
    o.teardown()
    assert abs(after - before) <= 1e-12 * abs(before)


@pytest.mark.parametrize("kind", ["DIRICHLET", "NEUMANN"])
def test_wall_profiles_are_steady(tmp_path, kind):
    if kind == "DIRICHLET":
        bc = {"B": ("DIRICHLET", 1.0), "T": ("DIRICHLET", 0.0)}
        prof = lambda z: 1.0 - z / L
    else:
        q = 0.25
        bc = {"B": ("NEUMANN", q), "T": ("NEUMANN", q)}
        prof = lambda z: 0.3 + q * z
    o, _ = periodic_case(tmp_path, D=0.1, fixed_dt=0.1, bc=bc, walls_z=True)
    z = centres(L)
    f = prof(z)[None, None, :] * np.ones((L + 2, L + 2, 1))
    set_s(o, f)
    for _ in range(3):
        o.step()
    s = np.array(o.field("s"))[1:-1, 1:-1, 1:-1]
    o.teardown()
    assert np.max(np.abs(s - f[1:-1, 1:-1, 1:-1])) <= 1e-14


def test_boussinesq_uniform_scalar(tmp_path):
    g, beta, sref, s1, dt = -3.0, 2.0, 0.25, 0.75, 0.01
    o, _ = periodic_case(tmp_path, D=0.0, g=g, beta=beta, sref=sref,
                         fixed_dt=dt)
    set_s(o, np.full((L + 2, L + 2, L + 2), s1))
    o.step()
    w = np.array(o.field("w"))[1:-1, 1:-1, 1:-1]
    u = np.array(o.field("u"))[1:-1, 1:-1, 1:-1]
    o.teardown()
    want = dt * (-g * beta * (s1 - sref))
    assert np.allclose(w, want, rtol=1e-13, atol=0)
    assert np.max(np.abs(u)) == 0.0


def test_time_step_follows_diffusivity(tmp_path):
    D = 0.2
    o, kw = periodic_case(tmp_path, D=D, fixed_dt=None)
    dt = o.scalar("dt")
    o.teardown()
    assert dt == pytest.approx(kw["CFL"] / (3 * 2 * D), rel=1e-15)


# ------------------------------ particle heat -------------------------------
def heat_case(tmp, a=6.0, n=32, srs_r=1.2, s_p=0.3, s_order=2, **scal):
    kw, parts = bbconfig.sediment(n=(n, n, n), nparts=1, a=a, seed=1)
    kw["g"] = ((0.0, 0.0), (0.0, 0.0), (0.0, 0.0))
    per = ("PERIODIC",)
    kw["scalar"] = dict(s_D=scal.get("s_D", 0.02), s_k=scal.get("s_k", 1.0),
                        s_beta=0.0, s_ref=0.0, init_s=(0.0, 0.0),
                        bc_s={f: per for f in "WESNBT"})
    c = n / 2.0
    parts[0].update(x=c + 0.2, y=c - 0.3, z=c + 0.1, s=s_p, update=1,
                    cp=scal.get("cp", 2.0), srs_r=srs_r, s_order=s_order)
    d = str(tmp)
    bbconfig.write_case(d, **kw)
    bbconfig.write_parts(d, parts)
    o = orc.Oracle(d)
    o.setup()
    return o, kw, parts[0]


def cell_grid(n):
    x = centres(n)
    return np.meshgrid(x, x, x, indexing="ij")


def test_heat_quadrature_linear_field(tmp_path):
    """s - s_p = c + g . (x - x_p) on the node sphere r = srs: the
    quadrature (exact to degree 7) and the interpolation (exact for a linear
    field) give a00 = c sqrt(4 pi) / (1 - a/srs),
    a10 = g_z srs sqrt(4 pi / 3) / (srs/a - (a/srs)^2),
    Re a11 = -g_x srs (4 pi / 3) N11 / (srs/a - (a/srs)^2),
    Im a11 = +g_y srs (4 pi / 3) N11 / (...), and 0 for m < 0 (the
    reference's Legendre table returns 0 there)."""
    a, srs_r, s_p = 6.0, 1.2, 0.3
    o, kw, P = heat_case(tmp_path, a=a, srs_r=srs_r, s_p=s_p)
    X, Y, Z = cell_grid(32)
    c, gx, gy, gz = 0.7, 0.011, -0.007, 0.013
    o.field("s")[...] = s_p + c + gx * (X - P["x"]) + gy * (Y - P["y"]) \
        + gz * (Z - P["z"])
    o.call("orc_scalar_lamb_once")
    s, q, are, aim = o.part_scalar(0)
    o.teardown()
    srs = srs_r * a
    A0 = 1.0 - a / srs
    A1 = srs / a - (a / srs) ** 2
    N11 = math.sqrt(3.0 / (8.0 * math.pi))
    want = np.zeros(25)
    want_im = np.zeros(25)
    want[0] = c * math.sqrt(4 * math.pi) / A0
    want[2] = gz * srs * math.sqrt(4 * math.pi / 3) / A1
    want[3] = -gx * srs * (4 * math.pi / 3) * N11 / A1
    want_im[3] = gy * srs * (4 * math.pi / 3) * N11 / A1
    assert np.allclose(are[:9], want[:9], rtol=0, atol=1e-10), (are[:9], want[:9])
    assert np.allclose(aim[:9], want_im[:9], rtol=0, atol=1e-10), aim[:9]
    assert are[1] == 0.0 and aim[1] == 0.0         # m = -1
    assert q == pytest.approx(2 * math.sqrt(math.pi) * a * want[0], rel=1e-11)


def test_heat_flux_of_conducting_sphere(tmp_path):
    """Steady conduction around a sphere held at s_p in a medium at s_inf:
    s = s_inf + (s_p - s_inf) a / r, heat flux q = 4 pi a (s_inf - s_p)
    (within the interpolation error of the 1/r field, 2 %)."""
    a, s_p, s_inf = 10.0, 1.0, 0.2
    o, kw, P = heat_case(tmp_path, a=a, n=48, s_p=s_p, s_order=0)
    X, Y, Z = cell_grid(48)
    r = np.sqrt((X - P["x"]) ** 2 + (Y - P["y"]) ** 2 + (Z - P["z"]) ** 2)
    o.field("s")[...] = s_inf + (s_p - s_inf) * a / np.maximum(r, 1e-9)
    o.call("orc_scalar_lamb_once")
    s, q, are, aim = o.part_scalar(0)
    o.teardown()
    assert q == pytest.approx(4 * math.pi * a * (s_inf - s_p), rel=2e-2)


def test_heat_series_in_particle_cells(tmp_path):
    """The analytic scalar written into particle cells: s_p inside, and in
    shell cells s_p + (1 - a/r) a00 N00 + (r/a - (a/r)^2) X_1 with
    X_1 = N10 a10 cos(theta) - N11 (Re a11 sin(theta) cos(phi)
    - Im a11 sin(theta) sin(phi)); the m = -1 coefficient contributes
    nothing."""
    a, s_p = 6.0, 0.3
    o, kw, P = heat_case(tmp_path, a=a, s_p=s_p, s_order=1)
    are = np.zeros(25)
    aim = np.zeros(25)
    are[0], are[1], aim[1] = 0.4, 5.0, -3.0       # a00, a1,-1 (ignored)
    are[2], are[3], aim[3] = -0.25, 0.15, 0.05    # a10, a11
    o.set_part_scalar(0, s_p, 0.0, are, aim)
    o.field("s")[...] = -7.0
    o.call("orc_scalar_part_bc_s")
    s = np.array(o.field("s"))
    ph = np.array(o.field("phase"))
    sh = np.array(o.field("phase_shell"))
    o.teardown()
    X, Y, Z = cell_grid(32)
    dx, dy, dz = X - P["x"], Y - P["y"], Z - P["z"]
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    N00 = math.sqrt(1 / (4 * math.pi))
    N10 = math.sqrt(3 / (4 * math.pi))
    N11 = math.sqrt(3 / (8 * math.pi))
    X1 = N10 * are[2] * dz / r - N11 * (are[3] * dx / r - aim[3] * dy / r)
    want = s_p + (1 - a / r) * are[0] * N00 + (r / a - (a / r) ** 2) * X1
    inner = (ph >= 0) & (sh > 0)
    shell = (ph >= 0) & (sh < 1)
    inner[[0, -1], :, :] = inner[:, [0, -1], :] = inner[:, :, [0, -1]] = False
    shell[[0, -1], :, :] = shell[:, [0, -1], :] = shell[:, :, [0, -1]] = False
    fluid = ph < 0
    assert shell.sum() > 0 and inner.sum() > 0
    assert np.allclose(s[shell], want[shell], rtol=0, atol=1e-13)
    assert np.all(s[inner] == s_p)
    assert np.all(s[fluid] == -7.0)


def test_particle_scalar_update(tmp_path):
    """After the scalar iteration: s_p += q s_k dt / (rho (4/3) pi a^3 cp)."""
    a, s_p = 6.0, 0.3
    s_k, cp = 1.7, 2.5
    o, kw, P = heat_case(tmp_path, a=a, s_p=s_p, s_k=s_k, cp=cp)
    X, Y, Z = cell_grid(32)
    o.field("s")[...] = 0.9
    o.field("s0")[...] = 0.9
    o.step()
    s_new, q, _, _ = o.part_scalar(0)
    dt = o.scalar("dt0")
    o.teardown()
    m = 2.0 * 4.0 / 3.0 * math.pi * a ** 3
    assert q != 0.0
    assert s_new - s_p == pytest.approx(q * s_k * dt / m / cp, rel=1e-12)
