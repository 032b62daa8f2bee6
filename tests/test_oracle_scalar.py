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
