"""Pins of the oracle's particle / Physalis restatement (CPU only).

- the 26-node Lebedev rule: weights against their closed forms
  (4 pi/21, 16 pi/105, 9 pi/70) and exactness on products of spherical
  harmonics of total degree <= 7 (scipy's sph_harm_y, independent of the
  oracle's Legendre table);
- Stokes flow: imposing the analytic field of a sphere translating (or
  rotating) in unbounded fluid, the quadrature must return the Stokes drag
  -6 pi mu a U (torque -8 pi mu a^3 Omega) to the accuracy of the grid
  interpolation (measured 1.5-1.6 % at a = 6 dx, bound 5 %);
- cage geometry: phase against a brute-force nearest-image distance test,
  phase_shell and the face flags against their definitions in numpy;
- the particle SpMV on a vector that is zero in particle cells equals the
  plain 7-point operator in fluid cells and is zero in particle cells (the
  identity the GPU path relies on when it masks particle cells out of the
  PCG);
- determinism of a particle run.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy.special import sph_harm_y

import bbconfig
import orc


def write(tmp, kw, parts):
    d = str(tmp)
    bbconfig.write_case(d, **kw)
    bbconfig.write_parts(d, parts)
    return d


def test_lebedev_weights_closed_form():
    t, p, w = orc.lebedev()
    assert np.allclose(w[:6], 4 * math.pi / 21, rtol=1e-14)
    assert np.allclose(w[6:18], 16 * math.pi / 105, rtol=1e-14)
    assert np.allclose(w[18:], 9 * math.pi / 70, rtol=1e-14)
    assert abs(w.sum() - 4 * math.pi) < 1e-13
    # nodes: 6 axis points, 12 edge midpoints, 8 cube corners on the sphere
    xyz = np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], 1)
    ax = np.abs(xyz)
    assert np.allclose(np.sort(ax[:6], 1), [[0, 0, 1]] * 6, atol=1e-9)
    assert np.allclose(np.sort(ax[6:18], 1),
                       [[0, 2 ** -0.5, 2 ** -0.5]] * 12, atol=1e-9)
    assert np.allclose(ax[18:], 3 ** -0.5, atol=1e-9)


def test_lebedev_exact_to_degree_seven():
    t, p, w = orc.lebedev()
    worst = 0.0
    for n1 in range(0, 5):
        for n2 in range(0, 8 - n1):
            for m1 in range(-n1, n1 + 1):
                for m2 in range(-n2, n2 + 1):
                    y1 = sph_harm_y(n1, m1, t, p)
                    y2 = sph_harm_y(n2, m2, t, p)
                    q = np.sum(w * y1 * np.conj(y2))
                    exact = 1.0 if (n1 == n2 and m1 == m2) else 0.0
                    worst = max(worst, abs(q - exact))
    # the pole nodes sit 1e-10 outside [0, pi] (DIV_ST)
    assert worst < 1e-9, worst


def stokes_case(tmp, U=(0.0, 0.0, 0.0), Om=(0.0, 0.0, 0.0), n=48, a=6.0):
    kw, parts = bbconfig.sediment(n=(n, n, n), nparts=1, a=a, seed=1)
    kw["g"] = ((0.0, 0.0), (0.0, 0.0), (0.0, 0.0))
    c = n / 2.0
    parts[0].update(x=c, y=c, z=c, u=U[0], v=U[1], w=U[2])
    d = write(tmp, kw, parts)
    o = orc.Oracle(d)
    o.setup()
    nu = kw["nu"]
    mu = nu * kw["rho_f"]
    U = np.array(U)
    Om = np.array(Om)

    def field(X, Y, Z):
        x = np.stack([X - c, Y - c, Z - c], -1)
        r = np.linalg.norm(x, axis=-1)[..., None]
        r = np.maximum(r, 1e-12)
        udx = (x @ U)[..., None]
        vel = (0.75 * a * (U / r + udx * x / r ** 3)
               + 0.25 * a ** 3 * (U / r ** 3 - 3 * udx * x / r ** 5))
        # a sphere turning at Om in fluid at rest, seen from the sphere:
        # rotlet minus the rigid rotation (no slip at r = a)
        vel += np.cross(Om, x) * (a ** 3 / r ** 3 - 1.0)
        p = 1.5 * mu * a * udx[..., 0] / r[..., 0] ** 3
        return vel, p

    h = 1.0
    idx = np.arange(n + 3)
    cc = (np.arange(n + 2) - 0.5) * h
    fc = (idx - 1.0) * h
    # p (cell centres)
    X, Y, Z = np.meshgrid(cc, cc, cc, indexing="ij")
    _, p = field(X, Y, Z)
    o.field("p")[...] = p
    for comp, name in enumerate("uvw"):
        axes = [cc, cc, cc]
        axes[comp] = fc
        X, Y, Z = np.meshgrid(*axes, indexing="ij")
        vel, _ = field(X, Y, Z)
        o.field(name)[...] = vel[..., comp]
    return o, mu, a


@pytest.mark.parametrize("U", [(0.0, 0.0, 0.01), (0.01, 0.0, 0.0),
                               (0.0, 0.007, 0.004)])
def test_stokes_drag(tmp_path, U):
    o, mu, a = stokes_case(tmp_path, U=U)
    try:
        o.lib.orc_lamb_once()
        st = o.part_state(0)
        F = np.array([st["Fx"], st["Fy"], st["Fz"]])
        F0 = -6 * math.pi * mu * a * np.array(U)
        err = np.linalg.norm(F - F0) / np.linalg.norm(F0)
        print("F", F, "Stokes", F0, "rel err", err)
        assert err < 0.05
    finally:
        o.teardown()


def test_rotlet_torque(tmp_path):
    Om = (0.0, 0.0, 0.002)
    o, mu, a = stokes_case(tmp_path, Om=Om)
    try:
        o.lib.orc_lamb_once()
        st = o.part_state(0)
        L = np.array([st["Lx"], st["Ly"], st["Lz"]])
        L0 = -8 * math.pi * mu * a ** 3 * np.array(Om)
        err = np.linalg.norm(L - L0) / np.linalg.norm(L0)
        print("L", L, "expected", L0, "rel err", err)
        assert err < 0.05
        F = np.array([st["Fx"], st["Fy"], st["Fz"]])
        # a pure rotation gives no net force, to the same accuracy
        assert np.linalg.norm(F) * a < 0.05 * np.linalg.norm(L0)
    finally:
        o.teardown()


def test_cage_geometry_brute_force(tmp_path):
    kw, parts = bbconfig.sediment(n=(32, 24, 28), nparts=3, a=4.0, seed=11)
    kw["hi"] = (32.0, 24.0, 28.0)
    parts[0].update(x=0.3, y=23.9, z=14.0)   # across two periodic faces
    d = write(tmp_path, kw, parts)
    with orc.Oracle(d) as o:
        ph = np.array(o.field("phase"))
        sh = np.array(o.field("phase_shell"))
        fu, fv, fw = (np.array(o.field(f)) for f in ("flag_u", "flag_v",
                                                      "flag_w"))
    L = np.array([32.0, 24.0, 28.0])
    nx, ny, nz = 32, 24, 28
    ii, jj, kk = np.meshgrid(np.arange(nx + 2), np.arange(ny + 2),
                             np.arange(nz + 2), indexing="ij")
    ctr = np.stack([ii - 0.5, jj - 0.5, kk - 0.5], -1)
    ref = -np.ones(ph.shape, dtype=int)
    for n, p in enumerate(parts):
        dd = ctr - np.array([p["x"], p["y"], p["z"]])
        dd -= L * np.round(dd / L)
        inside = np.linalg.norm(dd, axis=-1) < p["r"]
        ref[inside] = n
    assert np.array_equal(ph, ref)
    # shell: particle cells with a 6-neighbour centre outside that sphere
    refsh = np.ones(ph.shape, dtype=int)
    for n, p in enumerate(parts):
        for off in np.eye(3):
            for s in (1, -1):
                dd = ctr + s * off - np.array([p["x"], p["y"], p["z"]])
                dd -= L * np.round(dd / L)
                out = np.linalg.norm(dd, axis=-1) >= p["r"]
                refsh[(ph == n) & out] = 0
    assert np.array_equal(sh, refsh)
    # flags on interior-plane faces
    def flag(p0, p1, s0, s1):
        return 1 - 2 * (((p0 < 0) & (p1 > -1)) | ((p0 > -1) & (p1 < 0)) |
                        ((s1 < 1) & (s0 < 1)))
    assert np.array_equal(fu[1:nx + 2], flag(ph[:-1], ph[1:], sh[:-1], sh[1:]))
    assert np.array_equal(fv[:, 1:ny + 2], flag(ph[:, :-1], ph[:, 1:],
                                                sh[:, :-1], sh[:, 1:]))
    assert np.array_equal(fw[:, :, 1:nz + 2], flag(ph[:, :, :-1], ph[:, :, 1:],
                                                   sh[:, :, :-1], sh[:, :, 1:]))


def test_particle_operator_equals_masked_plain(tmp_path):
    kw, parts = bbconfig.sediment(n=(24, 24, 24), nparts=2, a=4.0, seed=5)
    d = write(tmp_path, kw, parts)
    rng = np.random.default_rng(3)
    with orc.Oracle(d) as o:
        ph = np.array(o.field("phase"))
        x = rng.standard_normal(ph.shape)
        x[ph > -1] = 0.0
        xr = np.ascontiguousarray(x.transpose(2, 1, 0))
        o.lib.orc_exchange(xr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0)
        x = xr.transpose(2, 1, 0).copy()
        x[ph > -1] = 0.0
        xr = np.ascontiguousarray(x.transpose(2, 1, 0))
        Ax = np.zeros_like(xr)
        P = ctypes.POINTER(ctypes.c_double)
        o.lib.orc_apply_A(xr.ctypes.data_as(P), Ax.ctypes.data_as(P))
        Ax = Ax.transpose(2, 1, 0)
    # plain 7-point operator (periodic box, dx = 1) in numpy
    xi = x
    K = 6 * xi[1:-1, 1:-1, 1:-1] - (
        xi[2:, 1:-1, 1:-1] + xi[:-2, 1:-1, 1:-1] + xi[1:-1, 2:, 1:-1] +
        xi[1:-1, :-2, 1:-1] + xi[1:-1, 1:-1, 2:] + xi[1:-1, 1:-1, :-2])
    pin = ph[1:-1, 1:-1, 1:-1] > -1
    Ai = Ax[1:-1, 1:-1, 1:-1]
    assert np.all(Ai[pin] == 0.0)
    assert np.allclose(Ai[~pin], K[~pin], rtol=0, atol=1e-12)


def test_particle_run_deterministic(tmp_path):
    kw, parts = bbconfig.sediment(n=(24, 24, 24), nparts=2, a=4.0, seed=9)
    d = write(tmp_path, kw, parts)
    res = []
    for _ in range(2):
        with orc.Oracle(d) as o:
            o.step(2)
            res.append((np.array(o.field("p")), o.part_state(0)["w"],
                        o.part_state(1)["Fz"]))
    assert np.array_equal(res[0][0].view(np.uint64), res[1][0].view(np.uint64))
    assert res[0][1] == res[1][1] and res[0][2] == res[1][2]
