"""ctypes front end of the FP64 host oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Arrays are returned as [i, j, k]
numpy views of the oracle's x-fastest ghosted storage.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
GCC, GFX, GFY, GFZ = 0, 1, 2, 3

_GRID = {"p": GCC, "p0": GCC, "phi": GCC, "rhs": GCC, "invM": GCC,
         "phase": GCC, "phase_shell": GCC}
for _c, _g in (("u", GFX), ("v", GFY), ("w", GFZ)):
    for _pre in ("", "0", "_star"):
        _GRID[_c + _pre] = _g
    for _pre in ("conv_", "conv0_", "diff_", "diff0_"):
        _GRID[_pre + _c] = _g
    _GRID["flag_" + _c] = _g
_GRID.update({"f_x": GFX, "f_y": GFY, "f_z": GFZ, "s": GCC, "s0": GCC,
              "s_conv": GCC, "s_diff": GCC})

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_setup.argtypes = [ctypes.c_char_p]
        L.orc_setup.restype = ctypes.c_int
        L.orc_step.restype = ctypes.c_int
        L.orc_pcg.restype = ctypes.c_int
        L.orc_field.argtypes = [ctypes.c_char_p]
        L.orc_field.restype = ctypes.POINTER(ctypes.c_double)
        L.orc_ifield.argtypes = [ctypes.c_char_p]
        L.orc_ifield.restype = ctypes.POINTER(ctypes.c_int)
        L.orc_scalar.argtypes = [ctypes.c_char_p]
        L.orc_scalar.restype = ctypes.c_double
        L.orc_set_scalar.argtypes = [ctypes.c_char_p, ctypes.c_double]
        L.orc_canon_sum.argtypes = [ctypes.POINTER(ctypes.c_double),
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_canon_sum.restype = ctypes.c_double
        L.orc_canon_linear.argtypes = [ctypes.POINTER(ctypes.c_double),
                                       ctypes.c_long]
        L.orc_canon_linear.restype = ctypes.c_double
        L.orc_bc_p.argtypes = [ctypes.POINTER(ctypes.c_double)]
        L.orc_exchange.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.orc_apply_A.argtypes = [ctypes.POINTER(ctypes.c_double)] * 2
        L.orc_dims.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 3
        D = ctypes.POINTER(ctypes.c_double)
        L.orc_nparts.restype = ctypes.c_int
        L.orc_part_state.argtypes = [ctypes.c_int, D]
        L.orc_part_state.restype = ctypes.c_int
        L.orc_part_set_coeffs.argtypes = [ctypes.c_int, D]
        L.orc_part_set_coeffs.restype = ctypes.c_int
        L.orc_lamb_eval.argtypes = [ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, D]
        L.orc_lamb_once.restype = ctypes.c_double
        L.orc_part_set_motion.argtypes = [ctypes.c_int, D]
        L.orc_part_set_motion.restype = ctypes.c_int
        L.orc_part_contacts.argtypes = [ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int), D,
                                        ctypes.POINTER(ctypes.c_int)]
        L.orc_part_contacts.restype = ctypes.c_int
        L.orc_lebedev.argtypes = [D, D, D]
        L.orc_lamb_eval_p.argtypes = [ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double]
        L.orc_lamb_eval_p.restype = ctypes.c_double
        L.orc_part_nodes.argtypes = [ctypes.c_int, D, D, D, D,
                                     ctypes.POINTER(ctypes.c_int)]
        L.orc_lamb_from_nodes.argtypes = [ctypes.c_int, D, D, D, D]
        _lib = L
    return _lib


class Oracle:
    def __init__(self, case_dir):
        self.lib = load()
        self.dir = os.path.abspath(case_dir)

    def __enter__(self):
        self.setup()
        return self

    def __exit__(self, *exc):
        self.teardown()

    def setup(self):
        rc = self.lib.orc_setup(self.dir.encode())
        if rc != 0:
            raise RuntimeError(f"orc_setup failed ({rc})")

    def step(self, nsteps=1):
        q = 0
        for _ in range(nsteps):
            q = self.lib.orc_step()
            if q < 0:
                raise RuntimeError("oracle PCG failed")
        return q

    def dims(self, grid):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.lib.orc_dims(grid, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def field(self, name):
        grid = _GRID[name]
        ni, nj, nk = self.dims(grid)
        size = (ni + 2) * (nj + 2) * (nk + 2)
        if name.startswith("flag_") or name.startswith("phase"):
            ptr = self.lib.orc_ifield(name.encode())
        else:
            ptr = self.lib.orc_field(name.encode())
        flat = np.ctypeslib.as_array(ptr, shape=(size,))
        return flat.reshape(nk + 2, nj + 2, ni + 2).transpose(2, 1, 0)

    def raw(self, name):
        return self.lib.orc_field(name.encode())

    def scalar(self, name):
        return self.lib.orc_scalar(name.encode())

    def set_scalar(self, name, v):
        self.lib.orc_set_scalar(name.encode(), float(v))

    def call(self, name, *args):
        return getattr(self.lib, name)(*args)

    def teardown(self):
        self.lib.orc_teardown()

    # -------------------------------- particles --------------------------
    def nparts(self):
        return self.lib.orc_nparts()

    def part_state(self, n):
        """Dict of one particle's state (names as PART_FIELDS)."""
        out = np.zeros(126)
        self.lib.orc_part_state(n, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)))
        return unpack_state(out)

    def set_coeffs(self, n, coeffs):
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
        assert c.shape == (90,)
        self.lib.orc_part_set_coeffs(n, c.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)))

    def set_motion(self, n, pos, vel, om, vel0=None, om0=None):
        """Position, velocity and rotation of particle n (the previous
        velocity and rotation default to the current ones)."""
        vel0 = vel if vel0 is None else vel0
        om0 = om if om0 is None else om0
        m = np.ascontiguousarray(np.concatenate(
            [pos, vel, om, vel0, om0]).astype(np.float64))
        assert m.shape == (15,)
        self.lib.orc_part_set_motion(n, m.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)))

    def collide(self):
        """Pair, spring and wall forces on the current particle state."""
        self.lib.orc_collide_once()

    def part_scalar(self, n):
        """(s, q, anm_re[25], anm_im[25]) of particle n."""
        out = np.zeros(52)
        self.lib.orc_part_scalar(n, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)))
        return out[0], out[1], out[2:27].copy(), out[27:52].copy()

    def set_part_scalar(self, n, s, q, anm_re, anm_im):
        v = np.zeros(52)
        v[0], v[1] = s, q
        v[2:27] = anm_re
        v[27:52] = anm_im
        self.lib.orc_part_set_scalar(n, v.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)))

    def move_sub(self):
        """Collisions, velocity update, collisions (one Lamb sub-step)."""
        self.lib.orc_move_parts_sub()

    def move_b(self):
        """End-of-step update: positions, body axes, previous values."""
        self.lib.orc_move_parts_b()

    def contacts(self, n):
        """(iSt list, St list, ncoll_part, ncoll_wall) of particle n."""
        i = np.zeros(12, dtype=np.int32)
        s = np.zeros(12)
        c = np.zeros(2, dtype=np.int32)
        self.lib.orc_part_contacts(
            n, i.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return list(i), list(s), int(c[0]), int(c[1])

    def lamb_eval_p(self, n, x, y, z):
        """Lamb pressure series of particle n at (x, y, z) from its centre
        (no rest-frame terms)."""
        return self.lib.orc_lamb_eval_p(n, x, y, z)

    def nodes(self, n):
        """Rest-frame (p, u_r, u_theta, u_phi, code) at the 26 nodes of
        particle n on the current fields."""
        P = ctypes.POINTER(ctypes.c_double)
        v = [np.zeros(26) for _ in range(4)]
        code = np.zeros(26, dtype=np.int32)
        self.lib.orc_part_nodes(n, *[a.ctypes.data_as(P) for a in v],
                                code.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return v[0], v[1], v[2], v[3], code

    def lamb_from_nodes(self, n, pp, ur, ut, up):
        """Quadrature, coefficients and forces of particle n from given
        rest-frame node values."""
        P = ctypes.POINTER(ctypes.c_double)
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float64))
                for a in (pp, ur, ut, up)]
        self.lib.orc_lamb_from_nodes(n, *[a.ctypes.data_as(P) for a in arrs])

    def part_bc_once(self):
        """Particle boundary values (faces next to particles, p inside)
        from the current coefficients."""
        self.lib.orc_part_bc_once()

    def lamb_eval(self, n, x, y, z):
        U = np.zeros(4)
        self.lib.orc_lamb_eval(n, x, y, z,
                               U.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return U


PART_SCALARS = ("x y z u v w ox oy oz udot vdot wdot oxdot oydot ozdot Fx Fy "
                "Fz Lx Ly Lz iFx iFy iFz iLx iLy iLz axx axy axz ayx ayy ayz "
                "azx azy azz").split()
PART_COEFFS = ("pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re",
               "chinm_im")


def unpack_state(v):
    d = {k: float(v[i]) for i, k in enumerate(PART_SCALARS)}
    for i, k in enumerate(PART_COEFFS):
        d[k] = np.array(v[36 + 15 * i:36 + 15 * (i + 1)])
    return d


def lebedev():
    """Nodes (theta, phi) and weights of the 26-node rule."""
    L = load()
    t, p, w = np.zeros(26), np.zeros(26), np.zeros(26)
    P = ctypes.POINTER(ctypes.c_double)
    L.orc_lebedev(t.ctypes.data_as(P), p.ctypes.data_as(P), w.ctypes.data_as(P))
    return t, p, w


def canon_sum(terms):
    """Canonical sum of an (nx, ny, nz) array of per-cell terms."""
    L = load()
    t = np.ascontiguousarray(np.asarray(terms, dtype=np.float64).transpose(2, 1, 0))
    nz, ny, nx = t.shape
    return L.orc_canon_sum(t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                           nx, ny, nz)
