"""ctypes front end of libbluebottle.so (the gfx950 flow path).

The library keeps the reference's global-variable C ABI; this module only
loads it, runs the driver phases (bluebottle_setup / bluebottle_step) from a
case directory, and exposes the host arrays (reference layouts) as numpy
views indexed [i, j, k] over the ghosted grid.

There is no fallback: if libbluebottle.so has not been built, loading fails
with an error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbluebottle.so")

GCC, GFX, GFY, GFZ = 0, 1, 2, 3

_FIELDS = {
    "p": GCC, "p0": GCC, "phi": GCC,
    "u": GFX, "u0": GFX, "u_star": GFX, "conv_u": GFX, "conv0_u": GFX,
    "diff_u": GFX, "diff0_u": GFX, "f_x": GFX,
    "v": GFY, "v0": GFY, "v_star": GFY, "conv_v": GFY, "conv0_v": GFY,
    "diff_v": GFY, "diff0_v": GFY, "f_y": GFY,
    "w": GFZ, "w0": GFZ, "w_star": GFZ, "conv_w": GFZ, "conv0_w": GFZ,
    "diff_w": GFZ, "diff0_w": GFZ, "f_z": GFZ,
    "s": GCC, "s0": GCC, "s_conv": GCC, "s_conv0": GCC, "s_diff": GCC,
    "s_diff0": GCC,
}
_IFIELDS = {"flag_u": GFX, "flag_v": GFY, "flag_w": GFZ, "phase": GCC,
            "phase_shell": GCC}


class PPStats(ctypes.Structure):
    _fields_ = [("niter", ctypes.c_int), ("nouter", ctypes.c_int),
                ("resid", ctypes.c_double), ("true_resid", ctypes.c_double),
                ("sp_rhs", ctypes.c_double), ("etime", ctypes.c_double)]


def _part_fields():
    d, i = ctypes.c_double, ctypes.c_int
    f = [("N", i), ("r", d)]
    f += [(n, d) for n in "x y z u v w u0 v0 w0 udot vdot wdot udot0 vdot0 "
          "wdot0 axx axy axz ayx ayy ayz azx azy azz ox oy oz ox0 oy0 oz0 "
          "oxdot oydot ozdot oxdot0 oydot0 ozdot0 Fx Fy Fz Lx Ly Lz aFx aFy "
          "aFz aLx aLy aLz kFx kFy kFz iFx iFy iFz iLx iLy iLz".split()]
    f += [("nodes", i * 26), ("rho", d), ("E", d), ("sigma", d),
          ("order", i), ("rs", d), ("ncoeff", i), ("spring_k", d),
          ("spring_x", d), ("spring_y", d), ("spring_z", d), ("spring_l", d),
          ("translating", i), ("rotating", i), ("St", d * 12),
          ("iSt", i * 12), ("e_dry", d), ("coeff_fric", d)]
    for n in ("pnm_re", "pnm_im", "phinm_re", "phinm_im", "chinm_re",
              "chinm_im", "pnm_re0", "pnm_im0", "phinm_re0", "phinm_im0",
              "chinm_re0", "chinm_im0"):
        f.append((n, d * 15))
    f += [("ncoll_part", i), ("ncoll_wall", i), ("s", d), ("update", i),
          ("srs", d), ("q", d), ("cp", d), ("sorder", i), ("sncoeff", i),
          ("anm_re", d * 25), ("anm_im", d * 25), ("anm_re0", d * 25),
          ("anm_im0", d * 25)]
    return f


class PartStruct(ctypes.Structure):
    """The reference's part_struct, member for member."""
    _fields_ = _part_fields()


_lib = None


def load():
    """Load libbluebottle.so once; raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name in ("bluebottle_setup", "bluebottle_step"):
            getattr(_lib, name).restype = ctypes.c_int
        _lib.cuda_device_count.restype = ctypes.c_int
    return _lib


def grid_dims(grid, n):
    nx, ny, nz = n
    ni, nj, nk = nx, ny, nz
    if grid == GFX:
        ni += 1
    elif grid == GFY:
        nj += 1
    elif grid == GFZ:
        nk += 1
    return ni, nj, nk


def host_to_ijk(flat, grid, n):
    """View a reference-layout host array as [i, j, k] (ghosts included)."""
    ni, nj, nk = grid_dims(grid, n)
    inb, jnb, knb = ni + 2, nj + 2, nk + 2
    if grid in (GCC, GFZ):
        return flat.reshape(knb, jnb, inb).transpose(2, 1, 0)
    if grid == GFX:                      # j + k*jnb + i*jnb*knb
        return flat.reshape(inb, knb, jnb).transpose(0, 2, 1)
    return flat.reshape(jnb, inb, knb).transpose(1, 0, 2)  # GFY


class Simulation:
    """One case directory driven through the library's C ABI."""

    def __init__(self, case_dir, n, verbose=False, record=False):
        self.lib = load()
        self.dir = os.path.abspath(case_dir)
        self.n = tuple(int(x) for x in n)
        self._cwd = None
        ctypes.c_int.in_dll(self.lib, "bb_verbose").value = 1 if verbose else 0
        ctypes.c_int.in_dll(self.lib, "recorder_enabled").value = \
            1 if record else 0

    def __enter__(self):
        self.setup()
        return self

    def __exit__(self, *exc):
        self.teardown()

    def setup(self):
        self._cwd = os.getcwd()
        os.chdir(self.dir)
        try:
            rc = self.lib.bluebottle_setup()
        finally:
            os.chdir(self._cwd)
        if rc != 0:
            raise RuntimeError(f"bluebottle_setup failed ({rc})")

    def step(self, nsteps=1):
        cwd = os.getcwd()
        os.chdir(self.dir)
        try:
            for _ in range(nsteps):
                self.lib.bluebottle_step()
        finally:
            os.chdir(cwd)

    def sync(self):
        self.lib.cuda_block()

    def pull(self):
        self.lib.cuda_dom_pull()
        if ctypes.c_int.in_dll(self.lib, "SCALAR").value >= 1:
            self.lib.cuda_scalar_pull()

    def push(self):
        self.lib.cuda_dom_push()

    def field(self, name):
        """Host array `name` as an [i, j, k] numpy view (call pull() first)."""
        if name in _FIELDS:
            grid, ct = _FIELDS[name], ctypes.c_double
        else:
            grid, ct = _IFIELDS[name], ctypes.c_int
        ni, nj, nk = grid_dims(grid, self.n)
        size = (ni + 2) * (nj + 2) * (nk + 2)
        ptr = ctypes.POINTER(ct).in_dll(self.lib, name)
        flat = np.ctypeslib.as_array(ptr, shape=(size,))
        return host_to_ijk(flat, grid, self.n)

    def parts(self):
        """Device particle array copied to the host, as a list of dicts."""
        n = ctypes.c_int.in_dll(self.lib, "nparts").value
        if n == 0:
            return []
        self.lib.cuda_part_pull()
        ptr = ctypes.POINTER(PartStruct).in_dll(self.lib, "parts")
        out = []
        for k in range(n):
            P = ptr[k]
            d = {}
            for name, ct in PartStruct._fields_:
                v = getattr(P, name)
                d[name] = np.array(v[:]) if hasattr(v, "_length_") else v
            out.append(d)
        return out

    def lamb_stats(self):
        return (ctypes.c_int.in_dll(self.lib, "bb_lamb_iters").value,
                ctypes.c_double.in_dll(self.lib, "bb_lamb_err").value)

    def scalar(self, name, ctype=ctypes.c_double):
        return ctype.in_dll(self.lib, name).value

    def set_scalar(self, name, value, ctype=ctypes.c_double):
        ctype.in_dll(self.lib, name).value = value

    def pp_stats(self):
        s = PPStats.in_dll(self.lib, "pp_last")
        return {k: getattr(s, k) for k, _ in PPStats._fields_}

    def call(self, name, *args):
        return getattr(self.lib, name)(*args)

    def teardown(self):
        self.lib.bluebottle_teardown()
