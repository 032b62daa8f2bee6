"""The allocation-free plain pressure solve (cuda_PP_cg_noalloc) against
the library's own cuda_PP_cg: the same right-hand side (the u* of one
timestep) solved by both, twice in a row on the same device, the second
time on a smaller grid, must give the same bits in phi, the same iteration
count and the same residual."""
import ctypes

import numpy as np
import pytest

import bbconfig
import bbhip
from helpers import bitwise_equal, interior

pytestmark = pytest.mark.gpu


def solve_both(case_dir, kw):
    bbconfig.write_case(case_dir, **kw)
    with bbhip.Simulation(case_dir, kw["n"]) as s:
        s.step()                          # u* of a real timestep
        s.lib.cuda_PP_cg()
        s.pull()
        want = np.array(s.field("phi"))
        want_stats = s.pp_stats()
        s.lib.cuda_PP_cg_noalloc.restype = ctypes.c_int
        rc = s.lib.cuda_PP_cg_noalloc()
        assert rc == 0
        s.pull()
        got = np.array(s.field("phi"))
        got_stats = s.pp_stats()
    print("iterations", want_stats["niter"], got_stats["niter"])
    assert want_stats["niter"] > 0
    assert got_stats["niter"] == want_stats["niter"]
    assert got_stats["resid"] == want_stats["resid"]
    assert bitwise_equal(interior(got), interior(want))


def test_noalloc_matches_cg_twice(case_dir):
    solve_both(case_dir, bbconfig.channel(n=(96, 48, 40)))
    solve_both(case_dir, bbconfig.channel(n=(64, 32, 24)))


def test_noalloc_rejects_large_grid(case_dir):
    """A grid whose scratch exceeds the static array is refused with an
    error code before any work."""
    kw = bbconfig.channel(n=(256, 256, 160))
    bbconfig.write_case(case_dir, **kw)
    with bbhip.Simulation(case_dir, kw["n"]) as s:
        s.lib.cuda_PP_cg_noalloc.restype = ctypes.c_int
        assert s.lib.cuda_PP_cg_noalloc() == 2
