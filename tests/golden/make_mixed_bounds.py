"""Extreme eigenvalues of the masked pressure operators of the particle
cases that compare a solve started elsewhere (learned guess, mixed
precision) with the FP64 oracle: tests/golden/mixed_bounds.json.

For each case the FP64 host oracle builds the particle cages (CPU); the
operator is (-A) over the fluid cells, periodic or Neumann per direction,
particle cells as fixed zeros (the fluid rows keep their full diagonal),
the system the PCG solves.  lam_min comes from ARPACK's shift-invert mode
with sigma = 0, the inverse applied by conjugate gradients to 1e-12;
lam_max from the Lanczos iteration directly.  The tests turn these into
bounds on the difference of two solves that pass the same stopping test
(helpers.solve_difference_bounds).

    python tests/golden/make_mixed_bounds.py
"""
import json
import os
import sys
import tempfile

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "bluebottle-3.0_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import bbconfig  # noqa: E402
import orc  # noqa: E402


def axis_op(n, h, periodic):
    """1-D -d2/dx2 on n cells: periodic wrap, or zero-flux walls."""
    main = np.full(n, 2.0)
    if not periodic:
        main[0] = main[-1] = 1.0
    L = sp.diags([-np.ones(n - 1), main, -np.ones(n - 1)], [-1, 0, 1],
                 format="lil")
    if periodic:
        L[0, n - 1] = -1.0
        L[n - 1, 0] = -1.0
    return L.tocsr() / (h * h)


def masked_operator(kw, fluid):
    n = kw["n"]
    h = [(kw["hi"][a] - kw["lo"][a]) / n[a] for a in range(3)]
    per = [kw["p_bc"][f][0] == "PERIODIC" for f in "WSB"]
    Ls = [axis_op(n[a], h[a], per[a]) for a in range(3)]
    # unknown (i, j, k) at i + nx (j + ny k): x fastest
    Ix, Iy, Iz = (sp.identity(n[a], format="csr") for a in range(3))
    A = (sp.kron(Iz, sp.kron(Iy, Ls[0])) + sp.kron(Iz, sp.kron(Ls[1], Ix))
         + sp.kron(Ls[2], sp.kron(Iy, Ix))).tocsr()
    idx = np.flatnonzero(fluid.transpose(2, 1, 0).ravel())
    return A[idx][:, idx].tocsc()


def extremes(S):
    def solve(v):
        x, info = sla.cg(S, v, rtol=1e-12, maxiter=20000)
        assert info == 0
        return x
    Op = sla.LinearOperator(S.shape, matvec=solve, dtype=np.float64)
    lmin = sla.eigsh(S, k=1, sigma=0.0, OPinv=Op, which="LM",
                     return_eigenvectors=False, tol=1e-8)[0]
    lmax = sla.eigsh(S, k=1, which="LA", return_eigenvectors=False,
                     tol=1e-6)[0]
    return float(lmin), float(lmax), float(S.diagonal().max())


def cases():
    import test_gpu_configs as C
    yield "rb_particles", C.rb_particles()
    yield "sediment32_4", bbconfig.sediment(n=(32, 32, 32), nparts=4, a=4.0,
                                            seed=11)
    yield "sediment32_3", bbconfig.sediment(n=(32, 32, 32), nparts=3, a=4.0,
                                            seed=7)


def main():
    out = {}
    for name, (kw, parts) in cases():
        d = tempfile.mkdtemp()
        bbconfig.write_case(d, **kw)
        bbconfig.write_parts(d, parts)
        with orc.Oracle(d) as o:
            phase = np.array(o.field("phase"))[1:-1, 1:-1, 1:-1]
        S = masked_operator(kw, phase < 0)
        lmin, lmax, dmax = extremes(S)
        out[name] = {"lam_min": lmin, "lam_max": lmax, "diag_max": dmax,
                     "fluid_cells": int(S.shape[0])}
        print(name, out[name], flush=True)
    with open(os.path.join(HERE, "mixed_bounds.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
