/*
 * Scalar transport (temperature or concentration) and its Boussinesq
 * coupling: src/scalar_kernel.cu:4-301, src/cuda_scalar.cu:149-257,
 * 949-957.
 *
 * - cuda_scalar_BC: ghost cells of a Dirichlet (2 s_b - s) or Neumann
 *   (s -+ h q) face over the face's interior rows; periodic faces come from
 *   cuda_exchange as for the flow.
 * - cuda_scalar_solve: one explicit Adams-Bashforth step of
 *   ds/dt + div(u s) = D lap s with the velocities of the start of the
 *   timestep, one thread per cell, the reference's arithmetic term by term.
 *   Reads s0 (7 points), u0, v0, w0 (2 faces each), conv0, diff0, s and
 *   phase; writes conv, diff, s: 84 B per cell, HBM-bound.
 * - cuda_compute_boussinesq: f += -g s_beta (s_face - s_ref) on every face.
 *   The u*, v*, w* kernels form it from the two cells beside the face when
 *   this call has followed the step's cuda_compute_forcing (they hold the
 *   constant part of f in registers; no f arrays exist on the device).
 * The particle heat exchange (cuda_scalar_part_BC, _part_fill, _lamb,
 * _lamb_err, _update_part) is with the particle kernels (bb_particle.hip).
 */
#include "bb_common.h"

/* Set by cuda_compute_boussinesq, cleared by cuda_compute_forcing. */
int bb_bous_on = 0;

extern "C" void cuda_compute_boussinesq(void)
{
  bb_bous_on = 1;
}

/* ---------------------------- boundary cells ---------------------------- */
/* face 0..5 = W E S N B T; kind 1 Dirichlet, 2 Neumann; h the grid spacing
 * across the face. */
__global__ void scalar_bc_kernel(DevDom D, double *a, int face, int kind,
  double val, double h)
{
  const Grid &C = D.cc;
  int dir = face / 2, hi = face & 1;
  int n1 = dir == 0 ? D.ny : D.nx;
  int n2 = dir == 2 ? D.ny : D.nz;
  int t1 = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int t2 = blockIdx.y + 1;
  if (t1 > n1 || t2 > n2) return;
  int e = dir == 0 ? D.nx : (dir == 1 ? D.ny : D.nz);
  long step = dir == 0 ? 1 : (dir == 1 ? C.sy : C.sz);
  long base = dir == 0 ? gidx(C, 0, t1, t2)
            : (dir == 1 ? gidx(C, t1, 0, t2) : gidx(C, t1, t2, 0));
  long g = base + (hi ? (long) (e + 1) * step : 0);
  long in = hi ? g - step : g + step;
  if (kind == DIRICHLET) a[g] = 2. * val - a[in];
  else a[g] = hi ? a[in] + h*val : a[in] - val*h;
}

extern "C" void cuda_scalar_BC(real *array)
{
  /* src/cuda_scalar.cu:157-237 */
  const int type[6] = {bc_s.sW, bc_s.sE, bc_s.sS, bc_s.sN, bc_s.sB, bc_s.sT};
  const double dv[6] = {bc_s.sWD, bc_s.sED, bc_s.sSD, bc_s.sND, bc_s.sBD,
                        bc_s.sTD};
  const double nv[6] = {bc_s.sWN, bc_s.sEN, bc_s.sSN, bc_s.sNN, bc_s.sBN,
                        bc_s.sTN};
  const double h[3] = {dom[0].dx, dom[0].dy, dom[0].dz};
  const DevDom &D = ddom;
  for (int f = 0; f < 6; f++) {
    if (type[f] != DIRICHLET && type[f] != NEUMANN) continue;
    int dir = f / 2;
    int n1 = dir == 0 ? D.ny : D.nx;
    int n2 = dir == 2 ? D.ny : D.nz;
    hipLaunchKernelGGL(scalar_bc_kernel, dim3((n1 + 255) / 256, n2),
      dim3(256), 0, bb_stream, D, array, f, type[f],
      type[f] == DIRICHLET ? dv[f] : nv[f], h[dir]);
  }
  BB_LAUNCH_CHECK();
}

/* ------------------------------ transport ------------------------------- */
struct ScalarArgs {
  double D, dt, dt0, ab0, ab, ddx, ddy, ddz;
};

__global__ void __launch_bounds__(256)
scalar_solve_kernel(DevDom Dm, ScalarArgs A, const int *__restrict__ phase,
  const double *__restrict__ s0, double *__restrict__ s,
  double *__restrict__ conv, double *__restrict__ diff,
  const double *__restrict__ conv0, const double *__restrict__ diff0,
  const double *__restrict__ u0, const double *__restrict__ v0,
  const double *__restrict__ w0)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > Dm.nx || j > Dm.ny) return;
  const Grid &C = Dm.cc, &X = Dm.fx, &Y = Dm.fy, &Z = Dm.fz;
  long c = gidx(C, i, j, k);
  double s111 = s0[c];
  double s011 = s0[c - 1], s211 = s0[c + 1];
  double s101 = s0[c - C.sy], s121 = s0[c + C.sy];
  double s110 = s0[c - C.sz], s112 = s0[c + C.sz];
  long cu = gidx(X, i, j, k), cv = gidx(Y, i, j, k), cw = gidx(Z, i, j, k);
  double uu0 = u0[cu], uu1 = u0[cu + 1];
  double vv0 = v0[cv], vv1 = v0[cv + Y.sy];
  double ww0 = w0[cw], ww1 = w0[cw + Z.sz];

  /* src/scalar_kernel.cu:270-298 */
  double dusdx = uu1 * 0.5 * (s211 + s111) - uu0 * 0.5 * (s111 + s011);
  dusdx *= A.ddx;
  double dvsdy = vv1 * 0.5 * (s121 + s111) - vv0 * 0.5 * (s111 + s101);
  dvsdy *= A.ddy;
  double dwsdz = ww1 * 0.5 * (s112 + s111) - ww0 * 0.5 * (s111 + s110);
  dwsdz *= A.ddz;
  double sconv = dusdx + dvsdy + dwsdz;
  conv[c] = sconv;
  double ddsddx = A.D * (s211 - 2.*s111 + s011) * A.ddx * A.ddx;
  double ddsddy = A.D * (s121 - 2.*s111 + s101) * A.ddy * A.ddy;
  double ddsddz = A.D * (s112 - 2.*s111 + s110) * A.ddz * A.ddz;
  double sdiff = ddsddx + ddsddy + ddsddz;
  diff[c] = sdiff;
  double ss;
  if (A.dt0 > 0)
    ss = (A.ab*sdiff - A.ab0*diff0[c]) - (A.ab*sconv - A.ab0*conv0[c]);
  else
    ss = sdiff - sconv;
  ss = s111 + A.dt * ss;
  double cur = s[c];
  s[c] = cur + (ss - cur) * (double) (phase[c] == -1);
}

extern "C" void cuda_scalar_solve(void)
{
  /* src/cuda_scalar.cu:253-257 */
  ScalarArgs A;
  A.D = s_D; A.dt = dt; A.dt0 = dt0;
  A.ab0 = 0.5 * dt / dt0;
  A.ab = 1. + A.ab0;
  A.ddx = 1. / dom[0].dx;
  A.ddy = 1. / dom[0].dy;
  A.ddz = 1. / dom[0].dz;
  const DevDom &D = ddom;
  hipLaunchKernelGGL(scalar_solve_kernel, dim3((D.nx + 63) / 64,
    (D.ny + 3) / 4, D.nz), dim3(64, 4), 0, bb_stream, D, A, _phase, _s0, _s,
    _s_conv, _s_diff, _s_conv0, _s_diff0, _u0, _v0, _w0);
  BB_LAUNCH_CHECK();
}

extern "C" void cuda_store_s(void)
{
  /* src/cuda_scalar.cu:949-957 */
  size_t n = ddom.cc.n * sizeof(real);
  BB_CHECK(hipMemcpyAsync(_s0, _s, n, hipMemcpyDeviceToDevice, bb_stream));
  BB_CHECK(hipMemcpyAsync(_s_conv0, _s_conv, n, hipMemcpyDeviceToDevice,
    bb_stream));
  BB_CHECK(hipMemcpyAsync(_s_diff0, _s_diff, n, hipMemcpyDeviceToDevice,
    bb_stream));
}
