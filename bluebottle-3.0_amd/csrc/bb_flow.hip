/*
 * Flow-path kernels: forcing, the u* advection-diffusion stencils, the
 * solvability correction, the projection, the pressure update, the time-step
 * estimate and the AB2 history rotation.
 *
 * Arithmetic follows the reference kernels operation by operation (same
 * operand order, no FMA contraction), so results are bit-identical to a host
 * evaluation of the same expressions:
 *   u*, v*, w*      src/bluebottle_kernel.cu:1637-2133
 *   surface sums    src/bluebottle_kernel.cu:2135-2229, cuda_bluebottle.cu:2313-2492
 *   projection      src/bluebottle_kernel.cu:2303-2355
 *   pressure update src/bluebottle_kernel.cu:2385-2402, cuda_bluebottle.cu:2506-2535
 *   time step       src/cuda_bluebottle.cu:1753-1808
 *
 * The face-flag magnitudes |flag| used by the projection are 0 exactly on
 * the faces of a non-periodic external boundary and 1 everywhere else
 * (src/particle_kernel.cu:481-576 only ever writes +-1 or, on external
 * boundaries, 0), so the kernels derive them from the face index instead of
 * reading the flag arrays.
 */
#include <math.h>

#include "bb_common.h"

extern double bb_force[3];
int bb_history_rotated = 0;
extern int bb_bous_on;                 /* bb_scalar.hip */

/* ------------------------------ forcing --------------------------------- */
extern "C" void cuda_compute_forcing(void)
{
  /* src/cuda_bluebottle.cu:1811-1982 without the PID branch (rejected at
   * parse time): ramped constant pressure gradient, oscillating z part. */
  real delta = ttime - p_bc_tdelay;
  if (delta >= 0) {
    if (gradP.xa == 0) gradP.x = gradP.xm;
    else if (fabs(delta*gradP.xa) > fabs(gradP.xm)) gradP.x = gradP.xm;
    else gradP.x = delta*gradP.xa;
    if (gradP.ya == 0) gradP.y = gradP.ym;
    else if (fabs(delta*gradP.ya) > fabs(gradP.ym)) gradP.y = gradP.ym;
    else gradP.y = delta*gradP.ya;
    if (gradP.za == 0) gradP.z = gradP.zm;
    else if (fabs(delta*gradP.za) > fabs(gradP.zm)) gradP.z = gradP.zm;
    else gradP.z = delta*gradP.za;
  }
  gradP.z = gradP.z * cos(osci_f*ttime);
  delta = ttime - g_bc_tdelay;
  if (delta >= 0) {
    if (g.xa == 0) g.x = g.xm;
    else if (fabs(delta*g.xa) > fabs(g.xm)) g.x = g.xm;
    else g.x = delta*g.xa;
    if (g.ya == 0) g.y = g.ym;
    else if (fabs(delta*g.ya) > fabs(g.ym)) g.y = g.ym;
    else g.y = delta*g.ya;
    if (g.za == 0) g.z = g.zm;
    else if (fabs(delta*g.za) > fabs(g.zm)) g.z = g.zm;
    else g.z = delta*g.za;
  }
  /* f = 0; f += -gradP/rho_f  (forcing_reset_* then forcing_add_*_const) */
  bb_bous_on = 0;                      /* f = 0 ... (forcing_reset_*) */
  bb_force[0] = 0. + (-gradP.x/rho_f);
  bb_force[1] = 0. + (-gradP.y/rho_f);
  bb_force[2] = 0. + (-gradP.z/rho_f);
}

/* --------------------------------- u* ----------------------------------- */
/* f: the constant part of the forcing (-gradP / rho_f); with the
 * Boussinesq term on (sb != NULL), f + bg (s_face - s_ref) with
 * bg = -g s_beta and s_face the mean of the two cells beside the face
 * (src/scalar_kernel.cu:160-206, added after the constant as the
 * reference adds it). */
struct StarArgs {
  double rho_f, nu, dt, dt0, ab0, ab, ddx, ddy, ddz, f, bg, sref;
};

/* The arithmetic of u*, v* and w* at one face from its loaded
 * neighbourhood: the star value s, the convective and diffusive terms sc
 * and sd.  Names follow the reference: u111 is the face itself, u011 /
 * u211 its x neighbours, u101 / u121 y, u110 / u112 z; the other two
 * components at the four faces around it; pm, pc the two cells beside it;
 * sbm, sbc the scalar there (Boussinesq); fluid = both cells are fluid. */
struct StarOut { double s, sc, sd; };

__device__ inline StarOut u_star_math(const StarArgs &A, double u011,
  double u111, double u211, double u101, double u121, double u110,
  double u112, double v011, double v111, double v021, double v121,
  double w011, double w111, double w012, double w112, double pm, double pc,
  double conv0c, double diff0c, bool bous, double sbm, double sbc,
  bool fluid)
{
  double s = (pm - pc) * A.ddx/A.rho_f;
  double duudx = (u211 + u111)*(u211 + u111) - (u111 + u011)*(u111 + u011);
  duudx *= 0.25 * A.ddx;
  double duvdy = (u121 + u111)*(v121 + v021) - (u111 + u101)*(v111 + v011);
  duvdy *= 0.25 * A.ddy;
  double duwdz = (u112 + u111)*(w112 + w012) - (u111 + u110)*(w111 + w011);
  duwdz *= 0.25 * A.ddz;
  double sc = duudx + duvdy + duwdz;
  if (A.dt0 > 0) s += (-A.ab * sc + A.ab0 * conv0c);
  else s += -sc;
  double dud1 = (u211 - u111) * A.ddx;
  double dud0 = (u111 - u011) * A.ddx;
  double ddudxx = (dud1 - dud0) * A.ddx;
  dud1 = (u121 - u111) * A.ddy;
  dud0 = (u111 - u101) * A.ddy;
  double ddudyy = (dud1 - dud0) * A.ddy;
  dud1 = (u112 - u111) * A.ddz;
  dud0 = (u111 - u110) * A.ddz;
  double ddudzz = (dud1 - dud0) * A.ddz;
  double sd = A.nu * (ddudxx + ddudyy + ddudzz);
  if (A.dt0 > 0) s += (A.ab * sd - A.ab0 * diff0c);
  else s += sd;
  double fo = A.f;
  if (bous) fo += A.bg * (0.5*(sbm + sbc) - A.sref);
  s += fo;
  s *= A.dt;
  s += u111;
  s *= (double) fluid;
  return {s, sc, sd};
}


__device__ inline StarOut v_star_math(const StarArgs &A, double v011,
  double v111, double v211, double v101, double v121, double v110,
  double v112, double u101, double u111, double u201, double u211,
  double w101, double w111, double w102, double w112, double pm, double pc,
  double conv0c, double diff0c, bool bous, double sbm, double sbc,
  bool fluid)
{
  double s = (pm - pc) * A.ddy/A.rho_f;
  double dvudx = (v211 + v111)*(u211 + u201) - (v111 + v011)*(u111 + u101);
  dvudx *= 0.25 * A.ddx;
  double dvvdy = (v121 + v111)*(v121 + v111) - (v111 + v101)*(v111 + v101);
  dvvdy *= 0.25 * A.ddy;
  double dvwdz = (v112 + v111)*(w112 + w102) - (v111 + v110)*(w111 + w101);
  dvwdz *= 0.25 * A.ddz;
  double sc = dvudx + dvvdy + dvwdz;
  if (A.dt0 > 0) s += (-A.ab * sc + A.ab0 * conv0c);
  else s += -sc;
  double dvd1 = (v211 - v111) * A.ddx;
  double dvd0 = (v111 - v011) * A.ddx;
  double ddvdxx = (dvd1 - dvd0) * A.ddx;
  dvd1 = (v121 - v111) * A.ddy;
  dvd0 = (v111 - v101) * A.ddy;
  double ddvdyy = (dvd1 - dvd0) * A.ddy;
  dvd1 = (v112 - v111) * A.ddz;
  dvd0 = (v111 - v110) * A.ddz;
  double ddvdzz = (dvd1 - dvd0) * A.ddz;
  double sd = A.nu * (ddvdxx + ddvdyy + ddvdzz);
  if (A.dt0 > 0) s += (A.ab * sd - A.ab0 * diff0c);
  else s += sd;
  double fo = A.f;
  if (bous) fo += A.bg * (0.5*(sbm + sbc) - A.sref);
  s += fo;
  s *= A.dt;
  s += v111;
  s *= (double) fluid;
  return {s, sc, sd};
}


__device__ inline StarOut w_star_math(const StarArgs &A, double w011,
  double w111, double w211, double w101, double w121, double w110,
  double w112, double u110, double u111, double u210, double u211,
  double v110, double v111, double v120, double v121, double pm, double pc,
  double conv0c, double diff0c, bool bous, double sbm, double sbc,
  bool fluid)
{
  double s = (pm - pc) * A.ddz/A.rho_f;
  double dwudx = (w211 + w111)*(u211 + u210) - (w111 + w011)*(u111 + u110);
  dwudx *= 0.25 * A.ddx;
  double dwvdy = (w121 + w111)*(v121 + v120) - (w111 + w101)*(v111 + v110);
  dwvdy *= 0.25 * A.ddy;
  double dwwdz = (w112 + w111)*(w112 + w111) - (w111 + w110)*(w111 + w110);
  dwwdz *= 0.25 * A.ddz;
  double sc = dwudx + dwvdy + dwwdz;
  if (A.dt0 > 0) s += (-A.ab * sc + A.ab0 * conv0c);
  else s += -sc;
  double dwd1 = (w211 - w111) * A.ddx;
  double dwd0 = (w111 - w011) * A.ddx;
  double ddwdxx = (dwd1 - dwd0) * A.ddx;
  dwd1 = (w121 - w111) * A.ddy;
  dwd0 = (w111 - w101) * A.ddy;
  double ddwdyy = (dwd1 - dwd0) * A.ddy;
  dwd1 = (w112 - w111) * A.ddz;
  dwd0 = (w111 - w110) * A.ddz;
  double ddwdzz = (dwd1 - dwd0) * A.ddz;
  double sd = A.nu * (ddwdxx + ddwdyy + ddwdzz);
  if (A.dt0 > 0) s += (A.ab * sd - A.ab0 * diff0c);
  else s += sd;
  double fo = A.f;
  if (bous) fo += A.bg * (0.5*(sbm + sbc) - A.sref);
  s += fo;
  s *= A.dt;
  s += w111;
  s *= (double) fluid;
  return {s, sc, sd};
}


/* u*, v* and w* over z-columns: a thread owns face column (i, j) for
 * STAR_KZ consecutive planes (the workgroup: 64 x 4 columns, numbered
 * through tile_of_block so that each XCD's L2 holds a contiguous slab).
 * Values a face shares with the face one plane up (its own component's z
 * neighbours, the other components on the planes k and k + 1, and for w*
 * the cells below) are carried in registers from plane to plane instead
 * of being loaded again; every face computes the same expression as
 * before (u_star_math etc.), so the results are unchanged bit for bit. */
#define STAR_KZ 4

struct StarCol { int i, j, k0, k1; bool ok; };

__device__ inline StarCol star_column(const Grid &G)
{
  const int nbx = (G.ni + 63) / 64, nby = (G.nj + 3) / 4;
  const int nbz = (G.nk + STAR_KZ - 1) / STAR_KZ;
  long t = tile_of_block((long) nbx * nby * nbz);
  StarCol S;
  S.i = (int) (t % nbx) * 64 + (int) (threadIdx.x & 63) + 1;
  S.j = (int) ((t / nbx) % nby) * 4 + (int) (threadIdx.x >> 6) + 1;
  S.k0 = (int) (t / ((long) nbx * nby)) * STAR_KZ + 1;
  S.k1 = S.k0 + STAR_KZ - 1 < G.nk ? S.k0 + STAR_KZ - 1 : G.nk;
  S.ok = S.i <= G.ni && S.j <= G.nj;
  return S;
}

static unsigned star_blocks(const Grid &G)
{
  return (unsigned) (((G.ni + 63) / 64) * ((G.nj + 3) / 4) *
                     ((G.nk + STAR_KZ - 1) / STAR_KZ));
}

__global__ void __launch_bounds__(256)
u_star_kernel(DevDom D, StarArgs A, const double *__restrict__ u0,
  const double *__restrict__ v0, const double *__restrict__ w0,
  const double *__restrict__ p, const double *__restrict__ conv0,
  const double *__restrict__ diff0, const int *__restrict__ phase,
  const double *__restrict__ sb,
  double *__restrict__ ust, double *__restrict__ conv,
  double *__restrict__ diff)
{
  const Grid &X = D.fx, &Y = D.fy, &Z = D.fz, &C = D.cc;
  const StarCol S = star_column(X);
  if (!S.ok) return;
  long c = gidx(X, S.i, S.j, S.k0), cv = gidx(Y, S.i, S.j, S.k0);
  long cw = gidx(Z, S.i, S.j, S.k0), cc = gidx(C, S.i, S.j, S.k0);
  double u110 = u0[c - X.sz], u111 = u0[c];
  double w011 = w0[cw - 1], w111 = w0[cw];
  for (int k = S.k0; k <= S.k1; k++) {
    const double u112 = u0[c + X.sz];
    const double w012 = w0[cw - 1 + Z.sz], w112 = w0[cw + Z.sz];
    StarOut o = u_star_math(A, u0[c - 1], u111, u0[c + 1], u0[c - X.sy],
      u0[c + X.sy], u110, u112, v0[cv - 1], v0[cv], v0[cv - 1 + Y.sy],
      v0[cv + Y.sy], w011, w111, w012, w112, p[cc - 1], p[cc], conv0[c],
      diff0[c], sb != NULL, sb ? sb[cc - 1] : 0., sb ? sb[cc] : 0.,
      phase[cc - 1] < 0 && phase[cc] < 0);
    ust[c] = o.s;
    conv[c] = o.sc;
    diff[c] = o.sd;
    u110 = u111; u111 = u112;
    w011 = w012; w111 = w112;
    c += X.sz; cv += Y.sz; cw += Z.sz; cc += C.sz;
  }
}

__global__ void __launch_bounds__(256)
v_star_kernel(DevDom D, StarArgs A, const double *__restrict__ u0,
  const double *__restrict__ v0, const double *__restrict__ w0,
  const double *__restrict__ p, const double *__restrict__ conv0,
  const double *__restrict__ diff0, const int *__restrict__ phase,
  const double *__restrict__ sb,
  double *__restrict__ vst, double *__restrict__ conv,
  double *__restrict__ diff)
{
  const Grid &X = D.fx, &Y = D.fy, &Z = D.fz, &C = D.cc;
  const StarCol S = star_column(Y);
  if (!S.ok) return;
  long c = gidx(Y, S.i, S.j, S.k0), cu = gidx(X, S.i, S.j, S.k0);
  long cw = gidx(Z, S.i, S.j, S.k0), cc = gidx(C, S.i, S.j, S.k0);
  double v110 = v0[c - Y.sz], v111 = v0[c];
  double w101 = w0[cw - Z.sy], w111 = w0[cw];
  for (int k = S.k0; k <= S.k1; k++) {
    const double v112 = v0[c + Y.sz];
    const double w102 = w0[cw - Z.sy + Z.sz], w112 = w0[cw + Z.sz];
    StarOut o = v_star_math(A, v0[c - 1], v111, v0[c + 1], v0[c - Y.sy],
      v0[c + Y.sy], v110, v112, u0[cu - X.sy], u0[cu], u0[cu - X.sy + 1],
      u0[cu + 1], w101, w111, w102, w112, p[cc - C.sy], p[cc], conv0[c],
      diff0[c], sb != NULL, sb ? sb[cc - C.sy] : 0., sb ? sb[cc] : 0.,
      phase[cc - C.sy] < 0 && phase[cc] < 0);
    vst[c] = o.s;
    conv[c] = o.sc;
    diff[c] = o.sd;
    v110 = v111; v111 = v112;
    w101 = w102; w111 = w112;
    c += Y.sz; cu += X.sz; cw += Z.sz; cc += C.sz;
  }
}

__global__ void __launch_bounds__(256)
w_star_kernel(DevDom D, StarArgs A, const double *__restrict__ u0,
  const double *__restrict__ v0, const double *__restrict__ w0,
  const double *__restrict__ p, const double *__restrict__ conv0,
  const double *__restrict__ diff0, const int *__restrict__ phase,
  const double *__restrict__ sb,
  double *__restrict__ wst, double *__restrict__ conv,
  double *__restrict__ diff)
{
  const Grid &X = D.fx, &Y = D.fy, &Z = D.fz, &C = D.cc;
  const StarCol S = star_column(Z);
  if (!S.ok) return;
  long c = gidx(Z, S.i, S.j, S.k0), cu = gidx(X, S.i, S.j, S.k0);
  long cv = gidx(Y, S.i, S.j, S.k0), cc = gidx(C, S.i, S.j, S.k0);
  double w110 = w0[c - Z.sz], w111 = w0[c];
  double u110 = u0[cu - X.sz], u210 = u0[cu - X.sz + 1];
  double v110 = v0[cv - Y.sz], v120 = v0[cv - Y.sz + Y.sy];
  double pm = p[cc - C.sz];
  int phm = phase[cc - C.sz];
  double sbm = sb ? sb[cc - C.sz] : 0.;
  for (int k = S.k0; k <= S.k1; k++) {
    const double w112 = w0[c + Z.sz];
    const double u111 = u0[cu], u211 = u0[cu + 1];
    const double v111 = v0[cv], v121 = v0[cv + Y.sy];
    const double pc = p[cc];
    const int phc = phase[cc];
    const double sbc = sb ? sb[cc] : 0.;
    StarOut o = w_star_math(A, w0[c - 1], w111, w0[c + 1], w0[c - Z.sy],
      w0[c + Z.sy], w110, w112, u110, u111, u210, u211, v110, v111, v120,
      v121, pm, pc, conv0[c], diff0[c], sb != NULL, sbm, sbc,
      phm < 0 && phc < 0);
    wst[c] = o.s;
    conv[c] = o.sc;
    diff[c] = o.sd;
    w110 = w111; w111 = w112;
    u110 = u111; u210 = u211;
    v110 = v111; v120 = v121;
    pm = pc; phm = phc; sbm = sbc;
    c += Z.sz; cu += X.sz; cv += Y.sz; cc += C.sz;
  }
}

/* (A kernel forming u*, v* and w* in one pass, reading u0, v0, w0, p and
 * phase once, took 884 us at 256^3 against 3 x 292 us for the three
 * kernels, which share those reads through L2.  At 256^3 the three
 * kernels take 303-321 us each whether a thread owns 1, 4 or 8 planes
 * and with or without non-temporal stores.) */
static StarArgs star_args(double f, double gc)
{
  StarArgs A;
  A.bg = -gc * s_beta;
  A.sref = s_ref;
  A.rho_f = rho_f; A.nu = nu; A.dt = dt; A.dt0 = dt0;
  A.ab0 = 0.5 * dt / dt0;
  A.ab = 1. + A.ab0;
  A.ddx = 1. / dom[0].dx;
  A.ddy = 1. / dom[0].dy;
  A.ddz = 1. / dom[0].dz;
  A.f = f;
  return A;
}

extern "C" void cuda_U_star(void)
{
  /* src/cuda_bluebottle.cu:2095-2108 */
  const DevDom &D = ddom;
  const double *sb = (SCALAR >= 1 && bb_bous_on) ? _s : NULL;
  hipLaunchKernelGGL(u_star_kernel, dim3(star_blocks(D.fx)), dim3(256), 0,
    bb_stream, D, star_args(bb_force[0], g.x), _u0, _v0, _w0, _p0, _conv0_u,
    _diff0_u, _phase, sb, _u_star, _conv_u, _diff_u);
  hipLaunchKernelGGL(v_star_kernel, dim3(star_blocks(D.fy)), dim3(256), 0,
    bb_stream, D, star_args(bb_force[1], g.y), _u0, _v0, _w0, _p0, _conv0_v,
    _diff0_v, _phase, sb, _v_star, _conv_v, _diff_v);
  hipLaunchKernelGGL(w_star_kernel, dim3(star_blocks(D.fz)), dim3(256), 0,
    bb_stream, D, star_args(bb_force[2], g.z), _u0, _v0, _w0, _p0, _conv0_w,
    _diff0_w, _phase, sb, _w_star, _conv_w, _diff_w);
  BB_LAUNCH_CHECK();
  bb_history_rotated = 0;
}

/* --------------------------- solvability ------------------------------- */
/* Gather one boundary-face plane into the reference's temporary order
 * (src/bluebottle_kernel.cu:2135-2229). */
__global__ void face_gather_kernel(const double *__restrict__ a, Grid G,
  int dir, int loc, double *__restrict__ out)
{
  int n1 = (dir == 0) ? G.nj : G.ni;
  int n2 = (dir == 2) ? G.nj : G.nk;
  int t1 = blockIdx.x * blockDim.x + threadIdx.x;
  int t2 = blockIdx.y * blockDim.y + threadIdx.y;
  if (t1 >= n1 || t2 >= n2) return;
  long idx;
  if (dir == 0) idx = gidx(G, loc, t1 + 1, t2 + 1);
  else if (dir == 1) idx = gidx(G, t1 + 1, loc, t2 + 1);
  else idx = gidx(G, t1 + 1, t2 + 1, loc);
  out[t1 + (long) t2 * n1] = a[idx];
}

__global__ void face_add_kernel(double *a, Grid G, int dir, int loc,
  const double *__restrict__ valp, int subtract)
{
  const double val = *valp;
  int n1 = (dir == 0) ? G.nj : G.ni;
  int n2 = (dir == 2) ? G.nj : G.nk;
  int t1 = blockIdx.x * blockDim.x + threadIdx.x;
  int t2 = blockIdx.y * blockDim.y + threadIdx.y;
  if (t1 >= n1 || t2 >= n2) return;
  long idx;
  if (dir == 0) idx = gidx(G, loc, t1 + 1, t2 + 1);
  else if (dir == 1) idx = gidx(G, t1 + 1, loc, t2 + 1);
  else idx = gidx(G, t1 + 1, t2 + 1, loc);
  if (subtract) a[idx] = a[idx] - val;
  else a[idx] = a[idx] + val;
}

static void face_sum(const double *a, const Grid &G, int dir, int loc,
  int slot)
{
  int n1 = (dir == 0) ? G.nj : G.ni;
  int n2 = (dir == 2) ? G.nj : G.nk;
  hipLaunchKernelGGL(face_gather_kernel, dim3((n1 + 63) / 64, (n2 + 3) / 4),
    dim3(64, 4), 0, bb_stream, a, G, dir, loc, _bb_rows + 0);
  bb_reduce_rows(_bb_rows, (long) n1 * n2, _bb_scalars + slot);
}

/* a += (or -=) the value in device slot `slot` of _bb_scalars on one
 * boundary face plane */
static void face_add(double *a, const Grid &G, int dir, int loc, int slot,
  int subtract)
{
  int n1 = (dir == 0) ? G.nj : G.ni;
  int n2 = (dir == 2) ? G.nj : G.nk;
  hipLaunchKernelGGL(face_add_kernel, dim3((n1 + 63) / 64, (n2 + 3) / 4),
    dim3(64, 4), 0, bb_stream, a, G, dir, loc,
    (const double *) _bb_scalars + slot, subtract);
  BB_LAUNCH_CHECK();
}

/* The net boundary flux of the six face sums in slots 0-5 and the
 * correction(s) it calls for, into slots 20 (one out-plane) or 20-22
 * (HOMOGENEOUS: x, y, z), with the host's operations in the host's order
 * (src/cuda_bluebottle.cu:2313-2492). */
__global__ void solv_eps_kernel(double *s, double dx, double dy, double dz,
  double xl, double yl, double zl, int plane)
{
  double eps_xs = s[0], eps_xe = s[1], eps_ys = s[2], eps_ye = s[3];
  double eps_zs = s[4], eps_ze = s[5];
  eps_xs *= dy * dz;
  eps_xe *= dy * dz;
  eps_ys *= dz * dx;
  eps_ye *= dz * dx;
  eps_zs *= dx * dy;
  eps_ze *= dx * dy;
  double eps[3];
  eps[0] = eps_xe - eps_xs;
  eps[1] = eps_ye - eps_ys;
  eps[2] = eps_ze - eps_zs;
  switch (plane) {
    case WEST: case EAST:
      s[20] = (eps[0] + eps[1] + eps[2])/(yl * zl);
      break;
    case SOUTH: case NORTH:
      s[20] = (eps[0] + eps[1] + eps[2])/(zl * xl);
      break;
    case BOTTOM: case TOP:
      s[20] = (eps[0] + eps[1] + eps[2])/(xl * yl);
      break;
    case HOMOGENEOUS:
      s[20] = 0.5*eps[0]/(yl * zl);
      s[21] = 0.5*eps[1]/(zl * xl);
      s[22] = 0.5*eps[2]/(xl * yl);
      break;
  }
}

extern "C" void cuda_solvability(void)
{
  /* src/cuda_bluebottle.cu:2313-2492 on a single subdomain, which is on
   * every global edge. */
  dom_struct *d = &dom[0];
  face_sum(_u_star, ddom.fx, 0, 1, 0);
  face_sum(_u_star, ddom.fx, 0, ddom.fx.ni, 1);
  face_sum(_v_star, ddom.fy, 1, 1, 2);
  face_sum(_v_star, ddom.fy, 1, ddom.fy.nj, 3);
  face_sum(_w_star, ddom.fz, 2, 1, 4);
  face_sum(_w_star, ddom.fz, 2, ddom.fz.nk, 5);
  /* the correction stays on the device: no host read */
  hipLaunchKernelGGL(solv_eps_kernel, dim3(1), dim3(1), 0, bb_stream,
    _bb_scalars, d->dx, d->dy, d->dz, DOM.xl, DOM.yl, DOM.zl, out_plane);
  switch (out_plane) {
    case WEST:
      face_add(_u_star, ddom.fx, 0, 1, 20, 0);
      break;
    case EAST:
      face_add(_u_star, ddom.fx, 0, ddom.fx.ni, 20, 1);
      break;
    case SOUTH:
      face_add(_v_star, ddom.fy, 1, 1, 20, 0);
      break;
    case NORTH:
      face_add(_v_star, ddom.fy, 1, ddom.fy.nj, 20, 1);
      break;
    case BOTTOM:
      face_add(_w_star, ddom.fz, 2, 1, 20, 0);
      break;
    case TOP:
      face_add(_w_star, ddom.fz, 2, ddom.fz.nk, 20, 1);
      break;
    case HOMOGENEOUS:
      face_add(_u_star, ddom.fx, 0, 1, 20, 0);
      face_add(_u_star, ddom.fx, 0, ddom.fx.ni, 20, 1);
      face_add(_v_star, ddom.fy, 1, 1, 21, 0);
      face_add(_v_star, ddom.fy, 1, ddom.fy.nj, 21, 1);
      face_add(_w_star, ddom.fz, 2, 1, 22, 0);
      face_add(_w_star, ddom.fz, 2, ddom.fz.nk, 22, 1);
      break;
  }
  BB_LAUNCH_CHECK();
}

/* ----------------------------- projection ------------------------------- */
__global__ void __launch_bounds__(256)
project_kernel(DevDom D, int comp, const double *__restrict__ ust,
  const double *__restrict__ ph, double *__restrict__ out, double ddh,
  double dt_rho)
{
  const Grid &G = comp == 0 ? D.fx : (comp == 1 ? D.fy : D.fz);
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > G.ni || j > G.nj) return;
  long c = gidx(G, i, j, k);
  long cc = gidx(D.cc, i, j, k);
  long sh;
  int wall;
  if (comp == 0) {
    sh = 1;
    wall = !D.perx && (i == 1 || i == G.ni);
  } else if (comp == 1) {
    sh = D.cc.sy;
    wall = !D.pery && (j == 1 || j == G.nj);
  } else {
    sh = D.cc.sz;
    wall = !D.perz && (k == 1 || k == G.nk);
  }
  double af = wall ? 0. : 1.;
  double gradPhi = af * ddh * (ph[cc] - ph[cc - sh]);
  out[c] = (ust[c] - dt_rho * gradPhi);
}

extern "C" void cuda_project(void)
{
  /* src/cuda_bluebottle.cu:2495-2503 */
  dim3 bl(64, 4, 1);
  const DevDom &D = ddom;
  double dt_rho = dt / rho_f;
  hipLaunchKernelGGL(project_kernel, dim3((D.fx.ni + 63) / 64,
    (D.fx.nj + 3) / 4, D.fx.nk), bl, 0, bb_stream, D, 0, _u_star, _phi, _u,
    1. / dom[0].dx, dt_rho);
  hipLaunchKernelGGL(project_kernel, dim3((D.fy.ni + 63) / 64,
    (D.fy.nj + 3) / 4, D.fy.nk), bl, 0, bb_stream, D, 1, _v_star, _phi, _v,
    1. / dom[0].dy, dt_rho);
  hipLaunchKernelGGL(project_kernel, dim3((D.fz.ni + 63) / 64,
    (D.fz.nj + 3) / 4, D.fz.nk), bl, 0, bb_stream, D, 2, _w_star, _phi, _w,
    1. / dom[0].dz, dt_rho);
  BB_LAUNCH_CHECK();
}

/* --------------------------- pressure update ---------------------------- */
/* p = (phase < 0) (p0 + phi) on interior cells, with the tile partials of
 * p for the mean. */
__global__ void __launch_bounds__(TILE_THREADS)
update_p_kernel(DevDom D, const double *__restrict__ p0,
  const double *__restrict__ ph, const int *__restrict__ phase,
  double *__restrict__ p, double *__restrict__ gpart)
{
  TILE_PROLOGUE
  double acc = 0.0;
  if (colok) {
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      double val = (double) (phase[c] < 0) * (p0[c] + ph[c]);
      p[c] = val;
      acc += val;
    }
  }
  tile_partial(acc, gpart, tile);
}

/* p -= (sum / n) on interior cells, sum read from device memory (the mean
 * formed as the host would form it: one division, then the subtraction
 * as the addition of its negation). */
__global__ void __launch_bounds__(256)
sub_mean_kernel(DevDom D, double *__restrict__ p,
  const double *__restrict__ sum, double n)
{
  const Grid &C = D.cc;
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > C.ni || j > C.nj) return;
  const double pmean = *sum / n;
  long c = gidx(C, i, j, k);
  p[c] += -pmean;
}

extern "C" void cuda_update_p(void)
{
  /* src/cuda_bluebottle.cu:2506-2535 (the unused Laplacian pass of
   * update_p_laplacian is not evaluated). */
  const DevDom &D = ddom;
  hipLaunchKernelGGL(update_p_kernel, tiles_grid(D), tiles_block(), 0,
    bb_stream, D, _p0, _phi, _phase, _p, _bb_rows);
  BB_LAUNCH_CHECK();
  bb_reduce_rows(_bb_rows, n_tiles(D), _bb_scalars + 0);
  /* the mean stays on the device: no host read between the two kernels */
  hipLaunchKernelGGL(sub_mean_kernel, dim3((D.cc.ni + 63) / 64,
    (D.cc.nj + 3) / 4, D.cc.nk), dim3(64, 4, 1), 0, bb_stream, D, _p,
    (const double *) _bb_scalars, (double) DOM.Gcc.s3);
  BB_LAUNCH_CHECK();
}

/* ------------------------------ time step ------------------------------- */
/* Row maxima of a*a over the interior points of one face grid.  The maximum
 * is exact, so any combination order gives the same value; the combine uses
 * the reference's (a < b) ? b : a starting from 0. */
__device__ inline double dmax(double a, double b) { return (a < b) ? b : a; }

__global__ void __launch_bounds__(256)
sqmax_rows_kernel(Grid G, const double *__restrict__ a,
  double *__restrict__ rows)
{
  int lane = threadIdx.x & (WAVE - 1);
  long row = (long) blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
  if (row >= (long) G.nj * G.nk) return;
  int j = (int) (row % G.nj) + 1;
  int k = (int) (row / G.nj) + 1;
  long base = gidx(G, 0, j, k);
  double m = 0.;
  for (int i = lane + 1; i <= G.ni; i += WAVE) {
    double x = a[base + i];
    m = dmax(m, x * x);
  }
  for (int off = 32; off >= 1; off >>= 1) m = dmax(m, __shfl_down(m, off, WAVE));
  if (lane == 0) rows[row] = m;
}

__global__ void __launch_bounds__(REDUCE_THREADS)
max_rows_kernel(const double *__restrict__ rows, long n, double *out)
{
  __shared__ double s[REDUCE_THREADS];
  int t = threadIdx.x;
  double m = 0.;
  for (long r = t; r < n; r += REDUCE_THREADS) m = dmax(m, rows[r]);
  s[t] = m;
  __syncthreads();
  for (int off = REDUCE_THREADS / 2; off >= 1; off >>= 1) {
    if (t < off) s[t] = dmax(s[t], s[t + off]);
    __syncthreads();
  }
  if (t == 0) *out = s[0];
}

static void sqmax(const Grid &G, const double *a, int slot)
{
  long nrows = (long) G.nj * G.nk;
  hipLaunchKernelGGL(sqmax_rows_kernel,
    dim3((unsigned) ((nrows + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK)),
    rows_block(), 0, bb_stream, G, a, _bb_rows);
  hipLaunchKernelGGL(max_rows_kernel, dim3(1), dim3(REDUCE_THREADS), 0,
    bb_stream, _bb_rows, (long) G.nj * G.nk, _bb_scalars + slot);
  BB_LAUNCH_CHECK();
}

extern "C" void cuda_find_dt(void)
{
  /* src/cuda_bluebottle.cu:1753-1808 */
  if (fixed_dt > 0.) {
    dt = fixed_dt;
    return;
  }
  sqmax(ddom.fx, _u, 0);
  sqmax(ddom.fy, _v, 1);
  sqmax(ddom.fz, _w, 2);
  bb_fetch_scalars(3);
  real u_max = sqrt(bb_host_scalars[0]);
  real v_max = sqrt(bb_host_scalars[1]);
  real w_max = sqrt(bb_host_scalars[2]);
  dom_struct *d = &dom[0];
  if (SCALAR >= 1 && s_D > nu) {
    dt = u_max/d->dx + 2.*s_D/(d->dx * d->dx);
    dt += v_max/d->dy + 2.*s_D/(d->dy * d->dy);
    dt += w_max/d->dz + 2.*s_D/(d->dz * d->dz);
  } else {
    dt = u_max/d->dx + 2.*nu/(d->dx * d->dx);
    dt += v_max/d->dy + 2.*nu/(d->dy * d->dy);
    dt += w_max/d->dz + 2.*nu/(d->dz * d->dz);
  }
  dt = CFL/dt;
}

/* ------------------------- turbulence forcing ---------------------------- */
extern "C" void cuda_compute_turb_forcing(void)
{
  /* src/cuda_bluebottle.cu:1985-2093 acts only when init_cond is TURBULENT
   * (linear forcing proportional to the velocity field, with kinetic energy
   * and dissipation records).  That forcing needs per-face force arrays,
   * which this library does not keep (its u* kernels add a constant force
   * per component), so such a case stops here with a message. */
  if (init_cond != TURBULENT) return;
  fprintf(stderr, "N%d >> init_cond TURBULENT (linear turbulence forcing) is "
    "not supported by this library\n", rank);
  exit(EXIT_FAILURE);
}

/* ------------------------- wall shear stress ----------------------------- */
/* Row sums over x of (u(i, j + 1, k) - u(i, j, k)) / dy for the interior
 * x faces i = 1..ni of plane k (one wavefront per k). */
__global__ void __launch_bounds__(256)
dudy_rows_kernel(Grid G, const double *__restrict__ u, int j, double idy,
  double *__restrict__ rows)
{
  int lane = threadIdx.x & (WAVE - 1);
  long row = (long) blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
  if (row >= G.nk) return;
  int k = (int) row + 1;
  double acc = 0.;
  for (int i = lane + 1; i <= G.ni; i += WAVE)
    acc += (u[gidx(G, i, j + 1, k)] - u[gidx(G, i, j, k)]) * idy;
  acc = wave_tree_sum(acc);
  if (lane == 0) rows[row] = acc;
}

extern "C" void cuda_wall_shear_stress(void)
{
  /* src/cuda_bluebottle.cu:2764-2835: the mean du/dy on the south face
   * (between the ghost row and the first row) and on the north face
   * (between the last row and the ghost row), over the in x kn interior x
   * faces, written as rho_f nu du/dy to output/wss.dat.  The reference's
   * per-face kernel maps (i, k) and (i + 1, 0) onto one slot for k = kn;
   * here each face is summed once. */
  const Grid &G = ddom.fx;
  long nrows = G.nk;
  dim3 grid((unsigned) ((nrows + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK));
  hipLaunchKernelGGL(dudy_rows_kernel, grid, rows_block(), 0, bb_stream, G,
    _u, 0, 1. / dom[0].dy, _bb_rows);
  bb_reduce_rows(_bb_rows, nrows, _bb_scalars + 12);
  hipLaunchKernelGGL(dudy_rows_kernel, grid, rows_block(), 0, bb_stream, G,
    _u, G.nj, 1. / dom[0].dy, _bb_rows2);
  bb_reduce_rows(_bb_rows2, nrows, _bb_scalars + 13);
  BB_LAUNCH_CHECK();
  bb_fetch_scalars(14);
  double nface = (double) dom[0].Gfx.in * dom[0].Gfx.kn;
  double dudy_s = bb_host_scalars[12] / nface;
  double dudy_n = bb_host_scalars[13] / nface;
  if (rank != 0) return;
  char fname[FILE_NAME_SIZE];
  snprintf(fname, sizeof fname, "%s/%s/wss.dat", ROOT_DIR, OUTPUT_DIR);
  FILE *file = fopen(fname, stepnum == 1 ? "w" : "a");
  if (file == NULL) {
    fprintf(stderr, "Could not open file %s\n", fname);
    exit(EXIT_FAILURE);
  }
  if (stepnum == 1)
    fprintf(file, "%-9s%-11s%-11s%-11s", "stepnum", "ttime", "wss-s", "wss-n");
  fprintf(file, "\n%-9d%-11.3e%-11.3e%-11.3e", stepnum, ttime,
    rho_f*nu*dudy_s, rho_f*nu*dudy_n);
  fclose(file);
}

/* ----------------------------- AB2 history ------------------------------ */
static void swap_ptr(real **a, real **b)
{
  real *t = *a;
  *a = *b;
  *b = t;
}

extern "C" void cuda_store_u(void)
{
  /* src/cuda_bluebottle.cu:2592-2616.  conv/diff histories are rotated by
   * pointer (conv is fully rewritten by the next u* before it is read);
   * u, v, w, p are copied so that both copies stay current. */
  swap_ptr(&_conv0_u, &_conv_u);
  swap_ptr(&_conv0_v, &_conv_v);
  swap_ptr(&_conv0_w, &_conv_w);
  swap_ptr(&_diff0_u, &_diff_u);
  swap_ptr(&_diff0_v, &_diff_v);
  swap_ptr(&_diff0_w, &_diff_w);
  BB_CHECK(hipMemcpyAsync(_p0, _p, ddom.cc.n * sizeof(real),
    hipMemcpyDeviceToDevice, bb_stream));
  BB_CHECK(hipMemcpyAsync(_u0, _u, ddom.fx.n * sizeof(real),
    hipMemcpyDeviceToDevice, bb_stream));
  BB_CHECK(hipMemcpyAsync(_v0, _v, ddom.fy.n * sizeof(real),
    hipMemcpyDeviceToDevice, bb_stream));
  BB_CHECK(hipMemcpyAsync(_w0, _w, ddom.fz.n * sizeof(real),
    hipMemcpyDeviceToDevice, bb_stream));
  bb_history_rotated = 1;
}
