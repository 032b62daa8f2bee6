/*
 * Particles and Physalis on gfx950: cages (phase, phase_shell, face flags),
 * the Lamb-series boundary conditions on u, v, w and p, the 26-node Lebedev
 * quadrature of the Lamb coefficients, hydrodynamic forces, the coefficient
 * error of the Lamb iteration, lubrication/contact between particles and the
 * particle motion (src/particle_kernel.cu, src/physalis_kernel.cu,
 * src/cuda_particle.cu, src/cuda_physalis.cu).
 *
 * One GPU holds every particle once, in input order, as an array of the
 * reference's part_struct.  Where the reference keeps ghost copies of a
 * particle near a periodic boundary (positions x +- L,
 * src/particle_kernel.cu:1166-1210), this code evaluates the same image
 * positions directly: a cage is built for each image whose cage meets the
 * ghosted grid, a cell or face near a particle uses the image nearest to
 * it, and a quadrature node outside the domain is read through the image
 * that brings it inside.  The phase of a cell holds the particle's number.
 *
 * Sweeps that the reference runs over the whole grid for every particle
 * (cage set-up) run over the particle's cage only; sweeps that touch every
 * face once (BCs, internal velocities) stay whole-grid passes, evaluating
 * the Lamb series only where the result is used.
 *
 * Per-particle work (quadrature, coefficients, forces, collisions, motion)
 * is latency-bound: one workgroup (64 lanes) per particle for the
 * quadrature, one lane per particle elsewhere.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>

#include "bb_common.h"

part_struct *_parts = NULL;
uint8_t *_bb_fluid = NULL;             /* 1: fluid cell, 0: particle cell */
static double *_perr = NULL;           /* per-particle Lamb error */
static int *_pnum = NULL;
static double *_err_out = NULL;        /* {err, num} of the last error */
static double *err_host = NULL;
static int *_pbin = NULL;              /* collision bin of each particle */

__constant__ int c_nn[NCOEFFS];
__constant__ int c_mm[NCOEFFS];
__constant__ double c_node_t[NNODES];
__constant__ double c_node_p[NNODES];
__constant__ double c_A1, c_A2, c_A3;

/* Domain geometry for particle kernels (passed by value). */
struct PDom {
  double xs, ys, zs, xe, ye, ze, xl, yl, zl;
  double dx, dy, dz;
  int nx, ny, nz;
  int perx, pery, perz;
  Grid cc, fx, fy, fz;
};

static PDom pdom(void)
{
  PDom P;
  P.xs = dom[0].xs; P.ys = dom[0].ys; P.zs = dom[0].zs;
  P.xe = dom[0].xe; P.ye = dom[0].ye; P.ze = dom[0].ze;
  P.xl = DOM.xl; P.yl = DOM.yl; P.zl = DOM.zl;
  P.dx = dom[0].dx; P.dy = dom[0].dy; P.dz = dom[0].dz;
  P.nx = ddom.nx; P.ny = ddom.ny; P.nz = ddom.nz;
  P.perx = ddom.perx; P.pery = ddom.pery; P.perz = ddom.perz;
  P.cc = ddom.cc; P.fx = ddom.fx; P.fy = ddom.fy; P.fz = ddom.fz;
  return P;
}

/* The six domain faces W E S N B T as the particle kernels see them
 * (passed by value): the code a quadrature node beyond the face takes
 * (src/physalis_kernel.cu:118-214: the _D code when any velocity component
 * has a Dirichlet condition there, else the _N code when any has a Neumann
 * condition, else -1), the face's Dirichlet velocities, whether its
 * pressure condition is Neumann (wall contact forces act only then,
 * src/particle_kernel.cu:2373) and its screen offset ds. */
struct PWalls {
  int code[6];
  int pneu[6];
  double vel[6][3];
  double ds[6];
};

static PWalls pwalls(void)
{
  PWalls W;
  const int vt[6][3] = {{bc.uW, bc.vW, bc.wW}, {bc.uE, bc.vE, bc.wE},
                        {bc.uS, bc.vS, bc.wS}, {bc.uN, bc.vN, bc.wN},
                        {bc.uB, bc.vB, bc.wB}, {bc.uT, bc.vT, bc.wT}};
  const double vd[6][3] = {{bc.uWD, bc.vWD, bc.wWD}, {bc.uED, bc.vED, bc.wED},
                           {bc.uSD, bc.vSD, bc.wSD}, {bc.uND, bc.vND, bc.wND},
                           {bc.uBD, bc.vBD, bc.wBD}, {bc.uTD, bc.vTD, bc.wTD}};
  const int pt[6] = {bc.pW, bc.pE, bc.pS, bc.pN, bc.pB, bc.pT};
  const double ds[6] = {bc.dsW, bc.dsE, bc.dsS, bc.dsN, bc.dsB, bc.dsT};
  for (int f = 0; f < 6; f++) {
    bool d = vt[f][0] == DIRICHLET || vt[f][1] == DIRICHLET ||
             vt[f][2] == DIRICHLET;
    bool n = vt[f][0] == NEUMANN || vt[f][1] == NEUMANN ||
             vt[f][2] == NEUMANN;
    W.code[f] = d ? WEST_WALL_D - 10 * f : (n ? WEST_WALL_N - 10 * f : -1);
    W.pneu[f] = pt[f] == NEUMANN;
    for (int c = 0; c < 3; c++) W.vel[f][c] = vd[f][c];
    W.ds[f] = ds[f];
  }
  return W;
}

/* Image position p + s L (s in {-1, 0, 1}), formed as the reference forms
 * the ghost copy's position. */
__host__ __device__ inline double img(double p, int s, double l)
{
  return s > 0 ? p + l : (s < 0 ? p - l : p);
}

/* The image of particle coordinate p nearest to the point q. */
__device__ inline double near_img(double q, double p, double l, int per)
{
  if (!per) return p;
  double d = q - p;
  if (d > 0.5 * l) return p + l;
  if (d < -0.5 * l) return p - l;
  return p;
}

/* ------------------------- spherical helpers ----------------------------- */
/* src/particle_kernel.cu:1905-1917 (the DIV_ST guards compare against a
 * negative number and never fire). */
__device__ inline void xyz2rtp(double x, double y, double z, double *r,
  double *theta, double *phi)
{
  double XY = x*x + y*y;
  double XYZ = XY + z*z;
  if (XYZ >= 0 && XYZ < DIV_ST) XYZ = DIV_ST;
  *r = sqrt(XYZ);
  *theta = acos(z / *r);
  if (XY >= 0 && XY < DIV_ST) XY = DIV_ST;
  *phi = acos(x / sqrt(XY));
  if (y < 0.) *phi = 2.*PI - *phi;
}

/* src/particle_kernel.cu:1919-1929 */
__device__ inline double Nnm(int n, int m)
{
  double fact_top = 1;
  double fact_bot = 1;
  for (int i = 1; i <= (n-m); i++) fact_top *= (double) i;
  for (int i = 1; i <= (n+m); i++) fact_bot *= (double) i;
  return sqrt((2.*n+1.) / 4. / PI * fact_top / fact_bot);
}

/* Associated Legendre functions P_n^m(cos theta), n <= 5, m >= 0
 * (src/particle_kernel.cu:1931-1991). */
__device__ inline double Pnm(int n, int m, double theta)
{
  double x = cos(theta);
  double y = sin(theta);
  switch (n) {
    case 0: return 1;
    case 1:
      if (m == 0) return x;
      if (m == 1) return -y;
      break;
    case 2:
      if (m == 0) return 0.5*(3.*x*x - 1.);
      if (m == 1) return -3.*x*y;
      if (m == 2) return 3.*y*y;
      break;
    case 3:
      if (m == 0) return 0.5*x*(5.*x*x - 3.);
      if (m == 1) return -1.5*(5.*x*x - 1.)*y;
      if (m == 2) return 15.*x*y*y;
      if (m == 3) return -15.*y*y*y;
      break;
    case 4:
      if (m == 0) return 0.125*(35.*x*x*x*x - 30.*x*x + 3.);
      if (m == 1) return -2.5*(7.*x*x - 3.)*x*y;
      if (m == 2) return 7.5*(7.*x*x - 1.)*y*y;
      if (m == 3) return -105.*x*y*y*y;
      if (m == 4) return 105.*y*y*y*y;
      break;
    case 5:
      if (m == 0) return 0.125*x*(63.*x*x*x*x - 70.*x*x + 15.);
      if (m == 1) return -1.875*y*(21.*x*x*x*x - 14.*x*x + 1.);
      if (m == 2) return 52.5*x*y*y*(3.*x*x - 1.);
      if (m == 3) return -52.5*y*y*y*(9.*x*x - 1.);
      if (m == 4) return 945.*x*y*y*y*y;
      if (m == 5) return -945.*y*y*y*y*y;
      break;
  }
  return 0;
}

/* The series terms of one order n for coefficient arrays (re, im)
 * (src/particle_kernel.cu:1993-2170): X = sum_m N P (c e^{i m phi}),
 * Y = d/dtheta form, Z = 1/sin(theta) d/dphi form. */
__device__ inline int coeff0(int n)
{
  int c = 0;
  for (int j = 0; j < n; j++) c += j + 1;
  return c;
}

__device__ inline double Xn(int n, double theta, double phi,
  const double *re, const double *im)
{
  int coeff = coeff0(n);
  int m = 0;
  double sum = Nnm(n,m)*Pnm(n,m,theta)*re[coeff];
  for (m = 1; m <= n; m++) {
    coeff++;
    sum += 2.*Nnm(n,m)*Pnm(n,m,theta)
      *(re[coeff]*cos(m*phi) - im[coeff]*sin(m*phi));
  }
  return sum;
}

__device__ inline double st_guard(double theta)
{
  double st = sin(theta);
  if (st >= 0 && st < DIV_ST) st = DIV_ST;
  else if (st < 0 && st > -DIV_ST) st = -DIV_ST;
  return st;
}

__device__ inline double Yn(int n, double theta, double phi,
  const double *re, const double *im)
{
  double ct = cos(theta);
  double st = st_guard(theta);
  int coeff = coeff0(n);
  int m = 0;
  double sum = Nnm(n,m)
    *(-(n+1)*ct/st*Pnm(n,m,theta)+(n-m+1)/st*Pnm(n+1,m,theta))
    *re[coeff];
  for (m = 1; m <= n; m++) {
    coeff++;
    sum += 2.*Nnm(n,m)
      *(-(n+1)*ct/st*Pnm(n,m,theta)+(n-m+1)/st*Pnm(n+1,m,theta))
      *(re[coeff]*cos(m*phi) - im[coeff]*sin(m*phi));
  }
  return sum;
}

__device__ inline double Zn(int n, double theta, double phi,
  const double *re, const double *im)
{
  double st = st_guard(theta);
  int coeff = coeff0(n);
  double sum = 0.;
  for (int m = 1; m <= n; m++) {
    coeff++;
    sum += -2.*m/st*Nnm(n,m)*Pnm(n,m,theta)
      *(re[coeff]*sin(m*phi) + im[coeff]*cos(m*phi));
  }
  return sum;
}

/* Lamb's velocity at (r, theta, phi) in Cartesian components
 * (src/particle_kernel.cu:1839-1903). */
__device__ inline void lamb_vel(int order, double a, double r, double theta,
  double phi, const part_struct *P, double nu, double *Ux, double *Uy,
  double *Uz)
{
  double ia = 1. / a;
  double ar = a / r;
  double ra = r * ia;
  double ur = 0.;
  double ut = 0.5*ra*Yn(0, theta, phi, P->pnm_re, P->pnm_im);
  double up = 0.5*ra*Zn(0, theta, phi, P->pnm_re, P->pnm_im);
  for (int n = 1; n <= order; n++) {
    double powranm1 = pow(ra,n-1.);
    double powranp1 = powranm1*ra*ra;
    double powarn = pow(ar,n);
    double powarnp1 = powarn*ar;
    double powarnp2 = powarnp1*ar;
    double od2np3 = 1./(2.*n+3.);
    double odnp1 = 1./(n+1.);
    ur += (0.5*n*od2np3*powranp1
      + 0.25*n*((2.*n+1.)*od2np3*ar*ar-1.)*powarn)
      * Xn(n, theta, phi, P->pnm_re, P->pnm_im);
    ur += (n*powranm1
      + 0.5*n*(2.*n-1.-(2.*n+1.)*ra*ra)*powarnp2)
      * Xn(n, theta, phi, P->phinm_re, P->phinm_im);
    ut += (0.5*(n+3.)*odnp1*od2np3*powranp1
      + 0.25*odnp1*(n-2.-n*(2.*n+1.)*od2np3*ar*ar)*powarn)
      * Yn(n, theta, phi, P->pnm_re, P->pnm_im);
    ut += (powranm1
      + 0.5*odnp1*((n-2.)*(2.*n+1.)*ra*ra-n*(2.*n-1.))*powarnp2)
      * Yn(n, theta, phi, P->phinm_re, P->phinm_im);
    ut += (powranm1
      - powarnp1)
      * Zn(n, theta, phi, P->chinm_re, P->chinm_im);
    up += (0.5*(n+3.)*odnp1*od2np3*powranp1
      + 0.25*odnp1*(n-2.-n*(2.*n+1.)*od2np3*ar*ar)*powarn)
      * Zn(n, theta, phi, P->pnm_re, P->pnm_im);
    up += (powranm1
      + 0.5*odnp1*((n-2.)*(2.*n+1.)*ra*ra-n*(2.*n-1.))*powarnp2)
      * Zn(n, theta, phi, P->phinm_re, P->phinm_im);
    up += (-powranm1
      + powarnp1)
      * Yn(n, theta, phi, P->chinm_re, P->chinm_im);
  }
  ur *= nu * ia;
  ut *= nu * ia;
  up *= nu * ia;
  double st = sin(theta);
  double ct = cos(theta);
  double sp = sin(phi);
  double cp = cos(phi);
  *Ux = ur*st*cp + ut*ct*cp - up*sp;
  *Uy = ur*st*sp + ut*ct*sp + up*cp;
  *Uz = ur*ct - ut*st;
}

/* ------------------------------ cages ------------------------------------ */
__global__ void reset_phase_kernel(long n, int *phase, int *shell)
{
  long c = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) {
    phase[c] = -1;
    shell[c] = 1;
  }
}

/* Cage extent of one particle image in one direction
 * (src/particle_kernel.cu:135-146, 155-226): start/end cell, clamped to
 * the ghosted range in a periodic direction, to the interior otherwise. */
__device__ inline void cage_range(double xp, double xs, double dx, double r,
  int ncell, int per, int *s, int *e)
{
  int cage = (int) (2. * ceil(r / dx)) + 2;
  cage -= (ncell % 2);
  int a = (int) (round((xp - xs) * (1./dx)) - 0.5*cage + DOM_BUF);
  int b = a + cage;
  int lo = per ? 0 : 1, hi = per ? ncell + 1 : ncell;
  a += (lo - a) * (a < lo) + (hi - a) * (a > hi);
  b += (lo - b) * (b < lo) + (hi - b) * (b > hi);
  *s = a;
  *e = b;
}

/* Image (sx, sy, sz) of workgroup y index 0..26; false if it needs a shift
 * in a non-periodic direction. */
__device__ inline bool image_of(int code, const PDom &P, int *sx, int *sy,
  int *sz)
{
  *sx = code % 3 - 1;
  *sy = (code / 3) % 3 - 1;
  *sz = code / 9 - 1;
  return !((*sx && !P.perx) || (*sy && !P.pery) || (*sz && !P.perz));
}

/* phase = n inside particle n (src/particle_kernel.cu:148-253). */
__global__ void __launch_bounds__(256)
build_phase_kernel(PDom P, const part_struct *__restrict__ parts, int *phase)
{
  int n = blockIdx.x, sx, sy, sz;
  if (!image_of(blockIdx.y, P, &sx, &sy, &sz)) return;
  const part_struct *Q = &parts[n];
  double X = img(Q->x, sx, P.xl), Y = img(Q->y, sy, P.yl),
         Z = img(Q->z, sz, P.zl);
  int is, ie, js, je, ks, ke;
  cage_range(X, P.xs, P.dx, Q->r, P.nx, P.perx, &is, &ie);
  cage_range(Y, P.ys, P.dy, Q->r, P.ny, P.pery, &js, &je);
  cage_range(Z, P.zs, P.dz, Q->r, P.nz, P.perz, &ks, &ke);
  if (is == ie || js == je || ks == ke) return;
  int ni = ie - is + 1, nj = je - js + 1, nk = ke - ks + 1;
  double irad = 1./Q->r;
  for (int t = threadIdx.x; t < ni * nj * nk; t += blockDim.x) {
    int ti = is + t % ni, tj = js + (t / ni) % nj, tk = ks + t / (ni * nj);
    double xx = (ti - 0.5)*P.dx - (X - P.xs);
    double yy = (tj - 0.5)*P.dy - (Y - P.ys);
    double zz = (tk - 0.5)*P.dz - (Z - P.zs);
    double d = sqrt(xx*xx + yy*yy + zz*zz);
    if (floor(d * irad) < 1) phase[gidx(P.cc, ti, tj, tk)] = n;
  }
}

/* phase_shell = 0 in cells of particle n with a 6-neighbour centre outside
 * the sphere (src/particle_kernel.cu:255-426). */
__global__ void __launch_bounds__(256)
build_shell_kernel(PDom P, const part_struct *__restrict__ parts,
  const int *__restrict__ phase, int *shell)
{
  int n = blockIdx.x, sx, sy, sz;
  if (!image_of(blockIdx.y, P, &sx, &sy, &sz)) return;
  const part_struct *Q = &parts[n];
  double X = img(Q->x, sx, P.xl), Y = img(Q->y, sy, P.yl),
         Z = img(Q->z, sz, P.zl);
  int is, ie, js, je, ks, ke;
  cage_range(X, P.xs, P.dx, Q->r, P.nx, P.perx, &is, &ie);
  cage_range(Y, P.ys, P.dy, Q->r, P.ny, P.pery, &js, &je);
  cage_range(Z, P.zs, P.dz, Q->r, P.nz, P.perz, &ks, &ke);
  if (is == ie || js == je || ks == ke) return;
  int ni = ie - is + 1, nj = je - js + 1, nk = ke - ks + 1;
  double irad = 1./Q->r;
  for (int t = threadIdx.x; t < ni * nj * nk; t += blockDim.x) {
    int ti = is + t % ni, tj = js + (t / ni) % nj, tk = ks + t / (ni * nj);
    long C = gidx(P.cc, ti, tj, tk);
    if (phase[C] != n) continue;
    double xx_w = (ti - 1 - 0.5)*P.dx - (X - P.xs);
    double xx = (ti - 0.5)*P.dx - (X - P.xs);
    double xx_e = (ti + 1 - 0.5)*P.dx - (X - P.xs);
    double yy_s = (tj - 1 - 0.5)*P.dy - (Y - P.ys);
    double yy = (tj - 0.5)*P.dy - (Y - P.ys);
    double yy_n = (tj + 1 - 0.5)*P.dy - (Y - P.ys);
    double zz_b = (tk - 1 - 0.5)*P.dz - (Z - P.zs);
    double zz = (tk - 0.5)*P.dz - (Z - P.zs);
    double zz_t = (tk + 1 - 0.5)*P.dz - (Z - P.zs);
    double d_w = sqrt(xx_w*xx_w + yy*yy + zz*zz);
    double d_e = sqrt(xx_e*xx_e + yy*yy + zz*zz);
    double d_s = sqrt(xx*xx + yy_s*yy_s + zz*zz);
    double d_n = sqrt(xx*xx + yy_n*yy_n + zz*zz);
    double d_b = sqrt(xx*xx + yy*yy + zz_b*zz_b);
    double d_t = sqrt(xx*xx + yy*yy + zz_t*zz_t);
    bool out = !(floor(d_w * irad) < 1) || !(floor(d_e * irad) < 1) ||
               !(floor(d_s * irad) < 1) || !(floor(d_n * irad) < 1) ||
               !(floor(d_b * irad) < 1) || !(floor(d_t * irad) < 1);
    if (out) shell[C] = 0;
  }
}

/* Face flags from phase and phase_shell (src/particle_kernel.cu:482-540):
 * -1 between fluid and solid or between two shell cells, else 1.  `dir`
 * selects the face grid; faces of index 1..n+1 along it, every ghost row
 * across it. */
template <int dir>
__global__ void cage_flag_kernel(PDom P, const int *__restrict__ phase,
  const int *__restrict__ shell, int *flag)
{
  const Grid G = dir == 0 ? P.fx : (dir == 1 ? P.fy : P.fz);
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.n) return;
  int i = (int) (t % G.sy), j = (int) ((t / G.sy) % (G.sz / G.sy)),
      k = (int) (t / G.sz);
  int a = dir == 0 ? i : (dir == 1 ? j : k);
  int amax = dir == 0 ? P.nx + 1 : (dir == 1 ? P.ny + 1 : P.nz + 1);
  if (a < 1 || a > amax) return;
  long c1 = gidx(P.cc, i, j, k);
  long c0 = c1 - (dir == 0 ? 1 : (dir == 1 ? P.cc.sy : P.cc.sz));
  flag[gidx(G, i, j, k)] = 1 - 2*((phase[c0] < 0 && phase[c1] > -1) ||
                                  (phase[c0] > -1 && phase[c1] < 0) ||
                                  (shell[c1] < 1 && shell[c0] < 1));
}

__global__ void fluid_mask_kernel(long n, const int *__restrict__ phase,
  uint8_t *fluid)
{
  long c = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) fluid[c] = phase[c] < 0;
}

static inline unsigned nblk(long n, int t) { return (unsigned) ((n + t - 1) / t); }

/* Called by cuda_build_cages (bb_bc.hip) between the flag reset and the
 * external-wall flags. */
void bb_build_particle_cages(void)
{
  PDom P = pdom();
  long n = ddom.cc.n;
  hipLaunchKernelGGL(reset_phase_kernel, dim3(nblk(n, 256)), dim3(256), 0,
    bb_stream, n, _phase, _phase_shell);
  hipLaunchKernelGGL(build_phase_kernel, dim3(nparts, 27), dim3(256), 0,
    bb_stream, P, _parts, _phase);
  hipLaunchKernelGGL(build_shell_kernel, dim3(nparts, 27), dim3(256), 0,
    bb_stream, P, _parts, _phase, _phase_shell);
  hipLaunchKernelGGL(cage_flag_kernel<0>, dim3(nblk(ddom.fx.n, 256)), dim3(256),
    0, bb_stream, P, _phase, _phase_shell, _flag_u);
  hipLaunchKernelGGL(cage_flag_kernel<1>, dim3(nblk(ddom.fy.n, 256)), dim3(256),
    0, bb_stream, P, _phase, _phase_shell, _flag_v);
  hipLaunchKernelGGL(cage_flag_kernel<2>, dim3(nblk(ddom.fz.n, 256)), dim3(256),
    0, bb_stream, P, _phase, _phase_shell, _flag_w);
  hipLaunchKernelGGL(fluid_mask_kernel, dim3(nblk(n, 256)), dim3(256), 0,
    bb_stream, n, _phase, _bb_fluid);
  BB_LAUNCH_CHECK();
}

/* ------------------------ particle velocity BCs -------------------------- */
/* u, v or w on faces next to a particle with flag < 1 gets the Lamb
 * velocity of the particle plus its rigid motion
 * (src/particle_kernel.cu:1370-1653).  Faces: index 1..n+1 along `dir`,
 * every row (ghosts included) across it, as the reference.
 *
 * Which faces qualify depends only on the cages (phase and flag), so the
 * cage build lists them once per direction (face_list_kernel) and every BC
 * call runs one thread per listed face: full wavefronts instead of the few
 * lanes per wavefront a sweep over all faces would keep busy.  Each face's
 * value depends on that face alone, so the list order is free. */
template <int dir>
__device__ inline bool bc_face(const PDom &P, long t, const int *phase,
  const int *flag, int *i_, int *j_, int *k_, int *np)
{
  const Grid G = dir == 0 ? P.fx : (dir == 1 ? P.fy : P.fz);
  int i = (int) (t % G.sy), j = (int) ((t / G.sy) % (G.sz / G.sy)),
      k = (int) (t / G.sz);
  if (dir == 0 && (i < 1 || i > P.nx + 1 || j > P.ny + 1 || k > P.nz + 1)) return false;
  if (dir == 1 && (j < 1 || j > P.ny + 1 || i > P.nx + 1 || k > P.nz + 1)) return false;
  if (dir == 2 && (k < 1 || k > P.nz + 1 || i > P.nx + 1 || j > P.ny + 1)) return false;
  long c1 = gidx(P.cc, i, j, k);
  long c0 = c1 - (dir == 0 ? 1 : (dir == 1 ? P.cc.sy : P.cc.sz));
  int p0 = phase[c0], p1 = phase[c1];
  if (!(flag[gidx(G, i, j, k)] < 1 && (p0 > -1 || p1 > -1))) return false;
  *i_ = i; *j_ = j; *k_ = k;
  *np = p0 > -1 ? p0 : p1;
  return true;
}

template <int dir>
__global__ void face_list_kernel(PDom P, const int *__restrict__ phase,
  const int *__restrict__ flag, int *__restrict__ count, int *__restrict__ list)
{
  const Grid G = dir == 0 ? P.fx : (dir == 1 ? P.fy : P.fz);
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.n) return;
  int i, j, k, np;
  if (!bc_face<dir>(P, t, phase, flag, &i, &j, &k, &np)) return;
  list[atomicAdd(&count[dir], 1)] = (int) t;
}

/* One wavefront per workgroup: the Lamb series needs more registers than
 * the default launch bound allows without spilling. */
template <int dir>
__global__ void __launch_bounds__(64) part_bc_kernel(PDom P, double *vel,
  const int *__restrict__ phase, const int *__restrict__ flag,
  const part_struct *__restrict__ parts, double nu,
  const int *__restrict__ list, int nlist, double *__restrict__ vals)
{
  const Grid G = dir == 0 ? P.fx : (dir == 1 ? P.fy : P.fz);
  long e = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nlist) return;
  long t = list[e];
  int i, j, k, np;
  if (!bc_face<dir>(P, t, phase, flag, &i, &j, &k, &np)) return;
  const part_struct *Q = &parts[np];
  long C = gidx(G, i, j, k);
  /* Position of the face. */
  double x = dir == 0 ? (i - DOM_BUF) * P.dx + P.xs : (i - 0.5) * P.dx + P.xs;
  double y = dir == 1 ? (j - DOM_BUF) * P.dy + P.ys : (j - 0.5) * P.dy + P.ys;
  double z = dir == 2 ? (k - DOM_BUF) * P.dz + P.zs : (k - 0.5) * P.dz + P.zs;
  x -= near_img(x, Q->x, P.xl, P.perx);
  y -= near_img(y, Q->y, P.yl, P.pery);
  z -= near_img(z, Q->z, P.zl, P.perz);
  double r, theta, phi;
  xyz2rtp(x, y, z, &r, &theta, &phi);
  double Ux, Uy, Uz;
  double a = Q->r;
  lamb_vel(Q->order, a, r, theta, phi, Q, nu, &Ux, &Uy, &Uz);
  double val;
  if (dir == 0) {
    double ocrossr_x = Q->oy*z - Q->oz*y;
    double odotcrossr_x = Q->oydot*z - Q->ozdot*y;
    Ux += Q->u + ocrossr_x;
    Ux += 0.1/nu *(r*r*r*r*r-a*a*a*a*a)/(r*r*r) * odotcrossr_x;
    val = Ux;
  } else if (dir == 1) {
    double ocrossr_y = -(Q->ox*z - Q->oz*x);
    double odotcrossr_y = -(Q->oxdot*z - Q->ozdot*x);
    Uy += Q->v + ocrossr_y;
    Uy += 0.1/nu *(r*r*r*r*r-a*a*a*a*a)/(r*r*r) * odotcrossr_y;
    val = Uy;
  } else {
    double ocrossr_z = Q->ox*y - Q->oy*x;
    double odotcrossr_z = Q->oxdot*y - Q->oydot*x;
    Uz += Q->w + ocrossr_z;
    Uz += 0.1/nu *(r*r*r*r*r-a*a*a*a*a)/(r*r*r) * odotcrossr_z;
    val = Uz;
  }
  vel[C] = val;
  vals[e] = val;
}

/* The values of the last part_bc_kernel pass again: a face's value depends
 * only on the particle state, which the four BC calls of a Lamb pass
 * share (face index t of the list is the face's array index). */
__global__ void part_bc_apply_kernel(double *vel, const int *__restrict__ list,
  int nlist, const double *__restrict__ vals)
{
  long e = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nlist) return;
  vel[list[e]] = vals[e];
}

/* Face lists of the current cages (device), their lengths (host), and the
 * BC values of the listed faces for the current particle state: valid
 * until a function that changes the particles or the cages runs. */
static int *_bc_list[3] = {NULL, NULL, NULL};
static int *_bc_count = NULL;
static int bc_n[3] = {0, 0, 0};
static double *_bc_val[3] = {NULL, NULL, NULL};
static int bc_vals_ok = 0;

/* Called by cuda_build_cages (bb_bc.hip) once every flag is final. */
void bb_particle_face_lists(void)
{
  PDom P = pdom();
  const long gn[3] = {ddom.fx.n, ddom.fy.n, ddom.fz.n};
  if (!_bc_count) {
    BB_CHECK(hipMalloc((void **) &_bc_count, 3 * sizeof(int)));
    for (int d = 0; d < 3; d++) {
      BB_CHECK(hipMalloc((void **) &_bc_list[d], gn[d] * sizeof(int)));
      BB_CHECK(hipMalloc((void **) &_bc_val[d], gn[d] * sizeof(double)));
    }
  }
  bc_vals_ok = 0;
  BB_CHECK(hipMemsetAsync(_bc_count, 0, 3 * sizeof(int), bb_stream));
  hipLaunchKernelGGL(face_list_kernel<0>, dim3(nblk(gn[0], 256)), dim3(256),
    0, bb_stream, P, _phase, _flag_u, _bc_count, _bc_list[0]);
  hipLaunchKernelGGL(face_list_kernel<1>, dim3(nblk(gn[1], 256)), dim3(256),
    0, bb_stream, P, _phase, _flag_v, _bc_count, _bc_list[1]);
  hipLaunchKernelGGL(face_list_kernel<2>, dim3(nblk(gn[2], 256)), dim3(256),
    0, bb_stream, P, _phase, _flag_w, _bc_count, _bc_list[2]);
  BB_LAUNCH_CHECK();
  BB_CHECK(hipMemcpyAsync(bc_n, _bc_count, 3 * sizeof(int),
    hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
}

static void free_face_lists(void)
{
  if (!_bc_count) return;
  for (int d = 0; d < 3; d++) {
    BB_CHECK(hipFree(_bc_list[d]));
    BB_CHECK(hipFree(_bc_val[d]));
    _bc_list[d] = NULL;
    _bc_val[d] = NULL;
    bc_n[d] = 0;
  }
  bc_vals_ok = 0;
  BB_CHECK(hipFree(_bc_count));
  _bc_count = NULL;
}

static void part_bc(double *u, double *v, double *w)
{
  PDom P = pdom();
  double *vel[3] = {u, v, w};
  if (bc_vals_ok) {
    for (int d = 0; d < 3; d++)
      if (bc_n[d] > 0)
        hipLaunchKernelGGL(part_bc_apply_kernel, dim3(nblk(bc_n[d], 256)),
          dim3(256), 0, bb_stream, vel[d], _bc_list[d], bc_n[d], _bc_val[d]);
    BB_LAUNCH_CHECK();
    return;
  }
  if (bc_n[0] > 0)
    hipLaunchKernelGGL(part_bc_kernel<0>, dim3(nblk(bc_n[0], 64)), dim3(64),
      0, bb_stream, P, u, _phase, _flag_u, _parts, nu, _bc_list[0], bc_n[0],
      _bc_val[0]);
  if (bc_n[1] > 0)
    hipLaunchKernelGGL(part_bc_kernel<1>, dim3(nblk(bc_n[1], 64)), dim3(64),
      0, bb_stream, P, v, _phase, _flag_v, _parts, nu, _bc_list[1], bc_n[1],
      _bc_val[1]);
  if (bc_n[2] > 0)
    hipLaunchKernelGGL(part_bc_kernel<2>, dim3(nblk(bc_n[2], 64)), dim3(64),
      0, bb_stream, P, w, _phase, _flag_w, _parts, nu, _bc_list[2], bc_n[2],
      _bc_val[2]);
  BB_LAUNCH_CHECK();
  bc_vals_ok = 1;
}

extern "C" void cuda_part_BC(void)
{
  if (nparts > 0) part_bc(_u, _v, _w);
}

extern "C" void cuda_part_BC_star(void)
{
  if (nparts > 0) part_bc(_u_star, _v_star, _w_star);
}

/* The pressure right-hand side in particle cells is zero
 * (src/particle_kernel.cu:1748-1753 masks it by phase < 0 and the
 * refinement of src/solver_kernel.cu:212-256 then adds only zeros); the
 * PCG applies that mask itself through _bb_fluid. */
extern "C" void cuda_part_BC_p(void) {}

/* p inside particles: the n = 0 Lamb term plus the rest-frame terms
 * (src/particle_kernel.cu:1758-1837). */
__global__ void part_p_fill_kernel(PDom P, double *p,
  const int *__restrict__ phase, const part_struct *__restrict__ parts,
  double mu, double nu, double rho_f, gradP_struct gradP, double s_beta,
  double s_ref, g_struct g)
{
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.cc.n) return;
  int i = (int) (t % P.cc.sy), j = (int) ((t / P.cc.sy) % (P.cc.sz / P.cc.sy)),
      k = (int) (t / P.cc.sz);
  if (i < 1 || i > P.nx || j < 1 || j > P.ny || k < 1 || k > P.nz) return;
  int np = phase[t];
  if (np < 0) return;
  const part_struct *Q = &parts[np];
  double x = (i-0.5) * P.dx + P.xs;
  double y = (j-0.5) * P.dy + P.ys;
  double z = (k-0.5) * P.dz + P.zs;
  x -= near_img(x, Q->x, P.xl, P.perx);
  y -= near_img(y, Q->y, P.yl, P.pery);
  z -= near_img(z, Q->z, P.zl, P.perz);
  double r, theta, phi;
  xyz2rtp(x, y, z, &r, &theta, &phi);
  double a = Q->r;
  double ox = Q->ox, oy = Q->oy, oz = Q->oz;
  double pp_tmp = Xn(0, theta, phi, Q->pnm_re, Q->pnm_im);
  pp_tmp *= mu*nu/(a*a);
  double ocrossr2 = (oy*z - oz*y) * (oy*z - oz*y);
  ocrossr2 += (ox*z - oz*x) * (ox*z - oz*x);
  ocrossr2 += (ox*y - oy*x) * (ox*y - oy*x);
  double rhoV = rho_f;
  double bousiq_x = -s_beta*(Q->s - s_ref)*g.x;
  double bousiq_y = -s_beta*(Q->s - s_ref)*g.y;
  double bousiq_z = -s_beta*(Q->s - s_ref)*g.z;
  double accdotr = (-gradP.x/rhoV - Q->udot + bousiq_x)*x +
                   (-gradP.y/rhoV - Q->vdot + bousiq_y)*y +
                   (-gradP.z/rhoV - Q->wdot + bousiq_z)*z;
  pp_tmp += 0.5 * rho_f * ocrossr2 + rho_f * accdotr;
  p[t] = pp_tmp;
}

extern "C" void cuda_part_p_fill(void)
{
  if (nparts == 0) return;
  hipLaunchKernelGGL(part_p_fill_kernel, dim3(nblk(ddom.cc.n, 256)),
    dim3(256), 0, bb_stream, pdom(), _p, _phase, _parts, mu, nu, rho_f, gradP,
    s_beta, s_ref, g);
  BB_LAUNCH_CHECK();
}

/* Rigid-body velocity on faces inside a particle
 * (src/particle_kernel.cu:2172-2288); interior faces only. */
template <int dir>
__global__ void internal_kernel(PDom P, double *vel,
  const int *__restrict__ phase, const int *__restrict__ flag,
  const part_struct *__restrict__ parts)
{
  const Grid G = dir == 0 ? P.fx : (dir == 1 ? P.fy : P.fz);
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.n) return;
  int i = (int) (t % G.sy), j = (int) ((t / G.sy) % (G.sz / G.sy)),
      k = (int) (t / G.sz);
  int ia = dir == 0 ? P.nx + 1 : P.nx, ja = dir == 1 ? P.ny + 1 : P.ny,
      ka = dir == 2 ? P.nz + 1 : P.nz;
  if (i < 1 || i > ia || j < 1 || j > ja || k < 1 || k > ka) return;
  long c1 = gidx(P.cc, i, j, k);
  long c0 = c1 - (dir == 0 ? 1 : (dir == 1 ? P.cc.sy : P.cc.sz));
  int p0 = phase[c0], p1 = phase[c1];
  long C = gidx(G, i, j, k);
  int f = flag[C];
  if (!(p0 > -1 && p1 > -1 && f != -1)) return;
  const part_struct *Q = &parts[p1];
  double x = (i-0.5) * P.dx + P.xs;
  double y = (j-0.5) * P.dy + P.ys;
  double z = (k-0.5) * P.dz + P.zs;
  if (dir == 0) {
    double ry = y - near_img(y, Q->y, P.yl, P.pery);
    double rz = z - near_img(z, Q->z, P.zl, P.perz);
    vel[C] = (Q->oy*rz - Q->oz*ry) + Q->u;
  } else if (dir == 1) {
    double rx = x - near_img(x, Q->x, P.xl, P.perx);
    double rz = z - near_img(z, Q->z, P.zl, P.perz);
    vel[C] = (Q->oz*rx - Q->ox*rz) + Q->v;
  } else {
    double rx = x - near_img(x, Q->x, P.xl, P.perx);
    double ry = y - near_img(y, Q->y, P.yl, P.pery);
    vel[C] = (Q->ox*ry - Q->oy*rx) + Q->w;
  }
}

extern "C" void cuda_parts_internal(void)
{
  if (nparts == 0) return;
  PDom P = pdom();
  hipLaunchKernelGGL(internal_kernel<0>, dim3(nblk(ddom.fx.n, 256)), dim3(256),
    0, bb_stream, P, _u, _phase, _flag_u, _parts);
  hipLaunchKernelGGL(internal_kernel<1>, dim3(nblk(ddom.fy.n, 256)), dim3(256),
    0, bb_stream, P, _v, _phase, _flag_v, _parts);
  hipLaunchKernelGGL(internal_kernel<2>, dim3(nblk(ddom.fz.n, 256)), dim3(256),
    0, bb_stream, P, _w, _phase, _flag_w, _parts);
  BB_LAUNCH_CHECK();
}

/* ------------------------------ Physalis --------------------------------- */
/* The 26-node Lebedev rule (src/cuda_physalis.cu:61-135): 6 axis nodes
 * (weight A1), 12 edge midpoints (A2), 8 cube corners (A3); the pole nodes
 * sit at theta = 0 + DIV_ST and PI - DIV_ST. */
extern "C" void cuda_init_physalis(void)
{
  bc_vals_ok = 0;                /* particle state changes */
  if (NPARTS == 0) return;
  int nn[NCOEFFS] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4};
  int mm[NCOEFFS] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0, 1, 2, 3, 4};
  double PI14 = 0.25 * PI, PI12 = 0.5 * PI, PI34 = 0.75 * PI;
  double PI54 = 1.25 * PI, PI32 = 1.5 * PI, PI74 = 1.75 * PI;
  double alph1 = 0.955316618124509;
  double alph2 = 2.186276035465284;
  double A1 = 0.598398600683775;
  double A2 = 0.478718880547015;
  double A3 = 0.403919055461543;
  double t[NNODES] = {PI12, PI12, PI12, PI12, 0.+DIV_ST, PI-DIV_ST,
                      PI12, PI12, PI12, PI12, PI14, PI14, PI14, PI14,
                      PI34, PI34, PI34, PI34,
                      alph1, alph1, alph1, alph1, alph2, alph2, alph2, alph2};
  double p[NNODES] = {0., PI12, PI, PI32, 0., 0.,
                      PI14, PI34, PI54, PI74, 0., PI12, PI, PI32,
                      0., PI12, PI, PI32,
                      PI14, PI34, PI54, PI74, PI14, PI34, PI54, PI74};
  BB_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_nn), nn, sizeof(nn)));
  BB_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_mm), mm, sizeof(mm)));
  BB_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_node_t), t, sizeof(t)));
  BB_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_node_p), p, sizeof(p)));
  BB_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_A1), &A1, sizeof(double)));
  BB_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_A2), &A2, sizeof(double)));
  BB_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_A3), &A3, sizeof(double)));
}

/* floor(arg), or the nearest integer when arg is within DBL_EPSILON of it
 * (src/physalis_kernel.cu:286-303). */
__device__ inline int cell_of(double arg)
{
  int i = (int) floor(arg);
  int rnd = (int) lrint(arg);
  if (rnd != i && fabs(rnd - arg) <= DBL_EPSILON) i = rnd;
  return i;
}

/* Node coordinate of one direction: the node position xp + X for the image
 * of the particle that puts the node's cell inside 1..n (periodic), and that
 * cell.  Between walls the position of the particle itself. */
__device__ inline double node_coord(double xp, double X, double s,
  double d, double l, int n, int per, int *cell)
{
  double x = xp + X;
  int i = cell_of((x - (s - d)) * (1. / d));
  if (per && i < 1) {
    x = xp + (X + l);
    i = cell_of((x - (s - d)) * (1. / d));
  } else if (per && i > n) {
    x = xp + (X - l);
    i = cell_of((x - (s - d)) * (1. / d));
  }
  *cell = i;
  return x;
}

/* Node codes and rest-frame field values at the 26 nodes of each particle
 * (src/physalis_kernel.cu:118-640).  One workgroup per particle, lanes
 * 0..25 are the nodes. */
__device__ inline void cart2sphere(double u, double v, double w, double theta,
  double phi, double *ur, double *ut, double *up)
{
  double st = sin(theta);
  double ct = cos(theta);
  double sp = sin(phi);
  double cp = cos(phi);
  *ur = st * (u * cp + v * sp) + w * ct;
  *ut = ct * (u * cp + v * sp) - w * st;
  *up = -u * sp + v * cp;
}

/* Value of a field at node (x, y, z) from cell/face (i, j, k) and its six
 * neighbours: f[C] + central gradient . (x - x_C)
 * (src/physalis_kernel.cu:334-340 and the u, v, w forms). */
__device__ inline double node_interp(const double *__restrict__ f,
  const Grid &G, int i, int j, int k, double x, double y, double z,
  double xx, double yy, double zz, double ddx, double ddy, double ddz)
{
  long C = gidx(G, i, j, k);
  double dfdx = 0.5*(f[C + 1] - f[C - 1]) * ddx;
  double dfdy = 0.5*(f[C + G.sy] - f[C - G.sy]) * ddy;
  double dfdz = 0.5*(f[C + G.sz] - f[C - G.sz]) * ddz;
  return f[C] + dfdx*(x - xx) + dfdy*(y - yy) + dfdz*(z - zz);
}

struct NodeVals {
  double pp[NNODES], ur[NNODES], ut[NNODES], up[NNODES];
};

__device__ void node_values(const PDom &P, const PWalls &W, part_struct *Q,
  int node,
  const double *__restrict__ p, const double *__restrict__ u,
  const double *__restrict__ v, const double *__restrict__ w, double rho_f,
  double nu, gradP_struct gradP, double s_beta, double s_ref, g_struct g,
  NodeVals *out)
{
  double ddx = 1. / P.dx, ddy = 1. / P.dy, ddz = 1. / P.dz;
  double irho_f = 1. / rho_f;
  double inu = 1. / nu;
  double ox = Q->ox, oy = Q->oy, oz = Q->oz;
  double oxdot = Q->oxdot, oydot = Q->oydot, ozdot = Q->ozdot;
  double udot = Q->udot, vdot = Q->vdot, wdot = Q->wdot;
  double rs2 = Q->rs * Q->rs;
  double rs3 = rs2 * Q->rs;
  double rs5 = rs3 * rs2;
  double irs3 = 1./rs3;
  double a5 = Q->r * Q->r;
  a5 *= a5 * Q->r;

  double th = c_node_t[node], ph = c_node_p[node];
  double xp = Q->rs * sin(th) * cos(ph);
  double yp = Q->rs * sin(th) * sin(ph);
  double zp = Q->rs * cos(th);
  int ic, jc, kc;
  double x = node_coord(xp, Q->x, P.xs, P.dx, P.xl, P.nx, P.perx, &ic);
  double y = node_coord(yp, Q->y, P.ys, P.dy, P.yl, P.ny, P.pery, &jc);
  double z = node_coord(zp, Q->z, P.zs, P.dz, P.zl, P.nz, P.perz, &kc);
  int oobi = ic < 1 || ic > P.nx, oobj = jc < 1 || jc > P.ny,
      oobk = kc < 1 || kc > P.nz;
  int oob = oobi || oobj || oobk;
  /* node code (src/physalis_kernel.cu:137-213): the first face, in the
   * order W E S N B T, that the node lies beyond; a periodic face has no
   * code, and a periodic coordinate is already inside the domain */
  {
    const bool beyond[6] = {x - P.xs < 0, x - P.xe >= 0, y - P.ys < 0,
                            y - P.ye >= 0, z - P.zs < 0, z - P.ze >= 0};
    int code = -1;
    for (int f = 0; f < 6; f++)
      if (code == -1 && beyond[f]) code = W.code[f];
    Q->nodes[node] = code;
  }
  const int ncode = Q->nodes[node];
  const int wall = ncode < -1;
  /* the value a node beyond a face takes: that face's Dirichlet velocity
   * (0 at a Neumann face) in the rest frame, when the node is out of
   * bounds in that one direction only (src/physalis_kernel.cu:437-460) */
  const int wi = oobi * (1-oobj) * (1-oobk) * wall;
  const int wj = (1-oobi) * oobj * (1-oobk) * wall;
  const int wk = (1-oobi) * (1-oobj) * oobk * wall;

  /* pressure: cell centre (src/physalis_kernel.cu:305-364) */
  int i = ic, j = jc, k = kc;
  i += (1 - i) * (i < 1); j += (1 - j) * (j < 1); k += (1 - k) * (k < 1);
  i += (P.nx - i) * (i > P.nx); j += (P.ny - j) * (j > P.ny);
  k += (P.nz - k) * (k > P.nz);
  double pp = node_interp(p, P.cc, i, j, k, x, y, z,
    (i - 0.5) * P.dx + P.xs, (j - 0.5) * P.dy + P.ys,
    (k - 0.5) * P.dz + P.zs, ddx, ddy, ddz);
  double ocrossr2 = (oy*zp - oz*yp) * (oy*zp - oz*yp);
  ocrossr2 += (ox*zp - oz*xp) * (ox*zp - oz*xp);
  ocrossr2 += (ox*yp - oy*xp) * (ox*yp - oy*xp);
  double bousiq_x = -s_beta*(Q->s - s_ref)*g.x;
  double bousiq_y = -s_beta*(Q->s - s_ref)*g.y;
  double bousiq_z = -s_beta*(Q->s - s_ref)*g.z;
  double accdotr = (-gradP.x * irho_f - udot + bousiq_x)*xp +
                   (-gradP.y * irho_f - vdot + bousiq_y)*yp +
                   (-gradP.z * irho_f - wdot + bousiq_z)*zp;
  pp -= 0.5 * rho_f * ocrossr2 + rho_f * accdotr;
  pp = pp * (1 - oob) * (Q->nodes[node] == -1);

  /* u: x-faces, the face index clamps to nx + 1 at ">=" as the reference */
  i = ic; j = jc; k = kc;
  i += (1 - i) * (i < 1); j += (1 - j) * (j < 1); k += (1 - k) * (k < 1);
  i += (P.nx + 1 - i) * (i >= P.nx + 1); j += (P.ny - j) * (j > P.ny);
  k += (P.nz - k) * (k > P.nz);
  double uu = node_interp(u, P.fx, i, j, k, x, y, z,
    (i - DOM_BUF) * P.dx + P.xs, (j - 0.5) * P.dy + P.ys,
    (k - 0.5) * P.dz + P.zs, ddx, ddy, ddz);
  double ocrossr_x = oy*zp - oz*yp;
  double odotcrossr_x = oydot*zp - ozdot*yp;
  double tmp_u = Q->u + ocrossr_x +
                 0.1 * inu * (rs5 - a5) * irs3 * odotcrossr_x;
  uu -= tmp_u;
  double uwi = (ncode == WEST_WALL_D) * W.vel[0][0]
             + (ncode == EAST_WALL_D) * W.vel[1][0];
  double uwj = (ncode == SOUTH_WALL_D) * W.vel[2][0]
             + (ncode == NORTH_WALL_D) * W.vel[3][0];
  double uwk = (ncode == BOTTOM_WALL_D) * W.vel[4][0]
             + (ncode == TOP_WALL_D) * W.vel[5][0];
  uwi -= tmp_u;
  uwj -= tmp_u;
  uwk -= tmp_u;
  uu = (1-oobi) * (1-oobj) * (1-oobk) * (ncode == -1) * uu
     + wi * uwi + wj * uwj + wk * uwk;

  /* v: y-faces */
  i = ic; j = jc; k = kc;
  i += (1 - i) * (i < 1); j += (1 - j) * (j < 1); k += (1 - k) * (k < 1);
  i += (P.nx - i) * (i > P.nx); j += (P.ny + 1 - j) * (j >= P.ny + 1);
  k += (P.nz - k) * (k > P.nz);
  double vv = node_interp(v, P.fy, i, j, k, x, y, z,
    (i - 0.5) * P.dx + P.xs, (j - DOM_BUF) * P.dy + P.ys,
    (k - 0.5) * P.dz + P.zs, ddx, ddy, ddz);
  double ocrossr_y = -(ox*zp - oz*xp);
  double odotcrossr_y = -(oxdot*zp - ozdot*xp);
  double tmp_v = Q->v + ocrossr_y +
                 0.1 * inu * (rs5 - a5) * irs3 * odotcrossr_y;
  vv -= tmp_v;
  double vwi = (ncode == WEST_WALL_D) * W.vel[0][1]
             + (ncode == EAST_WALL_D) * W.vel[1][1];
  double vwj = (ncode == SOUTH_WALL_D) * W.vel[2][1]
             + (ncode == NORTH_WALL_D) * W.vel[3][1];
  double vwk = (ncode == BOTTOM_WALL_D) * W.vel[4][1]
             + (ncode == TOP_WALL_D) * W.vel[5][1];
  vwi -= tmp_v;
  vwj -= tmp_v;
  vwk -= tmp_v;
  vv = (1-oobi) * (1-oobj) * (1-oobk) * (ncode == -1) * vv
     + wi * vwi + wj * vwj + wk * vwk;

  /* w: z-faces */
  i = ic; j = jc; k = kc;
  i += (1 - i) * (i < 1); j += (1 - j) * (j < 1); k += (1 - k) * (k < 1);
  i += (P.nx - i) * (i > P.nx); j += (P.ny - j) * (j > P.ny);
  k += (P.nz + 1 - k) * (k >= P.nz + 1);
  double ww = node_interp(w, P.fz, i, j, k, x, y, z,
    (i - 0.5) * P.dx + P.xs, (j - 0.5) * P.dy + P.ys,
    (k - DOM_BUF) * P.dz + P.zs, ddx, ddy, ddz);
  double ocrossr_z = ox*yp - oy*xp;
  double odotcrossr_z = oxdot*yp - oydot*xp;
  double tmp_w = Q->w + ocrossr_z +
                 0.1 * inu * (rs5 - a5) * irs3 * odotcrossr_z;
  ww -= tmp_w;
  double wwi = (ncode == WEST_WALL_D) * W.vel[0][2]
             + (ncode == EAST_WALL_D) * W.vel[1][2];
  double wwj = (ncode == SOUTH_WALL_D) * W.vel[2][2]
             + (ncode == NORTH_WALL_D) * W.vel[3][2];
  double wwk = (ncode == BOTTOM_WALL_D) * W.vel[4][2]
             + (ncode == TOP_WALL_D) * W.vel[5][2];
  wwi -= tmp_w;
  wwj -= tmp_w;
  wwk -= tmp_w;
  ww = (1-oobi) * (1-oobj) * (1-oobk) * (ncode == -1) * ww
     + wi * wwi + wj * wwj + wk * wwk;

  out->pp[node] = pp;
  cart2sphere(uu, vv, ww, th, ph, &out->ur[node], &out->ut[node],
    &out->up[node]);
}

/* Legendre function of the quadrature, same forms as Pnm
 * (src/physalis_kernel.cu:55-116). */
#define pnm Pnm
#define nnm Nnm

/* Quadrature (src/physalis_kernel.cu:642-722), Lamb coefficients
 * (:724-805) and forces (:807-841) for one particle per workgroup. */
__global__ void __launch_bounds__(64)
lamb_kernel(PDom P, PWalls W, part_struct *parts,
  const double *__restrict__ p,
  const double *__restrict__ u, const double *__restrict__ v,
  const double *__restrict__ w, double rho_f, double mu, double nu,
  gradP_struct gradP, double relax, double s_beta, double s_ref, g_struct g)
{
  __shared__ NodeVals N;
  __shared__ double sums[6][NCOEFFS];
  part_struct *Q = &parts[blockIdx.x];
  int t = threadIdx.x;
  if (t < NNODES)
    node_values(P, W, Q, t, p, u, v, w, rho_f, nu, gradP, s_beta, s_ref, g,
                &N);
  __syncthreads();

  int ncoeff = Q->ncoeff;
  if (t < ncoeff) {
    int n = c_nn[t], m = c_mm[t];
    double S[6];
    for (int node = 0; node < NNODES; node++) {
      double theta = c_node_t[node];
      double phi = c_node_p[node];
      double N_nm = nnm(n, m);
      double P_nm = pnm(n, m, theta);
      double P_n1m = pnm(n + 1, m, theta);
      double dPdt = (n - m + 1.)*P_n1m - (n + 1.)*cos(theta)*P_nm;
      double dPdp = m*P_nm;
      double isth = 1./sin(theta);
      double cmphi = cos(m * phi);
      double smphi = sin(m * phi);
      double I[6];
      I[0] = N_nm*P_nm*N.pp[node]*cmphi;
      I[1] = -N_nm*P_nm*N.pp[node]*smphi;
      I[2] = N_nm*isth*(dPdt * N.ut[node] * cmphi - dPdp * N.up[node] * smphi);
      I[3] = N_nm*isth*(-dPdt * N.ut[node] * smphi - dPdp * N.up[node] * cmphi);
      I[4] = N_nm*isth*(dPdp * N.ut[node] * smphi + dPdt * N.up[node] * cmphi);
      I[5] = N_nm*isth*(dPdp * N.ut[node] * cmphi - dPdt * N.up[node] * smphi);
      for (int q = 0; q < 6; q++) {
        if (node == 0) S[q] = I[q] * c_A1;
        else if (node < 6) S[q] += c_A1 * I[q];
        else if (node < 18) S[q] += c_A2 * I[q];
        else S[q] += c_A3 * I[q];
      }
    }
    for (int q = 0; q < 6; q++) sums[q][t] = S[q];

    /* Lamb coefficients with under-relaxation */
    double ars = Q->r / Q->rs;
    double rsa = Q->rs / Q->r;
    double r2 = Q->r * Q->r;
    double inu = 1./nu;
    double imunu = inu / mu;
    int coeff = t;
    double Yre = S[0], Yim = S[1], rDre = S[2], rDim = S[3], xXre = S[4],
           xXim = S[5];
    if (n == 0) {
      Q->pnm_re[coeff] = (1. - relax) * Q->pnm_re0[coeff] +
                         relax * r2 * imunu * Yre;
      Q->pnm_im[coeff] = (1. - relax) * Q->pnm_im0[coeff] +
                         relax * r2 * imunu * Yim;
      Q->phinm_re[coeff] = 0.;
      Q->phinm_im[coeff] = 0.;
      Q->chinm_re[coeff] = 0.;
      Q->chinm_im[coeff] = 0.;
    } else {
      double pow_ars_np1 = pow(ars, n + 1.);
      double pow_ars_2np1 = pow_ars_np1 * pow_ars_np1 * rsa;
      double pow_rsa_nm1 = pow(rsa, n - 1.);
      double pow_rsa_n = pow_rsa_nm1 * rsa;
      double pow_rsa_np1 = pow_rsa_n * rsa;
      double i_np1 = 1./(n + 1.);
      double i_2np3 = 1./(2.*n + 3.);
      double A = (1. - 0.5*n*(2.*n - 1.) * i_np1 * pow_ars_2np1) * pow_rsa_n;
      double B = n*(2.*n - 1.)*(2.*n + 1.) * i_np1*pow_ars_np1;
      double C = 0.25*n*(2.*(n + 3.)*i_2np3
        + (n - 2. - n*(2.*n + 1.)*i_2np3*ars*ars)*pow_ars_2np1)*pow_rsa_np1;
      double D = n*(n + 1. + 0.5*((n - 2.)*(2.*n + 1.)*rsa*rsa
        - n*(2.*n - 1.))*pow_ars_2np1)*pow_rsa_nm1;
      double idet = 1./ (A*D + B*C);
      double pre = (r2*imunu*Yre*D + Q->r*inu*rDre*B) * idet;
      double pim = (r2*imunu*Yim*D + Q->r*inu*rDim*B) * idet;
      double fre = (Q->r*inu*rDre*A - r2*imunu*Yre*C) * idet;
      double fim = (Q->r*inu*rDim*A - r2*imunu*Yim*C) * idet;
      double E = n*(n + 1.)*(pow_ars_2np1 - 1.)*pow_rsa_n;
      double iE = 1./ E;
      double cre = Q->r*inu*xXre * iE;
      double cim = Q->r*inu*xXim * iE;
      Q->pnm_re[coeff] = Q->pnm_re0[coeff]*(1. - relax) + relax*pre;
      Q->pnm_im[coeff] = Q->pnm_im0[coeff]*(1. - relax) + relax*pim;
      Q->phinm_re[coeff] = Q->phinm_re0[coeff]*(1. - relax) + relax*fre;
      Q->phinm_im[coeff] = Q->phinm_im0[coeff]*(1. - relax) + relax*fim;
      Q->chinm_re[coeff] = Q->chinm_re0[coeff]*(1. - relax) + relax*cre;
      Q->chinm_im[coeff] = Q->chinm_im0[coeff]*(1. - relax) + relax*cim;
    }
  }
  __syncthreads();
  if (t == 0) {
    double irho_f = 1./ rho_f;
    double vol = 4./3. * PI *  Q->r*Q->r*Q->r;
    double N10 = sqrt(3./4./PI);
    double N11 = sqrt(3./8./PI);
    double bousiq_x = -s_beta*(Q->s - s_ref)*g.x;
    double bousiq_y = -s_beta*(Q->s - s_ref)*g.y;
    double bousiq_z = -s_beta*(Q->s - s_ref)*g.z;
    Q->Fx = rho_f * vol * (Q->udot + gradP.x * irho_f - bousiq_x)
      - PI * mu * nu * 2.*N11 * (Q->pnm_re[2] + 6.*Q->phinm_re[2]);
    Q->Fy = rho_f * vol * (Q->vdot + gradP.y * irho_f - bousiq_y)
      + PI * mu * nu * 2.*N11 * (Q->pnm_im[2] + 6.*Q->phinm_im[2]);
    Q->Fz = rho_f * vol * (Q->wdot + gradP.z * irho_f - bousiq_z)
      + PI * mu * nu * N10 * (Q->pnm_re[1] + 6.*Q->phinm_re[1]);
    Q->Lx = rho_f * vol * Q->r*Q->r * Q->oxdot
      - 8. * PI * mu * nu * 2.*N11 * Q->r * Q->chinm_re[2];
    Q->Ly = rho_f * vol * Q->r*Q->r * Q->oydot
      + 8. * PI * mu * nu * 2.*N11 * Q->r * Q->chinm_im[2];
    Q->Lz = rho_f * vol * Q->r*Q->r * Q->ozdot
      + 8. * PI * mu * nu * N10 * Q->r * Q->chinm_re[1];
  }
}

extern "C" void cuda_lamb(void)
{
  bc_vals_ok = 0;                /* particle state changes */
  if (nparts == 0) return;
  hipLaunchKernelGGL(lamb_kernel, dim3(nparts), dim3(64), 0, bb_stream,
    pdom(), pwalls(), _parts, _p, _u, _v, _w, rho_f, mu, nu, gradP,
    lamb_relax, s_beta, s_ref, g);
  BB_LAUNCH_CHECK();
}

/* Largest relative coefficient change of each particle
 * (src/physalis_kernel.cu:1559-1633). */
__global__ void __launch_bounds__(64)
lamb_error_kernel(double lamb_cut, int ncoeffs_max, part_struct *parts,
  double *perr, int *pnum)
{
  __shared__ double s_max[MAX_COEFFS];
  part_struct *Q = &parts[blockIdx.x];
  int coeff = threadIdx.x;
  if (coeff < ncoeffs_max) {
    const double *cur[6] = {Q->pnm_re, Q->pnm_im, Q->phinm_re, Q->phinm_im,
                            Q->chinm_re, Q->chinm_im};
    const double *old[6] = {Q->pnm_re0, Q->pnm_im0, Q->phinm_re0,
                            Q->phinm_im0, Q->chinm_re0, Q->chinm_im0};
    double m = DBL_MIN;
    for (int i = 0; i < 6; i++) {
      double curr_val = cur[i][coeff];
      double zeroth_val = cur[i][0];
      int flag = (fabs(curr_val) > fabs(lamb_cut*zeroth_val)) *
                 (fabs(curr_val) > 1.e-16) *
                 (fabs(zeroth_val) > DBL_MIN);
      double div = fabs(curr_val);
      div += (1.e-16 - div) * (div < 1.e-16);
      double e = (double) flag * fabs(curr_val - old[i][coeff]) / div;
      m += (e - m) * (e > m);
    }
    s_max[coeff] = m;
  }
  __syncthreads();
  if (coeff == 0) {
    double mx = DBL_MIN;
    for (int i = 0; i < ncoeffs_max; i++) mx += (s_max[i] - mx) * (s_max[i] > mx);
    perr[blockIdx.x] = mx;
    pnum[blockIdx.x] = Q->N;
  }
}

/* Maximum over particles, first index on ties (as thrust::max_element),
 * then c0 <- c (src/physalis_kernel.cu:1635-1648). */
__global__ void __launch_bounds__(256)
lamb_max_store_kernel(int nparts, int ncoeffs_max, part_struct *parts,
  const double *__restrict__ perr, const int *__restrict__ pnum, double *out)
{
  if (threadIdx.x == 0) {
    int pos = 0;
    for (int n = 1; n < nparts; n++) if (perr[n] > perr[pos]) pos = n;
    out[0] = perr[pos];
    out[1] = (double) pnum[pos];
  }
  for (int t = threadIdx.x; t < nparts * ncoeffs_max; t += blockDim.x) {
    part_struct *Q = &parts[t / ncoeffs_max];
    int c = t % ncoeffs_max;
    Q->pnm_re0[c] = Q->pnm_re[c];
    Q->pnm_im0[c] = Q->pnm_im[c];
    Q->phinm_re0[c] = Q->phinm_re[c];
    Q->phinm_im0[c] = Q->phinm_im[c];
    Q->chinm_re0[c] = Q->chinm_re[c];
    Q->chinm_im0[c] = Q->chinm_im[c];
  }
}

extern "C" void cuda_lamb_err(real *error, int *number)
{
  bc_vals_ok = 0;                /* particle state changes */
  if (nparts == 0) {
    /* The reference leaves its result undefined here; no particles means
     * the Lamb loop has nothing to converge. */
    *error = 0.;
    *number = -1;
    return;
  }
  hipLaunchKernelGGL(lamb_error_kernel, dim3(nparts), dim3(64), 0, bb_stream,
    lamb_cut, ncoeffs_max, _parts, _perr, _pnum);
  hipLaunchKernelGGL(lamb_max_store_kernel, dim3(1), dim3(256), 0, bb_stream,
    nparts, ncoeffs_max, _parts, _perr, _pnum, _err_out);
  BB_LAUNCH_CHECK();
  BB_CHECK(hipMemcpyAsync(err_host, _err_out, 2 * sizeof(double),
    hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
  *error = err_host[0];
  *number = (int) err_host[1];
}

/* ---------------------------- particle heat ----------------------------- */
/* The scalar (heat) exchange of each particle with the fluid
 * (src/scalar_kernel.cu:303-577, 1143-1360; src/cuda_scalar.cu:866-969):
 * the scalar on a sphere of radius srs at the 26 nodes, its spherical
 * harmonic coefficients anm (n <= s_order, -n <= m <= n), the heat flux
 * q = 2 sqrt(pi) a a_00, the analytic scalar in particle cells, and the
 * particle's scalar from q.  The reference's Legendre table has its m < 0
 * cases commented out and returns 0 for them through its switch
 * (src/particle_kernel.cu:1928-1990): every m < 0 term is 0, here as
 * there (Pnm returns 0 for m < 0). */
__constant__ int c_snn[S_MAX_COEFFS] = {0, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3,
  3, 3, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4};
__constant__ int c_smm[S_MAX_COEFFS] = {0, -1, 0, 1, -2, -1, 0, 1, 2, -3, -2,
  -1, 0, 1, 2, 3, -4, -3, -2, -1, 0, 1, 2, 3, 4};

/* the scalar BC types and values of the six faces */
struct PScalarBC {
  int type[6];
  double val[6];
};

static PScalarBC pscalar_bc(void)
{
  PScalarBC B;
  const int t[6] = {bc_s.sW, bc_s.sE, bc_s.sS, bc_s.sN, bc_s.sB, bc_s.sT};
  const double d[6] = {bc_s.sWD, bc_s.sED, bc_s.sSD, bc_s.sND, bc_s.sBD,
                       bc_s.sTD};
  for (int f = 0; f < 6; f++) {
    B.type[f] = t[f];
    B.val[f] = d[f];
  }
  return B;
}

/* s - s_p at one node (src/scalar_kernel.cu:303-527) */
__device__ double scalar_node_value(const PDom &P, const PScalarBC &B,
  part_struct *Q, int node, const double *__restrict__ s)
{
  double ddx = 1. / P.dx, ddy = 1. / P.dy, ddz = 1. / P.dz;
  double th = c_node_t[node], ph = c_node_p[node];
  double xp = Q->srs * sin(th) * cos(ph);
  double yp = Q->srs * sin(th) * sin(ph);
  double zp = Q->srs * cos(th);
  int ic, jc, kc;
  double x = node_coord(xp, Q->x, P.xs, P.dx, P.xl, P.nx, P.perx, &ic);
  double y = node_coord(yp, Q->y, P.ys, P.dy, P.yl, P.ny, P.pery, &jc);
  double z = node_coord(zp, Q->z, P.zs, P.dz, P.zl, P.nz, P.perz, &kc);
  const bool beyond[6] = {x - P.xs < 0, x - P.xe >= 0, y - P.ys < 0,
                          y - P.ye >= 0, z - P.zs < 0, z - P.ze >= 0};
  int code = -1;
  for (int f = 0; f < 6; f++) {
    if (code != -1 || !beyond[f]) continue;
    if (B.type[f] == DIRICHLET) code = WEST_WALL_D - 10 * f;
    else if (B.type[f] == NEUMANN) code = WEST_WALL_N - 10 * f;
  }
  Q->nodes[node] = code;
  int oobi = ic < 1 || ic > P.nx, oobj = jc < 1 || jc > P.ny,
      oobk = kc < 1 || kc > P.nz;
  int i = ic, j = jc, k = kc;
  i += (1 - i) * (i < 1); j += (1 - j) * (j < 1); k += (1 - k) * (k < 1);
  i += (P.nx - i) * (i > P.nx); j += (P.ny - j) * (j > P.ny);
  k += (P.nz - k) * (k > P.nz);
  long C = gidx(P.cc, i, j, k);
  double ss = node_interp(s, P.cc, i, j, k, x, y, z,
    (i - 0.5) * P.dx + P.xs, (j - 0.5) * P.dy + P.ys,
    (k - 0.5) * P.dz + P.zs, ddx, ddy, ddz) - Q->s;
  double swi = (code == WEST_WALL_D) * B.val[0]
             + (code == EAST_WALL_D) * B.val[1]
             + (code == WEST_WALL_N) * s[C - 1]
             + (code == EAST_WALL_N) * s[C + 1]
             - Q->s;
  double swj = (code == SOUTH_WALL_D) * B.val[2]
             + (code == NORTH_WALL_D) * B.val[3]
             + (code == SOUTH_WALL_N) * s[C - P.cc.sy]
             + (code == NORTH_WALL_N) * s[C + P.cc.sy]
             - Q->s;
  double swk = (code == BOTTOM_WALL_D) * B.val[4]
             + (code == TOP_WALL_D) * B.val[5]
             + (code == BOTTOM_WALL_N) * s[C - P.cc.sz]
             + (code == TOP_WALL_N) * s[C + P.cc.sz]
             - Q->s;
  return (1-oobi) * (1-oobj) * (1-oobk) * (code == -1) * ss
       + oobi * (1-oobj) * (1-oobk) * (code < -1) * swi
       + (1-oobi) * oobj * (1-oobk) * (code < -1) * swj
       + (1-oobi) * (1-oobj) * oobk * (code < -1) * swk;
}

/* Node values, quadrature (nodes in the reference's order and weights),
 * anm = <Y_nm, s - s_p> / (rsa^n - ars^(n+1)) and q, one workgroup per
 * particle (src/scalar_kernel.cu:529-577, 1247-1272). */
__global__ void __launch_bounds__(64)
scalar_lamb_kernel(PDom P, PScalarBC B, part_struct *parts,
  const double *__restrict__ s)
{
  __shared__ double ss[NNODES];
  part_struct *Q = &parts[blockIdx.x];
  int t = threadIdx.x;
  if (t < NNODES) ss[t] = scalar_node_value(P, B, Q, t, s);
  __syncthreads();
  int sncoeff = Q->sncoeff;
  double ars = Q->r / Q->srs;
  double rsa = Q->srs / Q->r;
  if (t < sncoeff) {
    int n = c_snn[t], m = c_smm[t];
    double re = 0., im = 0.;
    for (int node = 0; node < NNODES; node++) {
      double theta = c_node_t[node];
      double phi = c_node_p[node];
      double N_nm = Nnm(n, m);
      double P_nm = m < 0 ? 0. : Pnm(n, m, theta);
      double Ire = N_nm*P_nm*ss[node]*cos(m * phi);
      double Iim = -N_nm*P_nm*ss[node]*sin(m * phi);
      if (node == 0) { re = Ire * c_A1; im = Iim * c_A1; }
      else if (node < 6) { re += c_A1 * Ire; im += c_A1 * Iim; }
      else if (node < 18) { re += c_A2 * Ire; im += c_A2 * Iim; }
      else { re += c_A3 * Ire; im += c_A3 * Iim; }
    }
    double A = pow(rsa, n) - pow(ars, n+1.);
    Q->anm_re[t] = re / A;
    Q->anm_im[t] = im / A;
  }
  __syncthreads();
  if (t == 0 && sncoeff > 0) Q->q = 2. * sqrt(PI) * Q->r * Q->anm_re[0];
}

/* Largest relative change of each particle's anm (SNSP = 2 arrays), then
 * the maximum over particles (thrust::reduce with 0 as the start) and
 * anm0 <- anm (src/scalar_kernel.cu:1274-1349). */
__global__ void __launch_bounds__(64)
scalar_error_kernel(double lamb_cut_s, int s_ncoeffs_max, part_struct *parts,
  double *perr)
{
  __shared__ double s_max[S_MAX_COEFFS];
  part_struct *Q = &parts[blockIdx.x];
  int coeff = threadIdx.x;
  if (coeff < s_ncoeffs_max) {
    const double *cur[2] = {Q->anm_re, Q->anm_im};
    const double *old[2] = {Q->anm_re0, Q->anm_im0};
    double m = DBL_MIN;
    for (int i = 0; i < 2; i++) {
      double curr_val = cur[i][coeff];
      double zeroth_val = cur[i][0];
      int flag = (fabs(curr_val) > fabs(lamb_cut_s*zeroth_val)) *
                 (fabs(curr_val) > 1.e-16) *
                 (fabs(zeroth_val) > DBL_MIN);
      double div = fabs(curr_val);
      div += (1.e-16 - div) * (div < 1.e-16);
      double e = (double) flag * fabs(curr_val - old[i][coeff]) / div;
      m += (e - m) * (e > m);
    }
    s_max[coeff] = m;
  }
  __syncthreads();
  if (coeff == 0) {
    double mx = DBL_MIN;
    for (int i = 0; i < s_ncoeffs_max; i++)
      mx += (s_max[i] - mx) * (s_max[i] > mx);
    perr[blockIdx.x] = mx;
  }
}

__global__ void __launch_bounds__(256)
scalar_max_store_kernel(int nparts, int s_ncoeffs_max, part_struct *parts,
  const double *__restrict__ perr, double *out)
{
  if (threadIdx.x == 0) {
    double best = 0.;
    for (int n = 0; n < nparts; n++) best = perr[n] > best ? perr[n] : best;
    out[0] = best;
  }
  for (int t = threadIdx.x; t < nparts * s_ncoeffs_max; t += blockDim.x) {
    part_struct *Q = &parts[t / s_ncoeffs_max];
    int c = t % s_ncoeffs_max;
    Q->anm_re0[c] = Q->anm_re[c];
    Q->anm_im0[c] = Q->anm_im[c];
  }
}

/* The analytic scalar in particle cells: the series in shell cells, the
 * particle's scalar inside; fluid cells keep theirs
 * (src/scalar_kernel.cu:1143-1227). */
__global__ void __launch_bounds__(256)
scalar_part_bc_kernel(PDom P, const part_struct *__restrict__ parts,
  const int *__restrict__ phase, const int *__restrict__ shell, double *s)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > P.nx) return;
  long C = gidx(P.cc, i, j, k);
  int pn = phase[C];
  if (pn < 0) return;
  const part_struct *Q = &parts[pn];
  if (shell[C] > 0) {
    s[C] = Q->s;
    return;
  }
  double x = (i - 0.5) * P.dx + P.xs;
  double y = (j - 0.5) * P.dy + P.ys;
  double z = (k - 0.5) * P.dz + P.zs;
  x -= near_img(x, Q->x, P.xl, P.perx);
  y -= near_img(y, Q->y, P.yl, P.pery);
  z -= near_img(z, Q->z, P.zl, P.perz);
  double r, theta, phi;
  xyz2rtp(x, y, z, &r, &theta, &phi);
  double ar = Q->r / r;
  double ra = r / Q->r;
  double v = Q->s;
  for (int n = 0; n <= Q->sorder; n++) {
    int coeff = 0;
    for (int jj = 0; jj < n; jj++) coeff += 2*jj + 1;
    double X = 0.;
    for (int m = -n; m <= n; m++) {
      X += Nnm(n, m) * (m < 0 ? 0. : Pnm(n, m, theta))
        * (Q->anm_re[coeff]*cos(m*phi) - Q->anm_im[coeff]*sin(m*phi));
      coeff++;
    }
    v += (pow(ra, n) - pow(ar, n + 1)) * X;
  }
  s[C] = v;
}

__global__ void __launch_bounds__(256)
scalar_part_fill_kernel(PDom P, const part_struct *__restrict__ parts,
  const int *__restrict__ phase, double *s)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > P.nx) return;
  long C = gidx(P.cc, i, j, k);
  int pn = phase[C];
  if (pn > -1) s[C] = parts[pn].s;
}

__global__ void update_part_scalar_kernel(int nparts, part_struct *parts,
  double dt, double s_k)
{
  int pp = threadIdx.x + blockIdx.x * blockDim.x;
  if (pp >= nparts) return;
  part_struct *Q = &parts[pp];
  double vol = 4./3. * PI * Q->r*Q->r*Q->r;
  double m = vol * Q->rho;
  Q->s += (float) Q->update * Q->q * s_k * dt / m / Q->cp;
}

extern "C" void cuda_scalar_part_BC(real *array)
{
  if (nparts == 0) return;
  PDom P = pdom();
  hipLaunchKernelGGL(scalar_part_bc_kernel, dim3((P.nx + 255) / 256, P.ny,
    P.nz), dim3(256), 0, bb_stream, P, _parts, _phase, _phase_shell, array);
  BB_LAUNCH_CHECK();
}

extern "C" void cuda_scalar_part_fill(void)
{
  if (nparts == 0) return;
  PDom P = pdom();
  hipLaunchKernelGGL(scalar_part_fill_kernel, dim3((P.nx + 255) / 256, P.ny,
    P.nz), dim3(256), 0, bb_stream, P, _parts, _phase, _s);
  BB_LAUNCH_CHECK();
}

extern "C" void cuda_scalar_lamb(void)
{
  if (nparts == 0) return;
  hipLaunchKernelGGL(scalar_lamb_kernel, dim3(nparts), dim3(64), 0, bb_stream,
    pdom(), pscalar_bc(), _parts, _s);
  BB_LAUNCH_CHECK();
}

extern "C" real cuda_scalar_lamb_err(void)
{
  /* src/cuda_scalar.cu:914-946 */
  if (nparts == 0) return DBL_MIN;
  hipLaunchKernelGGL(scalar_error_kernel, dim3(nparts), dim3(64), 0,
    bb_stream, lamb_cut_scalar, s_ncoeffs_max, _parts, _perr);
  hipLaunchKernelGGL(scalar_max_store_kernel, dim3(1), dim3(256), 0,
    bb_stream, nparts, s_ncoeffs_max, _parts, _perr, _err_out);
  BB_LAUNCH_CHECK();
  BB_CHECK(hipMemcpyAsync(err_host, _err_out, sizeof(double),
    hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
  return err_host[0];
}

extern "C" void cuda_scalar_update_part(void)
{
  if (nparts == 0) return;
  hipLaunchKernelGGL(update_part_scalar_kernel, dim3(nblk(nparts, 256)),
    dim3(256), 0,
    bb_stream, nparts, _parts, dt, s_k);
  BB_LAUNCH_CHECK();
}

/* ------------------------- collisions and motion ------------------------- */
__global__ void collision_init_kernel(int nparts, part_struct *parts)
{
  int j = threadIdx.x + blockIdx.x * blockDim.x;
  if (j < nparts) {
    part_struct *Q = &parts[j];
    Q->iFx = Q->iFy = Q->iFz = 0.;
    Q->iLx = Q->iLy = Q->iLz = 0.;
  }
}

/* Collision bin of each particle (src/particle_kernel.cu:716-789). */
__global__ void bin_kernel(int nparts, const part_struct *__restrict__ parts,
  bin_struct B, int *pbin)
{
  int pp = threadIdx.x + blockIdx.x * blockDim.x;
  if (pp >= nparts) return;
  double xsb = B.xs - B.dx, ysb = B.ys - B.dy, zsb = B.zs - B.dz;
  double ax = (parts[pp].x - xsb) / B.dx;
  double ay = (parts[pp].y - ysb) / B.dy;
  double az = (parts[pp].z - zsb) / B.dz;
  int ib = (int) floor(ax), jb = (int) floor(ay), kb = (int) floor(az);
  int rx = (int) lrint(ax), ry = (int) lrint(ay), rz = (int) lrint(az);
  if (rx != ib && fabs(rx - ax) <= 10.*DBL_EPSILON) ib = rx;
  if (ry != jb && fabs(ry - ay) <= 10.*DBL_EPSILON) jb = ry;
  if (rz != kb && fabs(rz - az) <= 10.*DBL_EPSILON) kb = rz;
  const bin_grid_info &G = B.Gcc;
  ib = ib < G._isb ? G._isb : (ib > G._ieb ? G._ieb : ib);
  jb = jb < G._jsb ? G._jsb : (jb > G._jeb ? G._jeb : jb);
  kb = kb < G._ksb ? G._ksb : (kb > G._keb ? G._keb : kb);
  pbin[pp] = ib + jb * G.s1b + kb * G.s2b;
}

/* Lubrication and contact forces on particle i from its neighbours
 * (src/particle_kernel.cu:3075-3366).  Neighbours are visited bin by bin
 * (z, then y, then x offsets -1..1, as the reference) and, inside a bin,
 * in particle order; a bin across a periodic face holds the neighbours'
 * images (positions x +- L). */
/* One wavefront per particle i: the lanes test 64 candidates j at a time
 * for membership of the target bin, and lane 0 applies the matches in
 * increasing j, so the forces add up in the order of a scan over j. */
__global__ void __launch_bounds__(WAVE)
collision_parts_kernel(int nparts, part_struct *parts,
  const int *__restrict__ pbin, bin_struct B, double eps, double mu,
  double rhof, double nu, int interaction_length_ratio, double dt,
  double xl, double yl, double zl, int perx, int pery, int perz)
{
  int i = blockIdx.x;
  int lane = threadIdx.x;
  if (i >= nparts) return;
  const bin_grid_info &G = B.Gcc;
  int bin = pbin[i];
  int kb = bin / G.s2b, jb = (bin - kb * G.s2b) / G.s1b,
      ib = bin - kb * G.s2b - jb * G.s1b;
  if (!(ib >= G._is && ib <= G._ie && jb >= G._js && jb <= G._je &&
        kb >= G._ks && kb <= G._ke)) return;
  part_struct *Pi = &parts[i];
  for (int n = -1; n <= 1; n++)
  for (int m = -1; m <= 1; m++)
  for (int l = -1; l <= 1; l++) {
    /* target bin and the image shift of its particles */
    int tb[3] = {ib + l, jb + m, kb + n};
    int lo[3] = {G._is, G._js, G._ks}, hi[3] = {G._ie, G._je, G._ke};
    int per[3] = {perx, pery, perz}, sh[3] = {0, 0, 0};
    bool ok = true;
    for (int d = 0; d < 3; d++) {
      if (tb[d] < lo[d]) {
        if (!per[d]) ok = false;
        tb[d] += hi[d] - lo[d] + 1;
        sh[d] = -1;
      } else if (tb[d] > hi[d]) {
        if (!per[d]) ok = false;
        tb[d] -= hi[d] - lo[d] + 1;
        sh[d] = 1;
      }
    }
    if (!ok) continue;
    int tbin = tb[0] + tb[1] * G.s1b + tb[2] * G.s2b;
    for (int j0 = 0; j0 < nparts; j0 += WAVE) {
      int jl = j0 + lane;
      unsigned long long hits =
        __ballot(jl < nparts && pbin[jl] == tbin && jl != i);
      while (hits) {
      int j = j0 + __ffsll((unsigned long long) hits) - 1;
      hits &= hits - 1;
      if (lane != 0) continue;
      const part_struct *Pj = &parts[j];
      double xj = img(Pj->x, sh[0], xl), yj = img(Pj->y, sh[1], yl),
             zj = img(Pj->z, sh[2], zl);
      double ai = Pi->r;
      double aj = Pj->r;
      double Bq = aj / ai;
      double hN = interaction_length_ratio * Pi->r;
      double ux, uy, uz, rx, ry, rz, r, h, ah, lnah, nx, ny, nz, udotn;
      double unx, uny, unz, utx, uty, utz, ut, tx, ty, tz, t, bx, by, bz, b;
      double omegax, omegay, omegaz, omega;
      double ocrossnx, ocrossny, ocrossnz, utcrossnx, utcrossny, utcrossnz;
      double opB, Fnx, Fny, Fnz, Ftx, Fty, Ftz, Lox, Loy, Loz;
      int q;
      rx = Pi->x - xj;
      ry = Pi->y - yj;
      rz = Pi->z - zj;
      ux = 0.5*((Pi->u - Pj->u) + (Pi->u0 - Pj->u0));
      uy = 0.5*((Pi->v - Pj->v) + (Pi->v0 - Pj->v0));
      uz = 0.5*((Pi->w - Pj->w) + (Pi->w0 - Pj->w0));
      r = sqrt(rx*rx + ry*ry + rz*rz);
      omegax = 0.5*((Pi->ox + Pj->ox) + (Pi->ox0 + Pj->ox0));
      omegay = 0.5*((Pi->oy + Pj->oy) + (Pi->oy0 + Pj->oy0));
      omegaz = 0.5*((Pi->oz + Pj->oz) + (Pi->oz0 + Pj->oz0));
      omega = sqrt(omegax*omegax + omegay*omegay + omegaz*omegaz);
      h = r - ai - aj;
      nx = rx / r;
      ny = ry / r;
      nz = rz / r;
      udotn = ux * nx + uy * ny + uz * nz;
      unx = udotn * nx;
      uny = udotn * ny;
      unz = udotn * nz;
      utx = ux - unx;
      uty = uy - uny;
      utz = uz - unz;
      ut = sqrt(utx*utx + uty*uty + utz*utz);
      if (ut > 0) {
        tx = utx / ut; ty = uty / ut; tz = utz / ut;
        bx = ny*tz - nz*ty; by = -nx*tz + nz*tx; bz = nx*ty - ny*tx;
        b = sqrt(bx*bx + by*by + bz*bz);
        bx = bx / b; by = by / b; bz = bz / b;
      } else if (omega > 0) {
        bx = omegax / omega; by = omegay / omega; bz = omegaz / omega;
        tx = by*nz - bz*ny; ty = -bx*nz + bz*nx; tz = bx*ny - by*nx;
        t = sqrt(tx*tx + ty*ty + tz*tz);
        tx = tx / t; ty = ty / t; tz = tz / t;
      } else {
        tx = 1.; ty = 0.; tz = 0.;
        bx = ny*tz - nz*ty; by = -nx*tz + nz*tx; bz = nx*ty - ny*tx;
        b = sqrt(bx*bx + by*by + bz*bz);
        bx = bx / b; by = by / b; bz = bz / b;
      }
      (void) bx; (void) by; (void) bz; (void) tx; (void) ty; (void) tz;
      opB = 1 + Bq;
      ocrossnx = omegay*nz - omegaz*ny;
      ocrossny = -omegax*nz + omegaz*nx;
      ocrossnz = omegax*ny - omegay*nx;
      utcrossnx = uty*nz - utz*ny;
      utcrossny = -utx*nz + utz*nx;
      utcrossnz = utx*ny - uty*nx;
      if (h < hN && h > 0) {
        q = 0;
        while (q < MAX_NEIGHBORS && Pi->iSt[q] != Pj->N) q++;
        if (q < MAX_NEIGHBORS && Pi->iSt[q] == Pj->N) {
          Pi->iSt[q] = -1;
          Pi->St[q] = 0.;
        }
        if (h < eps*Pi->r) h = eps*Pi->r;
        ah = ai/h - ai/hN;
        lnah = log(hN/h);
        Fnx = -1. * Bq*Bq / (opB*opB) * ah
          - Bq*(1.+7.*Bq+Bq*Bq)/(5.*opB*opB*opB)*lnah;
        Fny = Fnx;
        Fnz = Fnx;
        Fnx *= 6.*PI*mu*ai*unx;
        Fny *= 6.*PI*mu*ai*uny;
        Fnz *= 6.*PI*mu*ai*unz;
        Ftx = -6.*PI*mu*ai*utx*4.*Bq*(2.+Bq+2.*Bq*Bq)/(15.*opB*opB*opB)*lnah;
        Fty = -6.*PI*mu*ai*uty*4.*Bq*(2.+Bq+2.*Bq*Bq)/(15.*opB*opB*opB)*lnah;
        Ftz = -6.*PI*mu*ai*utz*4.*Bq*(2.+Bq+2.*Bq*Bq)/(15.*opB*opB*opB)*lnah;
        Ftx += 8.*PI*mu*ai*ai*ocrossnx*Bq*(4.+Bq)/(10.*opB*opB)*lnah;
        Fty += 8.*PI*mu*ai*ai*ocrossny*Bq*(4.+Bq)/(10.*opB*opB)*lnah;
        Ftz += 8.*PI*mu*ai*ai*ocrossnz*Bq*(4.+Bq)/(10.*opB*opB)*lnah;
        Lox = -8.*PI*mu*ai*ai*utcrossnx*Bq*(4.+Bq)/(10.*opB*opB)*lnah;
        Loy = -8.*PI*mu*ai*ai*utcrossny*Bq*(4.+Bq)/(10.*opB*opB)*lnah;
        Loz = -8.*PI*mu*ai*ai*utcrossnz*Bq*(4.+Bq)/(10.*opB*opB)*lnah;
        Lox += -8.*PI*mu*ai*ai*ai*omegax*2.*Bq/(5.*opB)*lnah;
        Loy += -8.*PI*mu*ai*ai*ai*omegay*2.*Bq/(5.*opB)*lnah;
        Loz += -8.*PI*mu*ai*ai*ai*omegaz*2.*Bq/(5.*opB)*lnah;
      } else {
        ah = 0; lnah = 0;
        Fnx = Fny = Fnz = 0;
        Ftx = Fty = Ftz = 0;
        Lox = Loy = Loz = 0;
      }
      if (h < 0) {
        q = 0;
        while (q < MAX_NEIGHBORS && Pi->iSt[q] != Pj->N) q++;
        if (q == MAX_NEIGHBORS) {
          q = 0;
          while (q < MAX_NEIGHBORS - 1 && Pi->iSt[q] != -1) q++;
          Pi->iSt[q] = Pj->N;
          Pi->St[q] = 1./9.*Pi->rho/rhof*2.*Pi->r*fabs(udotn)/nu;
          Pi->ncoll_part += 1;
        }
        double Vx = -utx + 0.5*(ai + aj + h)*ocrossnx;
        double Vy = -uty + 0.5*(ai + aj + h)*ocrossny;
        double Vz = -utz + 0.5*(ai + aj + h)*ocrossnz;
        double Hi = 0.5*Pi->E/(1.+Pi->sigma);
        double kt = 8./((1.-Pi->sigma*Pi->sigma)/Hi
          +(1.-Pi->sigma*Pi->sigma)/Hi)/sqrt(1./ai)*sqrt(-h);
        double Vdotn = Vx*nx + Vy*ny + Vz*nz;
        double sx = (Vx - Vdotn * nx) * dt;
        double sy = (Vy - Vdotn * ny) * dt;
        double sz = (Vz - Vdotn * nz) * dt;
        ah = 0;
        lnah = 0;
        double k = 4./3./((1.-Pi->sigma*Pi->sigma)/Pi->E
          + (1.-Pj->sigma*Pj->sigma)/Pj->E)/sqrt(1./ai + 1./aj);
        double xcx0 = 1.e-4;
        double e = Pi->e_dry + (1.+Pi->e_dry)/Pi->St[q]*log(xcx0);
        if (e < 0) e = 0;
        double alpha = -2.263*pow(e,0.3948)+2.22;
        double eta = alpha*sqrt(4./3.*PI*ai*ai*ai*Pi->rho*k*sqrt(-h));
        Fnx = (sqrt(-h*h*h)*k - eta*udotn)*nx;
        Fny = (sqrt(-h*h*h)*k - eta*udotn)*ny;
        Fnz = (sqrt(-h*h*h)*k - eta*udotn)*nz;
        double coeff_fric = 0.5 * (Pi->coeff_fric + Pj->coeff_fric);
        Ftx = -kt * sx;
        Fty = -kt * sy;
        Ftz = -kt * sz;
        double Ftdotn = Ftx*nx + Fty*ny + Ftz*nz;
        Ftx = Ftx - Ftdotn * nx;
        Fty = Fty - Ftdotn * ny;
        Ftz = Ftz - Ftdotn * nz;
        double Ft = sqrt(Ftx*Ftx + Fty*Fty + Ftz*Ftz);
        double Fn = sqrt(Fnx*Fnx + Fny*Fny + Fnz*Fnz);
        if (Ft > coeff_fric * Fn) {
          Ftx = coeff_fric * Fn * Ftx / Ft;
          Fty = coeff_fric * Fn * Fty / Ft;
          Ftz = coeff_fric * Fn * Ftz / Ft;
        }
        Lox = -(ai+0.5*h)*((Fny+Fty)*nz-(Fnz+Ftz)*ny);
        Loy =  (ai+0.5*h)*((Fnx+Ftx)*nz-(Fnz+Ftz)*nx);
        Loz = -(ai+0.5*h)*((Fnx+Ftx)*ny-(Fny+Fty)*nx);
      }
      (void) ah; (void) lnah;
      Pi->iFx += Fnx + Ftx;
      Pi->iFy += Fny + Fty;
      Pi->iFz += Fnz + Ftz;
      Pi->iLx += Lox;
      Pi->iLy += Loy;
      Pi->iLz += Loz;
      }
    }
  }
}

/* Spring force toward an anchor (src/particle_kernel.cu:2303-2345). */
__global__ void spring_kernel(int nparts, part_struct *parts, dom_struct D)
{
  int i = threadIdx.x + blockIdx.x * blockDim.x;
  if (i >= nparts || !(parts[i].spring_k > 0.)) return;
  part_struct *Q = &parts[i];
  double x = Q->x, y = Q->y, z = Q->z;
  x += D.xl * ((x < D.xs) - (x > D.xe));
  y += D.yl * ((y < D.ys) - (y > D.ye));
  z += D.zl * ((z < D.zs) - (z > D.ze));
  double nx = x - Q->spring_x;
  double ny = y - Q->spring_y;
  double nz = z - Q->spring_z;
  double n = sqrt(nx*nx+ny*ny+nz*nz);
  double lx = Q->spring_l * (nx / n);
  double ly = Q->spring_l * (ny / n);
  double lz = Q->spring_l * (nz / n);
  lx *= (nx*nx > 1e-16*1e-16);
  ly *= (ny*ny > 1e-16*1e-16);
  lz *= (nz*nz > 1e-16*1e-16);
  double dx = x - Q->spring_x - lx;
  double dy = y - Q->spring_y - ly;
  double dz = z - Q->spring_z - lz;
  Q->kFx = - Q->spring_k * dx;
  Q->kFy = - Q->spring_k * dy;
  Q->kFz = - Q->spring_k * dz;
}

/* Lubrication and contact forces between particle i and the six domain
 * faces (src/particle_kernel.cu:2347-3073), one thread per particle, faces
 * in the order W E S N B T.  Face f has normal axis a = f / 2 and side
 * sg = +1 (W, S, B: the particle must lie above the plane) or -1 (E, N,
 * T); t1 < t2 are the two tangential axes.  Every face runs through the
 * same code; the forms the reference writes out per face follow from
 * e_a x w (the coupling of the rotation w to the tangential forces and of
 * the tangential velocity to the torques) with these exceptions, kept as
 * the reference has them:
 *  - in contact, the wall's tangential velocity is always the south
 *    face's Dirichlet value (bc.uSD, vSD, wSD);
 *  - in contact with the W or E face, the z component of the contact
 *    point velocity takes w_x (not w_y);
 *  - in contact with the W face, the tangential spring takes the normal
 *    stiffness k (the other faces the Mindlin stiffness kt);
 *  - in contact with the E face, the z torque adds a variable that no
 *    earlier statement of the thread sets except the W-face lubrication
 *    branch, which stores 0: the term is 0 here.
 * The contact-list bookkeeping (codes -10 - f in iSt) and ncoll_wall run
 * for every face, periodic ones included; the forces are multiplied by
 * (pressure condition of the face is Neumann). */
#define CROSS_K ((const int[3][3]) {{0, 2, 1}, {2, 0, 0}, {1, 0, 0}})
#define CROSS_S ((const int[3][3]) {{0, -1, 1}, {1, 0, -1}, {-1, 1, 0}})

__global__ void collision_walls_kernel(int nparts, part_struct *parts,
  PWalls W, dom_struct D, double eps, double mu, double rhof, double nu,
  int ilr, double dt)
{
  int i = threadIdx.x + blockIdx.x * blockDim.x;
  if (i >= nparts) return;
  part_struct *Q = &parts[i];
  const double ai = Q->r;
  const double hN = ilr * Q->r;
  const double pos[3] = {Q->x, Q->y, Q->z};
  const double vel[3] = {Q->u, Q->v, Q->w};
  const double vel0[3] = {Q->u0, Q->v0, Q->w0};
  const double om[3] = {Q->ox, Q->oy, Q->oz};
  const double om0[3] = {Q->ox0, Q->oy0, Q->oz0};
  const double lo[3] = {D.xs, D.ys, D.zs}, hi[3] = {D.xe, D.ye, D.ze};
  double iF[3] = {Q->iFx, Q->iFy, Q->iFz};
  double iL[3] = {Q->iLx, Q->iLy, Q->iLz};
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const int a = f / 2;
    const double sg = (f & 1) ? -1. : 1.;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    const int code = -10 - f;
    double d = (f & 1) ? pos[a] - (hi[a] - W.ds[f])
                       : pos[a] - (lo[a] + W.ds[f]);
    double h = fabs(d) - ai;
    const int isTrue = W.pneu[f];
    const bool side = (f & 1) ? d < 0 : d > 0;
    /* (e_a x v)_t = xs[t] * v[xk[t]] on the tangential axes t */
    const int xk[3] = {CROSS_K[a][0], CROSS_K[a][1], CROSS_K[a][2]};
    const int xs[3] = {CROSS_S[a][0], CROSS_S[a][1], CROSS_S[a][2]};
    if (side && h < hN && h > 0) {
      int q = 0;
      while (q < MAX_NEIGHBORS && Q->iSt[q] != code) q++;
      if (q < MAX_NEIGHBORS && Q->iSt[q] == code) {
        Q->iSt[q] = -1;
        Q->St[q] = 0.;
      }
      if (h < eps*Q->r) h = eps*Q->r;
      double ah = ai/h - ai/hN;
      double lnah = log(hN/h);
      double F[3], L[3];
      double Un = vel[a] - W.vel[f][a];
      F[a] = -6.*PI*mu*ai*Un*ah;
      L[a] = 0.;
      const int ts[2] = {t1, t2};
      for (int m = 0; m < 2; m++) {
        int t = ts[m];
        double Ut = vel[t] - W.vel[f][t];
        F[t] = -6.*PI*mu*ai*Ut*8./15.*lnah;
        F[t] += -sg * xs[t] * (8.*PI*mu*ai*ai*om[xk[t]]*1./10.*lnah);
        double Utk = vel[xk[t]] - W.vel[f][xk[t]];
        L[t] = sg * xs[t] * (8.*PI*mu*ai*ai*Utk*1./10.*lnah);
        L[t] += -8.*PI*mu*ai*ai*ai*om[t]*2./5.*lnah;
      }
      for (int c = 0; c < 3; c++) {
        iF[c] += isTrue * F[c];
        iL[c] += isTrue * L[c];
      }
    }
    if (side && h < 0) {
      double Un = sg * (vel[a] - W.vel[f][a]);
      double Ut[3];
      Ut[t1] = 0.5*(vel[t1] + vel0[t1]) - W.vel[2][t1];
      Ut[t2] = 0.5*(vel[t2] + vel0[t2]) - W.vel[2][t2];
      int q = 0;
      while (q < MAX_NEIGHBORS && Q->iSt[q] != code) q++;
      if (q == MAX_NEIGHBORS) {
        q = 0;
        while (q < MAX_NEIGHBORS - 1 && Q->iSt[q] != -1) q++;
        Q->iSt[q] = code;
        Q->St[q] = 1./9.*Q->rho/rhof*2.*Q->r*fabs(Un)/nu;
        Q->ncoll_wall += 1;
      }
      double omh[3];
      for (int c = 0; c < 3; c++) omh[c] = 0.5*(om[c] + om0[c]);
      /* contact point velocity sg (Ut + (a + h/2) e_a x w) */
      double V[3];
      for (int m = 0; m < 2; m++) {
        int t = m ? t2 : t1;
        int k = (a == 0 && t == 2) ? 0 : xk[t];
        V[t] = sg * (Ut[t] + xs[t] * ((ai + 0.5*h)*omh[k]));
      }
      double Hi = 0.5*Q->E/(1.+Q->sigma);
      double kt = 8./((1.-Q->sigma*Q->sigma)/Hi
        +(1.-Q->sigma*Q->sigma)/Hi)/sqrt(1./ai)*sqrt(-h);
      double k = 4./3./((1.-Q->sigma*Q->sigma)/Q->E
        + (1.-Q->sigma*Q->sigma)/Q->E)/sqrt(1./ai);
      double e = Q->e_dry + (1.+Q->e_dry)/Q->St[q]*log(1.e-4);
      if (e < 0) e = 0;
      double alpha = -2.263*pow(e,0.3948)+2.22;
      double eta = alpha*sqrt(4./3.*PI*ai*ai*ai*Q->rho*k*sqrt(-h));
      double Ks = f == 0 ? k : kt;
      double Ft[3];
      Ft[t1] = -Ks * (V[t1] * dt);
      Ft[t2] = -Ks * (V[t2] * dt);
      double Fm = sqrt(Ft[t1]*Ft[t1] + Ft[t2]*Ft[t2]);
      double cf = Q->coeff_fric;
      if (Fm > fabs(cf * (sqrt(-h*h*h)*k - eta*Un))) {
        Ft[t1] = cf * (sqrt(-h*h*h)*k - eta*Un) * Ft[t1] / Fm;
        Ft[t2] = cf * (sqrt(-h*h*h)*k - eta*Un) * Ft[t2] / Fm;
      }
      iF[a] += sg * (isTrue * (sqrt(-h*h*h)*k - eta*Un));
      iF[t1] += isTrue * Ft[t1];
      iF[t2] += isTrue * Ft[t2];
      /* torque -sg (e_a x F_t) (a + h/2) on the tangential axes */
      for (int m = 0; m < 2; m++) {
        int t = m ? t2 : t1;
        double Fk = (f == 1 && t == 2) ? 0. : Ft[xk[t]];
        iL[t] += -sg * xs[t] * (isTrue * (ai+0.5*h) * Fk);
      }
    }
  }
  Q->iFx = iF[0]; Q->iFy = iF[1]; Q->iFz = iF[2];
  Q->iLx = iL[0]; Q->iLy = iL[1]; Q->iLz = iL[2];
}

/* Accelerations and velocities (src/particle_kernel.cu:3368-3410). */
__global__ void move_parts_a_kernel(int nparts, part_struct *parts, double dt,
  g_struct g, gradP_struct gradP, double rho_f)
{
  int pp = threadIdx.x + blockIdx.x * blockDim.x;
  if (pp >= nparts) return;
  part_struct *Q = &parts[pp];
  double vol = 4./3. * PI * Q->r*Q->r*Q->r;
  double m = vol * Q->rho;
  if (Q->translating) {
    Q->udot = (Q->Fx + Q->kFx + Q->iFx + Q->aFx - vol*gradP.x) / m
      + (Q->rho - rho_f) / Q->rho * g.x;
    Q->vdot = (Q->Fy + Q->kFy + Q->iFy + Q->aFy - vol*gradP.y) / m
      + (Q->rho - rho_f) / Q->rho * g.y;
    Q->wdot = (Q->Fz + Q->kFz + Q->iFz + Q->aFz - vol*gradP.z) / m
      + (Q->rho - rho_f) / Q->rho * g.z;
    Q->u = Q->u0 + 0.5*dt*(Q->udot + Q->udot0);
    Q->v = Q->v0 + 0.5*dt*(Q->vdot + Q->vdot0);
    Q->w = Q->w0 + 0.5*dt*(Q->wdot + Q->wdot0);
  }
  if (Q->rotating) {
    double I = 0.4 * m * Q->r*Q->r;
    Q->oxdot = (Q->Lx + Q->iLx + Q->aLx) / I;
    Q->oydot = (Q->Ly + Q->iLy + Q->aLy) / I;
    Q->ozdot = (Q->Lz + Q->iLz + Q->aLz) / I;
    Q->ox = Q->ox0 + 0.5*dt*(Q->oxdot + Q->oxdot0);
    Q->oy = Q->oy0 + 0.5*dt*(Q->oydot + Q->oydot0);
    Q->oz = Q->oz0 + 0.5*dt*(Q->ozdot + Q->ozdot0);
  }
}

__device__ inline void rotate(double qr, double qi, double qj, double qk,
  double *pi, double *pj, double *pk)
{
  double Pr = *pi*qi + *pj*qj + *pk*qk;
  double Pi = *pi*qr - *pj*qk + *pk*qj;
  double Pj = *pi*qk + *pj*qr - *pk*qi;
  double Pk = -*pi*qj + *pj*qi + *pk*qr;
  *pi = qr*Pi + qi*Pr + qj*Pk - qk*Pj;
  *pj = qr*Pj - qi*Pk + qj*Pr + qk*Pi;
  *pk = qr*Pk + qi*Pj - qj*Pi + qk*Pr;
}

/* Positions, body axes, stored *0 fields (src/particle_kernel.cu:3412-3470),
 * then the periodic wrap of a centre that left the domain (the reference's
 * particle transfer, src/particle_kernel.cu:1166-1210). */
__global__ void move_parts_b_kernel(int nparts, part_struct *parts, double dt,
  dom_struct D, int perx, int pery, int perz)
{
  int pp = threadIdx.x + blockIdx.x * blockDim.x;
  if (pp >= nparts) return;
  part_struct *Q = &parts[pp];
  if (Q->translating) {
    Q->x += 0.5*dt*(Q->u + Q->u0);
    Q->y += 0.5*dt*(Q->v + Q->v0);
    Q->z += 0.5*dt*(Q->w + Q->w0);
    Q->u0 = Q->u; Q->v0 = Q->v; Q->w0 = Q->w;
    Q->udot0 = Q->udot; Q->vdot0 = Q->vdot; Q->wdot0 = Q->wdot;
  }
  if (Q->rotating) {
    double mag = 0.5*sqrt(Q->ox*Q->ox + Q->oy*Q->oy + Q->oz*Q->oz);
    mag += 0.5*sqrt(Q->ox0*Q->ox0 + Q->oy0*Q->oy0 + Q->oz0*Q->oz0);
    double X = 0, Y = 0, Z = 0;
    if (mag > 0) {
      X = 0.5 * (Q->ox + Q->ox0) / mag;
      Y = 0.5 * (Q->oy + Q->oy0) / mag;
      Z = 0.5 * (Q->oz + Q->oz0) / mag;
    }
    double theta = mag * dt;
    double qr = cos(0.5*theta);
    double qi = X * sin(0.5*theta);
    double qj = Y * sin(0.5*theta);
    double qk = Z * sin(0.5*theta);
    rotate(qr, qi, qj, qk, &Q->axx, &Q->axy, &Q->axz);
    rotate(qr, qi, qj, qk, &Q->ayx, &Q->ayy, &Q->ayz);
    rotate(qr, qi, qj, qk, &Q->azx, &Q->azy, &Q->azz);
    Q->ox0 = Q->ox; Q->oy0 = Q->oy; Q->oz0 = Q->oz;
    Q->oxdot0 = Q->oxdot; Q->oydot0 = Q->oydot; Q->ozdot0 = Q->ozdot;
  }
  if (perx) { if (Q->x >= D.xe) Q->x = Q->x - D.xl; else if (Q->x < D.xs) Q->x = Q->x + D.xl; }
  if (pery) { if (Q->y >= D.ye) Q->y = Q->y - D.yl; else if (Q->y < D.ys) Q->y = Q->y + D.yl; }
  if (perz) { if (Q->z >= D.ze) Q->z = Q->z - D.zl; else if (Q->z < D.zs) Q->z = Q->z + D.zl; }
}

static dim3 pgrid(void) { return dim3(nblk(nparts, 256)); }

static void collide(void)
{
  hipLaunchKernelGGL(collision_init_kernel, pgrid(), dim3(256), 0, bb_stream,
    nparts, _parts);
  if (nparts > 1) {
    hipLaunchKernelGGL(bin_kernel, pgrid(), dim3(256), 0, bb_stream, nparts,
      _parts, bins, _pbin);
    hipLaunchKernelGGL(collision_parts_kernel, dim3(nparts), dim3(WAVE), 0,
      bb_stream, nparts, _parts, _pbin, bins, 0.01, mu, rho_f, nu,
      (int) interaction_length_ratio, dt, DOM.xl, DOM.yl, DOM.zl, ddom.perx,
      ddom.pery, ddom.perz);
  }
  hipLaunchKernelGGL(spring_kernel, pgrid(), dim3(256), 0, bb_stream, nparts,
    _parts, DOM);
  hipLaunchKernelGGL(collision_walls_kernel, pgrid(), dim3(256), 0, bb_stream,
    nparts, _parts, pwalls(), DOM, 0.01, mu, rho_f, nu,
    (int) interaction_length_ratio, dt);
  BB_LAUNCH_CHECK();
}

extern "C" void cuda_move_parts_sub(void)
{
  bc_vals_ok = 0;                /* particle state changes */
  if (nparts == 0) return;
  collide();
  hipLaunchKernelGGL(move_parts_a_kernel, pgrid(), dim3(256), 0, bb_stream,
    nparts, _parts, dt, g, gradP, rho_f);
  collide();
  BB_LAUNCH_CHECK();
}

extern "C" void cuda_update_part_velocity(void)
{
  bc_vals_ok = 0;                /* particle state changes */
  if (nparts == 0) return;
  hipLaunchKernelGGL(move_parts_a_kernel, pgrid(), dim3(256), 0, bb_stream,
    nparts, _parts, dt, g, gradP, rho_f);
  BB_LAUNCH_CHECK();
}

extern "C" void cuda_update_part_position(void)
{
  bc_vals_ok = 0;                /* particle state changes */
  if (nparts == 0) return;
  hipLaunchKernelGGL(move_parts_b_kernel, pgrid(), dim3(256), 0, bb_stream,
    nparts, _parts, dt, DOM, ddom.perx, ddom.pery, ddom.perz);
  BB_LAUNCH_CHECK();
}

/* ------------------------------ memory ----------------------------------- */
extern "C" void cuda_part_malloc_dev(void)
{
  bc_vals_ok = 0;                /* particle state changes */
  int n = nparts > 0 ? nparts : 1;
  BB_CHECK(hipMalloc((void **) &_parts, n * sizeof(part_struct)));
  BB_CHECK(hipMalloc((void **) &_perr, n * sizeof(double)));
  BB_CHECK(hipMalloc((void **) &_pnum, n * sizeof(int)));
  BB_CHECK(hipMalloc((void **) &_pbin, n * sizeof(int)));
  BB_CHECK(hipMalloc((void **) &_err_out, 2 * sizeof(double)));
  BB_CHECK(hipHostMalloc((void **) &err_host, 2 * sizeof(double),
    hipHostMallocDefault));
  BB_CHECK(hipMalloc((void **) &_bb_fluid, ddom.cc.n));
  BB_CHECK(hipMemset(_bb_fluid, 1, ddom.cc.n));
  gpumem += n * sizeof(part_struct) + ddom.cc.n;
}

extern "C" void cuda_part_push(void)
{
  bc_vals_ok = 0;                /* particle state changes */
  if (nparts > 0)
    BB_CHECK(hipMemcpy(_parts, parts, nparts * sizeof(part_struct),
      hipMemcpyHostToDevice));
}

extern "C" void cuda_part_pull(void)
{
  if (nparts > 0) {
    BB_CHECK(hipStreamSynchronize(bb_stream));
    BB_CHECK(hipMemcpy(parts, _parts, nparts * sizeof(part_struct),
      hipMemcpyDeviceToHost));
  }
}

extern "C" void cuda_part_free(void)
{
  bc_vals_ok = 0;                /* particle state changes */
  if (_parts) BB_CHECK(hipFree(_parts));
  if (_perr) BB_CHECK(hipFree(_perr));
  if (_pnum) BB_CHECK(hipFree(_pnum));
  if (_pbin) BB_CHECK(hipFree(_pbin));
  if (_err_out) BB_CHECK(hipFree(_err_out));
  if (err_host) BB_CHECK(hipHostFree(err_host));
  if (_bb_fluid) BB_CHECK(hipFree(_bb_fluid));
  _bb_fluid = NULL;
  free_face_lists();
  _parts = NULL; _perr = NULL; _pnum = NULL; _pbin = NULL; _err_out = NULL;
  err_host = NULL;
  /* the host phase arrays of cuda_part_malloc_host (src/cuda_particle.cu:
   * 1271-1272 frees them here) */
  if (phase) BB_CHECK(hipHostFree(phase));
  if (phase_shell) BB_CHECK(hipHostFree(phase_shell));
  phase = phase_shell = NULL;
}

extern "C" void cuda_part_pull_debug(void)
{
  /* src/cuda_particle.cu:254-265 pulls every particle, ghost copies
   * included; one GPU holds each particle once and keeps no copies */
  cuda_part_pull();
}

/* src/cuda_particle.cu:287-1186 move the particles that left a rank into
 * its neighbour and copy those within a bin of a periodic face as ghosts.
 * On one GPU no particle changes owner; cuda_update_part_position wraps a
 * centre that leaves through a periodic face, and every kernel evaluates
 * the periodic images x +- L itself, so there is nothing to move. */
extern "C" void cuda_transfer_parts_i(void) {}
extern "C" void cuda_transfer_parts_j(void) {}
extern "C" void cuda_transfer_parts_k(void) {}
