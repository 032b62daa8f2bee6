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
long bb_mask_gen = 0;                  /* bumped whenever the cages (and the
                                          mask) are rebuilt */
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

/* ------------------------- small vector algebra -------------------------- */
struct v3 { double x, y, z; };

__device__ inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline v3 operator*(double s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline double dot(v3 a, v3 b) { return a.x*b.x + a.y*b.y + a.z*b.z; }
__device__ inline v3 cross(v3 a, v3 b)
{
  return {a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x};
}
__device__ inline double norm(v3 a) { return sqrt(dot(a, a)); }

/* Acceleration of the fluid the particle frame sees from the imposed
 * pressure gradient and the Boussinesq buoyancy of the particle's scalar,
 * -gradP/rho_f - s_beta (s - s_ref) g. */
__device__ inline v3 frame_accel(gradP_struct gradP, double rho_f,
  double s_beta, double s, double s_ref, g_struct g)
{
  const double b = -s_beta * (s - s_ref);
  return (-1. / rho_f) * v3{gradP.x, gradP.y, gradP.z} + b * v3{g.x, g.y, g.z};
}

/* Velocity of a particle's rigid frame at offset rv (|rv| = r) from its
 * centre, with the rest frame's correction for unsteady rotation:
 * U + w x r + (r^5 - a^5)/(10 nu r^3) dw/dt x r. */
__device__ inline v3 frame_velocity(const part_struct *Q, v3 rv, double r,
  double nu)
{
  const double a2 = Q->r * Q->r, r2 = r * r;
  const double unsteady = 0.1 / nu * (r2*r2*r - a2*a2*Q->r) / (r2*r);
  return v3{Q->u, Q->v, Q->w} + cross(v3{Q->ox, Q->oy, Q->oz}, rv)
       + unsteady * cross(v3{Q->oxdot, Q->oydot, Q->ozdot}, rv);
}

/* ------------------------- spherical harmonics -------------------------- */
/* Spherical coordinates of (x, y, z) about a particle centre, phi in
 * [0, 2 pi) (the convention of src/particle_kernel.cu:1905-1917, whose
 * guards compare against the negative DIV_ST and never act). */
__device__ inline void to_spherical(double x, double y, double z, double *r,
  double *theta, double *phi)
{
  double rho2 = x*x + y*y;
  *r = sqrt(rho2 + z*z);
  *theta = acos(z / *r);
  *phi = acos(x / sqrt(rho2));
  if (y < 0.) *phi = 2.*PI - *phi;
}

/* Highest Lamb order: NCOEFFS = (LMAX + 1)(LMAX + 2)/2 coefficients. */
#define LMAX 4

/* Normalisation of the surface harmonic Y_n^m = N_n^m P_n^m(cos t) e^{imp}:
 * N_n^m = sqrt((2n+1)/(4 pi) (n-m)!/(n+m)!). */
__device__ inline double harmonic_norm(int n, int m)
{
  double ratio = 1.;
  for (int q = n - m + 1; q <= n + m; q++) ratio /= (double) q;
  return sqrt((2.*n+1.) / (4.*PI) * ratio);
}

/* Associated Legendre functions P_n^m(cos t), Condon-Shortley phase, for
 * 0 <= m <= n <= nmax, from the recurrences
 *   P_m^m = (-1)^m (2m-1)!! sin^m t,   P_{m+1}^m = (2m+1) cos t P_m^m,
 *   (n - m) P_n^m = (2n-1) cos t P_{n-1}^m - (n+m-1) P_{n-2}^m.
 * P is indexed [n][m]; entries with m > n are left untouched. */
template <int NMAX>
__device__ inline void legendre_table(double ct, double st,
  double P[NMAX + 1][NMAX + 1])
{
  double pmm = 1.;
#pragma unroll
  for (int m = 0; m <= NMAX; m++) {
    if (m > 0) pmm *= -(2.*m - 1.) * st;
    P[m][m] = pmm;
    if (m + 1 <= NMAX) P[m + 1][m] = (2.*m + 1.) * ct * pmm;
#pragma unroll
    for (int n = m + 2; n <= NMAX; n++)
      P[n][m] = ((2.*n - 1.) * ct * P[n - 1][m]
                 - (n + m - 1.) * P[n - 2][m]) / (double) (n - m);
  }
}

/* One P_n^m(cos theta), n <= LMAX + 1. */
__device__ inline double legendre_nm(int n, int m, double theta)
{
  double P[LMAX + 2][LMAX + 2];
  legendre_table<LMAX + 1>(cos(theta), sin(theta), P);
  return (m > n) ? 0. : P[n][m];
}

/* Velocity of Lamb's solution for the flow outside a sphere of radius a, at
 * (r, theta, phi) about its centre, in Cartesian components (the series of
 * src/particle_kernel.cu:1839-1903).  With rho = r/a, s = a/r and the three
 * harmonic coefficient sets p (pressure), f (phi) and c (chi) of each order
 * n, their surface fields at the point
 *   X_n = Re sum_m w_m c_nm Y_n^m,  T_n = Re sum_m w_m c_nm dY_n^m/dt,
 *   F_n = Re sum_m w_m c_nm (1/sin t) dY_n^m/dp
 * (w_0 = 1, w_m = 2: the m < 0 terms are the conjugates) give
 *   u_r = (nu/a) sum_n  Ar_n X_n[p] + Br_n X_n[f]
 *   u_t = (nu/a) sum_n  At_n T_n[p] + Bt_n T_n[f] + C_n F_n[c]
 *   u_p = (nu/a) sum_n  At_n F_n[p] + Bt_n F_n[f] - C_n T_n[c]
 * with the radial functions
 *   Ar_n = n rho^(n+1) / (2(2n+3)) + (n/4)((2n+1)/(2n+3) s^2 - 1) s^n,
 *   Br_n = n rho^(n-1) + (n/2)(2n-1 - (2n+1) rho^2) s^(n+2),
 *   At_n = (n+3) rho^(n+1) / (2(n+1)(2n+3))
 *          + (n-2 - n(2n+1)/(2n+3) s^2) s^n / (4(n+1)),
 *   Bt_n = rho^(n-1) + ((n-2)(2n+1) rho^2 - n(2n-1)) s^(n+2) / (2(n+1)),
 *   C_n  = rho^(n-1) - s^(n+1).
 * The n = 0 pressure term (a constant) drives no flow.  dP_n^m/dt is formed
 * as ((n-m+1) P_{n+1}^m - (n+1) cos t P_n^m) / sin t. */
__device__ inline void lamb_vel(int order, double a, double r, double theta,
  double phi, const part_struct *Q, double nu, double *Ux, double *Uy,
  double *Uz)
{
  const double ct = cos(theta), st = sin(theta);
  const double cp = cos(phi), sp = sin(phi);
  double P[LMAX + 2][LMAX + 2];
  legendre_table<LMAX + 1>(ct, st, P);
  const double rho = r / a, s = a / r;
  double ur = 0., ut = 0., up = 0.;
  double rho_nm1 = 1., s_n = s;             /* rho^(n-1), s^n at n = 1 */
  for (int n = 1; n <= order; n++) {
    double X[3] = {0., 0., 0.}, T[3] = {0., 0., 0.}, F[3] = {0., 0., 0.};
    const double *re[3] = {Q->pnm_re, Q->phinm_re, Q->chinm_re};
    const double *im[3] = {Q->pnm_im, Q->phinm_im, Q->chinm_im};
    const int c0 = n * (n + 1) / 2;
    double cm = 1., sm = 0.;                /* cos(m phi), sin(m phi) */
    for (int m = 0; m <= n; m++) {
      const double wN = (m ? 2. : 1.) * harmonic_norm(n, m);
      const double Y = wN * P[n][m];
      const double dY = wN * ((n - m + 1.) * P[n + 1][m]
                              - (n + 1.) * ct * P[n][m]) / st;
      const double mY = -m * Y / st;
      for (int c = 0; c < 3; c++) {
        const double cr = re[c][c0 + m] * cm - im[c][c0 + m] * sm;
        const double ci = re[c][c0 + m] * sm + im[c][c0 + m] * cm;
        X[c] += Y * cr;
        T[c] += dY * cr;
        F[c] += mY * ci;
      }
      const double cn = cm * cp - sm * sp;
      sm = sm * cp + cm * sp;
      cm = cn;
    }
    const double rho_np1 = rho_nm1 * rho * rho;
    const double s2 = s * s;
    const double i2n3 = 1. / (2.*n + 3.), in1 = 1. / (n + 1.);
    const double Ar = 0.5*n*i2n3*rho_np1 + 0.25*n*((2.*n+1.)*i2n3*s2 - 1.)*s_n;
    const double Br = n*rho_nm1 + 0.5*n*(2.*n-1. - (2.*n+1.)*rho*rho)*s_n*s2;
    const double At = 0.5*(n+3.)*in1*i2n3*rho_np1
                    + 0.25*in1*(n-2. - n*(2.*n+1.)*i2n3*s2)*s_n;
    const double Bt = rho_nm1
                    + 0.5*in1*((n-2.)*(2.*n+1.)*rho*rho - n*(2.*n-1.))*s_n*s2;
    const double Cc = rho_nm1 - s_n * s;
    ur += Ar * X[0] + Br * X[1];
    ut += At * T[0] + Bt * T[1] + Cc * F[2];
    up += At * F[0] + Bt * F[1] - Cc * T[2];
    rho_nm1 *= rho;
    s_n *= s;
  }
  const double scale = nu / a;
  ur *= scale;
  ut *= scale;
  up *= scale;
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
  if (i >= G.ib) return;                /* row padding */
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
  bb_mask_gen++;
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
  /* 32-bit index arithmetic: the face lists are built only for grids of
   * fewer than 2^31 elements (bb_particle_face_lists) */
  const unsigned tt = (unsigned) t, sy = (unsigned) G.sy;
  const unsigned row = tt / sy;
  int i = (int) (tt - row * sy), j = (int) (row % (unsigned) (G.sz / G.sy)),
      k = (int) (tt / (unsigned) G.sz);
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

/* The face lists are compacted without a shared counter: a counter that
 * every wavefront holding a listed face bumps serialises those atomics in
 * L2 (about 75 000 per direction at 256^3 with 500 spheres, 0.7 ms).
 * Instead each workgroup counts the listed faces of its FL_SPAN
 * consecutive faces (face_count_kernel), one workgroup turns the counts
 * into offsets (face_offsets_kernel), and each workgroup writes its faces
 * at its offset in increasing index order (face_write_kernel): the list
 * comes out sorted. */
#define FL_ITEMS 8
#define FL_SPAN (FL_ITEMS * 256)

/* Listed faces among this thread's item `it` of workgroup b, as the
 * ballot of its wavefront. */
template <int dir>
__device__ inline unsigned long long fl_ballot(const PDom &P, long n,
  const int *phase, const int *flag, long b, int it, long *t_out)
{
  const long t = b * FL_SPAN + (long) it * 256 + threadIdx.x;
  int i, j, k, np;
  const bool ok = t < n && bc_face<dir>(P, t, phase, flag, &i, &j, &k, &np);
  *t_out = ok ? t : -1;
  return __ballot(ok);
}

template <int dir>
__global__ void __launch_bounds__(256)
face_count_kernel(PDom P, long n, const int *__restrict__ phase,
  const int *__restrict__ flag, int *__restrict__ bcount)
{
  __shared__ int s_n[4];
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x >> 6;
  int c = 0;
  for (int it = 0; it < FL_ITEMS; it++) {
    long t;
    c += __popcll(fl_ballot<dir>(P, n, phase, flag, blockIdx.x, it, &t));
  }
  if (lane == 0) s_n[wv] = c;
  __syncthreads();
  if (threadIdx.x == 0) bcount[blockIdx.x] = (s_n[0] + s_n[1]) + (s_n[2] + s_n[3]);
}

/* Exclusive prefix sums of the nb workgroup counts (one workgroup of 1024
 * threads, each summing a run of consecutive counts) and the total. */
__global__ void __launch_bounds__(1024)
face_offsets_kernel(const int *__restrict__ bcount, int nb,
  int *__restrict__ boff, int *__restrict__ total)
{
  __shared__ int s[1024];
  const int per = (nb + 1023) / 1024;
  const int b0 = threadIdx.x * per;
  int sum = 0;
  for (int q = 0; q < per; q++)
    if (b0 + q < nb) sum += bcount[b0 + q];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {     /* inclusive scan */
    int v = threadIdx.x >= (unsigned) off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  int run = s[threadIdx.x] - sum;
  for (int q = 0; q < per; q++)
    if (b0 + q < nb) {
      boff[b0 + q] = run;
      run += bcount[b0 + q];
    }
  if (threadIdx.x == 1023) *total = s[1023];
}

template <int dir>
__global__ void __launch_bounds__(256)
face_write_kernel(PDom P, long n, const int *__restrict__ phase,
  const int *__restrict__ flag, const int *__restrict__ boff,
  int *__restrict__ list)
{
  __shared__ int s_n[4];
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x >> 6;
  int base = boff[blockIdx.x];
  for (int it = 0; it < FL_ITEMS; it++) {
    long t;
    const unsigned long long m =
      fl_ballot<dir>(P, n, phase, flag, blockIdx.x, it, &t);
    if (lane == 0) s_n[wv] = __popcll(m);
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wv; w++) before += s_n[w];
    if (t >= 0)
      list[base + before + __popcll(m & ((1ull << lane) - 1ull))] = (int) t;
    base += (s_n[0] + s_n[1]) + (s_n[2] + s_n[3]);
    __syncthreads();
  }
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
  to_spherical(x, y, z, &r, &theta, &phi);
  double U[3];
  lamb_vel(Q->order, Q->r, r, theta, phi, Q, nu, &U[0], &U[1], &U[2]);
  const v3 Uf = frame_velocity(Q, v3{x, y, z}, r, nu);
  const double frame[3] = {Uf.x, Uf.y, Uf.z};
  const double val = U[dir] + frame[dir];
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
static int *_bc_bcount = NULL, *_bc_boff = NULL;  /* per-workgroup counts,
                                                   offsets */
static int bc_n[3] = {0, 0, 0};
static double *_bc_val[3] = {NULL, NULL, NULL};
static int bc_vals_ok = 0;

/* Called by cuda_build_cages (bb_bc.hip) once every flag is final. */
void bb_particle_face_lists(void)
{
  PDom P = pdom();
  const long gn[3] = {ddom.fx.n, ddom.fy.n, ddom.fz.n};
  long nbmax = 0;
  for (int d = 0; d < 3; d++) {
    if (gn[d] >= (1L << 31)) {
      fprintf(stderr, "particle face lists: a face grid of %ld elements "
        "exceeds the 32-bit face index\n", gn[d]);
      exit(EXIT_FAILURE);
    }
    long nb = (gn[d] + FL_SPAN - 1) / FL_SPAN;
    nbmax = nb > nbmax ? nb : nbmax;
  }
  if (!_bc_count) {
    BB_CHECK(hipMalloc((void **) &_bc_count, 3 * sizeof(int)));
    BB_CHECK(hipMalloc((void **) &_bc_bcount, nbmax * sizeof(int)));
    BB_CHECK(hipMalloc((void **) &_bc_boff, nbmax * sizeof(int)));
    for (int d = 0; d < 3; d++) {
      BB_CHECK(hipMalloc((void **) &_bc_list[d], gn[d] * sizeof(int)));
      BB_CHECK(hipMalloc((void **) &_bc_val[d], gn[d] * sizeof(double)));
    }
  }
  bc_vals_ok = 0;
  const int *flags[3] = {_flag_u, _flag_v, _flag_w};
  for (int d = 0; d < 3; d++) {
    const unsigned nb = (unsigned) ((gn[d] + FL_SPAN - 1) / FL_SPAN);
#define FL(D) do {                                                          \
      hipLaunchKernelGGL(face_count_kernel<D>, dim3(nb), dim3(256), 0,      \
        bb_stream, P, gn[D], _phase, flags[D], _bc_bcount);                 \
      hipLaunchKernelGGL(face_offsets_kernel, dim3(1), dim3(1024), 0,        \
        bb_stream, _bc_bcount, (int) nb, _bc_boff, _bc_count + D);          \
      hipLaunchKernelGGL(face_write_kernel<D>, dim3(nb), dim3(256), 0,      \
        bb_stream, P, gn[D], _phase, flags[D], _bc_boff, _bc_list[D]);      \
    } while (0)
    if (d == 0) FL(0); else if (d == 1) FL(1); else FL(2);
#undef FL
  }
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
  BB_CHECK(hipFree(_bc_bcount));
  BB_CHECK(hipFree(_bc_boff));
  _bc_count = _bc_bcount = _bc_boff = NULL;
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
  to_spherical(x, y, z, &r, &theta, &phi);
  const double a = Q->r;
  const v3 rv = {x, y, z};
  const v3 om = {Q->ox, Q->oy, Q->oz};
  const v3 acc = {Q->udot, Q->vdot, Q->wdot};
  const v3 wxr = cross(om, rv);
  /* n = 0 pressure of the series plus the rigid-frame terms, as in the
   * node values below */
  double pc = harmonic_norm(0, 0) * Q->pnm_re[0] * (mu*nu/(a*a));
  pc += 0.5 * rho_f * dot(wxr, wxr)
      + rho_f * dot(frame_accel(gradP, rho_f, s_beta, Q->s, s_ref, g) - acc,
                    rv);
  p[t] = pc;
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

/* Index of a cell or face coordinate clamped to the grid's interior range
 * lo..hi. */
__device__ inline int clamp_index(int i, int lo, int hi)
{
  return i < lo ? lo : (i > hi ? hi : i);
}

__device__ void node_values(const PDom &P, const PWalls &W, part_struct *Q,
  int node,
  const double *__restrict__ p, const double *__restrict__ u,
  const double *__restrict__ v, const double *__restrict__ w, double rho_f,
  double nu, gradP_struct gradP, double s_beta, double s_ref, g_struct g,
  NodeVals *out)
{
  const double ddx = 1. / P.dx, ddy = 1. / P.dy, ddz = 1. / P.dz;
  const double th = c_node_t[node], ph = c_node_p[node];
  /* node offset from the centre on the sphere of radius rs */
  const v3 rn = {Q->rs * sin(th) * cos(ph), Q->rs * sin(th) * sin(ph),
                 Q->rs * cos(th)};
  int ic, jc, kc;
  const double x = node_coord(rn.x, Q->x, P.xs, P.dx, P.xl, P.nx, P.perx, &ic);
  const double y = node_coord(rn.y, Q->y, P.ys, P.dy, P.yl, P.ny, P.pery, &jc);
  const double z = node_coord(rn.z, Q->z, P.zs, P.dz, P.zl, P.nz, P.perz, &kc);
  const int oob[3] = {ic < 1 || ic > P.nx, jc < 1 || jc > P.ny,
                      kc < 1 || kc > P.nz};
  const int outside = oob[0] || oob[1] || oob[2];
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
  const int fluid = (1 - outside) * (ncode == -1);
  /* a node beyond a face in one direction only takes that face's Dirichlet
   * velocity (0 at a Neumann face) in the rest frame
   * (src/physalis_kernel.cu:437-460) */
  int beyond_only[3];
  for (int d = 0; d < 3; d++)
    beyond_only[d] = oob[d] * (1 - oob[(d+1)%3]) * (1 - oob[(d+2)%3]) * wall;

  /* the particle's rigid frame at the node: its velocity and the pressure
   * of the frame's centrifugal and linear accelerations */
  const v3 om = {Q->ox, Q->oy, Q->oz};
  const v3 acc = {Q->udot, Q->vdot, Q->wdot};
  const v3 wxr = cross(om, rn);
  const v3 rigid = frame_velocity(Q, rn, Q->rs, nu);

  /* pressure: cell centres (src/physalis_kernel.cu:305-364) */
  {
    const int i = clamp_index(ic, 1, P.nx), j = clamp_index(jc, 1, P.ny),
              k = clamp_index(kc, 1, P.nz);
    double pn = node_interp(p, P.cc, i, j, k, x, y, z,
      (i - 0.5) * P.dx + P.xs, (j - 0.5) * P.dy + P.ys,
      (k - 0.5) * P.dz + P.zs, ddx, ddy, ddz);
    pn -= 0.5 * rho_f * dot(wxr, wxr)
        + rho_f * dot(frame_accel(gradP, rho_f, s_beta, Q->s, s_ref, g) - acc,
                      rn);
    out->pp[node] = pn * fluid;
  }

  /* velocity components on their face grids (the face index of the
   * component's own direction clamps to n + 1) */
  const double *vel[3] = {u, v, w};
  const Grid *grid[3] = {&P.fx, &P.fy, &P.fz};
  const double rig[3] = {rigid.x, rigid.y, rigid.z};
  double un[3];
  for (int d = 0; d < 3; d++) {
    const int i = clamp_index(ic, 1, P.nx + (d == 0));
    const int j = clamp_index(jc, 1, P.ny + (d == 1));
    const int k = clamp_index(kc, 1, P.nz + (d == 2));
    double val = node_interp(vel[d], *grid[d], i, j, k, x, y, z,
      (i - (d == 0 ? DOM_BUF : 0.5)) * P.dx + P.xs,
      (j - (d == 1 ? DOM_BUF : 0.5)) * P.dy + P.ys,
      (k - (d == 2 ? DOM_BUF : 0.5)) * P.dz + P.zs, ddx, ddy, ddz);
    val -= rig[d];
    double at_wall[3];
    for (int e = 0; e < 3; e++) {
      at_wall[e] = (ncode == WEST_WALL_D - 10 * (2*e)) * W.vel[2*e][d]
                 + (ncode == WEST_WALL_D - 10 * (2*e + 1)) * W.vel[2*e + 1][d];
      at_wall[e] -= rig[d];
    }
    un[d] = fluid * val + beyond_only[0] * at_wall[0]
          + beyond_only[1] * at_wall[1] + beyond_only[2] * at_wall[2];
  }
  cart2sphere(un[0], un[1], un[2], th, ph, &out->ur[node], &out->ut[node],
    &out->up[node]);
}

/* Cartesian vector of an order-1 harmonic coefficient set: the
 * combination (-2 N_1^1 Re c_11, 2 N_1^1 Im c_11, N_1^0 Re c_10) that the
 * force and torque of Lamb's solution take (coefficients 1 = (1,0) and
 * 2 = (1,1)). */
__device__ inline v3 order1_vector(const double *re, const double *im)
{
  const double N10 = sqrt(3./4./PI), N11 = sqrt(3./8./PI);
  return {-2.*N11 * re[2], 2.*N11 * im[2], N10 * re[1]};
}

/* Per particle (one workgroup of 64 lanes): the rest-frame pressure and
 * velocity at the 26 nodes (lanes 0..25); then one lane per coefficient
 * (n, m) projects them on the sphere by the Lebedev rule and solves for
 * the Lamb coefficients of order n (with under-relaxation); then lane 0
 * forms the force and torque (src/physalis_kernel.cu:642-841).
 * Projections of one (n, m), with Y = N_n^m P_n^m(cos t) e^{-imp}:
 *   Pr = int p Y,
 *   Dv = int (u_t dY/dt + u_p (1/sin t) dY/dp)   (surface divergence),
 *   Cu = int (u_t (1/sin t) dY/dp - u_p dY/dt)   (surface curl),
 * each as (re, im).  On the node sphere rs, with s = a/rs and q = rs/a,
 * the order-n coefficients (p, f, c) of the pressure, phi and chi series
 * satisfy
 *   K11 p - K12 f = (a^2/(mu nu)) Pr,    K21 p + K22 f = (a/nu) Dv,
 *   E c = (a/nu) Cu,
 *   K11 = (1 - n(2n-1) s^(2n+1) / (2(n+1))) q^n,
 *   K12 = n(2n-1)(2n+1) s^(n+1) / (n+1),
 *   K21 = (n/4)(2(n+3)/(2n+3) + (n-2 - n(2n+1) s^2/(2n+3)) s^(2n+1)) q^(n+1),
 *   K22 = n(n+1 + ((n-2)(2n+1) q^2 - n(2n-1)) s^(2n+1) / 2) q^(n-1),
 *   E   = n(n+1)(s^(2n+1) - 1) q^n;
 * order 0 has the pressure only, p = (a^2/(mu nu)) Pr. */
__global__ void __launch_bounds__(64)
lamb_kernel(PDom P, PWalls W, part_struct *parts,
  const double *__restrict__ p,
  const double *__restrict__ u, const double *__restrict__ v,
  const double *__restrict__ w, double rho_f, double mu, double nu,
  gradP_struct gradP, double relax, double s_beta, double s_ref, g_struct g)
{
  __shared__ NodeVals N;
  part_struct *Q = &parts[blockIdx.x];
  int t = threadIdx.x;
  if (t < NNODES)
    node_values(P, W, Q, t, p, u, v, w, rho_f, nu, gradP, s_beta, s_ref, g,
                &N);
  __syncthreads();

  if (t < Q->ncoeff) {
    const int n = c_nn[t], m = c_mm[t];
    const double Nnm = harmonic_norm(n, m);
    double Pr[2], Dv[2], Cu[2];
    for (int node = 0; node < NNODES; node++) {
      const double theta = c_node_t[node];
      const double phi = c_node_p[node];
      const double ct = cos(theta), st = sin(theta);
      double Pl[LMAX + 2][LMAX + 2];
      legendre_table<LMAX + 1>(ct, st, Pl);
      const double Pn = Pl[n][m];
      const double dPt = (n - m + 1.)*Pl[n + 1][m] - (n + 1.)*ct*Pn;
      const double dPp = m*Pn;
      const double ist = 1./st;
      const double cm = cos(m * phi), sm = sin(m * phi);
      const double pn = N.pp[node], utn = N.ut[node], upn = N.up[node];
      const double term[6] = {
        Nnm*Pn*pn*cm, -Nnm*Pn*pn*sm,
        Nnm*ist*(dPt*utn*cm - dPp*upn*sm), Nnm*ist*(-dPt*utn*sm - dPp*upn*cm),
        Nnm*ist*(dPp*utn*sm + dPt*upn*cm), Nnm*ist*(dPp*utn*cm - dPt*upn*sm)};
      /* Lebedev weights: 6 axis nodes, 12 edge midpoints, 8 corners */
      const double wgt = node < 6 ? c_A1 : (node < 18 ? c_A2 : c_A3);
      double *acc[6] = {&Pr[0], &Pr[1], &Dv[0], &Dv[1], &Cu[0], &Cu[1]};
      for (int q = 0; q < 6; q++)
        *acc[q] = node == 0 ? term[q] * wgt : *acc[q] + wgt * term[q];
    }

    const double a = Q->r;
    const double sP = a * a / (mu * nu);       /* scale of Pr */
    const double sV = a / nu;                  /* scale of Dv and Cu */
    double cp[2], cf[2], cc[2];
    if (n == 0) {
      for (int c = 0; c < 2; c++) {
        cp[c] = sP * Pr[c];
        cf[c] = 0.;
        cc[c] = 0.;
      }
    } else {
      const double s = a / Q->rs, q = Q->rs / a;
      const double s_np1 = pow(s, n + 1.);
      const double s_2np1 = s_np1 * s_np1 * q;
      const double q_nm1 = pow(q, n - 1.);
      const double q_n = q_nm1 * q, q_np1 = q_n * q;
      const double in1 = 1./(n + 1.), i2n3 = 1./(2.*n + 3.);
      const double K11 = (1. - 0.5*n*(2.*n - 1.)*in1*s_2np1) * q_n;
      const double K12 = n*(2.*n - 1.)*(2.*n + 1.)*in1*s_np1;
      const double K21 = 0.25*n*(2.*(n + 3.)*i2n3
        + (n - 2. - n*(2.*n + 1.)*i2n3*s*s)*s_2np1)*q_np1;
      const double K22 = n*(n + 1. + 0.5*((n - 2.)*(2.*n + 1.)*q*q
        - n*(2.*n - 1.))*s_2np1)*q_nm1;
      const double E = n*(n + 1.)*(s_2np1 - 1.)*q_n;
      const double idet = 1./(K11*K22 + K12*K21);
      for (int c = 0; c < 2; c++) {
        const double rp = sP * Pr[c], rd = sV * Dv[c];
        cp[c] = (rp*K22 + rd*K12) * idet;
        cf[c] = (rd*K11 - rp*K21) * idet;
        cc[c] = sV * Cu[c] * (1./E);
      }
    }
    double *now[6] = {Q->pnm_re, Q->pnm_im, Q->phinm_re, Q->phinm_im,
                      Q->chinm_re, Q->chinm_im};
    const double *old[6] = {Q->pnm_re0, Q->pnm_im0, Q->phinm_re0,
                            Q->phinm_im0, Q->chinm_re0, Q->chinm_im0};
    const double fresh[6] = {cp[0], cp[1], cf[0], cf[1], cc[0], cc[1]};
    for (int q = 0; q < 6; q++)
      now[q][t] = (n == 0 && q >= 2) ? 0.
                                     : old[q][t]*(1. - relax) + relax*fresh[q];
  }
  __syncthreads();
  if (t == 0) {
    /* F = rho_f V (dU/dt - frame acceleration) + pi mu nu vec(p + 6 phi),
     * L = rho_f V a^2 dw/dt + 8 pi mu nu a vec(chi) */
    const double a = Q->r;
    const double vol = 4./3. * PI * a*a*a;
    double pf_re[3], pf_im[3];
    for (int c = 1; c <= 2; c++) {
      pf_re[c] = Q->pnm_re[c] + 6.*Q->phinm_re[c];
      pf_im[c] = Q->pnm_im[c] + 6.*Q->phinm_im[c];
    }
    const v3 Fh = (PI * mu * nu) * order1_vector(pf_re, pf_im);
    const v3 Lh = (8. * PI * mu * nu * a) * order1_vector(Q->chinm_re,
                                                        Q->chinm_im);
    const v3 acc = {Q->udot, Q->vdot, Q->wdot};
    const v3 omdot = {Q->oxdot, Q->oydot, Q->ozdot};
    const v3 F = (rho_f * vol)
               * (acc - frame_accel(gradP, rho_f, s_beta, Q->s, s_ref, g))
               + Fh;
    const v3 L = (rho_f * vol * a*a) * omdot + Lh;
    Q->Fx = F.x; Q->Fy = F.y; Q->Fz = F.z;
    Q->Lx = L.x; Q->Ly = L.y; Q->Lz = L.z;
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
 * there (legendre_nm is not called for m < 0). */
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
      double N_nm = harmonic_norm(n, m);
      double P_nm = m < 0 ? 0. : legendre_nm(n, m, theta);
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
  to_spherical(x, y, z, &r, &theta, &phi);
  double ar = Q->r / r;
  double ra = r / Q->r;
  double v = Q->s;
  for (int n = 0; n <= Q->sorder; n++) {
    int coeff = 0;
    for (int jj = 0; jj < n; jj++) coeff += 2*jj + 1;
    double X = 0.;
    for (int m = -n; m <= n; m++) {
      X += harmonic_norm(n, m) * (m < 0 ? 0. : legendre_nm(n, m, theta))
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

/* Velocity and rotation of a particle centred on the sub-step: the mean of
 * the current and the stored value. */
__device__ inline v3 mid_velocity(const part_struct *Q)
{
  return 0.5 * (v3{Q->u, Q->v, Q->w} + v3{Q->u0, Q->v0, Q->w0});
}
__device__ inline v3 mid_rotation(const part_struct *Q)
{
  return 0.5 * (v3{Q->ox, Q->oy, Q->oz} + v3{Q->ox0, Q->oy0, Q->oz0});
}

/* ------------------------- contact bookkeeping --------------------------- */
/* Each particle lists up to MAX_NEIGHBORS partners it is in contact with
 * (iSt: partner number, or -10 - face for a wall; -1 = free slot) and the
 * Stokes number of each contact at its start (St). */
__device__ inline int contact_find(const part_struct *Q, int partner)
{
  int q = 0;
  while (q < MAX_NEIGHBORS && Q->iSt[q] != partner) q++;
  return q;
}

/* A partner outside contact range leaves the list. */
__device__ inline void contact_forget(part_struct *Q, int partner)
{
  int q = contact_find(Q, partner);
  if (q < MAX_NEIGHBORS) {
    Q->iSt[q] = -1;
    Q->St[q] = 0.;
  }
}

/* Slot of a partner in contact; a new contact takes the first free slot
 * (the last one when all are taken), records St = (rho_p/rho_f)(2a)|U_n|/(9
 * nu) and counts one more collision in *count. */
__device__ inline int contact_slot(part_struct *Q, int partner, double Un,
  double rhof, double nu, int *count)
{
  int q = contact_find(Q, partner);
  if (q == MAX_NEIGHBORS) {
    q = 0;
    while (q < MAX_NEIGHBORS - 1 && Q->iSt[q] != -1) q++;
    Q->iSt[q] = partner;
    Q->St[q] = 1./9.*Q->rho/rhof*2.*Q->r*fabs(Un)/nu;
    *count += 1;
  }
  return q;
}

/* Coefficient of restitution of a wet contact from the dry one and the
 * contact's Stokes number, e = e_dry + (1 + e_dry) ln(x_c/x_0) / St with
 * x_c/x_0 = 1e-4 (clipped at 0), and the damping factor of the soft-sphere
 * dashpot fitted to it, alpha(e) = 2.22 - 2.263 e^0.3948. */
__device__ inline double dashpot_alpha(double e_dry, double St)
{
  double e = e_dry + (1.+e_dry)/St*log(1.e-4);
  if (e < 0) e = 0;
  return -2.263*pow(e,0.3948)+2.22;
}

struct PairForce { v3 F, L; };

/* Near-contact lubrication on sphere i (radius a) from sphere j (radius
 * beta a), surfaces h apart with 0 < h < h_N: the leading terms of the
 * two-sphere resistance functions (Jeffrey & Onishi 1984; Kim & Karrila
 * 1991, ch. 9) for the relative velocity U = U_n n + U_t of i against j and
 * the summed rotation w of the pair,
 *   F = -6 pi mu a (X1 g1 + X2 g2) U_n - 6 pi mu a Y1 g2 U_t
 *       + 8 pi mu a^2 Y2 g2 (w x n),
 *   L = -8 pi mu a^2 Y2 g2 (U_t x n) - 8 pi mu a^3 Y3 g2 w,
 * X1 = beta^2/(1+beta)^2, X2 = beta(1+7beta+beta^2)/(5(1+beta)^3),
 * Y1 = 4beta(2+beta+2beta^2)/(15(1+beta)^3), Y2 = beta(4+beta)/(10(1+beta)^2),
 * Y3 = 2beta/(5(1+beta)).  The singular gap functions are measured from the
 * cut-off, g1 = a/h - a/h_N and g2 = ln(h_N/h), so the force vanishes
 * continuously at h_N; h is floored at eps a. */
__device__ inline PairForce pair_lubrication(double a, double beta, double h,
  double hN, double eps, double mu, v3 n, v3 Un, v3 Ut, v3 w)
{
  if (h < eps*a) h = eps*a;
  const double g1 = a/h - a/hN;
  const double g2 = log(hN/h);
  const double b1 = 1. + beta;
  const double X1 = beta*beta / (b1*b1);
  const double X2 = beta*(1.+7.*beta+beta*beta) / (5.*b1*b1*b1);
  const double Y1 = 4.*beta*(2.+beta+2.*beta*beta) / (15.*b1*b1*b1);
  const double Y2 = beta*(4.+beta) / (10.*b1*b1);
  const double Y3 = 2.*beta / (5.*b1);
  const double stokes = 6.*PI*mu*a;           /* translational drag */
  const double rotlet = 8.*PI*mu*a*a;         /* rotational coupling */
  PairForce out;
  out.F = (-stokes*(X1*g1 + X2*g2)) * Un + (-stokes*Y1*g2) * Ut
        + (rotlet*Y2*g2) * cross(w, n);
  out.L = (-rotlet*Y2*g2) * cross(Ut, n) + (-rotlet*a*Y3*g2) * w;
  return out;
}

/* Soft-sphere contact on sphere i (radius a_i, density rho_i, Young's
 * modulus E_i, Poisson ratio s_i) from sphere j, overlapping by
 * delta = -h > 0:
 *  - normal: a Hertz spring k delta^(3/2) with k = (4/3) E* sqrt(R*),
 *    1/E* = (1-s_i^2)/E_i + (1-s_j^2)/E_j, 1/R* = 1/a_i + 1/a_j, and a
 *    dashpot -eta U.n with eta = alpha sqrt(m_i k delta^(1/2)),
 *    m_i = (4/3) pi a_i^3 rho_i (alpha from dashpot_alpha);
 *  - tangential: a Mindlin spring k_t = 8 G* sqrt(a_i delta) acting on the
 *    tangential slip V_t dt of the contact point, V = -U_t + ((a_i+a_j+h)/2)
 *    (w x n), capped at Coulomb friction cf |F_n| (cf the mean of the two
 *    friction coefficients); G* is formed from sphere i's shear modulus
 *    G_i = E_i/(2(1+s_i)) for both surfaces, 1/G* = 2(1-s_i^2)/G_i, and
 *    R from a_i alone, as the reference forms it;
 *  - torque: (a_i + h/2) n x F about the centre of i. */
__device__ inline PairForce pair_contact(const part_struct *Pi,
  const part_struct *Pj, double h, double alpha, double dt, v3 n, double un,
  v3 Ut, v3 w)
{
  const double ai = Pi->r, aj = Pj->r;
  const double delta = -h;
  const double Ei = (1.-Pi->sigma*Pi->sigma)/Pi->E;
  const double Ej = (1.-Pj->sigma*Pj->sigma)/Pj->E;
  const double k = 4./3./(Ei + Ej)/sqrt(1./ai + 1./aj);
  const double mass = 4./3.*PI*ai*ai*ai*Pi->rho;
  const double eta = alpha*sqrt(mass*k*sqrt(delta));
  const double Fn_mag = sqrt(delta*delta*delta)*k - eta*un;
  const v3 Fn = Fn_mag * n;

  const double Gi = 0.5*Pi->E/(1.+Pi->sigma);
  const double Gs = (1.-Pi->sigma*Pi->sigma)/Gi;
  const double kt = 8./(Gs + Gs)/sqrt(1./ai)*sqrt(delta);
  const v3 V = (0.5*(ai + aj + h)) * cross(w, n) - Ut;
  const v3 slip = dt * (V - dot(V, n) * n);
  v3 Ft = (-kt) * slip;
  Ft = Ft - dot(Ft, n) * n;
  const double cf = 0.5 * (Pi->coeff_fric + Pj->coeff_fric);
  const double Ft_mag = norm(Ft);
  const double Fn_abs = norm(Fn);
  if (Ft_mag > cf * Fn_abs) Ft = (cf * Fn_abs / Ft_mag) * Ft;

  PairForce out;
  out.F = Fn + Ft;
  out.L = (ai + 0.5*h) * cross(n, out.F);
  return out;
}

/* Lubrication and contact forces on particle i from its neighbours, the
 * physics of src/particle_kernel.cu:3075-3366 (forces above).  Neighbours
 * are visited bin by bin (z, then y, then x offsets -1..1, as the
 * reference) and, inside a bin, in particle order; a bin across a periodic
 * face holds the neighbours' images (positions x +- L).
 * One wavefront per particle i: the lanes test 64 candidates j at a time
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
  const v3 xi = {Pi->x, Pi->y, Pi->z};
  const v3 ui = mid_velocity(Pi), wi = mid_rotation(Pi);
  const double hN = interaction_length_ratio * Pi->r;
  v3 F = {Pi->iFx, Pi->iFy, Pi->iFz};
  v3 L = {Pi->iLx, Pi->iLy, Pi->iLz};
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
        const v3 xj = {img(Pj->x, sh[0], xl), img(Pj->y, sh[1], yl),
                       img(Pj->z, sh[2], zl)};
        const v3 sep = xi - xj;
        const double dist = norm(sep);
        const v3 nrm = (1. / dist) * sep;
        const double h = dist - Pi->r - Pj->r;
        const v3 U = ui - mid_velocity(Pj);
        const v3 w = wi + mid_rotation(Pj);
        const double un = dot(U, nrm);
        const v3 Un = un * nrm;
        const v3 Ut = U - Un;
        if (h < hN && h > 0) {
          contact_forget(Pi, Pj->N);
          PairForce f = pair_lubrication(Pi->r, Pj->r / Pi->r, h, hN, eps,
                                         mu, nrm, Un, Ut, w);
          F = F + f.F;
          L = L + f.L;
        }
        if (h < 0) {
          int q = contact_slot(Pi, Pj->N, un, rhof, nu, &Pi->ncoll_part);
          double alpha = dashpot_alpha(Pi->e_dry, Pi->St[q]);
          PairForce f = pair_contact(Pi, Pj, h, alpha, dt, nrm, un, Ut, w);
          F = F + f.F;
          L = L + f.L;
        }
      }
    }
  }
  if (lane == 0) {                     /* the lane that summed the pairs */
    Pi->iFx = F.x; Pi->iFy = F.y; Pi->iFz = F.z;
    Pi->iLx = L.x; Pi->iLy = L.y; Pi->iLz = L.z;
  }
}

/* Linear spring of stiffness k and rest length l between the particle
 * centre x (first wrapped back into the domain) and its anchor A:
 * kF = -k (d - l d/|d|) with d = x - A; a component of d below 1e-16 in
 * magnitude takes no rest-length part (the spring of
 * src/particle_kernel.cu:2303-2345). */
__global__ void spring_kernel(int nparts, part_struct *parts, dom_struct D)
{
  int i = threadIdx.x + blockIdx.x * blockDim.x;
  if (i >= nparts || !(parts[i].spring_k > 0.)) return;
  part_struct *Q = &parts[i];
  const v3 lo = {D.xs, D.ys, D.zs}, hi = {D.xe, D.ye, D.ze},
           len = {D.xl, D.yl, D.zl};
  const double xc[3] = {Q->x, Q->y, Q->z};
  const double l3[3] = {len.x, len.y, len.z}, lo3[3] = {lo.x, lo.y, lo.z},
               hi3[3] = {hi.x, hi.y, hi.z};
  double xw[3];
  for (int c = 0; c < 3; c++)
    xw[c] = xc[c] + l3[c] * ((xc[c] < lo3[c]) - (xc[c] > hi3[c]));
  const v3 d = v3{xw[0], xw[1], xw[2]}
             - v3{Q->spring_x, Q->spring_y, Q->spring_z};
  const v3 dir = (1. / norm(d)) * d;
  const double dd[3] = {d.x, d.y, d.z}, u3[3] = {dir.x, dir.y, dir.z};
  double F[3];
  for (int c = 0; c < 3; c++) {
    const double rest = (dd[c]*dd[c] > 1e-16*1e-16) ? Q->spring_l * u3[c] : 0.;
    F[c] = -Q->spring_k * (dd[c] - rest);
  }
  Q->kFx = F[0]; Q->kFy = F[1]; Q->kFz = F[2];
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
      contact_forget(Q, code);
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
      int q = contact_slot(Q, code, Un, rhof, nu, &Q->ncoll_wall);
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
      double alpha = dashpot_alpha(Q->e_dry, Q->St[q]);
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

/* Linear and angular acceleration from the hydrodynamic (F, L), spring
 * (kF), interaction (iF, iL) and applied (aF, aL) loads, the mean pressure
 * gradient's buoyancy on the particle volume and reduced gravity; velocities
 * advanced by the trapezoidal rule over the stored accelerations
 * (the motion of src/particle_kernel.cu:3368-3410). */
__global__ void move_parts_a_kernel(int nparts, part_struct *parts, double dt,
  g_struct g, gradP_struct gradP, double rho_f)
{
  int pp = threadIdx.x + blockIdx.x * blockDim.x;
  if (pp >= nparts) return;
  part_struct *Q = &parts[pp];
  const double vol = 4./3. * PI * Q->r*Q->r*Q->r;
  const double mass = vol * Q->rho;
  if (Q->translating) {
    const double g_red = (Q->rho - rho_f) / Q->rho;
    const v3 load = v3{Q->Fx, Q->Fy, Q->Fz} + v3{Q->kFx, Q->kFy, Q->kFz}
                  + v3{Q->iFx, Q->iFy, Q->iFz} + v3{Q->aFx, Q->aFy, Q->aFz}
                  - vol * v3{gradP.x, gradP.y, gradP.z};
    const v3 acc = (1. / mass) * load + g_red * v3{g.x, g.y, g.z};
    const v3 acc0 = {Q->udot0, Q->vdot0, Q->wdot0};
    const v3 vel = v3{Q->u0, Q->v0, Q->w0} + (0.5*dt) * (acc + acc0);
    Q->udot = acc.x; Q->vdot = acc.y; Q->wdot = acc.z;
    Q->u = vel.x; Q->v = vel.y; Q->w = vel.z;
  }
  if (Q->rotating) {
    const double inertia = 0.4 * mass * Q->r*Q->r;
    const v3 torque = v3{Q->Lx, Q->Ly, Q->Lz} + v3{Q->iLx, Q->iLy, Q->iLz}
                    + v3{Q->aLx, Q->aLy, Q->aLz};
    const v3 alpha = (1. / inertia) * torque;
    const v3 alpha0 = {Q->oxdot0, Q->oydot0, Q->ozdot0};
    const v3 om = v3{Q->ox0, Q->oy0, Q->oz0} + (0.5*dt) * (alpha + alpha0);
    Q->oxdot = alpha.x; Q->oydot = alpha.y; Q->ozdot = alpha.z;
    Q->ox = om.x; Q->oy = om.y; Q->oz = om.z;
  }
}

/* Unit quaternion s + v (v = (x, y, z)). */
struct quat { double s; v3 v; };

__device__ inline quat qmul(quat a, quat b)
{
  return {a.s*b.s - dot(a.v, b.v), a.s*b.v + b.s*a.v + cross(a.v, b.v)};
}

/* Vector p turned by the rotation q: the vector part of q (0, p) q*. */
__device__ inline v3 qrotate(quat q, v3 p)
{
  quat qc = {q.s, -1. * q.v};
  return qmul(qmul(q, quat{0., p}), qc).v;
}

/* Centre advanced by the trapezoidal rule; body axes turned through the
 * angle |w| dt about the axis of the mean rotation rate w = (w + w0)/2
 * (|w| taken as the mean of the two magnitudes); current velocities and
 * accelerations stored as the old ones (src/particle_kernel.cu:3412-3470);
 * then the periodic wrap of a centre that left the domain (the reference's
 * particle transfer, src/particle_kernel.cu:1166-1210). */
__global__ void move_parts_b_kernel(int nparts, part_struct *parts, double dt,
  dom_struct D, int perx, int pery, int perz)
{
  int pp = threadIdx.x + blockIdx.x * blockDim.x;
  if (pp >= nparts) return;
  part_struct *Q = &parts[pp];
  if (Q->translating) {
    const v3 vel = {Q->u, Q->v, Q->w}, vel0 = {Q->u0, Q->v0, Q->w0};
    const v3 x = v3{Q->x, Q->y, Q->z} + (0.5*dt) * (vel + vel0);
    Q->x = x.x; Q->y = x.y; Q->z = x.z;
    Q->u0 = vel.x; Q->v0 = vel.y; Q->w0 = vel.z;
    Q->udot0 = Q->udot; Q->vdot0 = Q->vdot; Q->wdot0 = Q->wdot;
  }
  if (Q->rotating) {
    const v3 om = {Q->ox, Q->oy, Q->oz}, om0 = {Q->ox0, Q->oy0, Q->oz0};
    const double rate = 0.5 * norm(om) + 0.5 * norm(om0);
    v3 axis = {0., 0., 0.};
    if (rate > 0) axis = (0.5 / rate) * (om + om0);
    const double half = 0.5 * (rate * dt);
    const quat q = {cos(half), sin(half) * axis};
    const v3 ex = qrotate(q, v3{Q->axx, Q->axy, Q->axz});
    const v3 ey = qrotate(q, v3{Q->ayx, Q->ayy, Q->ayz});
    const v3 ez = qrotate(q, v3{Q->azx, Q->azy, Q->azz});
    Q->axx = ex.x; Q->axy = ex.y; Q->axz = ex.z;
    Q->ayx = ey.x; Q->ayy = ey.y; Q->ayz = ey.z;
    Q->azx = ez.x; Q->azy = ez.y; Q->azz = ez.z;
    Q->ox0 = om.x; Q->oy0 = om.y; Q->oz0 = om.z;
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
  _bb_fluid = (uint8_t *) bb_grid_alloc(ddom.cc.n, 1);
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
  bb_grid_free(_bb_fluid);
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
