/*
 * Particles and Physalis in the FP64 host oracle (included by oracle.c).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows, function by function:
 *   part.config            src/particle.c:66-391 (grammar), :393-554 (init)
 *   bins                   src/particle.c:744-834, src/particle_kernel.cu:716-789
 *   cages                  src/cuda_particle.cu:1516-1646,
 *                          src/particle_kernel.cu:79-576
 *   particle BCs           src/particle_kernel.cu:1370-1653, 1758-1837,
 *                          1839-2170, 2172-2288
 *   particle PCG           src/cuda_solver.cu:39-300, src/solver_kernel.cu:
 *                          178-256 (refinement), 528-713 (SpMV with phases)
 *   Physalis               src/cuda_physalis.cu:61-220, 859-903,
 *                          src/physalis_kernel.cu:26-841, 1559-1648
 *   collisions and motion  src/particle_kernel.cu:2290-2345, 3075-3484,
 *                          src/cuda_particle.cu:1335-1513
 * Periodic images: the reference carries ghost copies of particles near a
 * periodic face at x +- L (src/particle_kernel.cu:1166-1210); the oracle
 * builds a cage for every image, reads a cell or face near a particle
 * through the nearest image, and takes each quadrature node from the image
 * (of the 27) that puts the node's cell inside the domain.  The quadrature
 * sums all 26 nodes of a particle in one sequence (the reference sums per
 * copy and then adds the copies' partial sums; the two differ only in
 * rounding, and only for particles whose shell crosses a periodic face).
 * Walls: a node beyond a wall face takes that face's code and value
 * (src/physalis_kernel.cu:118-214, 437-640), and the wall forces of
 * src/particle_kernel.cu:2347-3073 are restated face by face below.
 *
 * Pinning (each group of functions against mathematics that does not come
 * from the reference's source):
 *   Lebedev nodes/weights  closed-form weights, exactness on spherical
 *                          harmonics to degree 7 (test_oracle_particles.py)
 *   Lamb series (BCs, p fill, internal velocity)
 *                          Stokes equations and no slip of every p/phi
 *                          term, Lamb's toroidal radial function, the
 *                          reference's own cuda_lamb_test boundary values
 *                          (test_oracle_lamb.py)
 *   node interpolation     exact on linear fields, second order with
 *                          periodic images (test_oracle_lamb.py)
 *   quadrature + coefficient solve
 *                          round trip of chosen coefficients of orders
 *                          0-3, order-4 aliasing stays out of n <= 3,
 *                          Stokes drag, rotlet torque, Batchelor's strain
 *                          flow (test_oracle_lamb.py)
 *   cages, masked operator brute-force cage geometry, the masked-operator
 *                          identity (test_oracle_particles.py)
 *   pair lubrication/contact
 *                          Jeffrey-Onishi asymptotes (equal and unequal
 *                          spheres, reciprocity), Hertz normal force
 *                          (test_oracle_dynamics.py)
 *   wall forces            closed-form lubrication, Hertz contact, mirror
 *                          symmetries between faces (test_oracle_walls.py)
 *   springs                Hooke's law, periodic wrap, oscillator period
 *                          (test_oracle_dynamics.py)
 *   motion, rotation       free fall, Rodrigues' formula for the body axes
 *                          (test_oracle_dynamics.py)
 * Parity unpinned: the reference's thrust summation order (the canonical
 * order of the library is the definition here).
 */

#define O_NNODES 26
#define O_NCOEFF 15
#define O_SNCOEFF 25
#define O_MAXNB 12
#ifndef PI
#define PI 3.14159265358979323846
#endif
#define O_DIV_ST -1e-10
#define O_DOM_BUF 1

typedef struct {
  int N, order, ncoeff, translating, rotating;
  double r, x, y, z, u, v, w, u0, v0, w0;
  double udot, vdot, wdot, udot0, vdot0, wdot0;
  double axx, axy, axz, ayx, ayy, ayz, azx, azy, azz;
  double ox, oy, oz, ox0, oy0, oz0, oxdot, oydot, ozdot, oxdot0, oydot0, ozdot0;
  double Fx, Fy, Fz, Lx, Ly, Lz, aFx, aFy, aFz, aLx, aLy, aLz;
  double kFx, kFy, kFz, iFx, iFy, iFz, iLx, iLy, iLz;
  int nodes[O_NNODES];
  double rho, E, sigma, rs, spring_k, spring_x, spring_y, spring_z, spring_l;
  double St[O_MAXNB];
  int iSt[O_MAXNB];
  double e_dry, coeff_fric, s;
  int ncoll_part, ncoll_wall;
  /* scalar (heat) state, src/particle.h:426-445 */
  int sorder, sncoeff, update;
  double srs, q, cp;
  double anm_re[O_SNCOEFF], anm_im[O_SNCOEFF], anm_re0[O_SNCOEFF],
         anm_im0[O_SNCOEFF];
  double pnm_re[O_NCOEFF], pnm_im[O_NCOEFF], phinm_re[O_NCOEFF],
         phinm_im[O_NCOEFF], chinm_re[O_NCOEFF], chinm_im[O_NCOEFF];
  double pnm_re0[O_NCOEFF], pnm_im0[O_NCOEFF], phinm_re0[O_NCOEFF],
         phinm_im0[O_NCOEFF], chinm_re0[O_NCOEFF], chinm_im0[O_NCOEFF];
} OP;

static OP *P_ = NULL;
static int NP_ = 0;
static double ilr_;                     /* (l/a) */
static int ncmax_;
static struct {
  double xs, ys, zs, dx, dy, dz;
  int xn, yn, zn;
} BIN;

/* ----------------------------- input ------------------------------------- */
static int parts_parse(const char *dir)
{
  char path[1024];
  snprintf(path, sizeof(path), "%s/input/part.config", dir);
  NP_ = 0;
  FILE *f = fopen(path, "r");
  if (!f) return 0;
  if (fscanf(f, " n %d", &NP_) != 1) { fclose(f); return -5; }
  if (fscanf(f, " (l/a) %lf", &ilr_) != 1) { fclose(f); return -5; }
  P_ = (OP *) calloc(NP_ > 0 ? NP_ : 1, sizeof(OP));
  ncmax_ = -1;
  for (int i = 0; i < NP_; i++) {
    OP *Q = &P_[i];
    double rs_r, t_s, t_cp, t_rs;
    int t_up, t_so;
    int ok = 1;
    ok &= fscanf(f, " r %lf", &Q->r) == 1;
    ok &= fscanf(f, " (x, y, z) %lf %lf %lf", &Q->x, &Q->y, &Q->z) == 3;
    ok &= fscanf(f, " (u, v, w) %lf %lf %lf", &Q->u, &Q->v, &Q->w) == 3;
    ok &= fscanf(f, " (aFx, aFy, aFz) %lf %lf %lf", &Q->aFx, &Q->aFy,
                 &Q->aFz) == 3;
    ok &= fscanf(f, " (aLx, aLy, aLz) %lf %lf %lf", &Q->aLx, &Q->aLy,
                 &Q->aLz) == 3;
    ok &= fscanf(f, " rho %lf", &Q->rho) == 1;
    ok &= fscanf(f, " E %lf", &Q->E) == 1;
    ok &= fscanf(f, " sigma %lf", &Q->sigma) == 1;
    ok &= fscanf(f, " e_dry %lf", &Q->e_dry) == 1;
    ok &= fscanf(f, " coeff_fric %lf", &Q->coeff_fric) == 1;
    ok &= fscanf(f, " order %d", &Q->order) == 1;
    ok &= fscanf(f, " rs/r %lf", &rs_r) == 1;
    ok &= fscanf(f, " spring_k %lf", &Q->spring_k) == 1;
    ok &= fscanf(f, " spring (x, y, z) %lf %lf %lf", &Q->spring_x,
                 &Q->spring_y, &Q->spring_z) == 3;
    ok &= fscanf(f, " spring_l %lf", &Q->spring_l) == 1;
    ok &= fscanf(f, " translating %d", &Q->translating) == 1;
    ok &= fscanf(f, " rotating %d", &Q->rotating) == 1;
    ok &= fscanf(f, " s %lf", &t_s) == 1;
    ok &= fscanf(f, " update %d", &t_up) == 1;
    ok &= fscanf(f, " cp %lf", &t_cp) == 1;
    ok &= fscanf(f, " rs %lf", &t_rs) == 1;
    ok &= fscanf(f, " s_order %d", &t_so) == 1;
    if (!ok) { fclose(f); return -5; }
    /* periodic: a centre exactly at the end wraps to the start
     * (src/particle.c:176-194) */
    if (O.per[0]) Q->x -= O.xl * (Q->x == O.xe);
    if (O.per[1]) Q->y -= O.yl * (Q->y == O.ye);
    if (O.per[2]) Q->z -= O.zl * (Q->z == O.ze);
    Q->N = i;
    Q->rs = rs_r * Q->r;
    /* src/particle.c:131-137: the scalar state counts only with SCALAR */
    Q->s = t_s * (O_scalar >= 1);
    Q->update = t_up * (O_scalar >= 1);
    Q->cp = t_cp;
    Q->srs = t_rs;
    Q->sorder = t_so;
    Q->ncoeff = 0;
    for (int j = 0; j <= Q->order; j++) Q->ncoeff += j + 1;
    if (Q->ncoeff > ncmax_) ncmax_ = Q->ncoeff;
    Q->u0 = Q->u; Q->v0 = Q->v; Q->w0 = Q->w;
    Q->axx = 1.; Q->ayy = 1.; Q->azz = 1.;
    for (int j = 0; j < O_NNODES; j++) Q->nodes[j] = -1;
    for (int j = 0; j < O_MAXNB; j++) { Q->St[j] = 0.; Q->iSt[j] = -1; }
  }
  fclose(f);
  if (NP_ > 0) {
    double amax = -1.;
    for (int i = 0; i < NP_; i++) if (P_[i].r > amax) amax = P_[i].r;
    double bsz = 2.*amax + ilr_*amax;
    double ls[3] = {O.xl, O.yl, O.zl}, ss[3] = {O.xs, O.ys, O.zs};
    int nb[3];
    double db[3];
    for (int d = 0; d < 3; d++) {
      nb[d] = (int) floor(ls[d] / bsz);
      if (nb[d] == 0) nb[d] = 1;
      db[d] = ls[d] / nb[d];
    }
    BIN.xs = ss[0]; BIN.ys = ss[1]; BIN.zs = ss[2];
    BIN.xn = nb[0]; BIN.yn = nb[1]; BIN.zn = nb[2];
    BIN.dx = db[0]; BIN.dy = db[1]; BIN.dz = db[2];
  }
  return 0;
}

static double oimg(double p, int s, double l)
{
  return s > 0 ? p + l : (s < 0 ? p - l : p);
}

static double onear(double q, double p, double l, int per)
{
  if (!per) return p;
  double d = q - p;
  if (d > 0.5 * l) return p + l;
  if (d < -0.5 * l) return p - l;
  return p;
}

/* ------------------------------ cages ------------------------------------ */
static void ocage(double xp, double xs, double dx, double r, int ncell,
  int per, int *s, int *e)
{
  int cage = (int) (2. * ceil(r / dx)) + 2;
  cage -= (ncell % 2);
  int a = (int) (round((xp - xs) * (1./dx)) - 0.5*cage + O_DOM_BUF);
  int b = a + cage;
  int lo = per ? 0 : 1, hi = per ? ncell + 1 : ncell;
  a += (lo - a) * (a < lo) + (hi - a) * (a > hi);
  b += (lo - b) * (b < lo) + (hi - b) * (b > hi);
  *s = a;
  *e = b;
}

static int inball(double xx, double yy, double zz, double irad)
{
  double d = sqrt(xx*xx + yy*yy + zz*zz);
  return floor(d * irad) < 1;
}

void orc_build_cages(void)
{
  orc_build_flags();                    /* resets flags, external walls */
  if (NP_ == 0) return;
  /* with particles the cage flags below rewrite every face; the external
   * faces are set again after them */
  for (long c = 0; c < O.cc.n; c++) { phase_[c] = -1; shell_[c] = 1; }
  for (int pass = 0; pass < 2; pass++)
  for (int n = 0; n < NP_; n++)
  for (int im = 0; im < 27; im++) {
    int sx = im % 3 - 1, sy = (im / 3) % 3 - 1, sz = im / 9 - 1;
    if ((sx && !O.per[0]) || (sy && !O.per[1]) || (sz && !O.per[2])) continue;
    OP *Q = &P_[n];
    double X = oimg(Q->x, sx, O.xl), Y = oimg(Q->y, sy, O.yl),
           Z = oimg(Q->z, sz, O.zl);
    int is, ie, js, je, ks, ke;
    ocage(X, O.xs, O.dx, Q->r, O.nx, O.per[0], &is, &ie);
    ocage(Y, O.ys, O.dy, Q->r, O.ny, O.per[1], &js, &je);
    ocage(Z, O.zs, O.dz, Q->r, O.nz, O.per[2], &ks, &ke);
    if (is == ie || js == je || ks == ke) continue;
    double irad = 1./Q->r;
    for (int tk = ks; tk <= ke; tk++)
    for (int tj = js; tj <= je; tj++)
    for (int ti = is; ti <= ie; ti++) {
      long C = IDX(O.cc, ti, tj, tk);
      double xx = (ti - 0.5)*O.dx - (X - O.xs);
      double yy = (tj - 0.5)*O.dy - (Y - O.ys);
      double zz = (tk - 0.5)*O.dz - (Z - O.zs);
      if (pass == 0) {
        phase_[C] += inball(xx, yy, zz, irad) * (n - phase_[C]);
      } else {
        double xw = (ti - 1 - 0.5)*O.dx - (X - O.xs);
        double xe = (ti + 1 - 0.5)*O.dx - (X - O.xs);
        double ys = (tj - 1 - 0.5)*O.dy - (Y - O.ys);
        double yn = (tj + 1 - 0.5)*O.dy - (Y - O.ys);
        double zb = (tk - 1 - 0.5)*O.dz - (Z - O.zs);
        double zt = (tk + 1 - 0.5)*O.dz - (Z - O.zs);
        int out = !inball(xw, yy, zz, irad) || !inball(xe, yy, zz, irad) ||
                  !inball(xx, ys, zz, irad) || !inball(xx, yn, zz, irad) ||
                  !inball(xx, yy, zb, irad) || !inball(xx, yy, zt, irad);
        shell_[C] *= 1 - (phase_[C] == n && out);
      }
    }
  }
  /* face flags over every face index 1..n+1 and every ghost row */
  for (int k = 0; k <= O.nz + 1; k++) for (int j = 0; j <= O.ny + 1; j++)
  for (int i = 1; i <= O.nx + 1; i++) {
    long E = IDX(O.cc, i, j, k), W = IDX(O.cc, i - 1, j, k);
    flu_[IDX(O.fx, i, j, k)] = 1 - 2*((phase_[W] < 0 && phase_[E] > -1) ||
      (phase_[W] > -1 && phase_[E] < 0) || (shell_[E] < 1 && shell_[W] < 1));
  }
  for (int k = 0; k <= O.nz + 1; k++) for (int j = 1; j <= O.ny + 1; j++)
  for (int i = 0; i <= O.nx + 1; i++) {
    long N = IDX(O.cc, i, j, k), S = IDX(O.cc, i, j - 1, k);
    flv_[IDX(O.fy, i, j, k)] = 1 - 2*((phase_[S] < 0 && phase_[N] > -1) ||
      (phase_[S] > -1 && phase_[N] < 0) || (shell_[N] < 1 && shell_[S] < 1));
  }
  for (int k = 1; k <= O.nz + 1; k++) for (int j = 0; j <= O.ny + 1; j++)
  for (int i = 0; i <= O.nx + 1; i++) {
    long T = IDX(O.cc, i, j, k), B = IDX(O.cc, i, j, k - 1);
    flw_[IDX(O.fz, i, j, k)] = 1 - 2*((phase_[B] < 0 && phase_[T] > -1) ||
      (phase_[B] > -1 && phase_[T] < 0) || (shell_[T] < 1 && shell_[B] < 1));
  }
  /* the external faces last, as the reference orders the flag build */
  o_flags_external();
}

/* ---------------------------- Lamb series -------------------------------- */
static void o_xyz2rtp(double x, double y, double z, double *r, double *theta,
  double *phi)
{
  double XY = x*x + y*y;
  double XYZ = XY + z*z;
  if (XYZ >= 0 && XYZ < O_DIV_ST) XYZ = O_DIV_ST;
  *r = sqrt(XYZ);
  *theta = acos(z / *r);
  if (XY >= 0 && XY < O_DIV_ST) XY = O_DIV_ST;
  *phi = acos(x / sqrt(XY));
  if (y < 0.) *phi = 2.*PI - *phi;
}

static double o_Nnm(int n, int m)
{
  double ft = 1, fb = 1;
  for (int i = 1; i <= (n-m); i++) ft *= (double) i;
  for (int i = 1; i <= (n+m); i++) fb *= (double) i;
  return sqrt((2.*n+1.) / 4. / PI * ft / fb);
}

static double o_Pnm(int n, int m, double theta)
{
  double x = cos(theta), y = sin(theta);
  switch (10 * n + m) {
    case 0: return 1;
    case 10: return x;
    case 11: return -y;
    case 20: return 0.5*(3.*x*x - 1.);
    case 21: return -3.*x*y;
    case 22: return 3.*y*y;
    case 30: return 0.5*x*(5.*x*x - 3.);
    case 31: return -1.5*(5.*x*x - 1.)*y;
    case 32: return 15.*x*y*y;
    case 33: return -15.*y*y*y;
    case 40: return 0.125*(35.*x*x*x*x - 30.*x*x + 3.);
    case 41: return -2.5*(7.*x*x - 3.)*x*y;
    case 42: return 7.5*(7.*x*x - 1.)*y*y;
    case 43: return -105.*x*y*y*y;
    case 44: return 105.*y*y*y*y;
    case 50: return 0.125*x*(63.*x*x*x*x - 70.*x*x + 15.);
    case 51: return -1.875*y*(21.*x*x*x*x - 14.*x*x + 1.);
    case 52: return 52.5*x*y*y*(3.*x*x - 1.);
    case 53: return -52.5*y*y*y*(9.*x*x - 1.);
    case 54: return 945.*x*y*y*y*y;
    case 55: return -945.*y*y*y*y*y;
  }
  return 0;
}

static int o_c0(int n) { int c = 0; for (int j = 0; j < n; j++) c += j + 1; return c; }

static double o_X(int n, double th, double ph, const double *re, const double *im)
{
  int c = o_c0(n), m = 0;
  double sum = o_Nnm(n,m)*o_Pnm(n,m,th)*re[c];
  for (m = 1; m <= n; m++) {
    c++;
    sum += 2.*o_Nnm(n,m)*o_Pnm(n,m,th)*(re[c]*cos(m*ph) - im[c]*sin(m*ph));
  }
  return sum;
}

static double o_st(double th)
{
  double st = sin(th);
  if (st >= 0 && st < O_DIV_ST) st = O_DIV_ST;
  else if (st < 0 && st > -O_DIV_ST) st = -O_DIV_ST;
  return st;
}

static double o_Y(int n, double th, double ph, const double *re, const double *im)
{
  double ct = cos(th), st = o_st(th);
  int c = o_c0(n), m = 0;
  double sum = o_Nnm(n,m)*(-(n+1)*ct/st*o_Pnm(n,m,th)+(n-m+1)/st*o_Pnm(n+1,m,th))
    *re[c];
  for (m = 1; m <= n; m++) {
    c++;
    sum += 2.*o_Nnm(n,m)
      *(-(n+1)*ct/st*o_Pnm(n,m,th)+(n-m+1)/st*o_Pnm(n+1,m,th))
      *(re[c]*cos(m*ph) - im[c]*sin(m*ph));
  }
  return sum;
}

static double o_Z(int n, double th, double ph, const double *re, const double *im)
{
  double st = o_st(th);
  int c = o_c0(n);
  double sum = 0.;
  for (int m = 1; m <= n; m++) {
    c++;
    sum += -2.*m/st*o_Nnm(n,m)*o_Pnm(n,m,th)*(re[c]*sin(m*ph) + im[c]*cos(m*ph));
  }
  return sum;
}

static void o_lamb_vel(int order, double a, double r, double th, double ph,
  const OP *Q, double nu, double *Ux, double *Uy, double *Uz)
{
  double ia = 1. / a, ar = a / r, ra = r * ia;
  double ur = 0.;
  double ut = 0.5*ra*o_Y(0, th, ph, Q->pnm_re, Q->pnm_im);
  double up = 0.5*ra*o_Z(0, th, ph, Q->pnm_re, Q->pnm_im);
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
      * o_X(n, th, ph, Q->pnm_re, Q->pnm_im);
    ur += (n*powranm1
      + 0.5*n*(2.*n-1.-(2.*n+1.)*ra*ra)*powarnp2)
      * o_X(n, th, ph, Q->phinm_re, Q->phinm_im);
    ut += (0.5*(n+3.)*odnp1*od2np3*powranp1
      + 0.25*odnp1*(n-2.-n*(2.*n+1.)*od2np3*ar*ar)*powarn)
      * o_Y(n, th, ph, Q->pnm_re, Q->pnm_im);
    ut += (powranm1
      + 0.5*odnp1*((n-2.)*(2.*n+1.)*ra*ra-n*(2.*n-1.))*powarnp2)
      * o_Y(n, th, ph, Q->phinm_re, Q->phinm_im);
    ut += (powranm1 - powarnp1) * o_Z(n, th, ph, Q->chinm_re, Q->chinm_im);
    up += (0.5*(n+3.)*odnp1*od2np3*powranp1
      + 0.25*odnp1*(n-2.-n*(2.*n+1.)*od2np3*ar*ar)*powarn)
      * o_Z(n, th, ph, Q->pnm_re, Q->pnm_im);
    up += (powranm1
      + 0.5*odnp1*((n-2.)*(2.*n+1.)*ra*ra-n*(2.*n-1.))*powarnp2)
      * o_Z(n, th, ph, Q->phinm_re, Q->phinm_im);
    up += (-powranm1 + powarnp1) * o_Y(n, th, ph, Q->chinm_re, Q->chinm_im);
  }
  ur *= nu * ia; ut *= nu * ia; up *= nu * ia;
  double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
  *Ux = ur*st*cp + ut*ct*cp - up*sp;
  *Uy = ur*st*sp + ut*ct*sp + up*cp;
  *Uz = ur*ct - ut*st;
}

/* u, v or w on faces with flag < 1 next to a particle
 * (src/particle_kernel.cu:1370-1653). */
static void o_part_bc_dir(double *a, const OG *G, const int *fl, int dir)
{
  int ni = dir == 0 ? O.nx + 1 : O.nx + 1, nj = O.ny + 1, nk = O.nz + 1;
  (void) ni;
  long off = dir == 0 ? 1 : (dir == 1 ? O.cc.sy : O.cc.sz);
  int i0 = dir == 0 ? 1 : 0, j0 = dir == 1 ? 1 : 0, k0 = dir == 2 ? 1 : 0;
  #pragma omp parallel for schedule(static)
  for (int k = k0; k <= nk; k++)
  for (int j = j0; j <= nj; j++)
  for (int i = i0; i <= O.nx + 1; i++) {
    long c1 = IDX(O.cc, i, j, k), c0 = c1 - off;
    long C = IDX(*G, i, j, k);
    int p0 = phase_[c0], p1 = phase_[c1];
    if (!(fl[C] < 1 && (p0 > -1 || p1 > -1))) continue;
    const OP *Q = &P_[p0 > -1 ? p0 : p1];
    double x = dir == 0 ? (i - O_DOM_BUF) * O.dx + O.xs : (i - 0.5) * O.dx + O.xs;
    double y = dir == 1 ? (j - O_DOM_BUF) * O.dy + O.ys : (j - 0.5) * O.dy + O.ys;
    double z = dir == 2 ? (k - O_DOM_BUF) * O.dz + O.zs : (k - 0.5) * O.dz + O.zs;
    x -= onear(x, Q->x, O.xl, O.per[0]);
    y -= onear(y, Q->y, O.yl, O.per[1]);
    z -= onear(z, Q->z, O.zl, O.per[2]);
    double r, th, ph, Ux, Uy, Uz, A = Q->r, nu = O.nu;
    o_xyz2rtp(x, y, z, &r, &th, &ph);
    o_lamb_vel(Q->order, A, r, th, ph, Q, nu, &Ux, &Uy, &Uz);
    double v;
    if (dir == 0) {
      Ux += Q->u + (Q->oy*z - Q->oz*y);
      Ux += 0.1/nu *(r*r*r*r*r-A*A*A*A*A)/(r*r*r) * (Q->oydot*z - Q->ozdot*y);
      v = Ux;
    } else if (dir == 1) {
      Uy += Q->v + -(Q->ox*z - Q->oz*x);
      Uy += 0.1/nu *(r*r*r*r*r-A*A*A*A*A)/(r*r*r) * -(Q->oxdot*z - Q->ozdot*x);
      v = Uy;
    } else {
      Uz += Q->w + (Q->ox*y - Q->oy*x);
      Uz += 0.1/nu *(r*r*r*r*r-A*A*A*A*A)/(r*r*r) * (Q->oxdot*y - Q->oydot*x);
      v = Uz;
    }
    a[C] = v;
  }
}

static void o_part_bc(double *u, double *v, double *w)
{
  if (NP_ == 0) return;
  o_part_bc_dir(u, &O.fx, flu_, 0);
  o_part_bc_dir(v, &O.fy, flv_, 1);
  o_part_bc_dir(w, &O.fz, flw_, 2);
}

/* p inside particles (src/particle_kernel.cu:1758-1837). */
static void o_part_p_fill(void)
{
  if (NP_ == 0) return;
  FOR_CC {
    long c = IDX(O.cc, i, j, k);
    int n = phase_[c];
    if (n < 0) continue;
    const OP *Q = &P_[n];
    double x = (i-0.5) * O.dx + O.xs, y = (j-0.5) * O.dy + O.ys,
           z = (k-0.5) * O.dz + O.zs;
    x -= onear(x, Q->x, O.xl, O.per[0]);
    y -= onear(y, Q->y, O.yl, O.per[1]);
    z -= onear(z, Q->z, O.zl, O.per[2]);
    double r, th, ph;
    o_xyz2rtp(x, y, z, &r, &th, &ph);
    double A = Q->r, ox = Q->ox, oy = Q->oy, oz = Q->oz;
    double pp = o_X(0, th, ph, Q->pnm_re, Q->pnm_im);
    pp *= O.mu*O.nu/(A*A);
    double oc2 = (oy*z - oz*y) * (oy*z - oz*y);
    oc2 += (ox*z - oz*x) * (ox*z - oz*x);
    oc2 += (ox*y - oy*x) * (ox*y - oy*x);
    double bx = -O_sbeta*(Q->s - O_sref)*O_g[0];
    double by = -O_sbeta*(Q->s - O_sref)*O_g[1];
    double bz = -O_sbeta*(Q->s - O_sref)*O_g[2];
    double acc = (-O.gP[0]/O.rho_f - Q->udot + bx)*x +
                 (-O.gP[1]/O.rho_f - Q->vdot + by)*y +
                 (-O.gP[2]/O.rho_f - Q->wdot + bz)*z;
    pp += 0.5 * O.rho_f * oc2 + O.rho_f * acc;
    p_[c] = pp;
  }
}

/* Rigid velocity on faces inside a particle (src/particle_kernel.cu:2172-2288). */
static void o_internal(void)
{
  if (NP_ == 0) return;
  for (int k = 1; k <= O.nz; k++) for (int j = 1; j <= O.ny; j++)
  for (int i = 1; i <= O.nx + 1; i++) {
    long C = IDX(O.fx, i, j, k);
    int pw = phase_[IDX(O.cc, i - 1, j, k)], pe = phase_[IDX(O.cc, i, j, k)];
    if (!(pw > -1 && pe > -1 && flu_[C] != -1)) continue;
    const OP *Q = &P_[pe];
    double y = (j-0.5) * O.dy + O.ys, z = (k-0.5) * O.dz + O.zs;
    double ry = y - onear(y, Q->y, O.yl, O.per[1]);
    double rz = z - onear(z, Q->z, O.zl, O.per[2]);
    u_[C] = (Q->oy*rz - Q->oz*ry) + Q->u;
  }
  for (int k = 1; k <= O.nz; k++) for (int j = 1; j <= O.ny + 1; j++)
  for (int i = 1; i <= O.nx; i++) {
    long C = IDX(O.fy, i, j, k);
    int ps = phase_[IDX(O.cc, i, j - 1, k)], pn = phase_[IDX(O.cc, i, j, k)];
    if (!(ps > -1 && pn > -1 && flv_[C] != -1)) continue;
    const OP *Q = &P_[pn];
    double x = (i-0.5) * O.dx + O.xs, z = (k-0.5) * O.dz + O.zs;
    double rx = x - onear(x, Q->x, O.xl, O.per[0]);
    double rz = z - onear(z, Q->z, O.zl, O.per[2]);
    v_[C] = (Q->oz*rx - Q->ox*rz) + Q->v;
  }
  for (int k = 1; k <= O.nz + 1; k++) for (int j = 1; j <= O.ny; j++)
  for (int i = 1; i <= O.nx; i++) {
    long C = IDX(O.fz, i, j, k);
    int pb = phase_[IDX(O.cc, i, j, k - 1)], pt = phase_[IDX(O.cc, i, j, k)];
    if (!(pb > -1 && pt > -1 && flw_[C] != -1)) continue;
    const OP *Q = &P_[pt];
    double x = (i-0.5) * O.dx + O.xs, y = (j-0.5) * O.dy + O.ys;
    double rx = x - onear(x, Q->x, O.xl, O.per[0]);
    double ry = y - onear(y, Q->y, O.yl, O.per[1]);
    w_[C] = (Q->ox*ry - Q->oy*rx) + Q->w;
  }
}

/* ------------------------------- Physalis -------------------------------- */
static const int o_nn[O_NCOEFF] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4};
static const int o_mm[O_NCOEFF] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0, 1, 2, 3, 4};
static double o_nt[O_NNODES], o_np[O_NNODES];
static const double o_A1 = 0.598398600683775, o_A2 = 0.478718880547015,
                    o_A3 = 0.403919055461543;

static void o_init_physalis(void)
{
  double PI14 = 0.25 * PI, PI12 = 0.5 * PI, PI34 = 0.75 * PI;
  double PI54 = 1.25 * PI, PI32 = 1.5 * PI, PI74 = 1.75 * PI;
  double a1 = 0.955316618124509, a2 = 2.186276035465284;
  double t[O_NNODES] = {PI12, PI12, PI12, PI12, 0.+O_DIV_ST, PI-O_DIV_ST,
    PI12, PI12, PI12, PI12, PI14, PI14, PI14, PI14, PI34, PI34, PI34, PI34,
    a1, a1, a1, a1, a2, a2, a2, a2};
  double p[O_NNODES] = {0., PI12, PI, PI32, 0., 0.,
    PI14, PI34, PI54, PI74, 0., PI12, PI, PI32, 0., PI12, PI, PI32,
    PI14, PI34, PI54, PI74, PI14, PI34, PI54, PI74};
  memcpy(o_nt, t, sizeof(t));
  memcpy(o_np, p, sizeof(p));
}

static int o_cell(double arg)
{
  int i = (int) floor(arg);
  int r = (int) lrint(arg);
  if (r != i && fabs(r - arg) <= DBL_EPSILON) i = r;
  return i;
}

static double o_interp(const double *f, const OG *G, int i, int j, int k,
  double x, double y, double z, double xx, double yy, double zz)
{
  long C = IDX(*G, i, j, k);
  double dfdx = 0.5*(f[C + 1] - f[C - 1]) * (1. / O.dx);
  double dfdy = 0.5*(f[C + G->sy] - f[C - G->sy]) * (1. / O.dy);
  double dfdz = 0.5*(f[C + G->sz] - f[C - G->sz]) * (1. / O.dz);
  return f[C] + dfdx*(x - xx) + dfdy*(y - yy) + dfdz*(z - zz);
}

static int o_clamp(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

/* Rest-frame p, u_r, u_theta, u_phi at one node (src/physalis_kernel.cu:
 * 216-640). */
static void o_node(OP *Q, int node, double *pp, double *ur, double *ut,
  double *up)
{
  double inu = 1. / O.nu, irho = 1. / O.rho_f;
  double rs2 = Q->rs * Q->rs, rs3 = rs2 * Q->rs, rs5 = rs3 * rs2;
  double irs3 = 1./rs3;
  double a5 = Q->r * Q->r;
  a5 *= a5 * Q->r;
  Q->nodes[node] = -1;
  double th = o_nt[node], ph = o_np[node];
  double xp = Q->rs * sin(th) * cos(ph);
  double yp = Q->rs * sin(th) * sin(ph);
  double zp = Q->rs * cos(th);
  /* the image that puts the node's cell inside 1..n */
  double x = xp + Q->x, y = yp + Q->y, z = zp + Q->z;
  int ic = 0, jc = 0, kc = 0;
  for (int im = 0; im < 27; im++) {
    int sx = im % 3 - 1, sy = (im / 3) % 3 - 1, sz = im / 9 - 1;
    if ((sx && !O.per[0]) || (sy && !O.per[1]) || (sz && !O.per[2])) continue;
    double xi = xp + oimg(Q->x, sx, O.xl), yi = yp + oimg(Q->y, sy, O.yl),
           zi = zp + oimg(Q->z, sz, O.zl);
    int a = o_cell((xi - (O.xs - O.dx)) * (1. / O.dx));
    int b = o_cell((yi - (O.ys - O.dy)) * (1. / O.dy));
    int c = o_cell((zi - (O.zs - O.dz)) * (1. / O.dz));
    if (im == 13 || (a >= 1 && a <= O.nx && b >= 1 && b <= O.ny &&
                     c >= 1 && c <= O.nz)) {
      x = xi; y = yi; z = zi; ic = a; jc = b; kc = c;
      if (a >= 1 && a <= O.nx && b >= 1 && b <= O.ny && c >= 1 && c <= O.nz)
        break;
    }
  }
  int oobi = ic < 1 || ic > O.nx, oobj = jc < 1 || jc > O.ny,
      oobk = kc < 1 || kc > O.nz, oob = oobi || oobj || oobk;
  /* node code (src/physalis_kernel.cu:137-213): faces in the order W E S N
   * B T, the Dirichlet code before the Neumann one, the first that applies
   * wins */
  {
    const double gap[6] = {x - O.xs, x - O.xe, y - O.ys, y - O.ye, z - O.zs,
                           z - O.ze};
    for (int f6 = 0; f6 < 6 && Q->nodes[node] == -1; f6++) {
      int out = (f6 & 1) ? gap[f6] >= 0 : gap[f6] < 0;
      if (!out) continue;
      int anyD = 0, anyN = 0;
      for (int c = 0; c < 3; c++) {
        anyD |= O.vtype[c][f6] == DIRICHLET;
        anyN |= O.vtype[c][f6] == NEUMANN;
      }
      if (anyD) Q->nodes[node] = -100 - 10 * f6;
      else if (anyN) Q->nodes[node] = -101 - 10 * f6;
    }
  }
  int code = Q->nodes[node];
  /* Dirichlet velocity (component c) of the face a node lies beyond,
   * along axis d (0: W/E, 1: S/N, 2: B/T) */
  double wall[3][3];
  for (int c = 0; c < 3; c++)
    for (int d = 0; d < 3; d++)
      wall[c][d] = (code == -100 - 20 * d) * O.vD[c][2 * d]
                 + (code == -110 - 20 * d) * O.vD[c][2 * d + 1];
  int only[3] = {oobi * (1-oobj) * (1-oobk) * (code < -1),
                 (1-oobi) * oobj * (1-oobk) * (code < -1),
                 (1-oobi) * (1-oobj) * oobk * (code < -1)};
  double ox = Q->ox, oy = Q->oy, oz = Q->oz;
  int i = o_clamp(ic, 1, O.nx), j = o_clamp(jc, 1, O.ny), k = o_clamp(kc, 1, O.nz);
  double P = o_interp(p_, &O.cc, i, j, k, x, y, z, (i - 0.5) * O.dx + O.xs,
    (j - 0.5) * O.dy + O.ys, (k - 0.5) * O.dz + O.zs);
  double oc2 = (oy*zp - oz*yp) * (oy*zp - oz*yp);
  oc2 += (ox*zp - oz*xp) * (ox*zp - oz*xp);
  oc2 += (ox*yp - oy*xp) * (ox*yp - oy*xp);
  double bx = -O_sbeta*(Q->s - O_sref)*O_g[0];
  double by = -O_sbeta*(Q->s - O_sref)*O_g[1];
  double bz = -O_sbeta*(Q->s - O_sref)*O_g[2];
  double acc = (-O.gP[0] * irho - Q->udot + bx)*xp +
               (-O.gP[1] * irho - Q->vdot + by)*yp +
               (-O.gP[2] * irho - Q->wdot + bz)*zp;
  P -= 0.5 * O.rho_f * oc2 + O.rho_f * acc;
  P = P * (1 - oob) * (Q->nodes[node] == -1);
  int in = (1-oobi) * (1-oobj) * (1-oobk) * (Q->nodes[node] == -1);

  i = o_clamp(ic, 1, O.nx + 1); j = o_clamp(jc, 1, O.ny); k = o_clamp(kc, 1, O.nz);
  double uu = o_interp(u_, &O.fx, i, j, k, x, y, z, (i - O_DOM_BUF) * O.dx + O.xs,
    (j - 0.5) * O.dy + O.ys, (k - 0.5) * O.dz + O.zs);
  double tu = Q->u + (oy*zp - oz*yp) + 0.1 * inu * (rs5 - a5) * irs3 *
        (Q->oydot*zp - Q->ozdot*yp);
  uu -= tu;
  uu = in * uu + only[0] * (wall[0][0] - tu) + only[1] * (wall[0][1] - tu)
     + only[2] * (wall[0][2] - tu);
  i = o_clamp(ic, 1, O.nx); j = o_clamp(jc, 1, O.ny + 1); k = o_clamp(kc, 1, O.nz);
  double vv = o_interp(v_, &O.fy, i, j, k, x, y, z, (i - 0.5) * O.dx + O.xs,
    (j - O_DOM_BUF) * O.dy + O.ys, (k - 0.5) * O.dz + O.zs);
  double tv = Q->v + -(ox*zp - oz*xp) + 0.1 * inu * (rs5 - a5) * irs3 *
        -(Q->oxdot*zp - Q->ozdot*xp);
  vv -= tv;
  vv = in * vv + only[0] * (wall[1][0] - tv) + only[1] * (wall[1][1] - tv)
     + only[2] * (wall[1][2] - tv);
  i = o_clamp(ic, 1, O.nx); j = o_clamp(jc, 1, O.ny); k = o_clamp(kc, 1, O.nz + 1);
  double ww = o_interp(w_, &O.fz, i, j, k, x, y, z, (i - 0.5) * O.dx + O.xs,
    (j - 0.5) * O.dy + O.ys, (k - O_DOM_BUF) * O.dz + O.zs);
  double tw = Q->w + (ox*yp - oy*xp) + 0.1 * inu * (rs5 - a5) * irs3 *
        (Q->oxdot*yp - Q->oydot*xp);
  ww -= tw;
  ww = in * ww + only[0] * (wall[2][0] - tw) + only[1] * (wall[2][1] - tw)
     + only[2] * (wall[2][2] - tw);
  *pp = P;
  double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
  *ur = st * (uu * cp + vv * sp) + ww * ct;
  *ut = ct * (uu * cp + vv * sp) - ww * st;
  *up = -uu * sp + vv * cp;
}

/* Quadrature, Lamb coefficients (with relaxation) and forces of one
 * particle from its rest-frame node values (src/physalis_kernel.cu:
 * 642-841). */
static void o_lamb_solve(OP *Q, const double *pp, const double *ur,
  const double *ut, const double *up)
{
  double relax = O_lamb_relax, mu = O.mu, nu = O.nu;
  (void) ur;
  {
    for (int coeff = 0; coeff < Q->ncoeff; coeff++) {
      int nn = o_nn[coeff], m = o_mm[coeff];
      double S[6] = {0, 0, 0, 0, 0, 0};
      for (int node = 0; node < O_NNODES; node++) {
        double th = o_nt[node], ph = o_np[node];
        double N_nm = o_Nnm(nn, m), P_nm = o_Pnm(nn, m, th);
        double P_n1m = o_Pnm(nn + 1, m, th);
        double dPdt = (nn - m + 1.)*P_n1m - (nn + 1.)*cos(th)*P_nm;
        double dPdp = m*P_nm;
        double isth = 1./sin(th), cm = cos(m * ph), sm = sin(m * ph);
        double I[6] = {
          N_nm*P_nm*pp[node]*cm,
          -N_nm*P_nm*pp[node]*sm,
          N_nm*isth*(dPdt * ut[node] * cm - dPdp * up[node] * sm),
          N_nm*isth*(-dPdt * ut[node] * sm - dPdp * up[node] * cm),
          N_nm*isth*(dPdp * ut[node] * sm + dPdt * up[node] * cm),
          N_nm*isth*(dPdp * ut[node] * cm - dPdt * up[node] * sm)};
        double wgt = node < 6 ? o_A1 : (node < 18 ? o_A2 : o_A3);
        for (int q = 0; q < 6; q++)
          S[q] = node == 0 ? I[q] * o_A1 : S[q] + wgt * I[q];
      }
      double ars = Q->r / Q->rs, rsa = Q->rs / Q->r, r2 = Q->r * Q->r;
      double inu = 1./nu, imunu = inu / mu;
      if (nn == 0) {
        Q->pnm_re[coeff] = (1. - relax) * Q->pnm_re0[coeff] + relax * r2 * imunu * S[0];
        Q->pnm_im[coeff] = (1. - relax) * Q->pnm_im0[coeff] + relax * r2 * imunu * S[1];
        Q->phinm_re[coeff] = 0.; Q->phinm_im[coeff] = 0.;
        Q->chinm_re[coeff] = 0.; Q->chinm_im[coeff] = 0.;
      } else {
        double pa1 = pow(ars, nn + 1.);
        double pa2 = pa1 * pa1 * rsa;
        double pr0 = pow(rsa, nn - 1.);
        double pr1 = pr0 * rsa;
        double pr2 = pr1 * rsa;
        double inp1 = 1./(nn + 1.), i2np3 = 1./(2.*nn + 3.);
        double A = (1. - 0.5*nn*(2.*nn - 1.) * inp1 * pa2) * pr1;
        double B = nn*(2.*nn - 1.)*(2.*nn + 1.) * inp1*pa1;
        double C = 0.25*nn*(2.*(nn + 3.)*i2np3
          + (nn - 2. - nn*(2.*nn + 1.)*i2np3*ars*ars)*pa2)*pr2;
        double D = nn*(nn + 1. + 0.5*((nn - 2.)*(2.*nn + 1.)*rsa*rsa
          - nn*(2.*nn - 1.))*pa2)*pr0;
        double idet = 1./ (A*D + B*C);
        double pre = (r2*imunu*S[0]*D + Q->r*inu*S[2]*B) * idet;
        double pim = (r2*imunu*S[1]*D + Q->r*inu*S[3]*B) * idet;
        double fre = (Q->r*inu*S[2]*A - r2*imunu*S[0]*C) * idet;
        double fim = (Q->r*inu*S[3]*A - r2*imunu*S[1]*C) * idet;
        double E = nn*(nn + 1.)*(pa2 - 1.)*pr1;
        double iE = 1./ E;
        double cre = Q->r*inu*S[4] * iE, cim = Q->r*inu*S[5] * iE;
        Q->pnm_re[coeff] = Q->pnm_re0[coeff]*(1. - relax) + relax*pre;
        Q->pnm_im[coeff] = Q->pnm_im0[coeff]*(1. - relax) + relax*pim;
        Q->phinm_re[coeff] = Q->phinm_re0[coeff]*(1. - relax) + relax*fre;
        Q->phinm_im[coeff] = Q->phinm_im0[coeff]*(1. - relax) + relax*fim;
        Q->chinm_re[coeff] = Q->chinm_re0[coeff]*(1. - relax) + relax*cre;
        Q->chinm_im[coeff] = Q->chinm_im0[coeff]*(1. - relax) + relax*cim;
      }
    }
    /* forces (src/physalis_kernel.cu:807-841) */
    double irho = 1./ O.rho_f;
    double vol = 4./3. * PI *  Q->r*Q->r*Q->r;
    double N10 = sqrt(3./4./PI), N11 = sqrt(3./8./PI);
    double bx = -O_sbeta*(Q->s - O_sref)*O_g[0];
    double by = -O_sbeta*(Q->s - O_sref)*O_g[1];
    double bz = -O_sbeta*(Q->s - O_sref)*O_g[2];
    Q->Fx = O.rho_f * vol * (Q->udot + O.gP[0] * irho - bx)
      - PI * mu * nu * 2.*N11 * (Q->pnm_re[2] + 6.*Q->phinm_re[2]);
    Q->Fy = O.rho_f * vol * (Q->vdot + O.gP[1] * irho - by)
      + PI * mu * nu * 2.*N11 * (Q->pnm_im[2] + 6.*Q->phinm_im[2]);
    Q->Fz = O.rho_f * vol * (Q->wdot + O.gP[2] * irho - bz)
      + PI * mu * nu * N10 * (Q->pnm_re[1] + 6.*Q->phinm_re[1]);
    Q->Lx = O.rho_f * vol * Q->r*Q->r * Q->oxdot
      - 8. * PI * mu * nu * 2.*N11 * Q->r * Q->chinm_re[2];
    Q->Ly = O.rho_f * vol * Q->r*Q->r * Q->oydot
      + 8. * PI * mu * nu * 2.*N11 * Q->r * Q->chinm_im[2];
    Q->Lz = O.rho_f * vol * Q->r*Q->r * Q->ozdot
      + 8. * PI * mu * nu * N10 * Q->r * Q->chinm_re[1];
  }
}

static void o_lamb(void)
{
  for (int n = 0; n < NP_; n++) {
    OP *Q = &P_[n];
    double pp[O_NNODES], ur[O_NNODES], ut[O_NNODES], up[O_NNODES];
    for (int node = 0; node < O_NNODES; node++)
      o_node(Q, node, &pp[node], &ur[node], &ut[node], &up[node]);
    o_lamb_solve(Q, pp, ur, ut, up);
  }
}

/* Largest relative coefficient change over particles, then c0 <- c
 * (src/physalis_kernel.cu:1559-1648, src/cuda_physalis.cu:859-903). */
static double o_lamb_err(int *num)
{
  double best = DBL_MIN;
  int pos = -1;
  for (int n = 0; n < NP_; n++) {
    OP *Q = &P_[n];
    double *cur[6] = {Q->pnm_re, Q->pnm_im, Q->phinm_re, Q->phinm_im,
                      Q->chinm_re, Q->chinm_im};
    double *old[6] = {Q->pnm_re0, Q->pnm_im0, Q->phinm_re0, Q->phinm_im0,
                      Q->chinm_re0, Q->chinm_im0};
    double mx = DBL_MIN;
    for (int coeff = 0; coeff < ncmax_; coeff++) {
      double m = DBL_MIN;
      for (int i = 0; i < 6; i++) {
        double cv = cur[i][coeff], zv = cur[i][0];
        int flag = (fabs(cv) > fabs(O_lamb_cut*zv)) * (fabs(cv) > 1.e-16) *
                   (fabs(zv) > DBL_MIN);
        double div = fabs(cv);
        div += (1.e-16 - div) * (div < 1.e-16);
        double e = (double) flag * fabs(cv - old[i][coeff]) / div;
        m += (e - m) * (e > m);
      }
      mx += (m - mx) * (m > mx);
    }
    if (pos < 0 || mx > best) { best = mx; pos = n; }
    for (int i = 0; i < 6; i++)
      for (int coeff = 0; coeff < ncmax_; coeff++) old[i][coeff] = cur[i][coeff];
  }
  *num = pos >= 0 ? P_[pos].N : -1;
  return best;
}

/* --------------------------- collisions, motion -------------------------- */
static int o_bin_axis(double p, double s, double d)
{
  double arg = (p - (s - d)) / d;
  int b = (int) floor(arg);
  int r = (int) lrint(arg);
  if (r != b && fabs(r - arg) <= 10.*DBL_EPSILON) b = r;
  return b;
}

static void o_collide_pair(OP *Pi, const OP *Pj, double xj, double yj,
  double zj)
{
  double eps = 0.01, mu = O.mu, rhof = O.rho_f, nu = O.nu, dt = O.dt;
  double ai = Pi->r, aj = Pj->r, B = aj / ai;
  double hN = (int) ilr_ * Pi->r;
  double rx = Pi->x - xj, ry = Pi->y - yj, rz = Pi->z - zj;
  double ux = 0.5*((Pi->u - Pj->u) + (Pi->u0 - Pj->u0));
  double uy = 0.5*((Pi->v - Pj->v) + (Pi->v0 - Pj->v0));
  double uz = 0.5*((Pi->w - Pj->w) + (Pi->w0 - Pj->w0));
  double r = sqrt(rx*rx + ry*ry + rz*rz);
  double omx = 0.5*((Pi->ox + Pj->ox) + (Pi->ox0 + Pj->ox0));
  double omy = 0.5*((Pi->oy + Pj->oy) + (Pi->oy0 + Pj->oy0));
  double omz = 0.5*((Pi->oz + Pj->oz) + (Pi->oz0 + Pj->oz0));
  double h = r - ai - aj;
  double nx = rx / r, ny = ry / r, nz = rz / r;
  double udotn = ux * nx + uy * ny + uz * nz;
  double unx = udotn * nx, uny = udotn * ny, unz = udotn * nz;
  double utx = ux - unx, uty = uy - uny, utz = uz - unz;
  double opB = 1 + B;
  double ocnx = omy*nz - omz*ny, ocny = -omx*nz + omz*nx, ocnz = omx*ny - omy*nx;
  double utcnx = uty*nz - utz*ny, utcny = -utx*nz + utz*nx,
         utcnz = utx*ny - uty*nx;
  double Fnx, Fny, Fnz, Ftx, Fty, Ftz, Lox, Loy, Loz;
  int q;
  if (h < hN && h > 0) {
    q = 0;
    while (q < O_MAXNB && Pi->iSt[q] != Pj->N) q++;
    if (q < O_MAXNB && Pi->iSt[q] == Pj->N) { Pi->iSt[q] = -1; Pi->St[q] = 0.; }
    if (h < eps*Pi->r) h = eps*Pi->r;
    double ah = ai/h - ai/hN;
    double lnah = log(hN/h);
    Fnx = -1. * B*B / (opB*opB) * ah - B*(1.+7.*B+B*B)/(5.*opB*opB*opB)*lnah;
    Fny = Fnx; Fnz = Fnx;
    Fnx *= 6.*PI*mu*ai*unx;
    Fny *= 6.*PI*mu*ai*uny;
    Fnz *= 6.*PI*mu*ai*unz;
    Ftx = -6.*PI*mu*ai*utx*4.*B*(2.+B+2.*B*B)/(15.*opB*opB*opB)*lnah;
    Fty = -6.*PI*mu*ai*uty*4.*B*(2.+B+2.*B*B)/(15.*opB*opB*opB)*lnah;
    Ftz = -6.*PI*mu*ai*utz*4.*B*(2.+B+2.*B*B)/(15.*opB*opB*opB)*lnah;
    Ftx += 8.*PI*mu*ai*ai*ocnx*B*(4.+B)/(10.*opB*opB)*lnah;
    Fty += 8.*PI*mu*ai*ai*ocny*B*(4.+B)/(10.*opB*opB)*lnah;
    Ftz += 8.*PI*mu*ai*ai*ocnz*B*(4.+B)/(10.*opB*opB)*lnah;
    Lox = -8.*PI*mu*ai*ai*utcnx*B*(4.+B)/(10.*opB*opB)*lnah;
    Loy = -8.*PI*mu*ai*ai*utcny*B*(4.+B)/(10.*opB*opB)*lnah;
    Loz = -8.*PI*mu*ai*ai*utcnz*B*(4.+B)/(10.*opB*opB)*lnah;
    Lox += -8.*PI*mu*ai*ai*ai*omx*2.*B/(5.*opB)*lnah;
    Loy += -8.*PI*mu*ai*ai*ai*omy*2.*B/(5.*opB)*lnah;
    Loz += -8.*PI*mu*ai*ai*ai*omz*2.*B/(5.*opB)*lnah;
  } else {
    Fnx = Fny = Fnz = Ftx = Fty = Ftz = Lox = Loy = Loz = 0;
  }
  if (h < 0) {
    q = 0;
    while (q < O_MAXNB && Pi->iSt[q] != Pj->N) q++;
    if (q == O_MAXNB) {
      q = 0;
      while (q < O_MAXNB - 1 && Pi->iSt[q] != -1) q++;
      Pi->iSt[q] = Pj->N;
      Pi->St[q] = 1./9.*Pi->rho/rhof*2.*Pi->r*fabs(udotn)/nu;
      Pi->ncoll_part += 1;
    }
    double Vx = -utx + 0.5*(ai + aj + h)*ocnx;
    double Vy = -uty + 0.5*(ai + aj + h)*ocny;
    double Vz = -utz + 0.5*(ai + aj + h)*ocnz;
    double Hi = 0.5*Pi->E/(1.+Pi->sigma);
    double kt = 8./((1.-Pi->sigma*Pi->sigma)/Hi+(1.-Pi->sigma*Pi->sigma)/Hi)
      /sqrt(1./ai)*sqrt(-h);
    double Vdotn = Vx*nx + Vy*ny + Vz*nz;
    double sx = (Vx - Vdotn * nx) * dt, sy = (Vy - Vdotn * ny) * dt,
           sz = (Vz - Vdotn * nz) * dt;
    double k = 4./3./((1.-Pi->sigma*Pi->sigma)/Pi->E
      + (1.-Pj->sigma*Pj->sigma)/Pj->E)/sqrt(1./ai + 1./aj);
    double e = Pi->e_dry + (1.+Pi->e_dry)/Pi->St[q]*log(1.e-4);
    if (e < 0) e = 0;
    double alpha = -2.263*pow(e,0.3948)+2.22;
    double eta = alpha*sqrt(4./3.*PI*ai*ai*ai*Pi->rho*k*sqrt(-h));
    Fnx = (sqrt(-h*h*h)*k - eta*udotn)*nx;
    Fny = (sqrt(-h*h*h)*k - eta*udotn)*ny;
    Fnz = (sqrt(-h*h*h)*k - eta*udotn)*nz;
    double cf = 0.5 * (Pi->coeff_fric + Pj->coeff_fric);
    Ftx = -kt * sx; Fty = -kt * sy; Ftz = -kt * sz;
    double Ftdn = Ftx*nx + Fty*ny + Ftz*nz;
    Ftx = Ftx - Ftdn * nx; Fty = Fty - Ftdn * ny; Ftz = Ftz - Ftdn * nz;
    double Ft = sqrt(Ftx*Ftx + Fty*Fty + Ftz*Ftz);
    double Fn = sqrt(Fnx*Fnx + Fny*Fny + Fnz*Fnz);
    if (Ft > cf * Fn) {
      Ftx = cf * Fn * Ftx / Ft; Fty = cf * Fn * Fty / Ft; Ftz = cf * Fn * Ftz / Ft;
    }
    Lox = -(ai+0.5*h)*((Fny+Fty)*nz-(Fnz+Ftz)*ny);
    Loy =  (ai+0.5*h)*((Fnx+Ftx)*nz-(Fnz+Ftz)*nx);
    Loz = -(ai+0.5*h)*((Fnx+Ftx)*ny-(Fny+Fty)*nx);
  }
  Pi->iFx += Fnx + Ftx; Pi->iFy += Fny + Fty; Pi->iFz += Fnz + Ftz;
  Pi->iLx += Lox; Pi->iLy += Loy; Pi->iLz += Loz;
}

/* ------------------------- wall forces --------------------------------- */
/* src/particle_kernel.cu:2347-3073, restated face by face.  Each face: the
 * centre-to-plane distance d (planes at xs + dsW, xe - dsE, ...), gap
 * h = |d| - a, lubrication for 0 < h < hN on the fluid side, contact for
 * h < 0.  Contact-list codes -10 (W) .. -15 (T).  The forces count only when
 * the face's pressure condition is Neumann; the bookkeeping runs always. */
typedef struct {
  double h, ah, lnah;
} OLub;

static int o_wall_lub_begin(OP *Q, int code, double h, double hN, OLub *L)
{
  int q = 0;
  while (q < O_MAXNB && Q->iSt[q] != code) q++;
  if (q < O_MAXNB && Q->iSt[q] == code) { Q->iSt[q] = -1; Q->St[q] = 0.; }
  double eps = 0.01, ai = Q->r;
  if (h < eps*Q->r) h = eps*Q->r;
  L->h = h;
  L->ah = ai/h - ai/hN;
  L->lnah = log(hN/h);
  return 0;
}

/* the contact slot of a face: found, or a new one (Stokes number from the
 * normal approach speed) */
static int o_wall_contact_slot(OP *Q, int code, double Un)
{
  int q = 0;
  while (q < O_MAXNB && Q->iSt[q] != code) q++;
  if (q == O_MAXNB) {
    q = 0;
    while (q < O_MAXNB - 1 && Q->iSt[q] != -1) q++;
    Q->iSt[q] = code;
    Q->St[q] = 1./9.*Q->rho/O.rho_f*2.*Q->r*fabs(Un)/O.nu;
    Q->ncoll_wall += 1;
  }
  return q;
}

typedef struct {
  double k, kt, eta, Fn;   /* Hertz and Mindlin stiffness, damping, normal */
} OCon;

static void o_wall_contact_coef(const OP *Q, int q, double h, double Un,
  OCon *C)
{
  double ai = Q->r, sg = Q->sigma;
  double Hi = 0.5*Q->E/(1.+sg);
  C->kt = 8./((1.-sg*sg)/Hi + (1.-sg*sg)/Hi)/sqrt(1./ai)*sqrt(-h);
  C->k = 4./3./((1.-sg*sg)/Q->E + (1.-sg*sg)/Q->E)/sqrt(1./ai);
  double e = Q->e_dry + (1.+Q->e_dry)/Q->St[q]*log(1.e-4);
  if (e < 0) e = 0;
  double alpha = -2.263*pow(e,0.3948)+2.22;
  C->eta = alpha*sqrt(4./3.*PI*ai*ai*ai*Q->rho*C->k*sqrt(-h));
  C->Fn = sqrt(-h*h*h)*C->k - C->eta*Un;
}

/* Coulomb limit on the tangential pair (f1, f2) */
static void o_fric(const OP *Q, double Fn, double *f1, double *f2)
{
  double Ft = sqrt(*f1 * *f1 + *f2 * *f2);
  double cf = Q->coeff_fric;
  if (Ft > fabs(cf * Fn)) {
    *f1 = cf * Fn * *f1 / Ft;
    *f2 = cf * Fn * *f2 / Ft;
  }
}

static void o_collide_walls(void)
{
  const double mu = O.mu, dt = O.dt;
  const double *uD = O.vD[0], *vD = O.vD[1], *wD = O.vD[2];
  for (int i = 0; i < NP_; i++) {
    OP *Q = &P_[i];
    double ai = Q->r, hN = (int) ilr_ * Q->r;
    double c6 = 6.*PI*mu*ai;            /* Stokes drag factor */
    OLub L;
    OCon C;
    double d, h;
    int on;
    /* West: n = +x */
    d = Q->x - (O.xs + O.ds[0]); h = fabs(d) - ai;
    on = O.ptype[0] == NEUMANN;
    if (d > 0 && h < hN && h > 0) {
      o_wall_lub_begin(Q, -10, h, hN, &L);
      double Un = Q->u - uD[0], Uy = Q->v - vD[0], Uz = Q->w - wD[0];
      double Fx = -c6*Un*L.ah;
      double Fy = -c6*Uy*8./15.*L.lnah + 8.*PI*mu*ai*ai*Q->oz*1./10.*L.lnah;
      double Fz = -c6*Uz*8./15.*L.lnah - 8.*PI*mu*ai*ai*Q->oy*1./10.*L.lnah;
      double Ly = -8.*PI*mu*ai*ai*Uz*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->oy*2./5.*L.lnah;
      double Lz = 8.*PI*mu*ai*ai*Uy*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->oz*2./5.*L.lnah;
      Q->iFx += on * Fx; Q->iFy += on * Fy; Q->iFz += on * Fz;
      Q->iLy += on * Ly; Q->iLz += on * Lz;
    }
    if (d > 0 && h < 0) {
      double Un = Q->u - uD[0];
      double Uy = 0.5*(Q->v + Q->v0) - vD[2], Uz = 0.5*(Q->w + Q->w0) - wD[2];
      int q = o_wall_contact_slot(Q, -10, Un);
      double ox = 0.5*(Q->ox + Q->ox0), oz = 0.5*(Q->oz + Q->oz0);
      double lev = ai + 0.5*h;
      double Vy = Uy - lev*oz, Vz = Uz + lev*ox;
      o_wall_contact_coef(Q, q, h, Un, &C);
      double Fy = -C.k * (Vy * dt), Fz = -C.k * (Vz * dt);
      o_fric(Q, C.Fn, &Fy, &Fz);
      Q->iFx += on * C.Fn; Q->iFy += on * Fy; Q->iFz += on * Fz;
      Q->iLy += on * lev * Fz; Q->iLz -= on * lev * Fy;
    }
    /* East: n = -x */
    d = Q->x - (O.xe - O.ds[1]); h = fabs(d) - ai;
    on = O.ptype[1] == NEUMANN;
    if (d < 0 && h < hN && h > 0) {
      o_wall_lub_begin(Q, -11, h, hN, &L);
      double Un = Q->u - uD[1], Uy = Q->v - vD[1], Uz = Q->w - wD[1];
      double Fx = -c6*Un*L.ah;
      double Fy = -c6*Uy*8./15.*L.lnah - 8.*PI*mu*ai*ai*Q->oz*1./10.*L.lnah;
      double Fz = -c6*Uz*8./15.*L.lnah + 8.*PI*mu*ai*ai*Q->oy*1./10.*L.lnah;
      double Ly = 8.*PI*mu*ai*ai*Uz*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->oy*2./5.*L.lnah;
      double Lz = -8.*PI*mu*ai*ai*Uy*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->oz*2./5.*L.lnah;
      Q->iFx += on * Fx; Q->iFy += on * Fy; Q->iFz += on * Fz;
      Q->iLy += on * Ly; Q->iLz += on * Lz;
    }
    if (d < 0 && h < 0) {
      double Un = -(Q->u - uD[1]);
      double Uy = 0.5*(Q->v + Q->v0) - vD[2], Uz = 0.5*(Q->w + Q->w0) - wD[2];
      int q = o_wall_contact_slot(Q, -11, Un);
      double ox = 0.5*(Q->ox + Q->ox0), oz = 0.5*(Q->oz + Q->oz0);
      double lev = ai + 0.5*h;
      double Vy = -(Uy - lev*oz), Vz = -(Uz + lev*ox);
      o_wall_contact_coef(Q, q, h, Un, &C);
      double Fy = -C.kt * (Vy * dt), Fz = -C.kt * (Vz * dt);
      o_fric(Q, C.Fn, &Fy, &Fz);
      Q->iFx -= on * C.Fn; Q->iFy += on * Fy; Q->iFz += on * Fz;
      Q->iLy -= on * lev * Fz;
      /* the z torque adds on * lev * 0 (see the note of this section) */
    }
    /* South: n = +y */
    d = Q->y - (O.ys + O.ds[2]); h = fabs(d) - ai;
    on = O.ptype[2] == NEUMANN;
    if (d > 0 && h < hN && h > 0) {
      o_wall_lub_begin(Q, -12, h, hN, &L);
      double Un = Q->v - vD[2], Ux = Q->u - uD[2], Uz = Q->w - wD[2];
      double Fy = -c6*Un*L.ah;
      double Fx = -c6*Ux*8./15.*L.lnah - 8.*PI*mu*ai*ai*Q->oz*1./10.*L.lnah;
      double Fz = -c6*Uz*8./15.*L.lnah + 8.*PI*mu*ai*ai*Q->ox*1./10.*L.lnah;
      double Lx = 8.*PI*mu*ai*ai*Uz*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->ox*2./5.*L.lnah;
      double Lz = -8.*PI*mu*ai*ai*Ux*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->oz*2./5.*L.lnah;
      Q->iFx += on * Fx; Q->iFy += on * Fy; Q->iFz += on * Fz;
      Q->iLx += on * Lx; Q->iLz += on * Lz;
    }
    if (d > 0 && h < 0) {
      double Un = Q->v - vD[2];
      double Ux = 0.5*(Q->u + Q->u0) - uD[2], Uz = 0.5*(Q->w + Q->w0) - wD[2];
      int q = o_wall_contact_slot(Q, -12, Un);
      double ox = 0.5*(Q->ox + Q->ox0), oz = 0.5*(Q->oz + Q->oz0);
      double lev = ai + 0.5*h;
      double Vx = Ux + lev*oz, Vz = Uz - lev*ox;
      o_wall_contact_coef(Q, q, h, Un, &C);
      double Fx = -C.kt * (Vx * dt), Fz = -C.kt * (Vz * dt);
      o_fric(Q, C.Fn, &Fx, &Fz);
      Q->iFx += on * Fx; Q->iFy += on * C.Fn; Q->iFz += on * Fz;
      Q->iLx -= on * lev * Fz; Q->iLz += on * lev * Fx;
    }
    /* North: n = -y */
    d = Q->y - (O.ye - O.ds[3]); h = fabs(d) - ai;
    on = O.ptype[3] == NEUMANN;
    if (d < 0 && h < hN && h > 0) {
      o_wall_lub_begin(Q, -13, h, hN, &L);
      double Un = Q->v - vD[3], Ux = Q->u - uD[3], Uz = Q->w - wD[3];
      double Fy = -c6*Un*L.ah;
      double Fx = -c6*Ux*8./15.*L.lnah + 8.*PI*mu*ai*ai*Q->oz*1./10.*L.lnah;
      double Fz = -c6*Uz*8./15.*L.lnah - 8.*PI*mu*ai*ai*Q->ox*1./10.*L.lnah;
      double Lx = -8.*PI*mu*ai*ai*Uz*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->ox*2./5.*L.lnah;
      double Lz = 8.*PI*mu*ai*ai*Ux*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->oz*2./5.*L.lnah;
      Q->iFx += on * Fx; Q->iFy += on * Fy; Q->iFz += on * Fz;
      Q->iLx += on * Lx; Q->iLz += on * Lz;
    }
    if (d < 0 && h < 0) {
      double Un = -(Q->v - vD[3]);
      double Ux = 0.5*(Q->u + Q->u0) - uD[2], Uz = 0.5*(Q->w + Q->w0) - wD[2];
      int q = o_wall_contact_slot(Q, -13, Un);
      double ox = 0.5*(Q->ox + Q->ox0), oz = 0.5*(Q->oz + Q->oz0);
      double lev = ai + 0.5*h;
      double Vx = -(Ux + lev*oz), Vz = -(Uz - lev*ox);
      o_wall_contact_coef(Q, q, h, Un, &C);
      double Fx = -C.kt * (Vx * dt), Fz = -C.kt * (Vz * dt);
      o_fric(Q, C.Fn, &Fx, &Fz);
      Q->iFx += on * Fx; Q->iFy -= on * C.Fn; Q->iFz += on * Fz;
      Q->iLx += on * lev * Fz; Q->iLz -= on * lev * Fx;
    }
    /* Bottom: n = +z */
    d = Q->z - (O.zs + O.ds[4]); h = fabs(d) - ai;
    on = O.ptype[4] == NEUMANN;
    if (d > 0 && h < hN && h > 0) {
      o_wall_lub_begin(Q, -14, h, hN, &L);
      double Un = Q->w - wD[4], Ux = Q->u - uD[4], Uy = Q->v - vD[4];
      double Fz = -c6*Un*L.ah;
      double Fx = -c6*Ux*8./15.*L.lnah + 8.*PI*mu*ai*ai*Q->oy*1./10.*L.lnah;
      double Fy = -c6*Uy*8./15.*L.lnah - 8.*PI*mu*ai*ai*Q->ox*1./10.*L.lnah;
      double Lx = -8.*PI*mu*ai*ai*Uy*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->ox*2./5.*L.lnah;
      double Ly = 8.*PI*mu*ai*ai*Ux*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->oy*2./5.*L.lnah;
      Q->iFx += on * Fx; Q->iFy += on * Fy; Q->iFz += on * Fz;
      Q->iLx += on * Lx; Q->iLy += on * Ly;
    }
    if (d > 0 && h < 0) {
      double Un = Q->w - wD[4];
      double Ux = 0.5*(Q->u + Q->u0) - uD[2], Uy = 0.5*(Q->v + Q->v0) - vD[2];
      int q = o_wall_contact_slot(Q, -14, Un);
      double ox = 0.5*(Q->ox + Q->ox0), oy = 0.5*(Q->oy + Q->oy0);
      double lev = ai + 0.5*h;
      double Vx = Ux - lev*oy, Vy = Uy + lev*ox;
      o_wall_contact_coef(Q, q, h, Un, &C);
      double Fx = -C.kt * (Vx * dt), Fy = -C.kt * (Vy * dt);
      o_fric(Q, C.Fn, &Fx, &Fy);
      Q->iFx += on * Fx; Q->iFy += on * Fy; Q->iFz += on * C.Fn;
      Q->iLx += on * lev * Fy; Q->iLy -= on * lev * Fx;
    }
    /* Top: n = -z */
    d = Q->z - (O.ze - O.ds[5]); h = fabs(d) - ai;
    on = O.ptype[5] == NEUMANN;
    if (d < 0 && h < hN && h > 0) {
      o_wall_lub_begin(Q, -15, h, hN, &L);
      double Un = Q->w - wD[5], Ux = Q->u - uD[5], Uy = Q->v - vD[5];
      double Fz = -c6*Un*L.ah;
      double Fx = -c6*Ux*8./15.*L.lnah - 8.*PI*mu*ai*ai*Q->oy*1./10.*L.lnah;
      double Fy = -c6*Uy*8./15.*L.lnah + 8.*PI*mu*ai*ai*Q->ox*1./10.*L.lnah;
      double Lx = 8.*PI*mu*ai*ai*Uy*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->ox*2./5.*L.lnah;
      double Ly = -8.*PI*mu*ai*ai*Ux*1./10.*L.lnah
                  - 8.*PI*mu*ai*ai*ai*Q->oy*2./5.*L.lnah;
      Q->iFx += on * Fx; Q->iFy += on * Fy; Q->iFz += on * Fz;
      Q->iLx += on * Lx; Q->iLy += on * Ly;
    }
    if (d < 0 && h < 0) {
      double Un = -(Q->w - wD[5]);
      double Ux = 0.5*(Q->u + Q->u0) - uD[2], Uy = 0.5*(Q->v + Q->v0) - vD[2];
      int q = o_wall_contact_slot(Q, -15, Un);
      double ox = 0.5*(Q->ox + Q->ox0), oy = 0.5*(Q->oy + Q->oy0);
      double lev = ai + 0.5*h;
      double Vx = -(Ux - lev*oy), Vy = -(Uy + lev*ox);
      o_wall_contact_coef(Q, q, h, Un, &C);
      double Fx = -C.kt * (Vx * dt), Fy = -C.kt * (Vy * dt);
      o_fric(Q, C.Fn, &Fx, &Fy);
      Q->iFx += on * Fx; Q->iFy += on * Fy; Q->iFz -= on * C.Fn;
      Q->iLx -= on * lev * Fy; Q->iLy += on * lev * Fx;
    }
  }
}

static void o_collide(void)
{
  for (int i = 0; i < NP_; i++) {
    P_[i].iFx = P_[i].iFy = P_[i].iFz = 0.;
    P_[i].iLx = P_[i].iLy = P_[i].iLz = 0.;
  }
  if (NP_ > 1) {
    int *bx = (int *) malloc(3 * NP_ * sizeof(int));
    for (int i = 0; i < NP_; i++) {
      bx[3*i] = o_clamp(o_bin_axis(P_[i].x, BIN.xs, BIN.dx), 0, BIN.xn + 1);
      bx[3*i+1] = o_clamp(o_bin_axis(P_[i].y, BIN.ys, BIN.dy), 0, BIN.yn + 1);
      bx[3*i+2] = o_clamp(o_bin_axis(P_[i].z, BIN.zs, BIN.dz), 0, BIN.zn + 1);
    }
    int nbin[3] = {BIN.xn, BIN.yn, BIN.zn};
    double len[3] = {O.xl, O.yl, O.zl};
    for (int i = 0; i < NP_; i++) {
      int *b = &bx[3*i];
      if (b[0] < 1 || b[0] > nbin[0] || b[1] < 1 || b[1] > nbin[1] ||
          b[2] < 1 || b[2] > nbin[2]) continue;
      for (int n = -1; n <= 1; n++) for (int m = -1; m <= 1; m++)
      for (int l = -1; l <= 1; l++) {
        int t[3] = {b[0] + l, b[1] + m, b[2] + n}, sh[3] = {0, 0, 0}, ok = 1;
        for (int d = 0; d < 3; d++) {
          if (t[d] < 1) { t[d] += nbin[d]; sh[d] = -1; ok &= O.per[d]; }
          else if (t[d] > nbin[d]) { t[d] -= nbin[d]; sh[d] = 1; ok &= O.per[d]; }
        }
        if (!ok) continue;
        for (int j = 0; j < NP_; j++) {
          if (j == i || bx[3*j] != t[0] || bx[3*j+1] != t[1] || bx[3*j+2] != t[2])
            continue;
          o_collide_pair(&P_[i], &P_[j], oimg(P_[j].x, sh[0], len[0]),
            oimg(P_[j].y, sh[1], len[1]), oimg(P_[j].z, sh[2], len[2]));
        }
      }
    }
    free(bx);
  }
  for (int i = 0; i < NP_; i++) {
    OP *Q = &P_[i];
    if (!(Q->spring_k > 0.)) continue;
    double x = Q->x, y = Q->y, z = Q->z;
    x += O.xl * ((x < O.xs) - (x > O.xe));
    y += O.yl * ((y < O.ys) - (y > O.ye));
    z += O.zl * ((z < O.zs) - (z > O.ze));
    double nx = x - Q->spring_x, ny = y - Q->spring_y, nz = z - Q->spring_z;
    double n = sqrt(nx*nx+ny*ny+nz*nz);
    double lx = Q->spring_l * (nx / n), ly = Q->spring_l * (ny / n),
           lz = Q->spring_l * (nz / n);
    lx *= (nx*nx > 1e-16*1e-16); ly *= (ny*ny > 1e-16*1e-16);
    lz *= (nz*nz > 1e-16*1e-16);
    Q->kFx = - Q->spring_k * (x - Q->spring_x - lx);
    Q->kFy = - Q->spring_k * (y - Q->spring_y - ly);
    Q->kFz = - Q->spring_k * (z - Q->spring_z - lz);
  }
  o_collide_walls();
}

static void o_move_a(void)
{
  for (int i = 0; i < NP_; i++) {
    OP *Q = &P_[i];
    double vol = 4./3. * PI * Q->r*Q->r*Q->r, m = vol * Q->rho;
    if (Q->translating) {
      Q->udot = (Q->Fx + Q->kFx + Q->iFx + Q->aFx - vol*O.gP[0]) / m
        + (Q->rho - O.rho_f) / Q->rho * O_g[0];
      Q->vdot = (Q->Fy + Q->kFy + Q->iFy + Q->aFy - vol*O.gP[1]) / m
        + (Q->rho - O.rho_f) / Q->rho * O_g[1];
      Q->wdot = (Q->Fz + Q->kFz + Q->iFz + Q->aFz - vol*O.gP[2]) / m
        + (Q->rho - O.rho_f) / Q->rho * O_g[2];
      Q->u = Q->u0 + 0.5*O.dt*(Q->udot + Q->udot0);
      Q->v = Q->v0 + 0.5*O.dt*(Q->vdot + Q->vdot0);
      Q->w = Q->w0 + 0.5*O.dt*(Q->wdot + Q->wdot0);
    }
    if (Q->rotating) {
      double I = 0.4 * m * Q->r*Q->r;
      Q->oxdot = (Q->Lx + Q->iLx + Q->aLx) / I;
      Q->oydot = (Q->Ly + Q->iLy + Q->aLy) / I;
      Q->ozdot = (Q->Lz + Q->iLz + Q->aLz) / I;
      Q->ox = Q->ox0 + 0.5*O.dt*(Q->oxdot + Q->oxdot0);
      Q->oy = Q->oy0 + 0.5*O.dt*(Q->oydot + Q->oydot0);
      Q->oz = Q->oz0 + 0.5*O.dt*(Q->ozdot + Q->ozdot0);
    }
  }
}

static void o_rotate(double qr, double qi, double qj, double qk, double *pi,
  double *pj, double *pk)
{
  double Pr = *pi*qi + *pj*qj + *pk*qk;
  double Pi = *pi*qr - *pj*qk + *pk*qj;
  double Pj = *pi*qk + *pj*qr - *pk*qi;
  double Pk = -*pi*qj + *pj*qi + *pk*qr;
  *pi = qr*Pi + qi*Pr + qj*Pk - qk*Pj;
  *pj = qr*Pj - qi*Pk + qj*Pr + qk*Pi;
  *pk = qr*Pk + qi*Pj - qj*Pi + qk*Pr;
}

static void o_move_b(void)
{
  for (int i = 0; i < NP_; i++) {
    OP *Q = &P_[i];
    if (Q->translating) {
      Q->x += 0.5*O.dt*(Q->u + Q->u0);
      Q->y += 0.5*O.dt*(Q->v + Q->v0);
      Q->z += 0.5*O.dt*(Q->w + Q->w0);
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
      double th = mag * O.dt;
      double qr = cos(0.5*th), qi = X * sin(0.5*th), qj = Y * sin(0.5*th),
             qk = Z * sin(0.5*th);
      o_rotate(qr, qi, qj, qk, &Q->axx, &Q->axy, &Q->axz);
      o_rotate(qr, qi, qj, qk, &Q->ayx, &Q->ayy, &Q->ayz);
      o_rotate(qr, qi, qj, qk, &Q->azx, &Q->azy, &Q->azz);
      Q->ox0 = Q->ox; Q->oy0 = Q->oy; Q->oz0 = Q->oz;
      Q->oxdot0 = Q->oxdot; Q->oydot0 = Q->oydot; Q->ozdot0 = Q->ozdot;
    }
    if (O.per[0]) { if (Q->x >= O.xe) Q->x = Q->x - O.xl; else if (Q->x < O.xs) Q->x = Q->x + O.xl; }
    if (O.per[1]) { if (Q->y >= O.ye) Q->y = Q->y - O.yl; else if (Q->y < O.ys) Q->y = Q->y + O.yl; }
    if (O.per[2]) { if (Q->z >= O.ze) Q->z = Q->z - O.zl; else if (Q->z < O.zs) Q->z = Q->z + O.zl; }
  }
}

static void o_move_sub(void)
{
  if (NP_ == 0) return;
  o_collide();
  o_move_a();
  o_collide();
}

/* ------------------------------ accessors -------------------------------- */
int orc_nparts(void) { return NP_; }

/* Particle state in a fixed order: x y z u v w ox oy oz udot vdot wdot
 * oxdot oydot ozdot Fx Fy Fz Lx Ly Lz iFx iFy iFz iLx iLy iLz
 * axx axy axz ayx ayy ayz azx azy azz, then the 15 pnm_re, pnm_im,
 * phinm_re, phinm_im, chinm_re, chinm_im (126 values in all). */
int orc_part_state(int n, double *out)
{
  if (n < 0 || n >= NP_) return -1;
  const OP *Q = &P_[n];
  double head[36] = {Q->x, Q->y, Q->z, Q->u, Q->v, Q->w, Q->ox, Q->oy, Q->oz,
    Q->udot, Q->vdot, Q->wdot, Q->oxdot, Q->oydot, Q->ozdot, Q->Fx, Q->Fy,
    Q->Fz, Q->Lx, Q->Ly, Q->Lz, Q->iFx, Q->iFy, Q->iFz, Q->iLx, Q->iLy, Q->iLz,
    Q->axx, Q->axy, Q->axz, Q->ayx, Q->ayy, Q->ayz, Q->azx, Q->azy, Q->azz};
  memcpy(out, head, sizeof(head));
  const double *c[6] = {Q->pnm_re, Q->pnm_im, Q->phinm_re, Q->phinm_im,
                        Q->chinm_re, Q->chinm_im};
  for (int i = 0; i < 6; i++) memcpy(out + 36 + 15 * i, c[i], 15 * sizeof(double));
  return 0;
}

/* Set the Lamb coefficients of particle n (same 90-value order as above),
 * for closed-form tests of the series and the quadrature. */
int orc_part_set_coeffs(int n, const double *c)
{
  if (n < 0 || n >= NP_) return -1;
  OP *Q = &P_[n];
  double *d[6] = {Q->pnm_re, Q->pnm_im, Q->phinm_re, Q->phinm_im,
                  Q->chinm_re, Q->chinm_im};
  for (int i = 0; i < 6; i++) memcpy(d[i], c + 15 * i, 15 * sizeof(double));
  return 0;
}

/* Set the motion of particle n: x y z u v w ox oy oz, then the previous
 * u0 v0 w0 ox0 oy0 oz0 (15 values), for tests of the contact forces. */
int orc_part_set_motion(int n, const double *m)
{
  if (n < 0 || n >= NP_) return -1;
  OP *Q = &P_[n];
  Q->x = m[0]; Q->y = m[1]; Q->z = m[2];
  Q->u = m[3]; Q->v = m[4]; Q->w = m[5];
  Q->ox = m[6]; Q->oy = m[7]; Q->oz = m[8];
  Q->u0 = m[9]; Q->v0 = m[10]; Q->w0 = m[11];
  Q->ox0 = m[12]; Q->oy0 = m[13]; Q->oz0 = m[14];
  return 0;
}

/* Contact list of particle n: iSt and St (MAX_NEIGHBORS each), and the
 * particle and wall collision counts. */
int orc_part_contacts(int n, int *iSt, double *St, int *ncoll)
{
  if (n < 0 || n >= NP_) return -1;
  const OP *Q = &P_[n];
  memcpy(iSt, Q->iSt, sizeof(Q->iSt));
  memcpy(St, Q->St, sizeof(Q->St));
  ncoll[0] = Q->ncoll_part;
  ncoll[1] = Q->ncoll_wall;
  return 0;
}

/* One evaluation of the collision forces (pairs, springs, walls) on the
 * current particle state. */
void orc_collide_once(void)
{
  o_collide();
}

/* The sub-step particle update of the Lamb loop (collisions, velocities,
 * collisions again: src/cuda_particle.cu:1335-1473). */
void orc_move_parts_sub(void)
{
  o_move_sub();
}

/* The end-of-step particle update: positions, body axes, stored
 * previous values and the periodic wrap (src/particle_kernel.cu:
 * 3412-3484). */
void orc_move_parts_b(void)
{
  o_move_b();
}

/* Lamb velocity (Cartesian) of particle n's current coefficients at the
 * point (x, y, z) relative to its centre, and the n = 0 pressure term. */
void orc_lamb_eval(int n, double x, double y, double z, double *U)
{
  const OP *Q = &P_[n];
  double r, th, ph;
  o_xyz2rtp(x, y, z, &r, &th, &ph);
  o_lamb_vel(Q->order, Q->r, r, th, ph, Q, O.nu, &U[0], &U[1], &U[2]);
  U[3] = o_X(0, th, ph, Q->pnm_re, Q->pnm_im) * O.mu * O.nu / (Q->r * Q->r);
}

/* The Lamb pressure series of particle n's coefficients at (x, y, z)
 * relative to its centre, without the rest-frame terms (the series of
 * src/particle_kernel.cu:1716-1729). */
double orc_lamb_eval_p(int n, double x, double y, double z)
{
  const OP *Q = &P_[n];
  double r, th, ph;
  o_xyz2rtp(x, y, z, &r, &th, &ph);
  double a = Q->r, ar = a / r, ra = r / a;
  double pp = o_X(0, th, ph, Q->pnm_re, Q->pnm_im);
  for (int nn = 1; nn <= Q->order; nn++) {
    pp += (1.-0.5*nn*(2.*nn-1.)/(nn+1.)*pow(ar,2.*nn+1.))*pow(ra,nn)
      * o_X(nn, th, ph, Q->pnm_re, Q->pnm_im);
    pp -= nn*(2.*nn-1.)*(2.*nn+1.)/(nn+1.)*pow(ar,nn+1.)
      * o_X(nn, th, ph, Q->phinm_re, Q->phinm_im);
  }
  return pp * O.mu * O.nu / (a * a);
}

/* The rest-frame node values of particle n on the current fields
 * (pressure, u_r, u_theta, u_phi at the 26 nodes) and the node codes. */
void orc_part_nodes(int n, double *pp, double *ur, double *ut, double *up,
  int *code)
{
  OP *Q = &P_[n];
  for (int node = 0; node < O_NNODES; node++) {
    o_node(Q, node, &pp[node], &ur[node], &ut[node], &up[node]);
    code[node] = Q->nodes[node];
  }
}

/* Quadrature, coefficients and forces of particle n from given rest-frame
 * node values (the part of the Physalis update after the interpolation). */
void orc_lamb_from_nodes(int n, const double *pp, const double *ur,
  const double *ut, const double *up)
{
  o_lamb_solve(&P_[n], pp, ur, ut, up);
}

/* The particle boundary values on the current coefficients: u, v, w on
 * the faces next to particles (o_part_bc) and p inside them
 * (o_part_p_fill). */
void orc_part_bc_once(void)
{
  o_part_bc(u_, v_, w_);
  o_part_p_fill();
}

/* Run the Physalis update once on the current fields (quadrature, Lamb
 * coefficients with relaxation, forces); returns the Lamb error. */
double orc_lamb_once(void)
{
  int num;
  o_lamb();
  return o_lamb_err(&num);
}

/* The quadrature nodes and weights (for tests of the rule). */
void orc_lebedev(double *theta, double *phi, double *w)
{
  o_init_physalis();
  for (int i = 0; i < O_NNODES; i++) {
    theta[i] = o_nt[i];
    phi[i] = o_np[i];
    w[i] = i < 6 ? o_A1 : (i < 18 ? o_A2 : o_A3);
  }
}
