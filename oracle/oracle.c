/*
 * FP64 host oracle of the Bluebottle per-timestep flow path (C99 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libbluebottle.so.  Nothing in the
 * product loads it.  Each function restates the reference's algorithm and
 * cites the file:line it follows.  It keeps the reference's data flow
 * literally (flag arrays in the stencils, separate z / p / pb CG vectors,
 * D2D copies for the AB2 history, the halo exchange both before and after
 * each boundary-condition pass) so that the GPU library's shortcuts are
 * checked rather than shared.
 *
 * Reductions follow the canonical order documented in the library (tiles
 * of 64 x 8 x 8 cells: column sums along z, a halving tree over the 64
 * columns of a row, a fixed tree over the 8 rows, then 1024 interleaved
 * accumulators over the tiles + halving tree) because the reference's own
 * order is that of thrust, which is not specified; that order is the
 * definition here.
 * Compile with -ffp-contract=off.
 *
 * Pinning: tests/test_oracle.py (u* on exponential and sine fields with
 * second-order convergence, exact Dirichlet ghost formulas, periodic
 * planes, the operator against an independent sparse matrix, the Laplacian
 * on a quadratic, PCG against a dense solve, Taylor-Green decay, cavity
 * invariants); forcing by a steady Poiseuille channel and find_dt's
 * velocity branch (tests/test_oracle_dynamics.py).  Parity unpinned: the
 * reference's thrust summation order (the canonical order above is the
 * definition).
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "oracle.h"

#define PERIODIC 0
#define DIRICHLET 1
#define NEUMANN 2
#define PRECURSOR 3
#define HOMOGENEOUS 10

typedef struct { int ni, nj, nk; long sy, sz, n; } OG;

static struct {
  int nx, ny, nz;
  double xs, xe, ys, ye, zs, ze, xl, yl, zl, dx, dy, dz;
  double rho_f, nu, mu, CFL, duration;
  int pp_max_iter;
  double pp_residual;
  int ptype[6];                       /* W E S N B T */
  double ds[6];                       /* screen offsets of the p BCs */
  int vtype[3][6];                    /* [u,v,w][face] */
  double vDm[3][6], vDa[3][6], vD[3][6];
  double v_bc_tdelay, p_bc_tdelay, g_bc_tdelay;
  double gPm[3], gPa[3], gP[3], osci_f;
  int init_cond;                      /* 0 quiescent 1 shear 2 channel 4 tg3 */
  int out_plane;
  double fixed_dt, perturb_amp;
  unsigned long long perturb_seed;
  int per[3];
  OG cc, fx, fy, fz;
  double dt, dt0, ttime;
  int stepnum;
  int niter;
  double resid, true_resid, sp_rhs;
  int lamb_iters;
  double lamb_err;
} O;

static double *p_, *p0_, *phi_, *u_, *u0_, *v_, *v0_, *w_, *w0_;
static double *us_, *vs_, *ws_;
static double *cu_, *cv_, *cw_, *c0u_, *c0v_, *c0w_;
static double *du_, *dv_, *dw_, *d0u_, *d0v_, *d0w_;
static double *fx_, *fy_, *fz_;
static double *rhs_, *rq_, *zq_, *pq_, *pbq_, *Apq_, *invM_;
static int *flu_, *flv_, *flw_, *phase_, *shell_;

#define IDX(G, i, j, k) ((long) (i) + (long) (j) * (G).sy + (long) (k) * (G).sz)

/* ------------------------------- parser --------------------------------- */
/* Tolerant key/value reading of flow.config (the grammar of
 * src/domain.c:72-916): each line is "key values..." except the three
 * domain lines "(Xs, Xe, Xn) a b n". */
#define MAXL 256
static char lines[512][MAXL];
static int nlines;

static int find_line(const char *key, char *rest)
{
  size_t kl = strlen(key);
  for (int i = 0; i < nlines; i++) {
    const char *s = lines[i];
    while (*s == ' ' || *s == '\t') s++;
    if (strncmp(s, key, kl) == 0 && (s[kl] == ' ' || s[kl] == '\t' ||
        s[kl] == '\0' || s[kl] == '\n')) {
      if (rest) strcpy(rest, s + kl);
      return 1;
    }
  }
  return 0;
}

static double get_num(const char *key)
{
  char rest[MAXL];
  if (!find_line(key, rest)) {
    fprintf(stderr, "oracle: missing key %s\n", key);
    exit(1);
  }
  return atof(rest);
}

static int get_bc(const char *key, double *a, double *b)
{
  char rest[MAXL], word[64];
  if (!find_line(key, rest)) {
    fprintf(stderr, "oracle: missing key %s\n", key);
    exit(1);
  }
  double x = 0, y = 0;
  int n = sscanf(rest, "%63s %lf %lf", word, &x, &y);
  if (a) *a = (n >= 2) ? x : 0.;
  if (b) *b = (n >= 3) ? y : 0.;
  if (strcmp(word, "PERIODIC") == 0) return PERIODIC;
  if (strcmp(word, "DIRICHLET") == 0) return DIRICHLET;
  if (strcmp(word, "NEUMANN") == 0) return NEUMANN;
  if (strcmp(word, "PRECURSOR") == 0) return PRECURSOR;
  fprintf(stderr, "oracle: bad BC %s\n", key);
  exit(1);
}

static void o_scalar_parse(void);
static int O_scalar;                 /* SCALAR (scalar.c) */
static int O_lamb_max_iter;
/* 1: a solve stopped at pp_max_iter lets the timestep go on (projection,
 * pressure update, Physalis, motion) instead of ending it; used only to
 * time the components of a timestep on a bounded sample. */
static int O_pcg_continue = 0;
static double O_lamb_residual, O_lamb_relax, O_lamb_cut;
static double O_sbeta, O_sref, O_g[3], O_gm[3], O_ga[3];

static int parse(const char *dir)
{
  char path[1024];
  snprintf(path, sizeof(path), "%s/input/flow.config", dir);
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  nlines = 0;
  while (nlines < 512 && fgets(lines[nlines], MAXL, f)) nlines++;
  fclose(f);
  char rest[MAXL];
  if (!find_line("(Xs, Xe, Xn)", rest)) return -2;
  sscanf(rest, "%lf %lf %d", &O.xs, &O.xe, &O.nx);
  if (!find_line("(Ys, Ye, Yn)", rest)) return -2;
  sscanf(rest, "%lf %lf %d", &O.ys, &O.ye, &O.ny);
  if (!find_line("(Zs, Ze, Zn)", rest)) return -2;
  sscanf(rest, "%lf %lf %d", &O.zs, &O.ze, &O.nz);
  O.rho_f = get_num("rho_f");
  O.nu = get_num("nu");
  O.mu = O.nu * O.rho_f;
  O.duration = get_num("duration");
  O.CFL = get_num("CFL");
  O.pp_max_iter = (int) get_num("pp_max_iter");
  O.pp_residual = get_num("pp_residual");
  O.v_bc_tdelay = get_num("v_bc_tdelay");
  const char *F = "WESNBT";
  char key[32];
  for (int f6 = 0; f6 < 6; f6++) {
    snprintf(key, sizeof(key), "bc.p%c", F[f6]);
    O.ptype[f6] = get_bc(key, &O.ds[f6], NULL);
    for (int c = 0; c < 3; c++) {
      snprintf(key, sizeof(key), "bc.%c%c", "uvw"[c], F[f6]);
      O.vtype[c][f6] = get_bc(key, &O.vDm[c][f6], &O.vDa[c][f6]);
      if (O.vtype[c][f6] == NEUMANN) { O.vDm[c][f6] = 0.; O.vDa[c][f6] = 0.; }
      O.vD[c][f6] = 0.;
    }
  }
  char word[64] = "";
  if (find_line("init_cond", rest)) sscanf(rest, "%63s", word);
  if (strcmp(word, "QUIESCENT") == 0) O.init_cond = 0;
  else if (strcmp(word, "SHEAR") == 0) O.init_cond = 1;
  else if (strcmp(word, "CHANNEL") == 0) O.init_cond = 2;
  else if (strcmp(word, "TAYLOR_GREEN_3") == 0) O.init_cond = 4;
  else { fprintf(stderr, "oracle: init_cond %s\n", word); return -3; }
  if (find_line("out_plane", rest)) sscanf(rest, "%63s", word);
  const char *planes[] = {"WEST", "EAST", "SOUTH", "NORTH", "BOTTOM", "TOP"};
  O.out_plane = HOMOGENEOUS;
  for (int i = 0; i < 6; i++) if (strcmp(word, planes[i]) == 0) O.out_plane = i;
  O.p_bc_tdelay = get_num("p_bc_tdelay");
  O.g_bc_tdelay = get_num("g_bc_tdelay");
  const char *gk[] = {"gradP.x", "gradP.y", "gradP.z"};
  for (int c = 0; c < 3; c++) {
    if (!find_line(gk[c], rest)) return -2;
    sscanf(rest, "%lf %lf", &O.gPm[c], &O.gPa[c]);
    O.gP[c] = 0.;
  }
  O.osci_f = get_num("osci_f");
  O_lamb_max_iter = find_line("lamb_max_iter", rest) ? atoi(rest) : 1;
  O_lamb_residual = find_line("lamb_residual", rest) ? atof(rest) : 0.;
  O_lamb_relax = find_line("lamb_relax", rest) ? atof(rest) : 1.;
  O_lamb_cut = find_line("lamb_cut", rest) ? atof(rest) : 0.;
  O_sbeta = find_line("s_beta", rest) ? atof(rest) : 0.;
  O_sref = find_line("s_ref", rest) ? atof(rest) : 0.;
  const char *gkk[] = {"g.x", "g.y", "g.z"};
  for (int c = 0; c < 3; c++) {
    O_gm[c] = O_ga[c] = O_g[c] = 0.;
    if (find_line(gkk[c], rest)) sscanf(rest, "%lf %lf", &O_gm[c], &O_ga[c]);
  }
  O.fixed_dt = find_line("fixed_dt", rest) ? atof(rest) : 0.;
  O.perturb_amp = 0.;
  O.perturb_seed = 0;
  if (find_line("perturb", rest))
    sscanf(rest, "%llu %lf", &O.perturb_seed, &O.perturb_amp);
  for (int d = 0; d < 3; d++) O.per[d] = (O.ptype[2 * d] == PERIODIC);
  o_scalar_parse();
  return 0;
}

/* -------------------------------- grid ---------------------------------- */
static OG mkgrid(int ni, int nj, int nk)
{
  OG G;
  G.ni = ni; G.nj = nj; G.nk = nk;
  G.sy = ni + 2;
  G.sz = (long) (ni + 2) * (nj + 2);
  G.n = G.sz * (nk + 2);
  return G;
}

static double *dal(long n) { return (double *) calloc(n, sizeof(double)); }
static int *ial(long n) { return (int *) calloc(n, sizeof(int)); }

/* -------------------------------- RNG ----------------------------------- */
/* Ranq1 of Numerical Recipes 3rd ed. (the generator of src/rng.c:24-81). */
static unsigned long long rs_;
static unsigned long long r64(void)
{
  rs_ ^= rs_ >> 21;
  rs_ ^= rs_ << 35;
  rs_ ^= rs_ >> 4;
  return rs_ * 2685821657736338717ULL;
}
static void rseed(unsigned long long s)
{
  rs_ = 4101842887655102017ULL ^ s;
  rs_ = r64();
}
static double rdbl(void) { return 5.42101086242752217e-20 * (double) r64(); }

/* ------------------------------- init ----------------------------------- */
#define FORALL(G) \
  for (int k = 0; k <= (G).nk + 1; k++) \
  for (int j = 0; j <= (G).nj + 1; j++) \
  for (int i = 0; i <= (G).ni + 1; i++)

static void init_fields(void)
{
  /* src/domain.c:2083-2142, 2144-2277, 2808-2887 */
  double xs = O.xs, ys = O.ys, zs = O.zs;
  if (O.init_cond == 1) {
    FORALL(O.fx) {
      double y = ((j - 0.5) * O.dy) + ys, z = ((k - 0.5) * O.dz) + zs;
      long c = IDX(O.fx, i, j, k);
      u_[c] = (O.vDm[0][3] - O.vDm[0][2])*(y - O.ys)/O.yl + O.vDm[0][2];
      u_[c] += (O.vDm[0][5] - O.vDm[0][4])*(z - O.zs)/O.zl + O.vDm[0][4];
      u0_[c] = u_[c];
    }
    FORALL(O.fy) {
      double z = ((k - 0.5) * O.dz) + zs, x = ((i - 0.5) * O.dx) + xs;
      long c = IDX(O.fy, i, j, k);
      v_[c] = (O.vDm[1][1] - O.vDm[1][0])*(x - O.xs)/O.xl + O.vDm[1][0];
      v_[c] += (O.vDm[1][5] - O.vDm[1][4])*(z - O.zs)/O.zl + O.vDm[1][4];
      v0_[c] = v_[c];
    }
    FORALL(O.fz) {
      double x = ((i - 0.5) * O.dx) + xs, y = ((j - 0.5) * O.dy) + ys;
      long c = IDX(O.fz, i, j, k);
      w_[c] = (O.vDm[2][1] - O.vDm[2][0])*(x - O.xs)/O.xl + O.vDm[2][0];
      w_[c] += (O.vDm[2][3] - O.vDm[2][2])*(y - O.ys)/O.yl + O.vDm[2][2];
      w0_[c] = w_[c];
    }
  } else if (O.init_cond == 2) {
    double m = O.mu;
    FORALL(O.fx) {
      double y = ((j - 0.5) * O.dy) + ys, z = ((k - 0.5) * O.dz) + zs;
      long c = IDX(O.fx, i, j, k);
      u_[c] = 0.5/m*O.gPm[0]*(y*y - (O.ys + O.ye)*y + O.ys*O.ye)
        * (O.vtype[0][2] == DIRICHLET)
        + 0.5/m*O.gPm[0]*(z*z - (O.zs + O.ze)*z + O.zs*O.ze)
        * (O.vtype[0][4] == DIRICHLET);
      u0_[c] = u_[c];
    }
    FORALL(O.fy) {
      double z = ((k - 0.5) * O.dz) + zs, x = ((i - 0.5) * O.dx) + xs;
      long c = IDX(O.fy, i, j, k);
      v_[c] = 0.5/m*O.gPm[1]*(x*x - (O.xs + O.xe)*x + O.xs*O.xe)
        * (O.vtype[1][0] == DIRICHLET)
        + 0.5/m*O.gPm[1]*(z*z - (O.zs + O.ze)*z + O.zs*O.ze)
        * (O.vtype[1][4] == DIRICHLET);
      v0_[c] = v_[c];
    }
    FORALL(O.fz) {
      double x = ((i - 0.5) * O.dx) + xs, y = ((j - 0.5) * O.dy) + ys;
      long c = IDX(O.fz, i, j, k);
      w_[c] = 0.5/m*O.gPm[2]*(x*x - (O.xs + O.xe)*x + O.xs*O.xe)
        * (O.vtype[2][0] == DIRICHLET)
        + 0.5/m*O.gPm[2]*(y*y - (O.ys + O.ye)*y + O.ys*O.ye)
        * (O.vtype[2][2] == DIRICHLET);
      w0_[c] = w_[c];
    }
  } else if (O.init_cond == 4) {
    double Re = 1600., L = 1.;
    double vel0 = Re * O.nu / L;
    FORALL(O.fx) {
      double x = ((i - 1.0) * O.dx) + xs, y = ((j - 0.5) * O.dy) + ys,
             z = ((k - 0.5) * O.dz) + zs;
      long c = IDX(O.fx, i, j, k);
      u_[c] = vel0*sin(x/L)*cos(y/L)*cos(z/L);
      u0_[c] = u_[c];
    }
    FORALL(O.fy) {
      double x = ((i - 0.5) * O.dx) + xs, y = ((j - 1.0) * O.dy) + ys,
             z = ((k - 0.5) * O.dz) + zs;
      long c = IDX(O.fy, i, j, k);
      v_[c] = -vel0*cos(x/L)*sin(y/L)*cos(z/L);
      v0_[c] = v_[c];
    }
    FORALL(O.cc) {
      double x = ((i - 0.5) * O.dx) + xs, y = ((j - 0.5) * O.dy) + ys,
             z = ((k - 0.5) * O.dz) + zs;
      long c = IDX(O.cc, i, j, k);
      p_[c] = O.rho_f*vel0*vel0/16.*(cos(2.*x/L) + cos(2.*y/L))*(cos(2.*z/L) + 2.);
      p0_[c] = p_[c];
    }
  }
  if (O.perturb_amp > 0.) {
    /* Seeded perturbation: u, then v, then w; k, j, i; in a periodic
     * direction the duplicated last face copies the first. */
    double a = O.perturb_amp;
    rseed(O.perturb_seed);
    OG *G = &O.fx;
    for (int k = 1; k <= G->nk; k++) for (int j = 1; j <= G->nj; j++)
      for (int i = 1; i <= G->ni; i++) {
        if (i == G->ni && O.per[0]) continue;
        long c = IDX(*G, i, j, k);
        u_[c] += a * (2. * rdbl() - 1.);
        u0_[c] = u_[c];
      }
    if (O.per[0]) for (int k = 1; k <= G->nk; k++) for (int j = 1; j <= G->nj; j++) {
      u_[IDX(*G, G->ni, j, k)] = u_[IDX(*G, 1, j, k)];
      u0_[IDX(*G, G->ni, j, k)] = u_[IDX(*G, 1, j, k)];
    }
    G = &O.fy;
    for (int k = 1; k <= G->nk; k++) for (int j = 1; j <= G->nj; j++)
      for (int i = 1; i <= G->ni; i++) {
        if (j == G->nj && O.per[1]) continue;
        long c = IDX(*G, i, j, k);
        v_[c] += a * (2. * rdbl() - 1.);
        v0_[c] = v_[c];
      }
    if (O.per[1]) for (int k = 1; k <= G->nk; k++) for (int i = 1; i <= G->ni; i++) {
      v_[IDX(*G, i, G->nj, k)] = v_[IDX(*G, i, 1, k)];
      v0_[IDX(*G, i, G->nj, k)] = v_[IDX(*G, i, 1, k)];
    }
    G = &O.fz;
    for (int k = 1; k <= G->nk; k++) for (int j = 1; j <= G->nj; j++)
      for (int i = 1; i <= G->ni; i++) {
        if (k == G->nk && O.per[2]) continue;
        long c = IDX(*G, i, j, k);
        w_[c] += a * (2. * rdbl() - 1.);
        w0_[c] = w_[c];
      }
    if (O.per[2]) for (int j = 1; j <= G->nj; j++) for (int i = 1; i <= G->ni; i++) {
      w_[IDX(*G, i, j, G->nk)] = w_[IDX(*G, i, j, 1)];
      w0_[IDX(*G, i, j, G->nk)] = w_[IDX(*G, i, j, 1)];
    }
  }
  O.dt = 2.*O.nu / (O.dx * O.dx);
  O.dt += 2.*O.nu / (O.dy * O.dy);
  O.dt += 2.*O.nu / (O.dz * O.dz);
  O.dt = O.CFL / O.dt;
  if (O.fixed_dt > 0.) O.dt = O.fixed_dt;
  O.dt0 = -1.;
  O.stepnum = 0;
  O.ttime = 0.;
}

/* ------------------------- canonical reductions ------------------------- */
double orc_canon_linear(const double *rows, long n)
{
  double acc[1024];
  for (int t = 0; t < 1024; t++) acc[t] = 0.0;
  for (long r = 0; r < n; r++) acc[r % 1024] += rows[r];
  for (int off = 512; off >= 1; off >>= 1)
    for (int t = 0; t < off; t++) acc[t] += acc[t + off];
  return acc[0];
}

/* terms: nx*ny*nz values, x-fastest, interior cells only.  Tiles of 64 x 8
 * x 8 cells: each (x, y) column of a tile sums its planes in increasing z;
 * the 64 columns of a row by the halving tree; the 8 rows as
 * ((w0 + w1) + (w2 + w3)) + ((w4 + w5) + (w6 + w7)); the tiles, numbered
 * tc + ntx (tj + nty tk), over 1024 interleaved accumulators. */
double orc_canon_sum(const double *terms, int nx, int ny, int nz)
{
  int ntx = (nx + 63) / 64, nty = (ny + 7) / 8, ntz = (nz + 7) / 8;
  long nt = (long) ntx * nty * ntz;
  double *part = dal(nt);
  #pragma omp parallel for schedule(static)
  for (long t = 0; t < nt; t++) {
    int tc = (int) (t % ntx), tj = (int) ((t / ntx) % nty),
        tk = (int) (t / ((long) ntx * nty));
    double wv[8];
    for (int w = 0; w < 8; w++) {
      int j = tj * 8 + w;
      double lane[64];
      for (int l = 0; l < 64; l++) {
        int x = tc * 64 + l;
        double acc = 0.0;
        if (x < nx && j < ny)
          for (int k = tk * 8; k < tk * 8 + 8 && k < nz; k++)
            acc += terms[(long) x + (long) j * nx + (long) k * nx * ny];
        lane[l] = acc;
      }
      for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; l++) lane[l] += lane[l + off];
      wv[w] = lane[0];
    }
    part[t] = ((wv[0] + wv[1]) + (wv[2] + wv[3])) +
              ((wv[4] + wv[5]) + (wv[6] + wv[7]));
  }
  double s = orc_canon_linear(part, nt);
  free(part);
  return s;
}

static double *tbuf_;                  /* nx*ny*nz scratch for dot terms */

static double dot_cc(const double *a, const double *b)
{
  /* (a, b) over interior Gcc cells (the thrust::inner_product calls of
   * src/cuda_solver.cu:151-231). */
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= O.nz; k++)
    for (int j = 1; j <= O.ny; j++)
      for (int i = 1; i <= O.nx; i++) {
        long c = IDX(O.cc, i, j, k);
        tbuf_[(long) (i - 1) + (long) (j - 1) * O.nx +
              (long) (k - 1) * O.nx * O.ny] = a[c] * b[c];
      }
  return orc_canon_sum(tbuf_, O.nx, O.ny, O.nz);
}

/* ---------------------------- flags / phase ----------------------------- */
/* Flags 0 on non-periodic external boundary faces over the whole ghosted
 * plane (src/cuda_particle.cu:1606-1640, the last step of the flag build,
 * after the particle cages). */
static void o_flags_external(void)
{
  if (!O.per[0])
    for (int k = 0; k <= O.fx.nk + 1; k++) for (int j = 0; j <= O.fx.nj + 1; j++) {
      flu_[IDX(O.fx, 1, j, k)] = 0;
      flu_[IDX(O.fx, O.fx.ni, j, k)] = 0;
    }
  if (!O.per[1])
    for (int k = 0; k <= O.fy.nk + 1; k++) for (int i = 0; i <= O.fy.ni + 1; i++) {
      flv_[IDX(O.fy, i, 1, k)] = 0;
      flv_[IDX(O.fy, i, O.fy.nj, k)] = 0;
    }
  if (!O.per[2])
    for (int j = 0; j <= O.fz.nj + 1; j++) for (int i = 0; i <= O.fz.ni + 1; i++) {
      flw_[IDX(O.fz, i, j, 1)] = 0;
      flw_[IDX(O.fz, i, j, O.fz.nk)] = 0;
    }
}

void orc_build_flags(void)
{
  /* src/cuda_particle.cu:1516-1646 without particles: flags 1 everywhere,
   * then the external faces */
  for (long c = 0; c < O.fx.n; c++) flu_[c] = 1;
  for (long c = 0; c < O.fy.n; c++) flv_[c] = 1;
  for (long c = 0; c < O.fz.n; c++) flw_[c] = 1;
  o_flags_external();
}

/* ------------------------- boundary conditions -------------------------- */
/* One face of one field, src/bluebottle_kernel.cu:26-598.  face: 0..5 =
 * W E S N B T.  kind: 0 none, 1 Neumann, 2 normal Dirichlet, 3 tangential
 * Dirichlet. */
static void bc_face(double *a, const OG *G, int face, int kind, double bc)
{
  if (kind == 0) return;
  int dir = face / 2, hi = face % 2;
  int n1 = dir == 0 ? G->nj : G->ni;
  int n2 = dir == 2 ? G->nj : G->nk;
  int e = dir == 0 ? G->ni : (dir == 1 ? G->nj : G->nk);
  long S = dir == 0 ? 1 : (dir == 1 ? G->sy : G->sz);
  for (int t2 = 1; t2 <= n2; t2++)
    for (int t1 = 1; t1 <= n1; t1++) {
      long b;
      if (dir == 0) b = IDX(*G, 0, t1, t2);
      else if (dir == 1) b = IDX(*G, t1, 0, t2);
      else b = IDX(*G, t1, t2, 0);
      long g, s0, s1;
      if (!hi) { g = b; s0 = b + S; s1 = b + 2 * S; }
      else { g = b + (e + 1) * S; s0 = b + e * S; s1 = b + (e - 1) * S; }
      if (kind == 1) {
        a[g] = a[s0];
      } else if (kind == 2) {
        a[g] = 2.*bc - a[s1];
        a[s0] = bc;
      } else {
        a[g] = 8./3.*bc - 2.*a[s0] + 1./3.*a[s1];
      }
    }
}

static int vkind(int c, int face)
{
  int t = O.vtype[c][face];
  if (t == NEUMANN) return 1;
  if (t == DIRICHLET) return (face / 2 == c) ? 2 : 3;
  return 0;
}

void orc_bc_p(double *a)
{
  /* src/cuda_bluebottle.cu:2538-2589 */
  for (int f6 = 0; f6 < 6; f6++)
    bc_face(a, &O.cc, f6, O.ptype[f6] == NEUMANN ? 1 : 0, 0.);
}

static void bc_vel(double *a, double *b, double *c)
{
  for (int f6 = 0; f6 < 6; f6++) {
    bc_face(a, &O.fx, f6, vkind(0, f6), O.vD[0][f6]);
    bc_face(b, &O.fy, f6, vkind(1, f6), O.vD[1][f6]);
    bc_face(c, &O.fz, f6, vkind(2, f6), O.vD[2][f6]);
  }
}

void orc_bc_all(void)
{
  /* src/cuda_bluebottle.cu:1218-1438: face by face, p, u, v, w. */
  for (int f6 = 0; f6 < 6; f6++) {
    bc_face(p_, &O.cc, f6, O.ptype[f6] == NEUMANN ? 1 : 0, 0.);
    bc_face(u_, &O.fx, f6, vkind(0, f6), O.vD[0][f6]);
    bc_face(v_, &O.fy, f6, vkind(1, f6), O.vD[1][f6]);
    bc_face(w_, &O.fz, f6, vkind(2, f6), O.vD[2][f6]);
  }
}

void orc_bc_star(void)
{
  /* src/cuda_bluebottle.cu:2111-2311 */
  bc_vel(us_, vs_, ws_);
}

/* Halo exchange on one rank (src/mpi_comm.c:257-405 with the plane pairs of
 * src/bluebottle_kernel.cu:684-1466): pack all six planes, then unpack. */
void orc_exchange(double *a, int grid)
{
  const OG *G = grid == 0 ? &O.cc : grid == 1 ? &O.fx : grid == 2 ? &O.fy : &O.fz;
  for (int dir = 0; dir < 3; dir++) {
    if (!O.per[dir]) continue;
    int stag = (grid == dir + 1);
    int n1 = dir == 0 ? G->nj : G->ni;
    int n2 = dir == 2 ? G->nj : G->nk;
    int e = dir == 0 ? G->ni : (dir == 1 ? G->nj : G->nk);
    long S = dir == 0 ? 1 : (dir == 1 ? G->sy : G->sz);
    double *lo = dal((long) n1 * n2), *hi = dal((long) n1 * n2);
    for (int t2 = 1; t2 <= n2; t2++)
      for (int t1 = 1; t1 <= n1; t1++) {
        long b = dir == 0 ? IDX(*G, 0, t1, t2) :
                 dir == 1 ? IDX(*G, t1, 0, t2) : IDX(*G, t1, t2, 0);
        long q = (t1 - 1) + (long) (t2 - 1) * n1;
        /* send_e (plane e or e-1) -> recv_w of self -> ghost 0;
         * send_w (plane s or s+1) -> recv_e of self -> ghost e+1 */
        lo[q] = a[b + (stag ? (e - 1) : e) * S];
        hi[q] = a[b + (stag ? 2 : 1) * S];
      }
    for (int t2 = 1; t2 <= n2; t2++)
      for (int t1 = 1; t1 <= n1; t1++) {
        long b = dir == 0 ? IDX(*G, 0, t1, t2) :
                 dir == 1 ? IDX(*G, t1, 0, t2) : IDX(*G, t1, t2, 0);
        long q = (t1 - 1) + (long) (t2 - 1) * n1;
        a[b] = lo[q];
        a[b + (e + 1) * S] = hi[q];
      }
    free(lo);
    free(hi);
  }
}

static void exch_vel(double *a, double *b, double *c)
{
  orc_exchange(a, 1);
  orc_exchange(b, 2);
  orc_exchange(c, 3);
}

/* ------------------------------- forcing -------------------------------- */
static double rampv(double m, double a, double delta, double cur)
{
  if (delta < 0) return cur;
  if (a == 0) return m;
  if (fabs(delta*a) > fabs(m)) return m;
  return delta*a;
}

void orc_forcing(void)
{
  /* src/cuda_bluebottle.cu:1811-1982: f = 0 + (-gradP/rho) everywhere. */
  double delta = O.ttime - O.p_bc_tdelay;
  for (int c = 0; c < 3; c++) O.gP[c] = rampv(O.gPm[c], O.gPa[c], delta, O.gP[c]);
  O.gP[2] = O.gP[2] * cos(O.osci_f*O.ttime);
  double dg = O.ttime - O.g_bc_tdelay;
  for (int c = 0; c < 3; c++) O_g[c] = rampv(O_gm[c], O_ga[c], dg, O_g[c]);
  double fv[3];
  for (int c = 0; c < 3; c++) fv[c] = -O.gP[c]/O.rho_f;
  for (long q = 0; q < O.fx.n; q++) { fx_[q] = 0.; fx_[q] += fv[0]; }
  for (long q = 0; q < O.fy.n; q++) { fy_[q] = 0.; fy_[q] += fv[1]; }
  for (long q = 0; q < O.fz.n; q++) { fz_[q] = 0.; fz_[q] += fv[2]; }
}

void orc_vel_bc_ramp(void)
{
  /* src/domain.c:1931-2082 */
  double delta = O.ttime - O.v_bc_tdelay;
  for (int c = 0; c < 3; c++)
    for (int f6 = 0; f6 < 6; f6++)
      O.vD[c][f6] = rampv(O.vDm[c][f6], O.vDa[c][f6], delta, O.vD[c][f6]);
}

/* --------------------------------- u* ----------------------------------- */
void orc_u_star(void)
{
  /* src/bluebottle_kernel.cu:1637-2133, host arguments of
   * src/cuda_bluebottle.cu:2095-2108 (p = p0). */
  double dt = O.dt, dt0 = O.dt0, rho = O.rho_f, nu = O.nu;
  double ab0 = 0.5 * dt / dt0;
  double ab = 1. + ab0;
  double ddx = 1. / O.dx, ddy = 1. / O.dy, ddz = 1. / O.dz;
  const OG X = O.fx, Y = O.fy, Z = O.fz, C = O.cc;

  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= X.nk; k++)
    for (int j = 1; j <= X.nj; j++)
      for (int i = 1; i <= X.ni; i++) {
        double u011 = u0_[IDX(X, i - 1, j, k)], u111 = u0_[IDX(X, i, j, k)],
               u211 = u0_[IDX(X, i + 1, j, k)];
        double u101 = u0_[IDX(X, i, j - 1, k)], u121 = u0_[IDX(X, i, j + 1, k)];
        double u110 = u0_[IDX(X, i, j, k - 1)], u112 = u0_[IDX(X, i, j, k + 1)];
        double v011 = v0_[IDX(Y, i - 1, j, k)], v111 = v0_[IDX(Y, i, j, k)];
        double v021 = v0_[IDX(Y, i - 1, j + 1, k)], v121 = v0_[IDX(Y, i, j + 1, k)];
        double w011 = w0_[IDX(Z, i - 1, j, k)], w111 = w0_[IDX(Z, i, j, k)];
        double w012 = w0_[IDX(Z, i - 1, j, k + 1)], w112 = w0_[IDX(Z, i, j, k + 1)];
        long c = IDX(X, i, j, k);
        double s = (p0_[IDX(C, i - 1, j, k)] - p0_[IDX(C, i, j, k)]) * ddx/rho;
        double duudx = (u211 + u111)*(u211 + u111) - (u111 + u011)*(u111 + u011);
        duudx *= 0.25 * ddx;
        double duvdy = (u121 + u111)*(v121 + v021) - (u111 + u101)*(v111 + v011);
        duvdy *= 0.25 * ddy;
        double duwdz = (u112 + u111)*(w112 + w012) - (u111 + u110)*(w111 + w011);
        duwdz *= 0.25 * ddz;
        double cv = duudx + duvdy + duwdz;
        if (dt0 > 0) s += (-ab * cv + ab0 * c0u_[c]);
        else s += -cv;
        double d1 = (u211 - u111) * ddx, d0 = (u111 - u011) * ddx;
        double xx = (d1 - d0) * ddx;
        d1 = (u121 - u111) * ddy; d0 = (u111 - u101) * ddy;
        double yy = (d1 - d0) * ddy;
        d1 = (u112 - u111) * ddz; d0 = (u111 - u110) * ddz;
        double zz = (d1 - d0) * ddz;
        double dv = nu * (xx + yy + zz);
        if (dt0 > 0) s += (ab * dv - ab0 * d0u_[c]);
        else s += dv;
        s += fx_[c];
        s *= dt;
        s += u111;
        s *= (phase_[IDX(C, i - 1, j, k)] < 0 && phase_[IDX(C, i, j, k)] < 0);
        us_[c] = s; cu_[c] = cv; du_[c] = dv;
      }

  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= Y.nk; k++)
    for (int j = 1; j <= Y.nj; j++)
      for (int i = 1; i <= Y.ni; i++) {
        double v101 = v0_[IDX(Y, i, j - 1, k)], v111 = v0_[IDX(Y, i, j, k)],
               v121 = v0_[IDX(Y, i, j + 1, k)];
        double v110 = v0_[IDX(Y, i, j, k - 1)], v112 = v0_[IDX(Y, i, j, k + 1)];
        double v011 = v0_[IDX(Y, i - 1, j, k)], v211 = v0_[IDX(Y, i + 1, j, k)];
        double w101 = w0_[IDX(Z, i, j - 1, k)], w111 = w0_[IDX(Z, i, j, k)];
        double w102 = w0_[IDX(Z, i, j - 1, k + 1)], w112 = w0_[IDX(Z, i, j, k + 1)];
        double u101 = u0_[IDX(X, i, j - 1, k)], u111 = u0_[IDX(X, i, j, k)];
        double u201 = u0_[IDX(X, i + 1, j - 1, k)], u211 = u0_[IDX(X, i + 1, j, k)];
        long c = IDX(Y, i, j, k);
        double s = (p0_[IDX(C, i, j - 1, k)] - p0_[IDX(C, i, j, k)]) * ddy/rho;
        double dvudx = (v211 + v111)*(u211 + u201) - (v111 + v011)*(u111 + u101);
        dvudx *= 0.25 * ddx;
        double dvvdy = (v121 + v111)*(v121 + v111) - (v111 + v101)*(v111 + v101);
        dvvdy *= 0.25 * ddy;
        double dvwdz = (v112 + v111)*(w112 + w102) - (v111 + v110)*(w111 + w101);
        dvwdz *= 0.25 * ddz;
        double cv = dvudx + dvvdy + dvwdz;
        if (dt0 > 0) s += (-ab * cv + ab0 * c0v_[c]);
        else s += -cv;
        double d1 = (v211 - v111) * ddx, d0 = (v111 - v011) * ddx;
        double xx = (d1 - d0) * ddx;
        d1 = (v121 - v111) * ddy; d0 = (v111 - v101) * ddy;
        double yy = (d1 - d0) * ddy;
        d1 = (v112 - v111) * ddz; d0 = (v111 - v110) * ddz;
        double zz = (d1 - d0) * ddz;
        double dv = nu * (xx + yy + zz);
        if (dt0 > 0) s += (ab * dv - ab0 * d0v_[c]);
        else s += dv;
        s += fy_[c];
        s *= dt;
        s += v111;
        s *= (phase_[IDX(C, i, j - 1, k)] < 0 && phase_[IDX(C, i, j, k)] < 0);
        vs_[c] = s; cv_[c] = cv; dv_[c] = dv;
      }

  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= Z.nk; k++)
    for (int j = 1; j <= Z.nj; j++)
      for (int i = 1; i <= Z.ni; i++) {
        double w110 = w0_[IDX(Z, i, j, k - 1)], w111 = w0_[IDX(Z, i, j, k)],
               w112 = w0_[IDX(Z, i, j, k + 1)];
        double w011 = w0_[IDX(Z, i - 1, j, k)], w211 = w0_[IDX(Z, i + 1, j, k)];
        double w101 = w0_[IDX(Z, i, j - 1, k)], w121 = w0_[IDX(Z, i, j + 1, k)];
        double u110 = u0_[IDX(X, i, j, k - 1)], u111 = u0_[IDX(X, i, j, k)];
        double u210 = u0_[IDX(X, i + 1, j, k - 1)], u211 = u0_[IDX(X, i + 1, j, k)];
        double v110 = v0_[IDX(Y, i, j, k - 1)], v111 = v0_[IDX(Y, i, j, k)];
        double v120 = v0_[IDX(Y, i, j + 1, k - 1)], v121 = v0_[IDX(Y, i, j + 1, k)];
        long c = IDX(Z, i, j, k);
        double s = (p0_[IDX(C, i, j, k - 1)] - p0_[IDX(C, i, j, k)]) * ddz/rho;
        double dwudx = (w211 + w111)*(u211 + u210) - (w111 + w011)*(u111 + u110);
        dwudx *= 0.25 * ddx;
        double dwvdy = (w121 + w111)*(v121 + v120) - (w111 + w101)*(v111 + v110);
        dwvdy *= 0.25 * ddy;
        double dwwdz = (w112 + w111)*(w112 + w111) - (w111 + w110)*(w111 + w110);
        dwwdz *= 0.25 * ddz;
        double cv = dwudx + dwvdy + dwwdz;
        if (dt0 > 0) s += (-ab * cv + ab0 * c0w_[c]);
        else s += -cv;
        double d1 = (w211 - w111) * ddx, d0 = (w111 - w011) * ddx;
        double xx = (d1 - d0) * ddx;
        d1 = (w121 - w111) * ddy; d0 = (w111 - w101) * ddy;
        double yy = (d1 - d0) * ddy;
        d1 = (w112 - w111) * ddz; d0 = (w111 - w110) * ddz;
        double zz = (d1 - d0) * ddz;
        double dv = nu * (xx + yy + zz);
        if (dt0 > 0) s += (ab * dv - ab0 * d0w_[c]);
        else s += dv;
        s += fz_[c];
        s *= dt;
        s += w111;
        s *= (phase_[IDX(C, i, j, k - 1)] < 0 && phase_[IDX(C, i, j, k)] < 0);
        ws_[c] = s; cw_[c] = cv; dw_[c] = dv;
      }
}

/* ----------------------------- solvability ------------------------------ */
static double face_sum(const double *a, const OG *G, int dir, int loc)
{
  int n1 = dir == 0 ? G->nj : G->ni;
  int n2 = dir == 2 ? G->nj : G->nk;
  double *t = dal((long) n1 * n2);
  for (int t2 = 0; t2 < n2; t2++)
    for (int t1 = 0; t1 < n1; t1++) {
      long c = dir == 0 ? IDX(*G, loc, t1 + 1, t2 + 1) :
               dir == 1 ? IDX(*G, t1 + 1, loc, t2 + 1) :
                          IDX(*G, t1 + 1, t2 + 1, loc);
      t[t1 + (long) t2 * n1] = a[c];
    }
  double s = orc_canon_linear(t, (long) n1 * n2);
  free(t);
  return s;
}

static void face_add(double *a, const OG *G, int dir, int loc, double val,
  int sub)
{
  int n1 = dir == 0 ? G->nj : G->ni;
  int n2 = dir == 2 ? G->nj : G->nk;
  for (int t2 = 0; t2 < n2; t2++)
    for (int t1 = 0; t1 < n1; t1++) {
      long c = dir == 0 ? IDX(*G, loc, t1 + 1, t2 + 1) :
               dir == 1 ? IDX(*G, t1 + 1, loc, t2 + 1) :
                          IDX(*G, t1 + 1, t2 + 1, loc);
      a[c] = sub ? a[c] - val : a[c] + val;
    }
}

void orc_solvability(void)
{
  /* src/cuda_bluebottle.cu:2313-2492 */
  double exs = face_sum(us_, &O.fx, 0, 1) * (O.dy * O.dz);
  double exe = face_sum(us_, &O.fx, 0, O.fx.ni) * (O.dy * O.dz);
  double eys = face_sum(vs_, &O.fy, 1, 1) * (O.dz * O.dx);
  double eye = face_sum(vs_, &O.fy, 1, O.fy.nj) * (O.dz * O.dx);
  double ezs = face_sum(ws_, &O.fz, 2, 1) * (O.dx * O.dy);
  double eze = face_sum(ws_, &O.fz, 2, O.fz.nk) * (O.dx * O.dy);
  double eps[3] = {exe - exs, eye - eys, eze - ezs};
  double sum;
  switch (O.out_plane) {
    case 0: sum = (eps[0] + eps[1] + eps[2])/(O.yl * O.zl);
            face_add(us_, &O.fx, 0, 1, sum, 0); break;
    case 1: sum = (eps[0] + eps[1] + eps[2])/(O.yl * O.zl);
            face_add(us_, &O.fx, 0, O.fx.ni, sum, 1); break;
    case 2: sum = (eps[0] + eps[1] + eps[2])/(O.zl * O.xl);
            face_add(vs_, &O.fy, 1, 1, sum, 0); break;
    case 3: sum = (eps[0] + eps[1] + eps[2])/(O.zl * O.xl);
            face_add(vs_, &O.fy, 1, O.fy.nj, sum, 1); break;
    case 4: sum = (eps[0] + eps[1] + eps[2])/(O.xl * O.yl);
            face_add(ws_, &O.fz, 2, 1, sum, 0); break;
    case 5: sum = (eps[0] + eps[1] + eps[2])/(O.xl * O.yl);
            face_add(ws_, &O.fz, 2, O.fz.nk, sum, 1); break;
    default: {
      double sx = 0.5*eps[0]/(O.yl * O.zl);
      double sy = 0.5*eps[1]/(O.zl * O.xl);
      double sz = 0.5*eps[2]/(O.xl * O.yl);
      face_add(us_, &O.fx, 0, 1, sx, 0);
      face_add(us_, &O.fx, 0, O.fx.ni, sx, 1);
      face_add(vs_, &O.fy, 1, 1, sy, 0);
      face_add(vs_, &O.fy, 1, O.fy.nj, sy, 1);
      face_add(ws_, &O.fz, 2, 1, sz, 0);
      face_add(ws_, &O.fz, 2, O.fz.nk, sz, 1);
    }
  }
}

/* --------------------------------- PCG ---------------------------------- */
#define FOR_CC for (int k = 1; k <= O.nz; k++) for (int j = 1; j <= O.ny; j++) \
               for (int i = 1; i <= O.nx; i++)

#include "parts.c"
#include "scalar.c"

void orc_jacobi(void)
{
  /* src/solver_kernel.cu:26-87 (-DJACOBI) */
  double idx2 = 1./(O.dx * O.dx), idy2 = 1./(O.dy * O.dy),
         idz2 = 1./(O.dz * O.dz);
  for (int k = 1; k <= O.nz; k++)
    for (int j = 1; j <= O.ny; j++)
      for (int i = 1; i <= O.nx; i++) {
        int W = flu_[IDX(O.fx, i, j, k)], E = flu_[IDX(O.fx, i + 1, j, k)];
        int S = flv_[IDX(O.fy, i, j, k)], N = flv_[IDX(O.fy, i, j + 1, k)];
        int B = flw_[IDX(O.fz, i, j, k)], T = flw_[IDX(O.fz, i, j, k + 1)];
        double M = -idx2*(E*E + W*W) - idy2*(N*N + S*S) - idz2*(T*T + B*B);
        invM_[IDX(O.cc, i, j, k)] = -1./M;
      }
}

static void spmv(const double *x, double *Ax)
{
  /* src/solver_kernel.cu:715-836: (-A) x with face flags squared. */
  double idx2 = 1./(O.dx * O.dx), idy2 = 1./(O.dy * O.dy),
         idz2 = 1./(O.dz * O.dz);
  const OG C = O.cc;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= O.nz; k++)
    for (int j = 1; j <= O.ny; j++)
      for (int i = 1; i <= O.nx; i++) {
        long c = IDX(C, i, j, k);
        int fe = flu_[IDX(O.fx, i + 1, j, k)], fw = flu_[IDX(O.fx, i, j, k)];
        int fn = flv_[IDX(O.fy, i, j + 1, k)], fs = flv_[IDX(O.fy, i, j, k)];
        int ft = flw_[IDX(O.fz, i, j, k + 1)], fb = flw_[IDX(O.fz, i, j, k)];
        Ax[c] = -idx2 * (fe*fe*(x[c + 1] - x[c]) - fw*fw*(x[c] - x[c - 1]))
                -idy2 * (fn*fn*(x[c + C.sy] - x[c]) - fs*fs*(x[c] - x[c - C.sy]))
                -idz2 * (ft*ft*(x[c + C.sz] - x[c]) - fb*fb*(x[c] - x[c - C.sz]));
      }
}

/* src/solver_kernel.cu:528-713: (-A) x with the particle modification:
 * in a particle cell the diagonal factor is -dx^2/6 and the off-diagonals
 * vanish; in a fluid cell the link to a particle neighbour drops its
 * off-diagonal. */
static void spmv_parts(const double *x, double *Ax)
{
  double idx2 = 1./(O.dx * O.dx), idy2 = 1./(O.dy * O.dy),
         idz2 = 1./(O.dz * O.dz);
  const OG C = O.cc;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= O.nz; k++)
    for (int j = 1; j <= O.ny; j++)
      for (int i = 1; i <= O.nx; i++) {
        long c = IDX(C, i, j, k);
        int fe = flu_[IDX(O.fx, i + 1, j, k)], fw = flu_[IDX(O.fx, i, j, k)];
        int fn = flv_[IDX(O.fy, i, j + 1, k)], fs = flv_[IDX(O.fy, i, j, k)];
        int ft = flw_[IDX(O.fz, i, j, k + 1)], fb = flw_[IDX(O.fz, i, j, k)];
        int pc = phase_[c];
        double pfx = -(O.dx * O.dx)/6. * (pc > -1) + (double) (pc == -1);
        double pfy = -(O.dy * O.dy)/6. * (pc > -1) + (double) (pc == -1);
        double pfz = -(O.dz * O.dz)/6. * (pc > -1) + (double) (pc == -1);
        int fl = pc == -1;
        int pfe = fl * !(fl && phase_[c + 1] > -1);
        int pfw = fl * !(fl && phase_[c - 1] > -1);
        int pfn = fl * !(fl && phase_[c + C.sy] > -1);
        int pfs = fl * !(fl && phase_[c - C.sy] > -1);
        int pft = fl * !(fl && phase_[c + C.sz] > -1);
        int pfb = fl * !(fl && phase_[c - C.sz] > -1);
        double a = -idx2 * (fe*fe*(x[c + 1]*pfe - pfx*x[c])
                          - fw*fw*(x[c]*pfx - pfw*x[c - 1]));
        a += -idy2 * (fn*fn*(x[c + C.sy]*pfn - pfy*x[c])
                    - fs*fs*(x[c]*pfy - pfs*x[c - C.sy]));
        a += -idz2 * (ft*ft*(x[c + C.sz]*pft - pfz*x[c])
                    - fb*fb*(x[c]*pfz - pfb*x[c - C.sz]));
        Ax[c] = a;
      }
}

static void spmv_any(const double *x, double *Ax)
{
  if (NP_ > 0) spmv_parts(x, Ax);
  else spmv(x, Ax);
}

void orc_apply_A(const double *x, double *Ax)
{
  spmv_any(x, Ax);
}

static double now_s(void)
{
  struct timeval t;
  gettimeofday(&t, 0);
  return t.tv_sec + t.tv_usec * 1e-6;
}

int orc_pcg(void)
{
  /* src/cuda_solver.cu:574-761 with PP_rhs, PP_cg_init and the update
   * kernels of src/solver_kernel.cu:89-176, 258-283, 838-927. */
  double t0 = now_s();
  (void) t0;
  double idx = 1. / O.dx, idy = 1. / O.dy, idz = 1. / O.dz;
  double rho_idt = O.rho_f / O.dt;
  memset(rhs_, 0, O.cc.n * sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= O.nz; k++)
    for (int j = 1; j <= O.ny; j++)
      for (int i = 1; i <= O.nx; i++) {
        double s = (us_[IDX(O.fx, i + 1, j, k)] - us_[IDX(O.fx, i, j, k)]) * idx;
        s += (vs_[IDX(O.fy, i, j + 1, k)] - vs_[IDX(O.fy, i, j, k)]) * idy;
        s += (ws_[IDX(O.fz, i, j, k + 1)] - ws_[IDX(O.fz, i, j, k)]) * idz;
        s *= rho_idt;
        rhs_[IDX(O.cc, i, j, k)] = -s;
      }
  if (NP_ > 0) {
    /* src/particle_kernel.cu:1744-1753: the right-hand side keeps its value
     * in fluid cells (phase_shell is 1 there) and is multiplied by zero in
     * particle cells; src/solver_kernel.cu:212-256 then adds -rhs/dx^2 of
     * particle neighbours, which are zero. */
    FOR_CC {
      long c = IDX(O.cc, i, j, k);
      rhs_[c] = (double) (phase_[c] < 0 && shell_[c]) * rhs_[c];
    }
    orc_exchange(rhs_, 0);
    double idx2 = 1./(O.dx * O.dx), idy2 = 1./(O.dy * O.dy),
           idz2 = 1./(O.dz * O.dz);
    FOR_CC {
      long c = IDX(O.cc, i, j, k);
      int is_fluid = phase_[c] == -1;
      rhs_[c] += is_fluid * (phase_[c + 1] > -1) * idx2 * (-rhs_[c + 1]);
      rhs_[c] += is_fluid * (phase_[c - 1] > -1) * idx2 * (-rhs_[c - 1]);
      rhs_[c] += is_fluid * (phase_[c + O.cc.sy] > -1) * idy2 * (-rhs_[c + O.cc.sy]);
      rhs_[c] += is_fluid * (phase_[c - O.cc.sy] > -1) * idy2 * (-rhs_[c - O.cc.sy]);
      rhs_[c] += is_fluid * (phase_[c + O.cc.sz] > -1) * idz2 * (-rhs_[c + O.cc.sz]);
      rhs_[c] += is_fluid * (phase_[c - O.cc.sz] > -1) * idz2 * (-rhs_[c - O.cc.sz]);
    }
  }
  double sp_rhs = dot_cc(rhs_, rhs_);
  orc_exchange(rhs_, 0);
  FOR_CC {
    long c = IDX(O.cc, i, j, k);
    double t = rhs_[c];
    double t2 = t*invM_[c];
    rq_[c] = t; zq_[c] = t2; pq_[c] = t2; pbq_[c] = t2; phi_[c] = 0.;
  }
  orc_exchange(pbq_, 0);
  double sp_rq = dot_cc(rq_, zq_), sp_rq1 = 0.;
  O.sp_rhs = sp_rhs;
  int q = 0;
  if (sp_rhs < 1.e-8*1.e-8) {
    O.niter = 0; O.resid = 0.; O.true_resid = 0.;
    return 0;
  }
  while (q <= O.pp_max_iter) {
    ++q;
    spmv_any(pbq_, Apq_);
    double DENOM = dot_cc(pq_, Apq_);
    double alpha = sp_rq/DENOM;
    if (q % 50 == 0) {
      FOR_CC { long c = IDX(O.cc, i, j, k); phi_[c] += alpha*pq_[c]; }
      orc_exchange(phi_, 0);
      spmv_any(phi_, Apq_);
      FOR_CC {
        long c = IDX(O.cc, i, j, k);
        rq_[c] = rhs_[c] - Apq_[c];
        zq_[c] = rq_[c]*invM_[c];
      }
    } else {
      #pragma omp parallel for schedule(static)
      for (int k = 1; k <= O.nz; k++)
        for (int j = 1; j <= O.ny; j++)
          for (int i = 1; i <= O.nx; i++) {
            long c = IDX(O.cc, i, j, k);
            phi_[c] += alpha*pq_[c];
            rq_[c] -= alpha*Apq_[c];
            zq_[c] = rq_[c]*invM_[c];
          }
    }
    sp_rq1 = dot_cc(rq_, zq_);
    if (sp_rq1 <= O.pp_residual*O.pp_residual*sp_rhs) {
      break;
    } else if (isnan(sp_rq1)) {
      fprintf(stderr, "oracle: PCG residual is nan at iteration %d\n", q);
      O.niter = -1;
      return -1;
    } else {
      double beta = sp_rq1/sp_rq;
      #pragma omp parallel for schedule(static)
      for (int k = 1; k <= O.nz; k++)
        for (int j = 1; j <= O.ny; j++)
          for (int i = 1; i <= O.nx; i++) {
            long c = IDX(O.cc, i, j, k);
            double np = zq_[c] + beta*pq_[c];
            pq_[c] = np;
            pbq_[c] = np;
          }
      orc_exchange(pbq_, 0);
      sp_rq = sp_rq1;
    }
  }
  if (q > O.pp_max_iter && !O_pcg_continue) {
    fprintf(stderr, "oracle: PCG did not converge in %d iterations\n", q);
    O.niter = -1;
    return -1;
  }
  O.niter = q;
  O.resid = sqrt(sp_rq1)/sqrt(sp_rhs);
  /* true residual ||b - (-A) phi|| / ||b|| (diagnostic) */
  double *tmp = dal(O.cc.n), *ph = dal(O.cc.n);
  memcpy(ph, phi_, O.cc.n * sizeof(double));
  orc_exchange(ph, 0);
  spmv_any(ph, tmp);
  FOR_CC { long c = IDX(O.cc, i, j, k); tmp[c] = rhs_[c] - tmp[c]; }
  O.true_resid = sqrt(dot_cc(tmp, tmp))/sqrt(sp_rhs);
  free(tmp);
  free(ph);
  return q;
}

/* ------------------------- projection, pressure ------------------------- */
void orc_project(void)
{
  /* src/bluebottle_kernel.cu:2303-2355 */
  double dtr = O.dt / O.rho_f;
  double ddx = 1. / O.dx, ddy = 1. / O.dy, ddz = 1. / O.dz;
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= O.fx.nk; k++) for (int j = 1; j <= O.fx.nj; j++)
    for (int i = 1; i <= O.fx.ni; i++) {
      long c = IDX(O.fx, i, j, k);
      double g = abs(flu_[c]) * ddx * (phi_[IDX(O.cc, i, j, k)] -
                                      phi_[IDX(O.cc, i - 1, j, k)]);
      u_[c] = (us_[c] - dtr * g);
    }
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= O.fy.nk; k++) for (int j = 1; j <= O.fy.nj; j++)
    for (int i = 1; i <= O.fy.ni; i++) {
      long c = IDX(O.fy, i, j, k);
      double g = abs(flv_[c]) * ddy * (phi_[IDX(O.cc, i, j, k)] -
                                      phi_[IDX(O.cc, i, j - 1, k)]);
      v_[c] = (vs_[c] - dtr * g);
    }
  #pragma omp parallel for schedule(static)
  for (int k = 1; k <= O.fz.nk; k++) for (int j = 1; j <= O.fz.nj; j++)
    for (int i = 1; i <= O.fz.ni; i++) {
      long c = IDX(O.fz, i, j, k);
      double g = abs(flw_[c]) * ddz * (phi_[IDX(O.cc, i, j, k)] -
                                      phi_[IDX(O.cc, i, j, k - 1)]);
      w_[c] = (ws_[c] - dtr * g);
    }
}

void orc_update_p(void)
{
  /* src/bluebottle_kernel.cu:2385-2402, src/cuda_bluebottle.cu:2506-2535 */
  FOR_CC {
    long c = IDX(O.cc, i, j, k);
    p_[c] = (phase_[c] < 0) * (p0_[c] + phi_[c]);
    tbuf_[(long) (i - 1) + (long) (j - 1) * O.nx +
          (long) (k - 1) * O.nx * O.ny] = p_[c];
  }
  double pmean = orc_canon_sum(tbuf_, O.nx, O.ny, O.nz);
  pmean /= (double) ((long) O.nx * O.ny * O.nz);
  FOR_CC { long c = IDX(O.cc, i, j, k); p_[c] += -pmean; }
}

void orc_store_u(void)
{
  /* src/cuda_bluebottle.cu:2592-2616 */
  memcpy(c0u_, cu_, O.fx.n * sizeof(double));
  memcpy(c0v_, cv_, O.fy.n * sizeof(double));
  memcpy(c0w_, cw_, O.fz.n * sizeof(double));
  memcpy(d0u_, du_, O.fx.n * sizeof(double));
  memcpy(d0v_, dv_, O.fy.n * sizeof(double));
  memcpy(d0w_, dw_, O.fz.n * sizeof(double));
  memcpy(p0_, p_, O.cc.n * sizeof(double));
  memcpy(u0_, u_, O.fx.n * sizeof(double));
  memcpy(v0_, v_, O.fy.n * sizeof(double));
  memcpy(w0_, w_, O.fz.n * sizeof(double));
}

static double sqmax(const double *a, const OG *G)
{
  double m = 0.;
  for (int k = 1; k <= G->nk; k++) for (int j = 1; j <= G->nj; j++)
    for (int i = 1; i <= G->ni; i++) {
      double x = a[IDX(*G, i, j, k)];
      double s = x * x;
      m = (m < s) ? s : m;
    }
  return m;
}

void orc_find_dt(void)
{
  /* src/cuda_bluebottle.cu:1753-1808 */
  if (O.fixed_dt > 0.) { O.dt = O.fixed_dt; return; }
  double um = sqrt(sqmax(u_, &O.fx)), vm = sqrt(sqmax(v_, &O.fy)),
         wm = sqrt(sqmax(w_, &O.fz));
  double k = (O_scalar >= 1 && O_sD > O.nu) ? O_sD : O.nu;
  O.dt = um/O.dx + 2.*k/(O.dx * O.dx);
  O.dt += vm/O.dy + 2.*k/(O.dy * O.dy);
  O.dt += wm/O.dz + 2.*k/(O.dz * O.dz);
  O.dt = O.CFL/O.dt;
}

/* ------------------------------- driver --------------------------------- */
int orc_setup(const char *dir)
{
  memset(&O, 0, sizeof(O));
  int rc = parse(dir);
  if (rc) return rc;
  O.xl = O.xe - O.xs; O.yl = O.ye - O.ys; O.zl = O.ze - O.zs;
  O.dx = O.xl / O.nx; O.dy = O.yl / O.ny; O.dz = O.zl / O.nz;
  O.cc = mkgrid(O.nx, O.ny, O.nz);
  O.fx = mkgrid(O.nx + 1, O.ny, O.nz);
  O.fy = mkgrid(O.nx, O.ny + 1, O.nz);
  O.fz = mkgrid(O.nx, O.ny, O.nz + 1);
  p_ = dal(O.cc.n); p0_ = dal(O.cc.n); phi_ = dal(O.cc.n);
  rhs_ = dal(O.cc.n); rq_ = dal(O.cc.n); zq_ = dal(O.cc.n); pq_ = dal(O.cc.n);
  pbq_ = dal(O.cc.n); Apq_ = dal(O.cc.n); invM_ = dal(O.cc.n);
  u_ = dal(O.fx.n); u0_ = dal(O.fx.n); us_ = dal(O.fx.n); cu_ = dal(O.fx.n);
  c0u_ = dal(O.fx.n); du_ = dal(O.fx.n); d0u_ = dal(O.fx.n); fx_ = dal(O.fx.n);
  v_ = dal(O.fy.n); v0_ = dal(O.fy.n); vs_ = dal(O.fy.n); cv_ = dal(O.fy.n);
  c0v_ = dal(O.fy.n); dv_ = dal(O.fy.n); d0v_ = dal(O.fy.n); fy_ = dal(O.fy.n);
  w_ = dal(O.fz.n); w0_ = dal(O.fz.n); ws_ = dal(O.fz.n); cw_ = dal(O.fz.n);
  c0w_ = dal(O.fz.n); dw_ = dal(O.fz.n); d0w_ = dal(O.fz.n); fz_ = dal(O.fz.n);
  flu_ = ial(O.fx.n); flv_ = ial(O.fy.n); flw_ = ial(O.fz.n);
  phase_ = ial(O.cc.n); shell_ = ial(O.cc.n);
  tbuf_ = dal((long) O.nx * O.ny * O.nz);
  for (long c = 0; c < O.cc.n; c++) { phase_[c] = -1; shell_[c] = -1; }
  init_fields();
  if (O_scalar >= 1) {
    o_scalar_alloc();
    o_scalar_init();
  }
  rc = parts_parse(dir);
  if (rc) return rc;
  o_scalar_part_init();
  o_init_physalis();
  /* src/bluebottle.c:113-135 */
  orc_exchange(p_, 0); exch_vel(u_, v_, w_);
  if (O_scalar >= 1) orc_exchange(s_, 0);
  orc_build_cages();
  orc_find_dt();
  orc_bc_all();
  orc_exchange(p_, 0); exch_vel(u_, v_, w_);
  if (O_scalar >= 1) {
    o_scalar_bc(s_);
    o_scalar_bc(s0_);
    orc_exchange(s_, 0);
    orc_exchange(s0_, 0);
  }
  orc_jacobi();
  return 0;
}

int orc_step(void)
{
  /* src/bluebottle.c:168-407: the Lamb iteration repeats u* ... pressure
   * until the Lamb coefficients settle (one pass without particles). */
  O.ttime += O.dt;
  O.stepnum++;
  orc_forcing();
  if (O_scalar >= 1) o_boussinesq();
  orc_vel_bc_ramp();
  int q = 0, iter = 0;
  double err = DBL_MAX;
  while (err > O_lamb_residual) {
    orc_u_star();
    exch_vel(us_, vs_, ws_);
    o_part_bc(us_, vs_, ws_);
    orc_bc_star();
    exch_vel(us_, vs_, ws_);
    orc_solvability();
    o_part_bc(us_, vs_, ws_);
    orc_bc_star();
    exch_vel(us_, vs_, ws_);
    q = orc_pcg();
    if (q < 0) return q;
    orc_exchange(phi_, 0);
    orc_bc_p(phi_);
    orc_project();
    exch_vel(u_, v_, w_);
    o_part_bc(u_, v_, w_);
    orc_bc_all();                     /* dom_BC applies p, u, v, w */
    exch_vel(u_, v_, w_);
    orc_update_p();
    o_part_bc(u_, v_, w_);
    o_part_p_fill();
    orc_bc_p(p_);
    orc_exchange(p_, 0);
    if (NP_ > 0) {
      int num;
      o_move_sub();
      o_lamb();
      err = o_lamb_err(&num);
    } else {
      err = 0.;
    }
    iter++;
    if (iter == O_lamb_max_iter) break;
  }
  O.lamb_iters = iter;
  O.lamb_err = err;
  if (O_scalar >= 1) o_scalar_step();
  if (NP_ > 0) {
    o_move_sub();
    o_move_a();
    o_internal();
    orc_bc_all();
    exch_vel(u_, v_, w_);
  }
  orc_store_u();
  O.dt0 = O.dt;
  orc_find_dt();
  if (NP_ > 0) {
    o_move_b();
    orc_build_cages();
    orc_jacobi();
  }
  return q;
}

void orc_teardown(void)
{
  double *d[] = {p_, p0_, phi_, rhs_, rq_, zq_, pq_, pbq_, Apq_, invM_, u_, u0_,
    us_, cu_, c0u_, du_, d0u_, fx_, v_, v0_, vs_, cv_, c0v_, dv_, d0v_, fy_, w_,
    w0_, ws_, cw_, c0w_, dw_, d0w_, fz_, tbuf_};
  for (size_t i = 0; i < sizeof(d) / sizeof(d[0]); i++) free(d[i]);
  free(flu_); free(flv_); free(flw_); free(phase_); free(shell_);
  if (s_) o_scalar_free();
  free(P_);
  P_ = NULL;
  NP_ = 0;
  p_ = NULL;
}

double *orc_field(const char *n)
{
  struct { const char *n; double *a; } t[] = {
    {"p", p_}, {"p0", p0_}, {"phi", phi_}, {"u", u_}, {"u0", u0_}, {"v", v_},
    {"v0", v0_}, {"w", w_}, {"w0", w0_}, {"u_star", us_}, {"v_star", vs_},
    {"w_star", ws_}, {"conv_u", cu_}, {"conv_v", cv_}, {"conv_w", cw_},
    {"conv0_u", c0u_}, {"conv0_v", c0v_}, {"conv0_w", c0w_}, {"diff_u", du_},
    {"diff_v", dv_}, {"diff_w", dw_}, {"diff0_u", d0u_}, {"diff0_v", d0v_},
    {"diff0_w", d0w_}, {"rhs", rhs_}, {"invM", invM_}, {"f_x", fx_},
    {"f_y", fy_}, {"f_z", fz_}, {"s", s_}, {"s0", s0_}, {"s_conv", sc_},
    {"s_diff", sd_}};
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); i++)
    if (strcmp(t[i].n, n) == 0) return t[i].a;
  return NULL;
}

int *orc_ifield(const char *n)
{
  if (strcmp(n, "flag_u") == 0) return flu_;
  if (strcmp(n, "flag_v") == 0) return flv_;
  if (strcmp(n, "flag_w") == 0) return flw_;
  if (strcmp(n, "phase") == 0) return phase_;
  if (strcmp(n, "phase_shell") == 0) return shell_;
  return NULL;
}

void orc_dims(int grid, int *ni, int *nj, int *nk)
{
  const OG *G = grid == 0 ? &O.cc : grid == 1 ? &O.fx : grid == 2 ? &O.fy : &O.fz;
  *ni = G->ni; *nj = G->nj; *nk = G->nk;
}

double orc_scalar(const char *n)
{
  if (strcmp(n, "dt") == 0) return O.dt;
  if (strcmp(n, "dt0") == 0) return O.dt0;
  if (strcmp(n, "ttime") == 0) return O.ttime;
  if (strcmp(n, "stepnum") == 0) return O.stepnum;
  if (strcmp(n, "niter") == 0) return O.niter;
  if (strcmp(n, "resid") == 0) return O.resid;
  if (strcmp(n, "true_resid") == 0) return O.true_resid;
  if (strcmp(n, "sp_rhs") == 0) return O.sp_rhs;
  if (strcmp(n, "lamb_iters") == 0) return O.lamb_iters;
  if (strcmp(n, "lamb_err") == 0) return O.lamb_err;
  if (strcmp(n, "nparts") == 0) return NP_;
  if (strcmp(n, "scalar_iters") == 0) return O_scalar_iters;
  if (strcmp(n, "scalar_err") == 0) return O_scalar_err;
  return NAN;
}

void orc_set_scalar(const char *n, double v)
{
  if (strcmp(n, "dt") == 0) O.dt = v;
  else if (strcmp(n, "dt0") == 0) O.dt0 = v;
  else if (strcmp(n, "ttime") == 0) O.ttime = v;
  else if (strcmp(n, "pp_max_iter") == 0) O.pp_max_iter = (int) v;
  else if (strcmp(n, "lamb_max_iter") == 0) O_lamb_max_iter = (int) v;
  else if (strcmp(n, "pcg_continue") == 0) O_pcg_continue = (int) v;
}
