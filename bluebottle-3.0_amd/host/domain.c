/*
 * Domain set-up: command line, flow.config/decomp.config parser, grid
 * geometry, velocity-BC ramps and initial fields.
 *
 * The parser reads the grammar of src/domain.c:72-916 in the same rigid
 * order (so every file the reference reads, this reads), and additionally:
 *   - checks every line and stops with a message naming the expected text,
 *     where the reference would silently desynchronise;
 *   - accepts an optional trailing "EXTENSIONS" block after osci_f with the
 *     keys fixed_dt, perturb, pp_precision, pp_inner_tol and pp_warmstart.
 * One GPU holds the whole domain: decomp.config must describe 1 x 1 x 1.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bluebottle.h"

int use_restart;

void parse_cmdline_args(int argc, char *argv[])
{
  /* src/domain.c:33-70: "-r" selects restart input. */
  use_restart = 0;
  for (int a = 1; a < argc; a++) {
    if (strcmp(argv[a], "-r") == 0) {
      use_restart = 1;
    } else {
      printf("Usage: bluebottle command line arguments: \n");
      printf("                  -r\n");
      printf("                    Use restart file\n");
      printf("No command line arguments runs bluebottle normally.\n");
      exit(EXIT_FAILURE);
    }
  }
}

/* ---- rigid-order reader ---- */
static FILE *cfg_file;
static const char *cfg_path;

static void cfg_fail(const char *what)
{
  fprintf(stderr, "%s read error: expected \"%s\"\n", cfg_path, what);
  exit(EXIT_FAILURE);
}

/* Match `fmt` (fscanf syntax) and require `nconv` conversions. */
static void cfg_read(int nconv, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  int got = vfscanf(cfg_file, fmt, ap);
  va_end(ap);
  if (got != nconv) cfg_fail(fmt);
}

/* Match a literal header line such as "GLOBAL DOMAIN". */
static void cfg_header(const char *text)
{
  char line[CHAR_BUF_SIZE];
  int c;
  do { c = fgetc(cfg_file); } while (c == ' ' || c == '\n' || c == '\r' ||
                                     c == '\t');
  if (c == EOF) cfg_fail(text);
  ungetc(c, cfg_file);
  if (!fgets(line, sizeof(line), cfg_file)) cfg_fail(text);
  size_t n = strlen(line);
  while (n > 0 && (line[n-1] == '\n' || line[n-1] == '\r' || line[n-1] == ' '))
    line[--n] = '\0';
  if (strcmp(line, text) != 0) cfg_fail(text);
}

/* Pressure BC on one face: PERIODIC | NEUMANN <ds> (src/domain.c:216-287). */
static void read_pbc(const char *key, int *type, real *ds)
{
  char buf[CHAR_BUF_SIZE], fmt[64];
  snprintf(fmt, sizeof(fmt), " %s %%255s", key);
  cfg_read(1, fmt, buf);
  if (strcmp(buf, "PERIODIC") == 0) {
    *type = PERIODIC;
    *ds = 0;
  } else if (strcmp(buf, "NEUMANN") == 0) {
    *type = NEUMANN;
    cfg_read(1, "%lf", ds);
  } else {
    cfg_fail("PERIODIC | NEUMANN");
  }
}

/* Velocity BC on one face (src/domain.c:297-699). */
static void read_vbc(const char *key, int *type, real *Dm, real *D, real *Da)
{
  char buf[CHAR_BUF_SIZE], fmt[64];
  snprintf(fmt, sizeof(fmt), " %s %%255s", key);
  cfg_read(1, fmt, buf);
  if (strcmp(buf, "PERIODIC") == 0) {
    *type = PERIODIC;
  } else if (strcmp(buf, "DIRICHLET") == 0) {
    *type = DIRICHLET;
    cfg_read(2, "%lf %lf", Dm, Da);
    *D = 0;
  } else if (strcmp(buf, "NEUMANN") == 0) {
    *type = NEUMANN;
    *Dm = 0; *Da = 0; *D = 0;
  } else if (strcmp(buf, "PRECURSOR") == 0) {
    *type = PRECURSOR;
    cfg_read(2, "%lf %lf", Dm, Da);
    *D = 0;
  } else {
    cfg_fail("PERIODIC | DIRICHLET | NEUMANN | PRECURSOR");
  }
}

/* Scalar BC on one face (src/domain.c:701-818). */
static void read_sbc(const char *key, int *type, real *D, real *N)
{
  char buf[CHAR_BUF_SIZE], fmt[64];
  snprintf(fmt, sizeof(fmt), " %s %%255s", key);
  cfg_read(1, fmt, buf);
  if (strcmp(buf, "PERIODIC") == 0) {
    *type = PERIODIC;
  } else if (strcmp(buf, "DIRICHLET") == 0) {
    *type = DIRICHLET;
    cfg_read(1, "%lf", D);
  } else if (strcmp(buf, "NEUMANN") == 0) {
    *type = NEUMANN;
    cfg_read(1, "%lf", N);
  } else {
    cfg_fail("PERIODIC | DIRICHLET | NEUMANN");
  }
}

/* Optional trailing block (absent from the reference grammar). */
static void read_extensions(void)
{
  char key[CHAR_BUF_SIZE], val[CHAR_BUF_SIZE];
  int c;
  do { c = fgetc(cfg_file); } while (c == ' ' || c == '\n' || c == '\r' ||
                                     c == '\t');
  if (c == EOF) return;
  ungetc(c, cfg_file);
  cfg_header("EXTENSIONS");
  while (fscanf(cfg_file, " %255s", key) == 1) {
    if (strcmp(key, "fixed_dt") == 0) {
      cfg_read(1, "%lf", &fixed_dt);
    } else if (strcmp(key, "perturb") == 0) {
      cfg_read(2, "%llu %lf", &perturb_seed, &perturb_amp);
    } else if (strcmp(key, "pp_precision") == 0) {
      cfg_read(1, " %255s", val);
      if (strcmp(val, "PLAIN") == 0) pp_precision = PP_PLAIN;
      else if (strcmp(val, "MIXED") == 0) pp_precision = PP_MIXED;
      else cfg_fail("pp_precision PLAIN | MIXED");
    } else if (strcmp(key, "pp_inner_tol") == 0) {
      cfg_read(1, "%lf", &pp_inner_tol);
    } else if (strcmp(key, "pp_warmstart") == 0) {
      cfg_read(1, " %255s", val);
      if (strcmp(val, "OFF") == 0) pp_warmstart = PP_WARM_OFF;
      else if (strcmp(val, "LEARNED") == 0) pp_warmstart = PP_WARM_LEARNED;
      else cfg_fail("pp_warmstart OFF | LEARNED");
    } else {
      fprintf(stderr, "%s read error: unknown extension key \"%s\"\n",
        cfg_path, key);
      exit(EXIT_FAILURE);
    }
  }
}

void domain_read_input(void)
{
  cpumem = 0;
  gpumem = 0;
  fixed_dt = 0.;
  perturb_seed = 0ULL;
  perturb_amp = 0.;
  pp_precision = PP_PLAIN;
  pp_warmstart = PP_WARM_OFF;
  pp_inner_tol = 1e-3;

  char fname[FILE_NAME_SIZE], dname[FILE_NAME_SIZE];
  snprintf(fname, sizeof(fname), "%s/%s/flow.config", ROOT_DIR, INPUT_DIR);
  snprintf(dname, sizeof(dname), "%s/%s/decomp.config", ROOT_DIR, INPUT_DIR);

  cfg_file = fopen(fname, "r");
  if (cfg_file == NULL) {
    fprintf(stderr, "Could not open file %s\n", fname);
    exit(EXIT_FAILURE);
  }
  cfg_path = fname;

  cfg_header("GLOBAL DOMAIN");
  cfg_read(3, " (Xs, Xe, Xn) %lf %lf %d", &DOM.xs, &DOM.xe, &DOM.xn);
  cfg_read(3, " (Ys, Ye, Yn) %lf %lf %d", &DOM.ys, &DOM.ye, &DOM.yn);
  cfg_read(3, " (Zs, Ze, Zn) %lf %lf %d", &DOM.zs, &DOM.ze, &DOM.zn);

  cfg_header("MPI/GPU SUBDOMAIN DECOMPOSITION");
  cfg_read(3, " (In, Jn, Kn) %d %d %d", &DOM.In, &DOM.Jn, &DOM.Kn);
  DOM.S1 = DOM.In;
  DOM.S2 = DOM.S1 * DOM.Jn;
  DOM.S3 = DOM.S2 * DOM.Kn;
  if (DOM.S3 != nprocs) {
    /* src/domain.c:111-117 */
    fprintf(stderr, "  The number of processes (%d) does not\n   match the "
      "number specified in the flow.config file (%d); one GPU holds the "
      "whole domain, so use (In, Jn, Kn) 1 1 1\n", nprocs, DOM.S3);
    exit(EXIT_FAILURE);
  }
  if (DOM.xn < 2 || DOM.yn < 2 || DOM.zn < 2) {
    fprintf(stderr, "%s: every direction needs at least 2 cells\n", fname);
    exit(EXIT_FAILURE);
  }

  dom = (dom_struct *) calloc(DOM.S3, sizeof(dom_struct));
  cpumem += DOM.S3 * sizeof(dom_struct);

  FILE *dfile = fopen(dname, "r");
  if (dfile == NULL) {
    fprintf(stderr, "Could not open file %s\n", dname);
    exit(EXIT_FAILURE);
  }
  if (fscanf(dfile, " (I, J, K) %d %d %d", &dom[0].I, &dom[0].J,
        &dom[0].K) != 3 ||
      fscanf(dfile, " (Xs, Xe, Xn) %lf %lf %d", &dom[0].xs, &dom[0].xe,
        &dom[0].xn) != 3 ||
      fscanf(dfile, " (Ys, Ye, Yn) %lf %lf %d", &dom[0].ys, &dom[0].ye,
        &dom[0].yn) != 3 ||
      fscanf(dfile, " (Zs, Ze, Zn) %lf %lf %d", &dom[0].zs, &dom[0].ze,
        &dom[0].zn) != 3) {
    fprintf(stderr, "%s read error\n", dname);
    exit(EXIT_FAILURE);
  }
  fclose(dfile);
  if (dom[0].I != 0 || dom[0].J != 0 || dom[0].K != 0 ||
      dom[0].xn != DOM.xn || dom[0].yn != DOM.yn || dom[0].zn != DOM.zn) {
    fprintf(stderr, "%s: the single subdomain must be the global domain\n",
      dname);
    exit(EXIT_FAILURE);
  }
  dom[0].rank = 0;

  cfg_header("PHYSICAL PARAMETERS");
  cfg_read(1, " rho_f %lf", &rho_f);
  cfg_read(1, " nu %lf", &nu);
  cfg_read(1, " s_D %lf", &s_D);
  cfg_read(1, " s_k %lf", &s_k);
  cfg_read(1, " s_beta %lf", &s_beta);
  cfg_read(1, " s_ref %lf", &s_ref);
  mu = nu * rho_f;

  cfg_header("SIMULATION PARAMETERS");
  cfg_read(1, " duration %lf", &duration);
  cfg_read(1, " CFL %lf", &CFL);
  cfg_read(1, " SCALAR %d", &SCALAR);
  cfg_read(1, " pp_max_iter %d", &pp_max_iter);
  cfg_read(1, " pp_residual %lf", &pp_residual);
  cfg_read(1, " lamb_max_iter %d", &lamb_max_iter);
  cfg_read(1, " lamb_residual %lf", &lamb_residual);
  cfg_read(1, " lamb_relax %lf", &lamb_relax);
  cfg_read(1, " lamb_cut %lf", &lamb_cut);
  cfg_read(1, " s_lamb_cut %lf", &lamb_cut_scalar);

  cfg_header("BOUNDARY CONDITIONS");
  cfg_read(1, " v_bc_tdelay %lf", &v_bc_tdelay);
  cfg_header("PRESSURE");
  read_pbc("bc.pW", &bc.pW, &bc.dsW);
  read_pbc("bc.pE", &bc.pE, &bc.dsE);
  read_pbc("bc.pS", &bc.pS, &bc.dsS);
  read_pbc("bc.pN", &bc.pN, &bc.dsN);
  read_pbc("bc.pB", &bc.pB, &bc.dsB);
  read_pbc("bc.pT", &bc.pT, &bc.dsT);
  if (bc.dsE < 0 || bc.dsW < 0 || bc.dsN < 0 || bc.dsS < 0 ||
      bc.dsT < 0 || bc.dsB < 0) {
    fprintf(stderr, "flow.config read error -- screen offsets must be >= 0\n");
    exit(EXIT_FAILURE);
  }
  cfg_header("X-VELOCITY");
  read_vbc("bc.uW", &bc.uW, &bc.uWDm, &bc.uWD, &bc.uWDa);
  read_vbc("bc.uE", &bc.uE, &bc.uEDm, &bc.uED, &bc.uEDa);
  read_vbc("bc.uS", &bc.uS, &bc.uSDm, &bc.uSD, &bc.uSDa);
  read_vbc("bc.uN", &bc.uN, &bc.uNDm, &bc.uND, &bc.uNDa);
  read_vbc("bc.uB", &bc.uB, &bc.uBDm, &bc.uBD, &bc.uBDa);
  read_vbc("bc.uT", &bc.uT, &bc.uTDm, &bc.uTD, &bc.uTDa);
  cfg_header("Y-VELOCITY");
  read_vbc("bc.vW", &bc.vW, &bc.vWDm, &bc.vWD, &bc.vWDa);
  read_vbc("bc.vE", &bc.vE, &bc.vEDm, &bc.vED, &bc.vEDa);
  read_vbc("bc.vS", &bc.vS, &bc.vSDm, &bc.vSD, &bc.vSDa);
  read_vbc("bc.vN", &bc.vN, &bc.vNDm, &bc.vND, &bc.vNDa);
  read_vbc("bc.vB", &bc.vB, &bc.vBDm, &bc.vBD, &bc.vBDa);
  read_vbc("bc.vT", &bc.vT, &bc.vTDm, &bc.vTD, &bc.vTDa);
  cfg_header("Z-VELOCITY");
  read_vbc("bc.wW", &bc.wW, &bc.wWDm, &bc.wWD, &bc.wWDa);
  read_vbc("bc.wE", &bc.wE, &bc.wEDm, &bc.wED, &bc.wEDa);
  read_vbc("bc.wS", &bc.wS, &bc.wSDm, &bc.wSD, &bc.wSDa);
  read_vbc("bc.wN", &bc.wN, &bc.wNDm, &bc.wND, &bc.wNDa);
  read_vbc("bc.wB", &bc.wB, &bc.wBDm, &bc.wBD, &bc.wBDa);
  read_vbc("bc.wT", &bc.wT, &bc.wTDm, &bc.wTD, &bc.wTDa);
  cfg_header("SCALAR");
  read_sbc("bc_s.sW", &bc_s.sW, &bc_s.sWD, &bc_s.sWN);
  read_sbc("bc_s.sE", &bc_s.sE, &bc_s.sED, &bc_s.sEN);
  read_sbc("bc_s.sS", &bc_s.sS, &bc_s.sSD, &bc_s.sSN);
  read_sbc("bc_s.sN", &bc_s.sN, &bc_s.sND, &bc_s.sNN);
  read_sbc("bc_s.sB", &bc_s.sB, &bc_s.sBD, &bc_s.sBN);
  read_sbc("bc_s.sT", &bc_s.sT, &bc_s.sTD, &bc_s.sTN);

  char buf[CHAR_BUF_SIZE];
  cfg_header("INITIAL CONDITION");
  cfg_read(1, " init_cond %255s", buf);
  if (strcmp(buf, "QUIESCENT") == 0) init_cond = QUIESCENT;
  else if (strcmp(buf, "SHEAR") == 0) init_cond = SHEAR;
  else if (strcmp(buf, "CHANNEL") == 0) init_cond = CHANNEL;
  else if (strcmp(buf, "TAYLOR_GREEN") == 0) init_cond = TAYLOR_GREEN;
  else if (strcmp(buf, "TAYLOR_GREEN_3") == 0) init_cond = TAYLOR_GREEN_3;
  else if (strcmp(buf, "TURB_CHANNEL") == 0) init_cond = TURB_CHANNEL;
  else if (strcmp(buf, "TURBULENT") == 0) init_cond = TURBULENT;
  else cfg_fail("init_cond <QUIESCENT|SHEAR|CHANNEL|TAYLOR_GREEN|...>");
  cfg_read(2, " init_s %lf %lf", &s_init, &s_init_rand);

  cfg_header("SOLVABILITY ENFORCEMENT PLANE");
  cfg_read(1, " out_plane %255s", buf);
  if (strcmp(buf, "WEST") == 0) out_plane = WEST;
  else if (strcmp(buf, "EAST") == 0) out_plane = EAST;
  else if (strcmp(buf, "SOUTH") == 0) out_plane = SOUTH;
  else if (strcmp(buf, "NORTH") == 0) out_plane = NORTH;
  else if (strcmp(buf, "BOTTOM") == 0) out_plane = BOTTOM;
  else if (strcmp(buf, "TOP") == 0) out_plane = TOP;
  else if (strcmp(buf, "HOMOGENEOUS") == 0) out_plane = HOMOGENEOUS;
  else cfg_fail("out_plane <WEST|EAST|SOUTH|NORTH|BOTTOM|TOP|HOMOGENEOUS>");

  gradP.x = 0.; gradP.y = 0.; gradP.z = 0.;
  g.x = 0.; g.y = 0.; g.z = 0.;
  pid_int = 0.; pid_back = 0.;
  cfg_header("SIMULATION DRIVING CONDITIONS");
  cfg_read(1, " p_bc_tdelay %lf", &p_bc_tdelay);
  cfg_read(2, " gradP.x %lf %lf", &gradP.xm, &gradP.xa);
  cfg_read(2, " gradP.y %lf %lf", &gradP.ym, &gradP.ya);
  cfg_read(2, " gradP.z %lf %lf", &gradP.zm, &gradP.za);
  cfg_read(1, " g_bc_tdelay %lf", &g_bc_tdelay);
  cfg_read(2, " g.x %lf %lf", &g.xm, &g.xa);
  cfg_read(2, " g.y %lf %lf", &g.ym, &g.ya);
  cfg_read(2, " g.z %lf %lf", &g.zm, &g.za);

  cfg_header("PID CONTROLLER GAINS");
  cfg_read(1, " Kp %lf", &Kp);
  cfg_read(1, " Ki %lf", &Ki);
  cfg_read(1, " Kd %lf", &Kd);

  cfg_header("TURBULENT FORCING PARAMETERS");
  cfg_read(1, " turbA %lf", &turbA);
  cfg_read(1, " osci_f %lf", &osci_f);

  read_extensions();
  fclose(cfg_file);
  cfg_file = NULL;

  if (Kp > 0 || Ki > 0 || Kd > 0) {
    fprintf(stderr, "flow.config: the PID-controlled pressure gradient "
      "(Kp, Ki, Kd > 0) is not supported by this build\n");
    exit(EXIT_FAILURE);
  }
  if (init_cond == TURBULENT || init_cond == TURB_CHANNEL) {
    fprintf(stderr, "flow.config: init_cond TURBULENT/TURB_CHANNEL are not "
      "supported by this build; use CHANNEL with the `perturb` key\n");
    exit(EXIT_FAILURE);
  }
  if ((bc.pW == PERIODIC) != (bc.pE == PERIODIC) ||
      (bc.pS == PERIODIC) != (bc.pN == PERIODIC) ||
      (bc.pB == PERIODIC) != (bc.pT == PERIODIC)) {
    fprintf(stderr, "flow.config: a periodic direction must be periodic on "
      "both faces\n");
    exit(EXIT_FAILURE);
  }
}

/* Fill one grid_info for a grid with (ni, nj, nk) points per direction.
 * `stag` names the staggered direction (0 = none, 1 = x, 2 = y, 3 = z) and
 * decides the reference host strides (src/domain.c:934-1146). */
static void fill_grid(grid_info *G, int ni, int nj, int nk, int stag)
{
  G->is = DOM_BUF; G->isb = G->is - DOM_BUF; G->in = ni;
  G->inb = G->in + 2 * DOM_BUF; G->ie = G->isb + G->in; G->ieb = G->ie + DOM_BUF;
  G->js = DOM_BUF; G->jsb = G->js - DOM_BUF; G->jn = nj;
  G->jnb = G->jn + 2 * DOM_BUF; G->je = G->jsb + G->jn; G->jeb = G->je + DOM_BUF;
  G->ks = DOM_BUF; G->ksb = G->ks - DOM_BUF; G->kn = nk;
  G->knb = G->kn + 2 * DOM_BUF; G->ke = G->ksb + G->kn; G->keb = G->ke + DOM_BUF;
  G->_is = G->is; G->_isb = G->isb; G->_ie = G->ie; G->_ieb = G->ieb;
  G->_js = G->js; G->_jsb = G->jsb; G->_je = G->je; G->_jeb = G->jeb;
  G->_ks = G->ks; G->_ksb = G->ksb; G->_ke = G->ke; G->_keb = G->keb;
  int a, b, c, ab, bb, cb;            /* fastest, middle, slowest extents */
  switch (stag) {
    case 1:  a = nj; b = nk; c = ni; ab = G->jnb; bb = G->knb; cb = G->inb; break;
    case 2:  a = nk; b = ni; c = nj; ab = G->knb; bb = G->inb; cb = G->jnb; break;
    default: a = ni; b = nj; c = nk; ab = G->inb; bb = G->jnb; cb = G->knb; break;
  }
  G->s1 = a; G->s2 = a * b; G->s3 = a * b * c;
  G->s1b = ab; G->s2b = ab * bb; G->s3b = ab * bb * cb;
  G->s2_i = G->jn * G->kn; G->s2_j = G->in * G->kn; G->s2_k = G->in * G->jn;
  G->s2b_i = G->jnb * G->knb; G->s2b_j = G->inb * G->knb;
  G->s2b_k = G->inb * G->jnb;
}

void domain_fill(void)
{
  /* src/domain.c:918-1487 specialised to a 1 x 1 x 1 decomposition. */
  DOM.xl = DOM.xe - DOM.xs;
  DOM.yl = DOM.ye - DOM.ys;
  DOM.zl = DOM.ze - DOM.zs;
  DOM.dx = DOM.xl / DOM.xn;
  DOM.dy = DOM.yl / DOM.yn;
  DOM.dz = DOM.zl / DOM.zn;
  fill_grid(&DOM.Gcc, DOM.xn, DOM.yn, DOM.zn, 0);
  fill_grid(&DOM.Gfx, DOM.xn + 1, DOM.yn, DOM.zn, 1);
  fill_grid(&DOM.Gfy, DOM.xn, DOM.yn + 1, DOM.zn, 2);
  fill_grid(&DOM.Gfz, DOM.xn, DOM.yn, DOM.zn + 1, 3);
  DOM.Is = 0; DOM.Ie = DOM.In - 1;
  DOM.Js = 0; DOM.Je = DOM.Jn - 1;
  DOM.Ks = 0; DOM.Ke = DOM.Kn - 1;

  dom_struct *d = &dom[0];
  int I = d->I, J = d->J, K = d->K;
  real xs = d->xs, xe = d->xe, ys = d->ys, ye = d->ye, zs = d->zs, ze = d->ze;
  int xn = d->xn, yn = d->yn, zn = d->zn;
  *d = DOM;
  d->I = I; d->J = J; d->K = K;
  d->xs = xs; d->xe = xe; d->ys = ys; d->ye = ye; d->zs = zs; d->ze = ze;
  d->xn = xn; d->yn = yn; d->zn = zn;
  d->xl = d->xe - d->xs; d->yl = d->ye - d->ys; d->zl = d->ze - d->zs;
  d->dx = d->xl / d->xn; d->dy = d->yl / d->yn; d->dz = d->zl / d->zn;
  d->rank = 0;
  /* Neighbours (src/domain.c:1147-1202): self on periodic faces. */
  d->w = (bc.pW == PERIODIC) ? 0 : BB_PROC_NULL;
  d->e = (bc.pE == PERIODIC) ? 0 : BB_PROC_NULL;
  d->s = (bc.pS == PERIODIC) ? 0 : BB_PROC_NULL;
  d->n = (bc.pN == PERIODIC) ? 0 : BB_PROC_NULL;
  d->b = (bc.pB == PERIODIC) ? 0 : BB_PROC_NULL;
  d->t = (bc.pT == PERIODIC) ? 0 : BB_PROC_NULL;
}

/* Linear ramp of a Dirichlet value (src/domain.c:1935-1942 pattern). */
static real ramp(real Dm, real Da, real delta, real cur)
{
  if (delta < 0) return cur;
  if (Da == 0) return Dm;
  if (fabs(delta * Da) > fabs(Dm)) return Dm;
  return delta * Da;
}

void compute_vel_BC(void)
{
  /* src/domain.c:1931-2082 */
  real delta = ttime - v_bc_tdelay;
  bc.uWD = ramp(bc.uWDm, bc.uWDa, delta, bc.uWD);
  bc.uED = ramp(bc.uEDm, bc.uEDa, delta, bc.uED);
  bc.uSD = ramp(bc.uSDm, bc.uSDa, delta, bc.uSD);
  bc.uND = ramp(bc.uNDm, bc.uNDa, delta, bc.uND);
  bc.uBD = ramp(bc.uBDm, bc.uBDa, delta, bc.uBD);
  bc.uTD = ramp(bc.uTDm, bc.uTDa, delta, bc.uTD);
  bc.vWD = ramp(bc.vWDm, bc.vWDa, delta, bc.vWD);
  bc.vED = ramp(bc.vEDm, bc.vEDa, delta, bc.vED);
  bc.vSD = ramp(bc.vSDm, bc.vSDa, delta, bc.vSD);
  bc.vND = ramp(bc.vNDm, bc.vNDa, delta, bc.vND);
  bc.vBD = ramp(bc.vBDm, bc.vBDa, delta, bc.vBD);
  bc.vTD = ramp(bc.vTDm, bc.vTDa, delta, bc.vTD);
  bc.wWD = ramp(bc.wWDm, bc.wWDa, delta, bc.wWD);
  bc.wED = ramp(bc.wEDm, bc.wEDa, delta, bc.wED);
  bc.wSD = ramp(bc.wSDm, bc.wSDa, delta, bc.wSD);
  bc.wND = ramp(bc.wNDm, bc.wNDa, delta, bc.wND);
  bc.wBD = ramp(bc.wBDm, bc.wBDa, delta, bc.wBD);
  bc.wTD = ramp(bc.wTDm, bc.wTDa, delta, bc.wTD);
  cuda_update_bc();
}

/* ---- initial fields (src/domain.c:2144-2277, 2808-2990) ---- */
#define LOOP_G(G) \
  for (int _kk = (G)._ksb; _kk <= (G)._keb; _kk++) \
  for (int _jj = (G)._jsb; _jj <= (G)._jeb; _jj++) \
  for (int _ii = (G)._isb; _ii <= (G)._ieb; _ii++)

static void init_quiescent(void)
{
  dom_struct *d = &dom[0];
  for (int i = 0; i < d->Gcc.s3b; i++) { p[i] = 0.; p0[i] = 0.; phi[i] = 0.; }
  for (int i = 0; i < d->Gfx.s3b; i++) {
    u[i] = 0.; u0[i] = 0.; conv_u[i] = 0.; conv0_u[i] = 0.;
    diff_u[i] = 0.; diff0_u[i] = 0.; f_x[i] = 0.; u_star[i] = 0.;
  }
  for (int i = 0; i < d->Gfy.s3b; i++) {
    v[i] = 0.; v0[i] = 0.; conv_v[i] = 0.; conv0_v[i] = 0.;
    diff_v[i] = 0.; diff0_v[i] = 0.; f_y[i] = 0.; v_star[i] = 0.;
  }
  for (int i = 0; i < d->Gfz.s3b; i++) {
    w[i] = 0.; w0[i] = 0.; conv_w[i] = 0.; conv0_w[i] = 0.;
    diff_w[i] = 0.; diff0_w[i] = 0.; f_z[i] = 0.; w_star[i] = 0.;
  }
}

static void init_shear(void)
{
  /* src/domain.c:2184-2227 */
  dom_struct *d = &dom[0];
  LOOP_G(d->Gfx) {
    real y = ((_jj - 0.5) * d->dy) + d->ys;
    real z = ((_kk - 0.5) * d->dz) + d->zs;
    int c = GFX_LOC(_ii, _jj, _kk, d->Gfx.s1b, d->Gfx.s2b);
    u[c] = (bc.uNDm - bc.uSDm)*(y - DOM.ys)/DOM.yl + bc.uSDm;
    u[c] += (bc.uTDm - bc.uBDm)*(z - DOM.zs)/DOM.zl + bc.uBDm;
    u0[c] = u[c];
  }
  LOOP_G(d->Gfy) {
    real z = ((_kk - 0.5) * d->dz) + d->zs;
    real x = ((_ii - 0.5) * d->dx) + d->xs;
    int c = GFY_LOC(_ii, _jj, _kk, d->Gfy.s1b, d->Gfy.s2b);
    v[c] = (bc.vEDm - bc.vWDm)*(x - DOM.xs)/DOM.xl + bc.vWDm;
    v[c] += (bc.vTDm - bc.vBDm)*(z - DOM.zs)/DOM.zl + bc.vBDm;
    v0[c] = v[c];
  }
  LOOP_G(d->Gfz) {
    real x = ((_ii - 0.5) * d->dx) + d->xs;
    real y = ((_jj - 0.5) * d->dy) + d->ys;
    int c = GFZ_LOC(_ii, _jj, _kk, d->Gfz.s1b, d->Gfz.s2b);
    w[c] = (bc.wEDm - bc.wWDm)*(x - DOM.xs)/DOM.xl + bc.wWDm;
    w[c] += (bc.wNDm - bc.wSDm)*(y - DOM.ys)/DOM.yl + bc.wSDm;
    w0[c] = w[c];
  }
}

static void init_channel(void)
{
  /* Laminar Poiseuille profiles (src/domain.c:2229-2277). */
  dom_struct *d = &dom[0];
  LOOP_G(d->Gfx) {
    real y = ((_jj - 0.5) * d->dy) + d->ys;
    real z = ((_kk - 0.5) * d->dz) + d->zs;
    int c = GFX_LOC(_ii, _jj, _kk, d->Gfx.s1b, d->Gfx.s2b);
    u[c] = 0.5/mu*gradP.xm*(y*y - (DOM.ys + DOM.ye)*y + DOM.ys*DOM.ye)
      * (bc.uS == DIRICHLET)
      + 0.5/mu*gradP.xm*(z*z - (DOM.zs + DOM.ze)*z + DOM.zs*DOM.ze)
      * (bc.uB == DIRICHLET);
    u0[c] = u[c];
  }
  LOOP_G(d->Gfy) {
    real z = ((_kk - 0.5) * d->dz) + d->zs;
    real x = ((_ii - 0.5) * d->dx) + d->xs;
    int c = GFY_LOC(_ii, _jj, _kk, d->Gfy.s1b, d->Gfy.s2b);
    v[c] = 0.5/mu*gradP.ym*(x*x - (DOM.xs + DOM.xe)*x + DOM.xs*DOM.xe)
      * (bc.vW == DIRICHLET)
      + 0.5/mu*gradP.ym*(z*z - (DOM.zs + DOM.ze)*z + DOM.zs*DOM.ze)
      * (bc.vB == DIRICHLET);
    v0[c] = v[c];
  }
  LOOP_G(d->Gfz) {
    real x = ((_ii - 0.5) * d->dx) + d->xs;
    real y = ((_jj - 0.5) * d->dy) + d->ys;
    int c = GFZ_LOC(_ii, _jj, _kk, d->Gfz.s1b, d->Gfz.s2b);
    w[c] = 0.5/mu*gradP.zm*(x*x - (DOM.xs + DOM.xe)*x + DOM.xs*DOM.xe)
      * (bc.wW == DIRICHLET)
      + 0.5/mu*gradP.zm*(y*y - (DOM.ys + DOM.ye)*y + DOM.ys*DOM.ye)
      * (bc.wS == DIRICHLET);
    w0[c] = w[c];
  }
}

static void init_tg3(void)
{
  /* src/domain.c:2808-2887: 3-D Taylor-Green, Re = 1600 hard-coded. */
  dom_struct *d = &dom[0];
  real Re = 1600., L = 1.;
  real vel0 = Re * nu / L;
  LOOP_G(d->Gfx) {
    real x = ((_ii - 1.0) * d->dx) + d->xs;
    real y = ((_jj - 0.5) * d->dy) + d->ys;
    real z = ((_kk - 0.5) * d->dz) + d->zs;
    int c = GFX_LOC(_ii, _jj, _kk, d->Gfx.s1b, d->Gfx.s2b);
    u[c] = vel0*sin(x/L)*cos(y/L)*cos(z/L);
    u0[c] = u[c];
  }
  LOOP_G(d->Gfy) {
    real x = ((_ii - 0.5) * d->dx) + d->xs;
    real y = ((_jj - 1.0) * d->dy) + d->ys;
    real z = ((_kk - 0.5) * d->dz) + d->zs;
    int c = GFY_LOC(_ii, _jj, _kk, d->Gfy.s1b, d->Gfy.s2b);
    v[c] = -vel0*cos(x/L)*sin(y/L)*cos(z/L);
    v0[c] = v[c];
  }
  LOOP_G(d->Gfz) {
    int c = GFZ_LOC(_ii, _jj, _kk, d->Gfz.s1b, d->Gfz.s2b);
    w[c] = 0.;
    w0[c] = 0.;
  }
  LOOP_G(d->Gcc) {
    real x = ((_ii - 0.5) * d->dx) + d->xs;
    real y = ((_jj - 0.5) * d->dy) + d->ys;
    real z = ((_kk - 0.5) * d->dz) + d->zs;
    int c = GCC_LOC(_ii, _jj, _kk, d->Gcc.s1b, d->Gcc.s2b);
    p[c] = rho_f*vel0*vel0/16.*(cos(2.*x/L) + cos(2.*y/L))*(cos(2.*z/L) + 2.);
    p0[c] = p[c];
  }
}

/* Seeded uniform(-a, a) perturbation of the interior faces (the `perturb`
 * extension key; the reference's TURB_CHANNEL seeds with time(NULL),
 * src/domain.c:2302).  Draw order: u, then v, then w; within a field k, j, i
 * over the faces [s, e] of each direction.  In a periodic direction the
 * last face duplicates the first and takes its value, so the two copies of
 * that face stay bit-identical. */
static void init_perturb(void)
{
  dom_struct *d = &dom[0];
  real a = perturb_amp;
  rng_init(perturb_seed);
  grid_info *G = &d->Gfx;
  for (int k = G->_ks; k <= G->_ke; k++)
    for (int j = G->_js; j <= G->_je; j++)
      for (int i = G->_is; i <= G->_ie; i++) {
        if (i == G->_ie && bc.uE == PERIODIC) continue;
        int c = GFX_LOC(i, j, k, G->s1b, G->s2b);
        u[c] += a * (2. * rng_dbl() - 1.);
        u0[c] = u[c];
      }
  if (bc.uE == PERIODIC)
    for (int k = G->_ks; k <= G->_ke; k++)
      for (int j = G->_js; j <= G->_je; j++) {
        int c = GFX_LOC(G->_ie, j, k, G->s1b, G->s2b);
        u[c] = u[GFX_LOC(G->_is, j, k, G->s1b, G->s2b)];
        u0[c] = u[c];
      }
  G = &d->Gfy;
  for (int k = G->_ks; k <= G->_ke; k++)
    for (int j = G->_js; j <= G->_je; j++)
      for (int i = G->_is; i <= G->_ie; i++) {
        if (j == G->_je && bc.vN == PERIODIC) continue;
        int c = GFY_LOC(i, j, k, G->s1b, G->s2b);
        v[c] += a * (2. * rng_dbl() - 1.);
        v0[c] = v[c];
      }
  if (bc.vN == PERIODIC)
    for (int k = G->_ks; k <= G->_ke; k++)
      for (int i = G->_is; i <= G->_ie; i++) {
        int c = GFY_LOC(i, G->_je, k, G->s1b, G->s2b);
        v[c] = v[GFY_LOC(i, G->_js, k, G->s1b, G->s2b)];
        v0[c] = v[c];
      }
  G = &d->Gfz;
  for (int k = G->_ks; k <= G->_ke; k++)
    for (int j = G->_js; j <= G->_je; j++)
      for (int i = G->_is; i <= G->_ie; i++) {
        if (k == G->_ke && bc.wT == PERIODIC) continue;
        int c = GFZ_LOC(i, j, k, G->s1b, G->s2b);
        w[c] += a * (2. * rng_dbl() - 1.);
        w0[c] = w[c];
      }
  if (bc.wT == PERIODIC)
    for (int j = G->_js; j <= G->_je; j++)
      for (int i = G->_is; i <= G->_ie; i++) {
        int c = GFZ_LOC(i, j, G->_ke, G->s1b, G->s2b);
        w[c] = w[GFZ_LOC(i, j, G->_ks, G->s1b, G->s2b)];
        w0[c] = w[c];
      }
}

void domain_init_fields(void)
{
  /* src/domain.c:2083-2142 */
  if (init_cond != TURBULENT) { turbA = 0.; turbl = 0.; turb_k0 = 0.; }
  init_quiescent();
  if (init_cond == SHEAR) init_shear();
  if (init_cond == CHANNEL) init_channel();
  if (init_cond == TAYLOR_GREEN_3) init_tg3();
  if (init_cond == TAYLOR_GREEN) {
    fprintf(stderr, "flow.config: init_cond TAYLOR_GREEN is not supported by "
      "this build; use TAYLOR_GREEN_3\n");
    exit(EXIT_FAILURE);
  }
  if (perturb_amp > 0.) init_perturb();

  dt = 2.*nu / (dom[0].dx * dom[0].dx);
  dt += 2.*nu / (dom[0].dy * dom[0].dy);
  dt += 2.*nu / (dom[0].dz * dom[0].dz);
  dt = CFL / dt;
  if (fixed_dt > 0.) dt = fixed_dt;
  dt0 = -1.;
  stepnum = 0;
  ttime = 0.;
}

void domain_free(void)
{
  free(dom);
  dom = NULL;
}
