/*
 * Particle input, initial state and collision bins.
 *
 * parts_read_input reads input/part.config in the grammar of
 * src/particle.c:66-391 (same keys, same order), checking every line; a
 * missing part.config means no particles (the reference requires the file).
 * parts_init (src/particle.c:393-554) and init_bins (src/particle.c:744-834)
 * follow the reference.  One GPU holds the whole domain, so every particle
 * is held once (nparts == NPARTS) and keeps its input order as its number N.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bluebottle.h"

/* The host and device arrays keep the reference's record layout. */
_Static_assert(sizeof(part_struct) == 3128, "part_struct layout");

part_struct *parts = NULL;
bin_struct bins;
real interaction_length_ratio = 0.;
real bin_size = 0.;
int ncoeffs_max = 0;
int s_ncoeffs_max = 0;
real volume_fraction = 0., total_parts_mass = 0., total_parts_volume = 0.,
     rho_avg = 0.;

static FILE *pf;
static const char *pf_path;

static void pfail(const char *what, int n)
{
  fprintf(stderr, "%s read error (particle %d): expected \"%s\"\n", pf_path,
    n, what);
  exit(EXIT_FAILURE);
}

static void pread(int n, int nconv, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  int got = vfscanf(pf, fmt, ap);
  va_end(ap);
  if (got != nconv) pfail(fmt, n);
}

/* Particle centre inside the domain: [start, end) in a periodic direction
 * (a centre exactly at the end wraps to the start), [start, end] between
 * walls (src/particle.c:178-202). */
static int place(real *x, real s, real e, real l, int periodic)
{
  if (!periodic) return *x >= s && *x <= e;
  *x -= l * (*x == e);
  return *x >= s && *x < e;
}

void parts_read_input(void)
{
  char fname[FILE_NAME_SIZE];
  snprintf(fname, sizeof(fname), "%s/%s/part.config", ROOT_DIR, INPUT_DIR);
  pf_path = fname;
  NPARTS = 0;
  nparts = 0;
  pf = fopen(fname, "r");
  if (pf == NULL) return;

  pread(-1, 1, "n %d\n", &NPARTS);
  pread(-1, 1, "(l/a) %lf\n", &interaction_length_ratio);
  if (NPARTS < 0) pfail("n >= 0", -1);
  parts = (part_struct *) calloc(NPARTS > 0 ? NPARTS : 1, sizeof(part_struct));
  cpumem += NPARTS * sizeof(part_struct);

  real max_a = -1.;
  ncoeffs_max = -1;
  s_ncoeffs_max = -1;
  total_parts_volume = 0.;
  total_parts_mass = 0.;
  for (int i = 0; i < NPARTS; i++) {
    part_struct *P = &parts[i];
    real rs_r, tmp_s, tmp_cp, tmp_rs;
    int tmp_update, tmp_s_order;
    fscanf(pf, "\n");
    pread(i, 1, "r %lf\n", &P->r);
    pread(i, 3, "(x, y, z) %lf %lf %lf\n", &P->x, &P->y, &P->z);
    pread(i, 3, "(u, v, w) %lf %lf %lf\n", &P->u, &P->v, &P->w);
    pread(i, 3, "(aFx, aFy, aFz) %lf %lf %lf\n", &P->aFx, &P->aFy, &P->aFz);
    pread(i, 3, "(aLx, aLy, aLz) %lf %lf %lf\n", &P->aLx, &P->aLy, &P->aLz);
    pread(i, 1, "rho %lf\n", &P->rho);
    pread(i, 1, "E %lf\n", &P->E);
    pread(i, 1, "sigma %lf\n", &P->sigma);
    pread(i, 1, "e_dry %lf\n", &P->e_dry);
    pread(i, 1, "coeff_fric %lf\n", &P->coeff_fric);
    pread(i, 1, "order %d\n", &P->order);
    pread(i, 1, "rs/r %lf\n", &rs_r);
    pread(i, 1, "spring_k %lf\n", &P->spring_k);
    pread(i, 3, "spring (x, y, z) %lf %lf %lf\n", &P->spring_x, &P->spring_y,
      &P->spring_z);
    pread(i, 1, "spring_l %lf\n", &P->spring_l);
    pread(i, 1, "translating %d\n", &P->translating);
    pread(i, 1, "rotating %d\n", &P->rotating);
    pread(i, 1, "s %lf\n", &tmp_s);
    pread(i, 1, "update %d\n", &tmp_update);
    pread(i, 1, "cp %lf\n", &tmp_cp);
    pread(i, 1, "rs %lf\n", &tmp_rs);
    pread(i, 1, "s_order %d\n", &tmp_s_order);
    if (P->order < 0 || P->order > 4) pfail("order 0..4", i);
    if (P->r <= 0.) pfail("r > 0", i);

    if (P->r > max_a) max_a = P->r;
    real vol = 4./3. * PI * P->r * P->r * P->r;
    total_parts_volume += vol;
    total_parts_mass += P->rho * vol;

    int nc = 0;
    for (int j = 0; j <= P->order; j++) nc += j + 1;
    if (nc > ncoeffs_max) ncoeffs_max = nc;
    int snc = 0;
    for (int j = 0; j <= tmp_s_order; j++) snc += 2 * j + 1;
    if (snc > s_ncoeffs_max) s_ncoeffs_max = snc;

    int cx = place(&P->x, DOM.xs, DOM.xe, DOM.xl, bc.pE == PERIODIC);
    int cy = place(&P->y, DOM.ys, DOM.ye, DOM.yl, bc.pN == PERIODIC);
    int cz = place(&P->z, DOM.zs, DOM.ze, DOM.zl, bc.pT == PERIODIC);
    if (!(cx && cy && cz)) {
      printf("Total particles %d does not match parts placed in domains "
        "(particle %d at (%f, %f, %f) lies outside)\n", NPARTS, i, P->x,
        P->y, P->z);
      exit(EXIT_FAILURE);
    }
    P->N = i;
    P->rs = rs_r;
    P->s = tmp_s * (SCALAR >= 1);
    P->update = tmp_update * (SCALAR >= 1);
    P->cp = tmp_cp;
    P->srs = tmp_rs;
    P->sorder = tmp_s_order;
  }
  fclose(pf);
  nparts = NPARTS;
  if (NPARTS == 0) return;

  volume_fraction = total_parts_volume / (DOM.xl * DOM.yl * DOM.zl);
  rho_avg = total_parts_mass / total_parts_volume * volume_fraction
            + rho_f * (1. - volume_fraction);
  bin_size = 2.*max_a + interaction_length_ratio*max_a;
  /* src/particle.c:221-247 */
  const char *ax[3] = {"x", "y", "z"};
  real len[3] = {DOM.xl, DOM.yl, DOM.zl};
  int per[3] = {bc.pW == PERIODIC || bc.pE == PERIODIC,
                bc.pS == PERIODIC || bc.pN == PERIODIC,
                bc.pB == PERIODIC || bc.pT == PERIODIC};
  for (int d = 0; d < 3; d++) {
    if (len[d] < 2. * bin_size && per[d]) {
      printf("N%d >> Error: The domain size in the %s-direction must contain "
        "at least two bins with periodic boundaries. The bin length is %lf\n",
        rank, ax[d], bin_size);
      exit(EXIT_FAILURE);
    }
  }
}

void parts_init(void)
{
  for (int i = 0; i < nparts; i++) {
    part_struct *P = &parts[i];
    P->rs = P->rs * P->r;
    P->ncoeff = 0;
    for (int j = 0; j <= P->order; j++) P->ncoeff += j + 1;
    P->u0 = P->u;
    P->v0 = P->v;
    P->w0 = P->w;
    P->udot = P->vdot = P->wdot = 0.;
    P->udot0 = P->vdot0 = P->wdot0 = 0.;
    P->axx = 1.; P->axy = 0.; P->axz = 0.;
    P->ayx = 0.; P->ayy = 1.; P->ayz = 0.;
    P->azx = 0.; P->azy = 0.; P->azz = 1.;
    P->ox = P->oy = P->oz = 0.;
    P->ox0 = P->oy0 = P->oz0 = 0.;
    P->oxdot = P->oydot = P->ozdot = 0.;
    P->oxdot0 = P->oydot0 = P->ozdot0 = 0.;
    if (init_cond == SHEAR) {
      if (P->translating) {
        P->u = (bc.uNDm - bc.uSDm) * (P->y - DOM.ys)/DOM.yl + bc.uSDm;
        P->u += (bc.uTDm - bc.uBDm) * (P->z - DOM.zs)/DOM.zl + bc.uBDm;
        P->u0 = P->u;
        P->v = (bc.vEDm - bc.vWDm) * (P->x - DOM.xs)/DOM.xl + bc.vWDm;
        P->v += (bc.vTDm - bc.vBDm) * (P->z - DOM.zs)/DOM.zl + bc.vBDm;
        P->v0 = P->v;
        P->w = (bc.wEDm - bc.wWDm) * (P->x - DOM.xs)/DOM.xl + bc.wWDm;
        P->w += (bc.wNDm - bc.wSDm) * (P->y - DOM.ys)/DOM.yl + bc.wSDm;
        P->w0 = P->w;
      }
      if (P->rotating) {
        P->ox = -0.5*(bc.vTDm - bc.vBDm)/DOM.zl;
        P->ox += 0.5*(bc.wNDm - bc.wSDm)/DOM.yl;
        P->oy = 0.5*(bc.uTDm - bc.uBDm)/DOM.zl;
        P->oy += -0.5*(bc.wEDm - bc.wWDm)/DOM.xl;
        P->oz = -0.5*(bc.uNDm - bc.uSDm)/DOM.yl;
        P->oz += 0.5*(bc.vEDm - bc.vWDm)/DOM.xl;
        P->ox0 = P->ox;
        P->oy0 = P->oy;
        P->oz0 = P->oz;
      }
    } else if (init_cond == CHANNEL) {
      if (P->translating) {
        real x = P->x, y = P->y, z = P->z;
        P->u = 0.5/mu*gradP.xm*(y*y - (DOM.ys + DOM.ye)*y + DOM.ys*DOM.ye)
          * (bc.uS == DIRICHLET)
          + 0.5/mu*gradP.xm*(z*z - (DOM.zs + DOM.ze)*z + DOM.zs*DOM.ze)
          * (bc.uB == DIRICHLET);
        P->v = 0.5/mu*gradP.ym*(x*x - (DOM.xs + DOM.xe)*x + DOM.xs*DOM.xe)
          * (bc.vW == DIRICHLET)
          + 0.5/mu*gradP.ym*(z*z - (DOM.zs + DOM.ze)*z + DOM.zs*DOM.ze)
          * (bc.vB == DIRICHLET);
        P->w = 0.5/mu*gradP.zm*(x*x - (DOM.xs + DOM.xe)*x + DOM.xs*DOM.xe)
          * (bc.wW == DIRICHLET)
          + 0.5/mu*gradP.zm*(y*y - (DOM.ys + DOM.ye)*y + DOM.ys*DOM.ye)
          * (bc.wS == DIRICHLET);
        P->u0 = P->u;
        P->v0 = P->v;
        P->w0 = P->w;
      }
    }
    P->Fx = P->Fy = P->Fz = 0.;
    P->Lx = P->Ly = P->Lz = 0.;
    P->kFx = P->kFy = P->kFz = 0.;
    P->iFx = P->iFy = P->iFz = 0.;
    P->iLx = P->iLy = P->iLz = 0.;
    for (int j = 0; j < NNODES; j++) P->nodes[j] = -1;
    for (int j = 0; j < MAX_NEIGHBORS; j++) {
      P->St[j] = 0.;
      P->iSt[j] = -1;
    }
    for (int j = 0; j < MAX_COEFFS; j++) {
      P->pnm_re[j] = P->pnm_im[j] = 0.;
      P->phinm_re[j] = P->phinm_im[j] = 0.;
      P->chinm_re[j] = P->chinm_im[j] = 0.;
      P->pnm_re0[j] = P->pnm_im0[j] = 0.;
      P->phinm_re0[j] = P->phinm_im0[j] = 0.;
      P->chinm_re0[j] = P->chinm_im0[j] = 0.;
    }
    P->ncoll_part = 0;
    P->ncoll_wall = 0;
  }
}

static void bin_axis(real s, real e, real l, real *bs, real *be, real *bl,
  int *bn, real *bd)
{
  *bs = s;
  *be = e;
  *bl = l;
  *bn = (int) floor(l / bin_size);
  if (*bn == 0) *bn = 1;
  *bd = l / *bn;
}

void init_bins(void)
{
  bin_axis(dom[0].xs, dom[0].xe, dom[0].xl, &bins.xs, &bins.xe, &bins.xl,
    &bins.xn, &bins.dx);
  bin_axis(dom[0].ys, dom[0].ye, dom[0].yl, &bins.ys, &bins.ye, &bins.yl,
    &bins.yn, &bins.dy);
  bin_axis(dom[0].zs, dom[0].ze, dom[0].zl, &bins.zs, &bins.ze, &bins.zl,
    &bins.zn, &bins.dz);
  if (bins.xn < 2 || bins.yn < 2 || bins.zn < 2) {
    printf("N%d >> Error: Subdomains must contain at least two bins. The "
      "bin size is %lf\n", rank, bin_size);
    exit(EXIT_FAILURE);
  }
  bin_grid_info *G = &bins.Gcc;
  G->_is = DOM_BUF; G->_isb = G->_is - DOM_BUF; G->in = bins.xn;
  G->inb = bins.xn + 2 * DOM_BUF; G->_ie = G->_isb + G->in;
  G->_ieb = G->_ie + DOM_BUF;
  G->_js = DOM_BUF; G->_jsb = G->_js - DOM_BUF; G->jn = bins.yn;
  G->jnb = bins.yn + 2 * DOM_BUF; G->_je = G->_jsb + G->jn;
  G->_jeb = G->_je + DOM_BUF;
  G->_ks = DOM_BUF; G->_ksb = G->_ks - DOM_BUF; G->kn = bins.zn;
  G->knb = bins.zn + 2 * DOM_BUF; G->_ke = G->_ksb + G->kn;
  G->_keb = G->_ke + DOM_BUF;
  G->s1 = G->in; G->s2 = G->s1 * G->jn; G->s3 = G->s2 * G->kn;
  G->s1b = G->inb; G->s2b = G->s1b * G->jnb; G->s3b = G->s2b * G->knb;
  G->s2_i = G->jn * G->kn; G->s2_j = G->in * G->kn; G->s2_k = G->in * G->jn;
  G->s2b_i = G->jnb * G->knb; G->s2b_j = G->inb * G->knb;
  G->s2b_k = G->inb * G->jnb;
}

void parts_free(void)
{
  free(parts);
  parts = NULL;
  NPARTS = nparts = 0;
}

/* Driver hooks: set-up after the flow fields exist, and teardown. */
void bb_parts_setup(void)
{
  if (NPARTS == 0) return;
  init_bins();
  if (!use_restart) {               /* a restart reads the particle state */
    parts_init();
    scalar_part_init();
  }
  cuda_part_malloc_dev();
  cuda_part_push();
  cuda_init_physalis();
}

void bb_parts_teardown(void)
{
  if (NPARTS == 0) return;
  cuda_part_free();
  parts_free();
}
