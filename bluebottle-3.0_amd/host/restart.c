/*
 * Restart files and the record.config switches that write them
 * (src/domain.c:3005-3261, src/recorder.c:43-98, 338-391).
 *
 * The restart file is the reference's: a raw dump of the host arrays in the
 * reference layouts and of the part_struct array, in the reference's order,
 * to <case>/input/restart.config-<rank>.  A file written by either program
 * can be read by the other (same struct layouts, one rank).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>

#include "bluebottle.h"

/* record.config (src/recorder.c:43-98).  VTK output uses rec_vtk_*; there
 * are no CGNS writers (libcgns), their clocks are kept for the restart
 * file's layout. */
real rec_cgns_flow_dt = -1., rec_cgns_part_dt = -1., rec_vtk_dt = -1.;
real rec_restart_dt = -1.;
int rec_restart_complete = -1, restart_stop = 0;
int rec_vtk_stepnum_out = 0;
real rec_cgns_flow_ttime_out = 0., rec_cgns_part_ttime_out = 0.,
     rec_vtk_ttime_out = 0.;
extern int use_restart;

static double wall_seconds(void)
{
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static double startwalltime = 0.;

void recorder_read_config(void)
{
  /* The reference requires the file; here a case without it writes no
   * restart files. */
  rec_cgns_flow_dt = rec_cgns_part_dt = rec_vtk_dt = rec_restart_dt = -1.;
  rec_restart_complete = -1;
  restart_stop = 0;
  startwalltime = wall_seconds();
  char fname[FILE_NAME_SIZE];
  snprintf(fname, sizeof(fname), "%s/%s/record.config", ROOT_DIR, INPUT_DIR);
  FILE *in = fopen(fname, "r");
  if (in == NULL) return;
  int fret = 0;
  fret += fscanf(in, "=== OUTPUT FORMATS ===\n");
  fret += fscanf(in, "CGNS_FLOW_FIELD %lf\n\n", &rec_cgns_flow_dt);
  fret += fscanf(in, "CGNS_PARTICLES %lf\n\n", &rec_cgns_part_dt);
  fret += fscanf(in, "VTK_FLOW_FIELD %lf\n", &rec_vtk_dt);
  fret += fscanf(in, "\n");
  fret += fscanf(in, "=== RESTART CAPABILITY ===\n");
  fret += fscanf(in, "WRITE_RESTART_MINUTES %lf\n", &rec_restart_dt);
  fret += fscanf(in, "WRITE_RESTART_COMPLETE %d\n", &rec_restart_complete);
  fret += fscanf(in, "RESTART_STOP %d\n", &restart_stop);
  (void) fret;
  fclose(in);
}

static void restart_name(char *fname, size_t n)
{
  int sigfigs = DOM.S3 == 1 ? 1 : (int) floor(log10(DOM.S3 - 1)) + 1;
  snprintf(fname, n, "%s/%s/restart.config-%0*d", ROOT_DIR, INPUT_DIR,
    sigfigs, rank);
}

/* one block of the file: write or read n elements */
static void blk(FILE *f, int out, void *a, size_t size, size_t n)
{
  size_t got = out ? fwrite(a, size, n, f) : fread(a, size, n, f);
  if (got != n) {
    fprintf(stderr, "restart file: short %s\n", out ? "write" : "read");
    exit(EXIT_FAILURE);
  }
}

/* the blocks before the particles, in the file's order */
static void fields_io(FILE *f, int out)
{
  dom_struct *d = &dom[0];
  size_t fx = d->Gfx.s3b, fy = d->Gfy.s3b, fz = d->Gfz.s3b, cc = d->Gcc.s3b;
  blk(f, out, &ttime, sizeof(real), 1);
  blk(f, out, &dt0, sizeof(real), 1);
  blk(f, out, &dt, sizeof(real), 1);
  blk(f, out, &stepnum, sizeof(int), 1);
  blk(f, out, &rec_vtk_stepnum_out, sizeof(int), 1);
  blk(f, out, &rec_cgns_flow_ttime_out, sizeof(real), 1);
  blk(f, out, &rec_cgns_part_ttime_out, sizeof(real), 1);
  blk(f, out, &rec_vtk_ttime_out, sizeof(real), 1);
  real *xs[] = {u, u0, diff0_u, conv0_u, diff_u, conv_u, u_star};
  real *ys[] = {v, v0, diff0_v, conv0_v, diff_v, conv_v, v_star};
  real *zs[] = {w, w0, diff0_w, conv0_w, diff_w, conv_w, w_star};
  for (int i = 0; i < 7; i++) blk(f, out, xs[i], sizeof(real), fx);
  for (int i = 0; i < 7; i++) blk(f, out, ys[i], sizeof(real), fy);
  for (int i = 0; i < 7; i++) blk(f, out, zs[i], sizeof(real), fz);
  blk(f, out, p, sizeof(real), cc);
  blk(f, out, phi, sizeof(real), cc);
  blk(f, out, p0, sizeof(real), cc);
  blk(f, out, phase, sizeof(int), cc);
  blk(f, out, phase_shell, sizeof(int), cc);
  blk(f, out, flag_u, sizeof(int), fx);
  blk(f, out, flag_v, sizeof(int), fy);
  blk(f, out, flag_w, sizeof(int), fz);
}

/* the blocks after the particles */
static void tail_io(FILE *f, int out)
{
  size_t cc = dom[0].Gcc.s3b;
  blk(f, out, &ncoeffs_max, sizeof(int), 1);
  blk(f, out, &pid_int, sizeof(real), 1);
  blk(f, out, &pid_back, sizeof(real), 1);
  blk(f, out, &gradP.z, sizeof(real), 1);
  if (SCALAR >= 1) {
    real *ss[] = {s, s0, s_conv, s_conv0, s_diff, s_diff0};
    for (int i = 0; i < 6; i++) blk(f, out, ss[i], sizeof(real), cc);
    blk(f, out, &s_ncoeffs_max, sizeof(int), 1);
  }
}

void out_restart(void)
{
  /* src/domain.c:3005-3092 */
  char fname[FILE_NAME_SIZE];
  restart_name(fname, sizeof(fname));
  FILE *f = fopen(fname, "w");
  if (f == NULL) {
    fprintf(stderr, "File %s could not be opened.\n", fname);
    exit(EXIT_FAILURE);
  }
  fields_io(f, 1);
  int nsub = nparts;
  blk(f, 1, &nsub, sizeof(int), 1);
  if (nsub > 0) blk(f, 1, parts, sizeof(part_struct), nsub);
  tail_io(f, 1);
  fclose(f);
}

void in_restart(void)
{
  /* src/domain.c:3094-3261 */
  char fname[FILE_NAME_SIZE];
  restart_name(fname, sizeof(fname));
  printf("N%d >> Reading %s\n", rank, fname);
  FILE *f = fopen(fname, "r");
  if (f == NULL) {
    fprintf(stderr, "File %s could not be opened.\n", fname);
    exit(EXIT_FAILURE);
  }
  fields_io(f, 0);
  int nsub = 0;
  blk(f, 0, &nsub, sizeof(int), 1);
  /* part.config as parsed keeps the say on translating, rotating and E
   * (src/domain.c:3163-3233) */
  part_struct *cfg = parts;
  int ncfg = NPARTS;
  parts = (part_struct *) calloc(nsub > 0 ? nsub : 1, sizeof(part_struct));
  if (nsub > 0) blk(f, 0, parts, sizeof(part_struct), nsub);
  nparts = nsub;
  int wng = 1;
  for (int i = 0; i < nparts; i++) {
    int N = parts[i].N;
    if (N < 0 || N >= ncfg) continue;
    if (parts[i].translating != cfg[N].translating ||
        parts[i].rotating != cfg[N].rotating) {
      parts[i].translating = cfg[N].translating;
      parts[i].rotating = cfg[N].rotating;
      if (wng) { printf("N%d >> part moving updated\n", rank); wng = 0; }
    }
    /* src/domain.c:3224 tests abs() of the difference, the int function:
     * the difference is truncated toward zero first, so an E that changed
     * by less than 1 keeps its restart value; the same test here */
    if (fabs(trunc(parts[i].E - cfg[N].E)) > 1e-6) {
      parts[i].E = cfg[N].E;
      if (wng) { printf("N%d >> part E updated\n", rank); wng = 0; }
    }
  }
  free(cfg);
  tail_io(f, 0);
  fclose(f);
}

int restart_recorder_write(void)
{
  /* src/recorder.c:338-391 */
  double diffwalltime = wall_seconds() - startwalltime;
  int due = 0;
  if (rec_restart_dt > 0 && diffwalltime / 60. > rec_restart_dt) {
    startwalltime += diffwalltime;
    due = 1;
  } else if (rec_restart_complete > 0 && ttime > duration) {
    due = 1;
  }
  if (!due) return 0;
  cuda_dom_pull();
  if (SCALAR >= 1) cuda_scalar_pull();
  if (nparts > 0) cuda_part_pull();
  printf("N%d >> Writing restart file (t = %e)...\n", rank, ttime);
  out_restart();
  return restart_stop ? 1 : 0;
}
