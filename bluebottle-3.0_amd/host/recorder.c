/*
 * Logs under record/: the pressure solver's solver_expd.rec in the column
 * format of src/recorder.c:157-221 (stepnum, ttime, dt, niter, resid,
 * time), and the Lamb / scalar iteration log iter_etime.rec of
 * src/scalar.c:226-295.
 */
#include <stdio.h>
#include <stdlib.h>
#include <sys/stat.h>

#include "bluebottle.h"

int recorder_enabled = 1;

void recorder_PP_init(const char *name)
{
  if (!recorder_enabled) return;
  char buf[CHAR_BUF_SIZE];
  snprintf(buf, sizeof(buf), "%s/record", ROOT_DIR);
  struct stat st;
  if (stat(buf, &st) == -1) mkdir(buf, 0700);
  char path[FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/record/%s", ROOT_DIR, name);
  FILE *rec = fopen(path, "w");
  if (rec == NULL) {
    fprintf(stderr, "Could not open file %s\n", name);
    exit(EXIT_FAILURE);
  }
  fprintf(rec, "%-12s", "stepnum");
  fprintf(rec, "%-15s", "ttime");
  fprintf(rec, "%-15s", "dt");
  fprintf(rec, "%-8s", "niter");
  fprintf(rec, "%-15s", "resid");
  fprintf(rec, "%-15s", "time (s)");
  fclose(rec);
}

void recorder_PP(const char *name, int niter, real resid, real etime)
{
  if (!recorder_enabled) return;
  char path[FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/record/%s", ROOT_DIR, name);
  FILE *rec = fopen(path, "r+");
  if (rec == NULL) {
    recorder_PP_init(name);
    rec = fopen(path, "r+");
    if (rec == NULL) return;
  }
  fseek(rec, 0, SEEK_END);
  fprintf(rec, "\n");
  fprintf(rec, "%-12d", stepnum);
  fprintf(rec, "%-15e", ttime);
  fprintf(rec, "%-15e", dt);
  fprintf(rec, "%-8d", niter);
  fprintf(rec, "%-15e", resid);
  fprintf(rec, "%-15e", etime);
  fclose(rec);
}

void recorder_scalar_init(const char *name)
{
  /* src/scalar.c:226-262 */
  if (!recorder_enabled) return;
  char buf[CHAR_BUF_SIZE];
  snprintf(buf, sizeof(buf), "%s/record", ROOT_DIR);
  struct stat st;
  if (stat(buf, &st) == -1) mkdir(buf, 0700);
  char path[FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/record/%s", ROOT_DIR, name);
  FILE *rec = fopen(path, "w");
  if (rec == NULL) {
    fprintf(stderr, "Could not open file %s\n", name);
    exit(EXIT_FAILURE);
  }
  fprintf(rec, "%-10s", "stepnum");
  fprintf(rec, "%-12s", "ttime");
  fprintf(rec, "%-10s", "mniter");
  fprintf(rec, "%-12s", "mtimeiter");
  fprintf(rec, "%-12s", "merriter");
  fprintf(rec, "%-10s", "sniter");
  fprintf(rec, "%-12s", "stimeiter");
  fprintf(rec, "%-12s", "serriter");
  fclose(rec);
}

void recorder_scalar(const char *name, real time, int mniter, real mitertime,
  real merr, int sniter, real sitertime, real serr)
{
  /* src/scalar.c:264-295 */
  if (!recorder_enabled) return;
  char path[FILE_NAME_SIZE];
  snprintf(path, sizeof(path), "%s/record/%s", ROOT_DIR, name);
  FILE *rec = fopen(path, "r+");
  if (rec == NULL) {
    recorder_scalar_init(name);
    rec = fopen(path, "r+");
    if (rec == NULL) return;
  }
  fseek(rec, 0, SEEK_END);
  fprintf(rec, "\n");
  fprintf(rec, "%-10d", stepnum);
  fprintf(rec, "%-12.2e", time);
  fprintf(rec, "%-10d", mniter);
  fprintf(rec, "%-12.2e", mitertime);
  fprintf(rec, "%-12.2e", merr);
  fprintf(rec, "%-10d", sniter);
  fprintf(rec, "%-12.2e", sitertime);
  fprintf(rec, "%-12.2e", serr);
  fclose(rec);
}
