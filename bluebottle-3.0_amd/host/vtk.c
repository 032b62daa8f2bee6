/*
 * ParaView output (src/vtk.c:29-405) and the output part of recorder_out
 * (src/recorder.c:91-155).
 *
 * Files, under <case>/output, in the reference's formats so that its
 * ParaView state files and post-processing find the same names and arrays:
 * - out.pvd: a Collection listing, per output, out_<n>.pvtr (part 0) and
 *   out_<n>.pvtp (part 1) at the output's time;
 * - out_<n>.pvtr + out_<n>_0_of_1.vtr: the flow on the rectilinear grid of
 *   cell corners (extent 0..nx, 0..ny, 0..nz), cell data p, phase,
 *   phase_shell and vel (face velocities averaged to the cell centre), ASCII;
 * - out_<n>.pvtp + out_<n>_0_of_1.vtp: one point per particle at its centre
 *   with n (particle number), r and pvel.
 * One rank holds the domain, so every parallel file names one piece.
 *
 * The reference writes flow output only when built with CGNS_OUTPUT
 * (src/bluebottle.c:159-162, 376-379), and CGNS needs libcgns.  This library
 * has no CGNS writers and calls the VTK recorder unconditionally; with
 * VTK_FLOW_FIELD <= 0 in record.config (or no record.config) nothing is
 * pulled or written.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include "bluebottle.h"

extern int bb_verbose;

static const char PVD_FOOTER[] = "</Collection>\n</VTKFile>";

static FILE *open_or_die(const char *path, const char *mode)
{
  FILE *f = fopen(path, mode);
  if (f == NULL) {
    fprintf(stderr, "Could not open file %s\n", path);
    exit(EXIT_FAILURE);
  }
  return f;
}

static void output_path(char *buf, size_t n, const char *name)
{
  snprintf(buf, n, "%s/%s/%s", ROOT_DIR, OUTPUT_DIR, name);
}

void init_VTK(void)
{
  /* src/vtk.c:62-92: the output directory and an empty collection */
  if (rank != 0) return;
  char dir[FILE_NAME_SIZE];
  snprintf(dir, sizeof(dir), "%s/%s", ROOT_DIR, OUTPUT_DIR);
  struct stat st;
  if (stat(dir, &st) == -1) mkdir(dir, 0700);
  char path[FILE_NAME_SIZE];
  output_path(path, sizeof(path), "out.pvd");
  FILE *f = open_or_die(path, "w");
  fprintf(f, "<VTKFile type=\"Collection\">\n<Collection>\n%s", PVD_FOOTER);
  fclose(f);
}

void vtk_recorder_init(void)
{
  /* src/vtk.c:29-38 */
  if (rec_vtk_dt > 0) init_VTK();
}

/* Append this output's two data sets to out.pvd, in front of the footer. */
static void pvd_append(void)
{
  char path[FILE_NAME_SIZE];
  output_path(path, sizeof(path), "out.pvd");
  FILE *f = fopen(path, "r+");
  if (f == NULL) {
    init_VTK();
    f = open_or_die(path, "r+");
  }
  fseek(f, -(long) strlen(PVD_FOOTER), SEEK_END);
  static const char *const ext[2] = {"pvtr", "pvtp"};
  for (int part = 0; part < 2; part++)
    fprintf(f, "<DataSet timestep=\"%e\" part=\"%d\" file=\"out_%d.%s\"/>\n",
      ttime, part, rec_vtk_stepnum_out, ext[part]);
  fputs(PVD_FOOTER, f);
  fclose(f);
}

/* Interior cells of the Gcc grid in VTK order (x fastest). */
#define FOR_CELLS(G) \
  for (int k = (G)._ks; k <= (G)._ke; k++) \
    for (int j = (G)._js; j <= (G)._je; j++) \
      for (int i = (G)._is; i <= (G)._ie; i++)

static void write_flow_piece(const dom_struct *d, int nprocs_out)
{
  /* src/vtk.c:131-309, the rank's .vtr piece */
  const grid_info *C = &d->Gcc;
  char name[96], path[FILE_NAME_SIZE];
  snprintf(name, sizeof(name), "out_%d_%d_of_%d.vtr", rec_vtk_stepnum_out,
    rank, nprocs_out);
  output_path(path, sizeof(path), name);
  FILE *f = open_or_die(path, "w");
  char ext[128];
  snprintf(ext, sizeof(ext), "%d %d %d %d %d %d", C->is - DOM_BUF, C->ie,
    C->js - DOM_BUF, C->je, C->ks - DOM_BUF, C->ke);
  fprintf(f, "<VTKFile type=\"RectilinearGrid\">\n");
  fprintf(f, "<RectilinearGrid WholeExtent=\"%s\" GhostLevel=\"0\">\n", ext);
  fprintf(f, "<Piece Extent=\"%s\">\n", ext);
  fprintf(f, "<CellData Scalars=\"p phase phase_shell\" Vectors=\"vel\">\n");

  fprintf(f, "<DataArray type=\"Float32\" Name=\"p\">\n");
  FOR_CELLS(*C) fprintf(f, "%lf ", p[GCC_LOC(i, j, k, C->s1b, C->s2b)]);
  fprintf(f, "\n</DataArray>\n");

  /* phase is written as min(phase, 0): -1 in fluid, 0 inside particles */
  fprintf(f, "<DataArray type=\"Int32\" Name=\"phase\">\n");
  FOR_CELLS(*C) {
    int ph = phase[GCC_LOC(i, j, k, C->s1b, C->s2b)];
    fprintf(f, "%d ", ph < 0 ? ph : 0);
  }
  fprintf(f, "\n</DataArray>\n");

  fprintf(f, "<DataArray type=\"Int32\" Name=\"phase_shell\">\n");
  FOR_CELLS(*C)
    fprintf(f, "%d ", phase_shell[GCC_LOC(i, j, k, C->s1b, C->s2b)]);
  fprintf(f, "\n</DataArray>\n");

  const grid_info *X = &d->Gfx, *Y = &d->Gfy, *Z = &d->Gfz;
  fprintf(f, "<DataArray type=\"Float32\" Name=\"vel\"");
  fprintf(f, " NumberOfComponents=\"3\" format=\"ascii\">\n");
  FOR_CELLS(*C) {
    real uc = 0.5 * (u[GFX_LOC(i, j, k, X->s1b, X->s2b)]
                   + u[GFX_LOC(i + 1, j, k, X->s1b, X->s2b)]);
    real vc = 0.5 * (v[GFY_LOC(i, j, k, Y->s1b, Y->s2b)]
                   + v[GFY_LOC(i, j + 1, k, Y->s1b, Y->s2b)]);
    real wc = 0.5 * (w[GFZ_LOC(i, j, k, Z->s1b, Z->s2b)]
                   + w[GFZ_LOC(i, j, k + 1, Z->s1b, Z->s2b)]);
    fprintf(f, "%lf %lf %lf ", uc, vc, wc);
  }
  fprintf(f, "\n</DataArray>\n</CellData>\n<Coordinates>\n");

  /* cell-corner coordinates, index 0 .. n along each axis */
  const int lo[3] = {C->_is, C->_js, C->_ks}, hi[3] = {C->_ie, C->_je, C->_ke};
  const real h[3] = {d->dx, d->dy, d->dz}, s0[3] = {d->xs, d->ys, d->zs};
  static const char axis[3] = {'x', 'y', 'z'};
  for (int a = 0; a < 3; a++) {
    fprintf(f, "<DataArray type=\"Float32\" Name=\"%c\">\n", axis[a]);
    for (int n = lo[a] - DOM_BUF; n <= hi[a]; n++)
      fprintf(f, "%lf ", n * h[a] + s0[a]);
    fprintf(f, "\n</DataArray>\n");
  }
  fprintf(f, "</Coordinates>\n</Piece>\n</RectilinearGrid>\n</VTKFile>\n");
  fclose(f);
}

void dom_out_VTK(void)
{
  /* src/vtk.c:131-309 */
  const int np = 1;
  if (rank == 0) {
    char name[96], path[FILE_NAME_SIZE];
    snprintf(name, sizeof(name), "out_%d.pvtr", rec_vtk_stepnum_out);
    output_path(path, sizeof(path), name);
    FILE *f = open_or_die(path, "w");
    fprintf(f, "<VTKFile type=\"PRectilinearGrid\">\n");
    fprintf(f, "<PRectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\" "
      "GhostLevel=\"0\">\n", DOM.xn, DOM.yn, DOM.zn);
    fprintf(f, "<PCellData Scalars=\"p phase phase_shell\" Vectors=\"vel\">\n");
    fprintf(f, "<PDataArray type=\"Float32\" Name=\"p\"/>\n");
    fprintf(f, "<PDataArray type=\"Int32\" Name=\"phase\"/>\n");
    fprintf(f, "<PDataArray type=\"Int32\" Name=\"phase_shell\"/>\n");
    fprintf(f, "<PDataArray type=\"Float32\" Name=\"vel\""
      " NumberOfComponents=\"3\"/>\n");
    fprintf(f, "</PCellData>\n<PCoordinates>\n");
    for (int a = 0; a < 3; a++)
      fprintf(f, "<PDataArray type=\"Float32\" Name=\"%c\"/>\n", "xyz"[a]);
    fprintf(f, "</PCoordinates>\n");
    for (int l = 0; l < np; l++) {
      const grid_info *C = &dom[l].Gcc;
      fprintf(f, "<Piece Extent=\"%d %d %d %d %d %d\" ", C->is - DOM_BUF,
        C->ie, C->js - DOM_BUF, C->je, C->ks - DOM_BUF, C->ke);
      fprintf(f, "Source=\"out_%d_%d_of_%d.vtr\"/>\n", rec_vtk_stepnum_out,
        l, np);
    }
    fprintf(f, "</PRectilinearGrid>\n</VTKFile>\n");
    fclose(f);
  }
  write_flow_piece(&dom[rank], np);
}

void part_out_VTK(void)
{
  /* src/vtk.c:311-404 */
  const int np = 1;
  char name[96], path[FILE_NAME_SIZE];
  if (rank == 0) {
    snprintf(name, sizeof(name), "out_%d.pvtp", rec_vtk_stepnum_out);
    output_path(path, sizeof(path), name);
    FILE *f = open_or_die(path, "w");
    fprintf(f, "<VTKFile type=\"PPolyData\">\n<PPolyData GhostLevel=\"0\">\n");
    fprintf(f, "<PPointData Scalars=\"n r\" Vectors=\"pvel\">\n");
    fprintf(f, "<PDataArray type=\"Int32\" Name=\"n\"/>\n");
    fprintf(f, "<PDataArray type=\"Float32\" Name=\"r\"/>\n");
    fprintf(f, "<PDataArray type=\"Float32\" Name=\"pvel\" "
      "NumberOfComponents=\"3\"/>\n");
    fprintf(f, "</PPointData>\n<PPoints>\n");
    fprintf(f, "<PDataArray type=\"Float32\" NumberOfComponents=\"3\"/>\n");
    fprintf(f, "</PPoints>\n");
    for (int l = 0; l < np; l++)
      fprintf(f, "<Piece Source=\"out_%d_%d_of_%d.vtp\"/>\n",
        rec_vtk_stepnum_out, l, np);
    fprintf(f, "</PPolyData>\n</VTKFile>\n");
    fclose(f);
  }

  snprintf(name, sizeof(name), "out_%d_%d_of_%d.vtp", rec_vtk_stepnum_out,
    rank, np);
  output_path(path, sizeof(path), name);
  FILE *f = open_or_die(path, "w");
  int n = parts ? nparts : 0;
  fprintf(f, "<VTKFile type=\"PolyData\">\n<PolyData>\n");
  fprintf(f, "<Piece NumberOfPoints=\"%d\" NumberOfVerts=\"0\" ", n);
  fprintf(f, "NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n");
  fprintf(f, "<Points>\n<DataArray type=\"Float32\" NumberOfComponents=\"3\" "
    "format=\"ascii\">\n");
  for (int i = 0; i < n; i++)
    fprintf(f, "%e %e %e ", parts[i].x, parts[i].y, parts[i].z);
  fprintf(f, "\n</DataArray>\n</Points>\n");
  fprintf(f, "<PointData Scalars=\"n r\" Vectors=\"pvel\">\n");
  fprintf(f, "<DataArray type=\"Int32\" Name=\"n\" format=\"ascii\">\n");
  for (int i = 0; i < n; i++) fprintf(f, "%d ", parts[i].N);
  fprintf(f, "\n</DataArray>\n");
  fprintf(f, "<DataArray type=\"Float32\" Name=\"r\" format=\"ascii\">\n");
  for (int i = 0; i < n; i++) fprintf(f, "%e ", parts[i].r);
  fprintf(f, "\n</DataArray>\n");
  fprintf(f, "<DataArray type=\"Float32\" Name=\"pvel\" format=\"ascii\" "
    "NumberOfComponents=\"3\">\n");
  for (int i = 0; i < n; i++)
    fprintf(f, "%e %e %e ", parts[i].u, parts[i].v, parts[i].w);
  fprintf(f, "\n</DataArray>\n</PointData>\n</Piece>\n</PolyData>\n"
    "</VTKFile>\n");
  fclose(f);
}

void out_VTK(void)
{
  /* src/vtk.c:94-129 */
  if (rank == 0) pvd_append();
  dom_out_VTK();
  part_out_VTK();
}

static int vtk_due(void)
{
  return rec_vtk_dt > 0 && (rec_vtk_ttime_out >= rec_vtk_dt || stepnum == 0);
}

void vtk_recorder_write(void)
{
  /* src/vtk.c:40-60 */
  if (!vtk_due()) return;
  out_VTK();
  if (bb_verbose)
    printf("N%d >> Writing ParaView output file %d at t = %e... done.\n",
      rank, rec_vtk_stepnum_out, ttime);
  rec_vtk_stepnum_out++;
  if (stepnum != 0) rec_vtk_ttime_out -= rec_vtk_dt;
}

void recorder_out(void)
{
  /* src/recorder.c:91-155, the VTK part: pull what the writer reads (the
   * flow fields with phase and phase_shell, the scalar, the particles),
   * then write. */
  if (vtk_due()) {
    cuda_dom_pull();
    if (SCALAR >= 1) cuda_scalar_pull();
    if (NPARTS > 0) cuda_part_pull();
  }
  vtk_recorder_write();
}
