/*
 * Internal definitions shared by the gfx950 kernels of libbluebottle.so.
 *
 * Device layout.  Every grid (Gcc, Gfx, Gfy, Gfz) is stored x-fastest with
 * one ghost layer: element (i, j, k), 0 <= i < inb, ... sits at
 * i + j*inb + k*inb*jnb.  A 64-lane wavefront therefore reads 64 consecutive
 * doubles along x for every field, including the face-centred u, v and w
 * (the reference stores Gfx y-fastest and Gfy z-fastest).
 *
 * Canonical reduction order.  Every sum over interior cells (the PCG dot
 * products, the pressure mean) is formed the same way so that the result is
 * bit-reproducible and can be mirrored exactly on the host:
 *   1. row partial: for each row (j, k) lane l of a 64-lane wavefront adds
 *      the terms at x-offsets l, l+64, l+128, ... in increasing order,
 *      starting from +0.0; then the 64 lane values are combined by the
 *      halving tree  v[l] += v[l + off]  for off = 32, 16, ..., 1; the row
 *      partial is v[0];
 *   2. the row partials, numbered r = (j-1) + (k-1)*ny, are combined by one
 *      256-thread workgroup: thread t adds rows t, t+256, ... in increasing
 *      order from +0.0, then the halving tree over 256 threads (off = 128,
 *      ..., 1) leaves the total in thread 0.
 * All arithmetic is IEEE binary64 with FMA contraction disabled
 * (-ffp-contract=off), so one multiply and one add are rounded separately
 * exactly as a plain C compiler does on the host.
 */
#ifndef BB_COMMON_H
#define BB_COMMON_H

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include "bluebottle.h"

#define BB_CHECK(call)                                                        \
  do {                                                                        \
    hipError_t err_ = (call);                                                 \
    if (err_ != hipSuccess) {                                                 \
      fprintf(stderr, "HIP error %s at %s:%d: %s\n", hipGetErrorName(err_),    \
        __FILE__, __LINE__, hipGetErrorString(err_));                         \
      exit(EXIT_FAILURE);                                                     \
    }                                                                         \
  } while (0)

#define BB_LAUNCH_CHECK() BB_CHECK(hipGetLastError())

#define WAVE 64
#define ROWS_PER_BLOCK 4              /* one wavefront per row, 4 per block */
#define REDUCE_THREADS 256

/* A ghosted x-fastest grid.  ni/nj/nk count the interior points (the
 * reference's in/jn/kn); sy/sz are the strides of the ghosted array. */
struct Grid {
  int ni, nj, nk;
  int sy, sz;
  long n;                             /* elements including ghosts */
};

__host__ __device__ inline long gidx(const Grid &G, int i, int j, int k)
{
  return (long) i + (long) j * G.sy + (long) k * G.sz;
}

/* Everything a kernel needs to know about the domain, passed by value. */
struct DevDom {
  Grid cc, fx, fy, fz;
  int nx, ny, nz;                     /* cells per direction */
  double dx, dy, dz;
  /* 1 if the direction is periodic (x, y, z). */
  int perx, pery, perz;
};

extern DevDom ddom;                   /* host copy, filled at malloc time */
extern hipStream_t bb_stream;         /* the library's single stream */

/* Scratch used by reductions. */
extern double *_bb_rows;              /* ny*nz row partials */
extern double *_bb_rows2;             /* second row-partial buffer */
extern double *_bb_scalars;           /* small device result slots */
extern double *bb_host_scalars;       /* pinned host mirror of the slots */

/* Interior-row kernel shape: one wavefront per (j, k) row. */
inline dim3 rows_grid(int ny, int nz)
{
  long rows = (long) ny * nz;
  return dim3((unsigned) ((rows + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK));
}
inline dim3 rows_block() { return dim3(WAVE * ROWS_PER_BLOCK); }

/* Row partial tree over one wavefront (step 1 of the canonical order).  The
 * value returned in lane 0 is the row partial. */
__device__ inline double wave_tree_sum(double v)
{
  for (int off = 32; off >= 1; off >>= 1)
    v += __shfl_down(v, off, WAVE);
  return v;
}

/* Host-side launchers implemented in bb_reduce.hip. */
void bb_reduce_rows(const double *rows, long nrows, double *out_slot);
double bb_fetch_scalar(int slot);
void bb_fetch_scalars(int n);

/* Optional per-kernel timing with HIP events on bb_stream (bb_reduce.hip).
 * Ids of the timed kernels: */
enum { KT_SPMV = 0, KT_UPDATE, KT_SEARCH, KT_RHS, KT_USTAR, KT_SPMV32,
       KT_UPDATE32, KT_SEARCH32, KT_COUNT };
extern int bb_time_kernels;
void bb_kt_begin(int id);
void bb_kt_end(int id);
void bb_kt_collect(void);     /* call after a stream synchronisation */

/* Ghost refresh (periodic planes) and BCs, bb_bc.hip. */
void bb_wrap(double *a, int grid);
void bb_bc_p(double *a);
void bb_phase_fluid(void);
void bb_flags_reset(void);
void bb_flags_external(void);

#endif
