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
 *   2. group partial: the rows of one plane k are taken four at a time,
 *      j = 4g+1 .. 4g+4 (a row past ny counts as +0.0), and combined as
 *      (r0 + r1) + (r2 + r3); groups are numbered g + (k-1)*ceil(ny/4);
 *   3. the group partials are combined by one 1024-thread workgroup: thread
 *      t adds groups t, t+1024, ... in increasing order from +0.0, then the
 *      halving tree over 1024 threads (off = 512, ..., 1) leaves the total
 *      in thread 0.  Step 3 is its own small launch: the kernel boundary is
 *      what makes every workgroup's partial visible across the XCDs (a
 *      device-scope fence per workgroup would write back L2 each time).
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
#define REDUCE_THREADS 1024

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

/* Interior-row kernel shape: one workgroup per row group (four rows of one
 * plane), one wavefront per row. */
__host__ __device__ inline long n_groups(int ny, int nz)
{
  return (long) ((ny + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK) * nz;
}
inline dim3 rows_grid(int ny, int nz) { return dim3((unsigned) n_groups(ny, nz)); }
inline dim3 rows_block() { return dim3(WAVE * ROWS_PER_BLOCK); }

/* Row partial tree over one wavefront (step 1 of the canonical order).  The
 * value returned in lane 0 is the row partial. */
__device__ inline double wave_tree_sum(double v)
{
  for (int off = 32; off >= 1; off >>= 1)
    v += __shfl_down(v, off, WAVE);
  return v;
}

/* Step 3 over `n` partials by one REDUCE_THREADS workgroup; total in thread
 * 0.  Loads are issued eight at a time and added in index order. */
__device__ inline double block_canonical_sum(const double *__restrict__ part,
  long n)
{
  __shared__ double s[REDUCE_THREADS];
  const int RT = REDUCE_THREADS;
  int t = threadIdx.x;
  double acc = 0.0;
  long q = t;
  for (; q + 7L * RT < n; q += 8L * RT) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = part[q + (long) u * RT];
#pragma unroll
    for (int u = 0; u < 8; u++) acc += v[u];
  }
  for (; q < n; q += RT) acc += part[q];
  s[t] = acc;
  __syncthreads();
  for (int off = RT / 2; off >= WAVE; off >>= 1) {
    if (t < off) s[t] += s[t + off];
    __syncthreads();
  }
  double v = 0.;
  if (t < WAVE) v = wave_tree_sum(s[t]);
  return v;
}

/* Steps 1-2 for a 4-wavefront workgroup that owns one row group: thread 0
 * writes the group partial.  All threads must call it with their lane's row
 * accumulator (+0.0 for a row past the end). */
__device__ inline void group_partial(double rowacc, double *__restrict__ gpart)
{
  __shared__ double s_rows[ROWS_PER_BLOCK];
  int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
  double v = wave_tree_sum(rowacc);
  if (lane == 0) s_rows[w] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    gpart[blockIdx.x] = (s_rows[0] + s_rows[1]) + (s_rows[2] + s_rows[3]);
}

/* Group-kernel prologue: this workgroup's plane k and row group, this
 * wavefront's row j (rowok false past ny). */
#define GROUP_PROLOGUE                                                        \
  const Grid &C = D.cc;                                                       \
  int lane = threadIdx.x & (WAVE - 1);                                        \
  int ngy_ = (D.ny + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK;                    \
  int k = (int) (blockIdx.x / ngy_) + 1;                                      \
  int j = (int) (blockIdx.x % ngy_) * ROWS_PER_BLOCK + (threadIdx.x >> 6) + 1; \
  bool rowok = (j <= D.ny);                                                   \
  int nxr = rowok ? D.nx : 0;                                                 \
  long base = gidx(C, 0, rowok ? j : D.ny, k);

/* Host-side launchers implemented in bb_reduce.hip.  bb_reduce_rows runs
 * step 3 alone over `nrows` partials (face sums and group partials). */
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
