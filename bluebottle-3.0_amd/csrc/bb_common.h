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
 * bit-reproducible and can be mirrored exactly on the host.  The domain is
 * cut into tiles of 64 cells in x by 8 rows in y by 8 planes in z; tile
 * (tc, tj, tk) is numbered tc + ntx*(tj + nty*tk) with ntx = ceil(nx/64),
 * nty = ceil(ny/8):
 *   1. for each (x, y) column of a tile, one lane adds the terms of the
 *      tile's planes in increasing z, starting from +0.0;
 *   2. the 64 lane values of a row are combined by the halving tree
 *      v[l] += v[l + off], off = 32, 16, ..., 1 (a cell outside the domain
 *      counts as +0.0);
 *   3. the 8 row values of the tile are combined as
 *      ((w0 + w1) + (w2 + w3)) + ((w4 + w5) + (w6 + w7));
 *   4. the tile partials are combined by 1024 accumulators: accumulator a
 *      adds tiles a, a+1024, ... in increasing order from +0.0, then the
 *      halving tree over the 1024 accumulators (off = 512, ..., 1).
 * Steps 1-3 are the epilogue of the producing kernel (one 512-thread
 * workgroup per tile, one wavefront per row); step 4 runs either as its own
 * launch or redundantly at the start of every workgroup of the consuming
 * kernel, which then all hold the same bits.
 * All arithmetic is IEEE binary64 with FMA contraction disabled
 * (-ffp-contract=off), so one multiply and one add are rounded separately
 * exactly as a plain C compiler does on the host.
 */
#ifndef BB_COMMON_H
#define BB_COMMON_H

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>
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
 * reference's in/jn/kn), ib/jb/kb the points with ghosts; sy/sz are the
 * strides of the device array.  The row stride is ib rounded up to a
 * multiple of 32 elements and every device grid array starts 31 elements
 * into its allocation (bb_grid_alloc), so point x = 1 of every row sits on
 * a 128-byte boundary for 4-byte elements and a 256-byte boundary for
 * 8-byte ones: a wavefront's 64 consecutive elements of a row fill whole
 * cache lines (unaligned rows cost the streaming kernels a third of their
 * bandwidth at 256^3).  The points i >= ib of a row are padding. */
struct Grid {
  int ni, nj, nk;
  int ib, jb, kb;
  int sy, sz;
  long n;                             /* elements of the device array */
};

#define GRID_ROW_ALIGN 32             /* row stride multiple (elements) */
#define GRID_ORIGIN 31                /* leading elements of an array */
void *bb_grid_alloc(long n, size_t esize);
void bb_grid_free(void *p);

__host__ __device__ inline long gidx(const Grid &G, int i, int j, int k)
{
  return (long) i + (long) j * G.sy + (long) k * G.sz;
}

/* Element `i` of a grid array by a 32-bit byte offset from the (uniform)
 * base, so the load takes a scalar base and one vector offset register. */
template <typename V>
__device__ inline V ld32(const V *__restrict__ a, unsigned i)
{
  return *(const V *) ((const char *) a + i * (unsigned) sizeof(V));
}

template <typename V>
__device__ inline void st32(V *__restrict__ a, unsigned i, V v)
{
  *(V *) ((char *) a + i * (unsigned) sizeof(V)) = v;
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
extern uint8_t *_bb_fluid;            /* 1: fluid cell, 0: particle cell */
extern hipStream_t bb_stream;         /* the library's single stream */

/* Scratch used by reductions. */
extern double *_bb_rows;              /* ny*nz row partials */
extern double *_bb_rows2;             /* second row-partial buffer */
extern double *_bb_scalars;           /* small device result slots */
extern double *bb_host_scalars;       /* pinned host mirror of the slots */

/* Row-kernel shape (order-free work such as maxima): four rows per
 * workgroup, one wavefront per row. */
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

/* ----------------------------- tiles ------------------------------------ */
#define TILE_X 64
#define TILE_Y 8
#define TILE_Z 8
#define TILE_THREADS (TILE_X * TILE_Y)

struct Tiles {
  int ntx, nty, ntz;
};

inline Tiles tiles_of(const DevDom &D)
{
  Tiles T;
  T.ntx = (D.nx + TILE_X - 1) / TILE_X;
  T.nty = (D.ny + TILE_Y - 1) / TILE_Y;
  T.ntz = (D.nz + TILE_Z - 1) / TILE_Z;
  return T;
}
inline long n_tiles(const DevDom &D)
{
  Tiles T = tiles_of(D);
  return (long) T.ntx * T.nty * T.ntz;
}
inline dim3 tiles_grid(const DevDom &D)
{
  return dim3((unsigned) n_tiles(D));
}
inline dim3 tiles_block() { return dim3(TILE_THREADS); }

/* XCD-aware tile order.  Workgroups of a launch are handed to the 8 XCDs
 * round robin (workgroup b to XCD b mod 8); each XCD has its own L2.  Tile
 * kernels are launched as a 1-D grid of n_tiles workgroups and workgroup b
 * takes tile tile_of_block(b), so that XCD x owns one contiguous range of
 * tile numbers (whole rows of x-tiles, neighbouring rows and planes) and the
 * lines that neighbouring tiles share stay in one L2.  A bijection on
 * [0, nt) for every nt. */
#define BB_NXCD 8
__device__ inline long tile_of_block(long nt)
{
  long b = blockIdx.x;
  long q = nt / BB_NXCD, rem = nt % BB_NXCD;
  long xcd = b % BB_NXCD, s = b / BB_NXCD;
  return xcd * q + (xcd < rem ? xcd : rem) + s;
}

/* Tile-kernel prologue: the tile of this workgroup, this thread's cell
 * column (x, j), the plane range of the tile and the tile's first x and y
 * cell minus one (xb, jb). */
#define TILE_PROLOGUE                                                         \
  const Grid &C = D.cc;                                                       \
  const int ntx_ = (D.nx + TILE_X - 1) / TILE_X;                              \
  const int nty_ = (D.ny + TILE_Y - 1) / TILE_Y;                              \
  const int ntz_ = (D.nz + TILE_Z - 1) / TILE_Z;                              \
  long tile = tile_of_block((long) ntx_ * nty_ * ntz_);                       \
  int lane = threadIdx.x & (WAVE - 1);                                        \
  /* the row index is the same for all lanes of a wavefront */              \
  int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);                  \
  int xb = (int) (tile % ntx_) * TILE_X;                                      \
  int jb = (int) ((tile / ntx_) % nty_) * TILE_Y;                             \
  int k0 = (int) (tile / ((long) ntx_ * nty_)) * TILE_Z + 1;                  \
  int x = xb + lane + 1;                                                      \
  int j = jb + wv + 1;                                                        \
  int k1 = k0 + TILE_Z - 1 < D.nz ? k0 + TILE_Z - 1 : D.nz;                   \
  bool colok = (x <= D.nx) && (j <= D.ny);                                    \
  (void) tile; (void) colok; (void) xb; (void) jb;

/* Steps 2-3: thread 0 writes the tile partial.  Every thread calls it. */
__device__ inline void tile_partial(double acc, double *__restrict__ gpart,
  long tile)
{
  __shared__ double s_w[TILE_Y];
  int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
  double v = wave_tree_sum(acc);
  if (lane == 0) s_w[w] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    gpart[tile] = ((s_w[0] + s_w[1]) + (s_w[2] + s_w[3])) +
                  ((s_w[4] + s_w[5]) + (s_w[6] + s_w[7]));
}

/* tile_partial without a branch around the global store: every thread
 * forms the same row combination and stores the same value (same order,
 * same bits), so no memory operation depends on the thread index. */
__device__ inline void tile_partial_all(double acc, double *__restrict__ gpart,
  long tile)
{
  __shared__ double s_w2[TILE_Y];
  int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
  double v = wave_tree_sum(acc);
  if (lane == 0) s_w2[w] = v;
  __syncthreads();
  gpart[tile] = ((s_w2[0] + s_w2[1]) + (s_w2[2] + s_w2[3])) +
                ((s_w2[4] + s_w2[5]) + (s_w2[6] + s_w2[7]));
}

/* Step 4 inside a workgroup of NT >= TILE_THREADS (512) threads: thread
 * t < 512 plays accumulators t and t + 512, the others only take part in
 * the barriers.  Returns the total to every thread. */
template <int NT = TILE_THREADS>
__device__ inline double tile_block_sum(const double *__restrict__ part, long n)
{
  static_assert(NT >= TILE_THREADS, "tile_block_sum needs 512 threads");
  __shared__ double s[TILE_THREADS];
  __shared__ double s_tot;
  const long RT = REDUCE_THREADS;
  int t = threadIdx.x;
  if (NT == TILE_THREADS || t < TILE_THREADS) {
    double a = 0.0, b = 0.0;
#pragma unroll 4
    for (long q = t; q < n; q += RT) a += part[q];
#pragma unroll 4
    for (long q = t + TILE_THREADS; q < n; q += RT) b += part[q];
    s[t] = a + b;
  }
  __syncthreads();
  for (int off = TILE_THREADS / 2; off >= WAVE; off >>= 1) {
    if (t < off) s[t] += s[t + off];
    __syncthreads();
  }
  if (t < WAVE) {
    double v = wave_tree_sum(s[t]);
    if (t == 0) s_tot = v;
  }
  __syncthreads();
  return s_tot;
}

/* Host-side launchers implemented in bb_reduce.hip.  bb_reduce_rows runs
 * the final stage alone over `nrows` partials (face sums, tile partials). */
void bb_reduce_rows(const double *rows, long nrows, double *out_slot);
double bb_fetch_scalar(int slot);
void bb_fetch_scalars(int n);

/* Optional per-kernel timing with HIP events on bb_stream (bb_reduce.hip).
 * Ids of the timed kernels: */
enum { KT_SPMV = 0, KT_UPDATE, KT_SEARCH, KT_RHS, KT_USTAR, KT_SPMV32,
       KT_UPDATE32, KT_SEARCH32,
       /* update kernels of paired phi updates: the one that leaves phi
        * alone and the one that adds two terms */
       KT_UPDATE_SKIP, KT_UPDATE_PAIR, KT_UPDATE32_SKIP, KT_UPDATE32_PAIR,
       KT_COUNT };
extern int bb_time_kernels;
void bb_kt_begin(int id);
void bb_kt_end(int id);
/* The event pair of the timed launch in progress (NULL when this launch
 * is not timed).  A timed kernel is launched through BB_TIMED_LAUNCH, which
 * hands the pair to hipExtLaunchKernelGGL: the runtime records them at the
 * kernel's own start and end, so a sample is the kernel's duration without
 * the dispatch gap that events recorded on the stream around it add. */
extern hipEvent_t bb_kt_start, bb_kt_stop;
#define BB_TIMED_LAUNCH(...) BB_TIMED_LAUNCH_(__VA_ARGS__)
#define BB_TIMED_LAUNCH_(K, G, B, S, ST, ...)                                 \
  hipExtLaunchKernelGGL(K, G, B, S, ST, bb_kt_start, bb_kt_stop, 0, __VA_ARGS__)
void bb_kt_collect(void);     /* call after a stream synchronisation */
void bb_kt_collect_upto(long qmax);   /* drops samples tagged > qmax */
extern long bb_kt_q;

/* Ghost refresh (periodic planes), BCs and flags, bb_bc.hip. */
void bb_wrap(double *a, int grid);
void bb_bc_p(double *a);
void bb_phase_fluid(void);
void bb_flags_reset(void);
void bb_flags_external(void);

#endif
