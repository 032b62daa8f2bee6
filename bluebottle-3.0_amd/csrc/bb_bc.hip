/*
 * Domain boundary conditions and the periodic ghost refresh.
 *
 * Ghost-cell semantics follow the reference exactly:
 *   - BC kernels (src/bluebottle_kernel.cu:26-598) touch one face plane each,
 *     over the interior tangential points only, and are applied in the order
 *     W, E, S, N, B, T.  Normal Dirichlet: ghost = 2 bc - a[s+1], a[s] = bc;
 *     tangential Dirichlet: ghost = 8/3 bc - 2 a[s] + 1/3 a[s+1];
 *     Neumann: ghost = a[s].
 *   - The periodic refresh replaces the MPI self-exchange of
 *     src/mpi_comm.c:257-405 with the plane pairs of the pack/unpack kernels
 *     (src/bluebottle_kernel.cu:684-1466): in a cell-centred direction
 *     ghost[0] <- a[e], ghost[e+1] <- a[s]; in the staggered direction of a
 *     face grid ghost[0] <- a[e-1], ghost[e+1] <- a[s+1] (the duplicated
 *     boundary face is not exchanged).  Interior tangential points only.
 * A W/E pair (and S/N, B/T) is one launch: the two sides never touch the
 * same element when a direction has at least two cells.
 */
#include "bb_common.h"

enum { BCK_NONE = 0, BCK_NEUMANN, BCK_DIR_NORMAL, BCK_DIR_TANG };

/* Map (t1, t2) to the base index of a face plane normal to `dir`. */
__device__ inline long plane_base(const Grid &G, int dir, int t1, int t2)
{
  if (dir == 0) return (long) t1 * G.sy + (long) t2 * G.sz;   /* (j, k) */
  if (dir == 1) return (long) t1 + (long) t2 * G.sz;          /* (i, k) */
  return (long) t1 + (long) t2 * G.sy;                        /* (i, j) */
}

__device__ inline void apply_side(double *a, long b, long S, int e, int lo,
  int kind, double bcv)
{
  if (kind == BCK_NONE) return;
  if (lo) {
    long g = b, s0 = b + S, s1 = b + 2 * S;
    if (kind == BCK_NEUMANN) {
      a[g] = a[s0];
    } else if (kind == BCK_DIR_NORMAL) {
      a[g] = 2.*bcv - a[s1];
      a[s0] = bcv;
    } else {
      a[g] = 8./3.*bcv - 2.*a[s0] + 1./3.*a[s1];
    }
  } else {
    long g = b + (long) (e + 1) * S, s0 = b + (long) e * S,
         s1 = b + (long) (e - 1) * S;
    if (kind == BCK_NEUMANN) {
      a[g] = a[s0];
    } else if (kind == BCK_DIR_NORMAL) {
      a[g] = 2.*bcv - a[s1];
      a[s0] = bcv;
    } else {
      a[g] = 8./3.*bcv - 2.*a[s0] + 1./3.*a[s1];
    }
  }
}

__global__ void bc_pair_kernel(double *a, Grid G, int dir, int klo,
  double vlo, int khi, double vhi)
{
  int n1 = (dir == 0) ? G.nj : G.ni;
  int n2 = (dir == 2) ? G.nj : G.nk;
  int t1 = blockIdx.x * blockDim.x + threadIdx.x;
  int t2 = blockIdx.y * blockDim.y + threadIdx.y;
  if (t1 >= n1 || t2 >= n2) return;
  long b = plane_base(G, dir, t1 + 1, t2 + 1);
  long S = (dir == 0) ? 1 : (dir == 1 ? G.sy : G.sz);
  int e = (dir == 0) ? G.ni : (dir == 1 ? G.nj : G.nk);
  apply_side(a, b, S, e, 1, klo, vlo);
  apply_side(a, b, S, e, 0, khi, vhi);
}

/* The periodic planes of every periodic direction in one launch
 * (blockIdx.z picks the direction; `stag` = 1 if the grid is staggered in
 * that direction).  Each direction's copies read and write only points
 * whose other two indices are interior, so the directions are independent
 * and one launch gives what one launch per direction gives. */
struct WrapDirs { int n; int dir[3]; int stag[3]; };

__global__ void wrap_dirs_kernel(double *a, Grid G, WrapDirs W)
{
  const int dir = W.dir[blockIdx.z], stag = W.stag[blockIdx.z];
  int n1 = (dir == 0) ? G.nj : G.ni;
  int n2 = (dir == 2) ? G.nj : G.nk;
  int t1 = blockIdx.x * blockDim.x + threadIdx.x;
  int t2 = blockIdx.y * blockDim.y + threadIdx.y;
  if (t1 >= n1 || t2 >= n2) return;
  long b = plane_base(G, dir, t1 + 1, t2 + 1);
  long S = (dir == 0) ? 1 : (dir == 1 ? G.sy : G.sz);
  int e = (dir == 0) ? G.ni : (dir == 1 ? G.nj : G.nk);
  if (stag) {
    a[b] = a[b + (long) (e - 1) * S];
    a[b + (long) (e + 1) * S] = a[b + 2 * S];
  } else {
    a[b] = a[b + (long) e * S];
    a[b + (long) (e + 1) * S] = a[b + S];
  }
}

static void launch_plane(const Grid &G, int dir, dim3 &grid, dim3 &block)
{
  int n1 = (dir == 0) ? G.nj : G.ni;
  int n2 = (dir == 2) ? G.nj : G.nk;
  block = dim3(64, 4);
  grid = dim3((n1 + 63) / 64, (n2 + 3) / 4);
}

static const Grid &grid_of(int grid)
{
  switch (grid) {
    case GFX: return ddom.fx;
    case GFY: return ddom.fy;
    case GFZ: return ddom.fz;
    default:  return ddom.cc;
  }
}

void bb_wrap(double *a, int grid)
{
  const Grid &G = grid_of(grid);
  int per[3] = {ddom.perx, ddom.pery, ddom.perz};
  WrapDirs W;
  W.n = 0;
  unsigned gx = 0, gy = 0;
  for (int dir = 0; dir < 3; dir++) {
    if (!per[dir]) continue;
    W.dir[W.n] = dir;
    W.stag[W.n] = (grid == GFX && dir == 0) || (grid == GFY && dir == 1) ||
                  (grid == GFZ && dir == 2);
    W.n++;
    dim3 gr, bl;
    launch_plane(G, dir, gr, bl);
    gx = gr.x > gx ? gr.x : gx;
    gy = gr.y > gy ? gr.y : gy;
  }
  if (W.n == 0) return;
  hipLaunchKernelGGL(wrap_dirs_kernel, dim3(gx, gy, (unsigned) W.n),
    dim3(64, 4), 0, bb_stream, a, G, W);
  BB_LAUNCH_CHECK();
}

extern "C" void cuda_exchange(real *array, int grid)
{
  bb_wrap(array, grid);
}

/* The periodic planes of u, v and w in one launch: blockIdx.z picks the
 * (component, direction) pair, each pair copying as wrap_dirs_kernel does
 * (the three arrays are disjoint). */
struct WrapVel {
  int n;
  double *a[9];
  Grid G[9];
  int dir[9], stag[9];
};

__global__ void wrap_vel_kernel(WrapVel W)
{
  const int z = blockIdx.z;
  double *a = W.a[z];
  const Grid &G = W.G[z];
  const int dir = W.dir[z], stag = W.stag[z];
  int n1 = (dir == 0) ? G.nj : G.ni;
  int n2 = (dir == 2) ? G.nj : G.nk;
  int t1 = blockIdx.x * blockDim.x + threadIdx.x;
  int t2 = blockIdx.y * blockDim.y + threadIdx.y;
  if (t1 >= n1 || t2 >= n2) return;
  long b = plane_base(G, dir, t1 + 1, t2 + 1);
  long S = (dir == 0) ? 1 : (dir == 1 ? G.sy : G.sz);
  int e = (dir == 0) ? G.ni : (dir == 1 ? G.nj : G.nk);
  if (stag) {
    a[b] = a[b + (long) (e - 1) * S];
    a[b + (long) (e + 1) * S] = a[b + 2 * S];
  } else {
    a[b] = a[b + (long) e * S];
    a[b + (long) (e + 1) * S] = a[b + S];
  }
}

/* cuda_exchange of the three velocity components (on GFX, GFY, GFZ). */
extern "C" void cuda_exchange_vel(real *u, real *v, real *w)
{
  int per[3] = {ddom.perx, ddom.pery, ddom.perz};
  double *arr[3] = {u, v, w};
  const int grids[3] = {GFX, GFY, GFZ};
  WrapVel W;
  W.n = 0;
  unsigned gx = 0, gy = 0;
  for (int c = 0; c < 3; c++)
    for (int dir = 0; dir < 3; dir++) {
      if (!per[dir]) continue;
      const Grid &G = grid_of(grids[c]);
      W.a[W.n] = arr[c];
      W.G[W.n] = G;
      W.dir[W.n] = dir;
      W.stag[W.n] = c == dir;
      W.n++;
      dim3 gr, bl;
      launch_plane(G, dir, gr, bl);
      gx = gr.x > gx ? gr.x : gx;
      gy = gr.y > gy ? gr.y : gy;
    }
  if (W.n == 0) return;
  hipLaunchKernelGGL(wrap_vel_kernel, dim3(gx, gy, (unsigned) W.n),
    dim3(64, 4), 0, bb_stream, W);
  BB_LAUNCH_CHECK();
}

/* Kind of BC kernel for a velocity component on a face. */
static int vel_kind(int type, int normal)
{
  if (type == DIRICHLET) return normal ? BCK_DIR_NORMAL : BCK_DIR_TANG;
  if (type == NEUMANN) return BCK_NEUMANN;
  return BCK_NONE;                    /* PERIODIC and PRECURSOR */
}

static void bc_pair(double *a, int grid, int dir, int klo, double vlo,
  int khi, double vhi)
{
  if (klo == BCK_NONE && khi == BCK_NONE) return;
  const Grid &G = grid_of(grid);
  dim3 gr, bl;
  launch_plane(G, dir, gr, bl);
  hipLaunchKernelGGL(bc_pair_kernel, gr, bl, 0, bb_stream, a, G, dir, klo, vlo,
    khi, vhi);
  BB_LAUNCH_CHECK();
}

static int pkind(int t) { return t == NEUMANN ? BCK_NEUMANN : BCK_NONE; }

void bb_bc_p(double *a)
{
  bc_pair(a, GCC, 0, pkind(bc.pW), 0., pkind(bc.pE), 0.);
  bc_pair(a, GCC, 1, pkind(bc.pS), 0., pkind(bc.pN), 0.);
  bc_pair(a, GCC, 2, pkind(bc.pB), 0., pkind(bc.pT), 0.);
}

static void bc_u(double *a)
{
  bc_pair(a, GFX, 0, vel_kind(bc.uW, 1), bc.uWD, vel_kind(bc.uE, 1), bc.uED);
  bc_pair(a, GFX, 1, vel_kind(bc.uS, 0), bc.uSD, vel_kind(bc.uN, 0), bc.uND);
  bc_pair(a, GFX, 2, vel_kind(bc.uB, 0), bc.uBD, vel_kind(bc.uT, 0), bc.uTD);
}

static void bc_v(double *a)
{
  bc_pair(a, GFY, 0, vel_kind(bc.vW, 0), bc.vWD, vel_kind(bc.vE, 0), bc.vED);
  bc_pair(a, GFY, 1, vel_kind(bc.vS, 1), bc.vSD, vel_kind(bc.vN, 1), bc.vND);
  bc_pair(a, GFY, 2, vel_kind(bc.vB, 0), bc.vBD, vel_kind(bc.vT, 0), bc.vTD);
}

static void bc_w(double *a)
{
  bc_pair(a, GFZ, 0, vel_kind(bc.wW, 0), bc.wWD, vel_kind(bc.wE, 0), bc.wED);
  bc_pair(a, GFZ, 1, vel_kind(bc.wS, 0), bc.wSD, vel_kind(bc.wN, 0), bc.wND);
  bc_pair(a, GFZ, 2, vel_kind(bc.wB, 1), bc.wBD, vel_kind(bc.wT, 1), bc.wTD);
}

extern "C" void cuda_dom_BC(void)
{
  /* src/cuda_bluebottle.cu:1218-1438: p, u, v, w on every external face. */
  bb_bc_p(_p);
  bc_u(_u);
  bc_v(_v);
  bc_w(_w);
}

extern "C" void cuda_dom_BC_star(void)
{
  /* src/cuda_bluebottle.cu:2111-2311 */
  bc_u(_u_star);
  bc_v(_v_star);
  bc_w(_w_star);
}

extern "C" void cuda_dom_BC_p(real *array)
{
  /* src/cuda_bluebottle.cu:2538-2589 */
  bb_bc_p(array);
}

/* ---- flags and phase without particles (src/cuda_particle.cu:1516-1646,
 *      src/particle_kernel.cu:79-576, src/particle.c:556-598) ---- */
__global__ void fill_int_kernel(int *a, long n, int val)
{
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = val;
}

__global__ void flag_external_kernel(int *flag, Grid G, int dir, int loc)
{
  /* Whole ghosted plane, as flag_external_{u,v,w} do. */
  int m1 = (dir == 0) ? G.nj + 2 : G.ni + 2;
  int m2 = (dir == 2) ? G.nj + 2 : G.nk + 2;
  int t1 = blockIdx.x * blockDim.x + threadIdx.x;
  int t2 = blockIdx.y * blockDim.y + threadIdx.y;
  if (t1 >= m1 || t2 >= m2) return;
  long b = plane_base(G, dir, t1, t2);
  long S = (dir == 0) ? 1 : (dir == 1 ? G.sy : G.sz);
  flag[b + (long) loc * S] = 0;
}

static void fill_int(int *a, long n, int val)
{
  hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned) ((n + 255) / 256)),
    dim3(256), 0, bb_stream, a, n, val);
  BB_LAUNCH_CHECK();
}

static void flag_external(int *flag, const Grid &G, int dir, int loc)
{
  int m1 = (dir == 0) ? G.nj + 2 : G.ni + 2;
  int m2 = (dir == 2) ? G.nj + 2 : G.nk + 2;
  hipLaunchKernelGGL(flag_external_kernel, dim3((m1 + 63) / 64, (m2 + 3) / 4),
    dim3(64, 4), 0, bb_stream, flag, G, dir, loc);
  BB_LAUNCH_CHECK();
}

void bb_flags_external(void)
{
  if (bc.pW != PERIODIC && bc.pE != PERIODIC) {
    flag_external(_flag_u, ddom.fx, 0, 1);
    flag_external(_flag_u, ddom.fx, 0, ddom.fx.ni);
  }
  if (bc.pS != PERIODIC && bc.pN != PERIODIC) {
    flag_external(_flag_v, ddom.fy, 1, 1);
    flag_external(_flag_v, ddom.fy, 1, ddom.fy.nj);
  }
  if (bc.pB != PERIODIC && bc.pT != PERIODIC) {
    flag_external(_flag_w, ddom.fz, 2, 1);
    flag_external(_flag_w, ddom.fz, 2, ddom.fz.nk);
  }
}

void bb_flags_reset(void)
{
  fill_int(_flag_u, ddom.fx.n, 1);
  fill_int(_flag_v, ddom.fy.n, 1);
  fill_int(_flag_w, ddom.fz.n, 1);
}

void bb_phase_fluid(void)
{
  fill_int(_phase, ddom.cc.n, -1);
  fill_int(_phase_shell, ddom.cc.n, -1);
}

/* Particle cages are built in bb_particle.hip when particles exist. */
void bb_build_particle_cages(void);
void bb_particle_face_lists(void);

extern "C" void cuda_build_cages(void)
{
  bb_flags_reset();
  if (NPARTS > 0) bb_build_particle_cages();
  bb_flags_external();
  if (NPARTS > 0) bb_particle_face_lists();
}
