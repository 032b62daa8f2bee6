/*
 * Pressure-Poisson solve: Jacobi-preconditioned CG on -A phi = -b, the
 * reference's algorithm (src/cuda_solver.cu:32-36, 574-761,
 * src/solver_kernel.cu:26-927) with its recurrences, stopping test
 * (r, M^-1 r) <= pp_residual^2 (b, b), and true-residual refresh every 50
 * iterations.  Every rounded value is the reference's; what changes is where
 * values live and how many passes over HBM an iteration takes.
 *
 * One iteration q is two launches over 64 x 8 x 8 cell tiles (one 512-thread
 * workgroup per tile, one wavefront per row, each lane marching z):
 *   search_spmv   [start: (r, z) of the last update -> stopping test, beta]
 *                 p = z + beta p_old, Ap = (-A) p   + tile partials (p, Ap)
 *   update        [start: (p, Ap) -> alpha]
 *                 phi += alpha p, r -= alpha Ap     + tile partials (r, z)
 * The final reduction stage runs redundantly at the start of every
 * workgroup of the consuming launch (all get the same bits), so alpha, beta
 * and the stopping test cost no launch of their own.  z = M^-1 r is never
 * stored; M^-1 comes from a 27-entry table indexed by the number of
 * non-boundary faces per direction; search_spmv forms p_new for the tile,
 * its halo rows, columns and planes in LDS (periodic neighbours by wrapped
 * index), so the search direction needs no ghost refresh.  HBM traffic per
 * iteration: 80 bytes per cell (the reference's kernel sequence moves 136).
 *
 * The 7-point weights flag^2 are 0 exactly on faces of a non-periodic
 * external boundary and 1 elsewhere (src/particle_kernel.cu:481-576 writes
 * only +-1 or, at external boundaries, 0), so they are computed from the
 * cell index.
 *
 * The host enqueues iterations in batches and reads the device-resident CG
 * state once per batch; every launch returns at once when the state says
 * done, so the iterates stop exactly where the reference's host loop breaks.
 */
#include <math.h>
#include <sys/time.h>

#include <stdint.h>

#include <limits.h>
#include <type_traits>

#include "bb_common.h"

/* Stencil weights 1/dx^2, ... in the arithmetic type V of the vectors. */
template <typename V> struct CoefT {
  V idx2, idy2, idz2;
};
typedef CoefT<double> Coef;

static Coef coef(void)
{
  Coef K;
  K.idx2 = 1./(dom[0].dx * dom[0].dx);
  K.idy2 = 1./(dom[0].dy * dom[0].dy);
  K.idz2 = 1./(dom[0].dz * dom[0].dz);
  return K;
}

static CoefT<float> coef32(const Coef &K)
{
  CoefT<float> F;
  F.idx2 = (float) K.idx2;
  F.idy2 = (float) K.idy2;
  F.idz2 = (float) K.idz2;
  return F;
}

/* Jacobi diagonal by the count of non-boundary faces per direction (0..2):
 * invM = -1/M, M = -idx2 wx - idy2 wy - idz2 wz (src/solver_kernel.cu:66-77),
 * evaluated once on the host with the same operand order. */
template <typename V> struct InvMT {
  V v[27];
};
typedef InvMT<double> InvMTab;

static InvMTab invm_table(const Coef &K)
{
  InvMTab T;
  for (int wx = 0; wx < 3; wx++)
    for (int wy = 0; wy < 3; wy++)
      for (int wz = 0; wz < 3; wz++) {
        double M = -K.idx2*wx - K.idy2*wy - K.idz2*wz;
        T.v[wx + 3 * wy + 9 * wz] = -1./M;
      }
  return T;
}

static InvMT<float> invm_table32(const InvMTab &T)
{
  InvMT<float> F;
  for (int i = 0; i < 27; i++) F.v[i] = (float) T.v[i];
  return F;
}

/* Non-boundary face count of cell index `a` in a direction with n cells. */
__device__ inline int wcount(int a, int n, int per)
{
  if (per) return 2;
  return (a == 1 ? 0 : 1) + (a == n ? 0 : 1);
}

template <typename V>
__device__ inline V invm_at(const InvMT<V> &T, const DevDom &D, int x, int j,
  int k)
{
  return T.v[wcount(x, D.nx, D.perx) + 3 * wcount(j, D.ny, D.pery) +
             9 * wcount(k, D.nz, D.perz)];
}

/* Particle cells take no part in the solve: their right-hand side is zero
 * (src/particle_kernel.cu:1748-1753) and their rows of the particle SpMV
 * (src/solver_kernel.cu:652-712) give exactly zero while the search
 * direction is zero there.  Every kernel that writes the residual (CG
 * init, the periodic recompute, the update, the mixed form's corrections)
 * writes 0 in particle cells (`fluid` = 0) and the Jacobi factor is 0
 * there, so z = invM r and the search direction stay zero in them: the
 * fluid rows see the same neighbours and every dot product the same terms.
 * The z-column search kernels therefore need no mask.  fluid == NULL: no
 * particles. */
__device__ inline bool solid(const uint8_t *__restrict__ fluid, long c)
{
  return fluid != NULL && fluid[c] == 0;
}

/* Neighbour index with periodic wrap (ghost index in a wall direction). */
__device__ inline int wrap_lo(int a, int n, int per) { return (per && a == 1) ? n : a - 1; }
__device__ inline int wrap_hi(int a, int n, int per) { return (per && a == n) ? 1 : a + 1; }

/* Face weight flag^2 is 0 on a non-periodic boundary face and 1 elsewhere;
 * these say whether the low/high face of cell index a is such a wall. */
__device__ inline bool wall_lo(int a, int per) { return !per && a == 1; }
__device__ inline bool wall_hi(int a, int n, int per) { return !per && a == n; }

/* w * d for a face weight w in {0, 1}: d itself (1 * d is exact) or 0 * d
 * (which keeps the sign of zero and NaN of the product). */
template <typename V>
__device__ inline V wmul(bool wall, V d) { return wall ? (V) 0 * d : d; }

/* The device-resident CG state.  rzs[q & 1] holds (r, z) as it stands at
 * the start of iteration q; launches of iteration q read rzs[(q-1) & 1] and
 * write rzs[q & 1], never the slot they read. */
struct CGState {
  double sp_rhs, tol2, rz1;
  double rzs[2];
  double rz_ref;      /* mixed form: (r, z) at the last FP64 correction */
  double relup;       /* mixed form: inner reduction of (r, z) per sweep */
  int q, done, nan_flag, inner_done;
  double alpha;       /* alpha of the last update that left phi alone */
};

/* ------------------------------ Jacobi ---------------------------------- */
__global__ void __launch_bounds__(TILE_THREADS)
jacobi_kernel(DevDom D, InvMTab T, double *__restrict__ invM)
{
  /* src/solver_kernel.cu:26-87: the stored preconditioner of the C ABI. */
  TILE_PROLOGUE
  if (!colok) return;
  for (int k = k0; k <= k1; k++)
    invM[gidx(C, x, j, k)] = invm_at(T, D, x, j, k);
}

extern "C" void cuda_PP_init_jacobi_preconditioner(void)
{
  hipLaunchKernelGGL(jacobi_kernel, tiles_grid(ddom), tiles_block(), 0,
    bb_stream, ddom, invm_table(coef()), _invM);
  BB_LAUNCH_CHECK();
}

/* ------------------------------- RHS ------------------------------------ */
/* rhs = -(rho/dt) div u* (src/solver_kernel.cu:89-176) + tile partials of
 * (b, b). */
__global__ void __launch_bounds__(TILE_THREADS)
rhs_kernel(DevDom D, double idx, double idy, double idz, double rho_idt,
  const double *__restrict__ us, const double *__restrict__ vs,
  const double *__restrict__ ws, const uint8_t *__restrict__ fluid,
  double *__restrict__ rhs, double *__restrict__ gpart)
{
  TILE_PROLOGUE
  double acc = 0.0;
  if (colok) {
    for (int k = k0; k <= k1; k++) {
      long cu = gidx(D.fx, x, j, k), cv = gidx(D.fy, x, j, k),
           cw = gidx(D.fz, x, j, k);
      double s = (us[cu + 1] - us[cu]) * idx;
      s += (vs[cv + D.fy.sy] - vs[cv]) * idy;
      s += (ws[cw + D.fz.sz] - ws[cw]) * idz;
      s *= rho_idt;
      double b = -s;
      if (solid(fluid, gidx(C, x, j, k))) b = 0.;
      rhs[gidx(C, x, j, k)] = b;
      acc += b * b;
    }
  }
  tile_partial(acc, gpart, tile);
}

/* ------------------------------ CG init --------------------------------- */
/* r = b, phi = 0 (src/solver_kernel.cu:258-283) + tile partials of (r, z);
 * p = z is formed by the first search_spmv pass. */
__global__ void __launch_bounds__(TILE_THREADS)
cg_init_kernel(DevDom D, InvMTab T, const double *__restrict__ rhs,
  const uint8_t *__restrict__ fluid, double *__restrict__ r,
  double *__restrict__ phi, double *__restrict__ gpart)
{
  TILE_PROLOGUE
  double acc = 0.0;
  if (colok) {
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      double t = rhs[c];
      double t2 = t * (solid(fluid, c) ? 0. : invm_at(T, D, x, j, k));
      r[c] = solid(fluid, c) ? 0. : t;
      phi[c] = 0.;
      acc += t * t2;
    }
  }
  tile_partial(acc, gpart, tile);
}

/* (b, b) and (r, z) -> a fresh CG state. */
__global__ void __launch_bounds__(TILE_THREADS)
cg_state_init_kernel(const double *__restrict__ bpart,
  const double *__restrict__ rpart, long n, CGState *st, double tol2,
  double relup)
{
  double sb = tile_block_sum(bpart, n);
  double sr = tile_block_sum(rpart, n);
  if (threadIdx.x == 0) {
    st->sp_rhs = sb;
    st->tol2 = tol2;
    st->rz1 = sr;
    st->rzs[0] = 0.;
    st->rzs[1] = sr;
    st->rz_ref = sr;
    st->relup = relup;
    st->q = 0;
    st->done = 0;
    st->nan_flag = 0;
    st->inner_done = 0;
    st->alpha = 0.;
  }
}

/* ------------------------ stopping test and beta ------------------------- */
/* (src/cuda_solver.cu:229-267) run by every workgroup of a launch on the
 * same bits; only workgroup 0 writes.  Returns false when the solve is over
 * (converged or NaN), else sets *beta.  In the mixed form the test ends the
 * inner FP32 sweep instead (inner_done), either at the target or once (r, z)
 * has fallen by the factor relup since the last FP64 correction; the host
 * then forms the FP64 residual and decides. */
template <bool MIXED, int NT = TILE_THREADS>
__device__ inline bool cg_continue(CGState *st, const double *__restrict__ rpart,
  long n, int q, double *beta)
{
  double s = tile_block_sum<NT>(rpart, n);
  bool first = blockIdx.x == 0 && threadIdx.x == 0;
  bool stop = s <= st->tol2*st->sp_rhs || isnan(s);
  if (MIXED && !stop) stop = s <= st->relup*st->rz_ref;
  if (stop) {
    if (first) {
      st->rz1 = s;
      st->q = q - 1;
      st->nan_flag = isnan(s) ? 1 : 0;
      if (MIXED) st->inner_done = 1;
      else st->done = 1;
    }
    return false;
  }
  *beta = s/st->rzs[(q - 1) & 1];
  if (first) {
    st->rzs[q & 1] = s;
    st->rz1 = s;
  }
  return true;
}

/* ------------------------------ search_spmv ------------------------------ */
/* The vector kernels are templates on the storage and arithmetic type V of
 * the CG vectors: double for the plain form, float for the inner sweeps of
 * the mixed form.  Dot-product terms are always formed and summed in FP64
 * (for V = double the code below is the plain form's arithmetic). */

/* p_new = z + beta p_old with z = r invM (src/solver_kernel.cu:906-927). */
template <bool FIRST, typename V>
__device__ inline V p_new(const V *__restrict__ r, const V *__restrict__ pold,
  long c, V invm, V beta)
{
  V z = r[c]*invm;
  if (FIRST) return z;
  return z + beta*pold[c];
}

#define LX (TILE_X + 2)
#define LY (TILE_Y + 2)
#define LPLANE (LX * LY)

/* The LDS plane of the search kernel holds p_new over the tile's rows and
 * columns plus a one-cell halo, LX x LY entries; thread t fills entries t
 * and t + TILE_THREADS.  A wall ghost holds 0 (the never-written ghost of the
 * search direction); a periodic ghost holds p_new of the wrapped cell. */
struct PlaneEntry {
  int cxy;            /* x + y*sy of the source column, -1: the entry is 0
                         (wall ghost or outside the domain) */
  int wxy;            /* invM table index, x and y parts */
};

__device__ inline PlaneEntry plane_entry(const DevDom &D, int xb, int jb,
  int it)
{
  PlaneEntry E;
  int rr = it / LX, xx = it - rr * LX;
  int jj = jb + rr, xg = xb + xx;
  int js = jj, xs = xg;
  if (jj < 1) js = D.pery ? D.ny : 0;
  else if (jj > D.ny) js = (D.pery && jj == D.ny + 1) ? 1 : 0;
  if (xg < 1) xs = D.perx ? D.nx : 0;
  else if (xg > D.nx) xs = (D.perx && xg == D.nx + 1) ? 1 : 0;
  bool ok = (it < LPLANE) && js && xs;
  E.cxy = ok ? xs + js * D.cc.sy : -1;
  E.wxy = ok ? wcount(xs, D.nx, D.perx) + 3 * wcount(js, D.ny, D.pery) : 0;
  return E;
}

/* Source plane of LDS plane kk (0 = a wall ghost plane). */
__device__ inline int plane_src(const DevDom &D, int kk)
{
  if (kk < 1) return D.perz ? D.nz : 0;
  if (kk > D.nz) return D.perz ? 1 : 0;
  return kk;
}

/* Loads with the non-temporal hint: p_old in the FP64 search kernel and Ap
 * in the update kernel are read once per iteration and not again before
 * they are rewritten (measured at 256^3: FP64 search 150 -> 146 us, update
 * 150 -> 145 us; the hint on 4-byte p_old loads slowed the FP32 search from
 * 70 to 87 us, so those keep the ordinary load). */
template <typename V>
__device__ inline V ld32_nt(const V *__restrict__ a, unsigned i)
{
  return __builtin_nontemporal_load(
    (const V *) ((const char *) a + i * (unsigned) sizeof(V)));
}

template <typename V>
__device__ inline V ld32_pold(const V *__restrict__ a, unsigned i)
{
  return sizeof(V) == 8 ? ld32_nt(a, i) : ld32(a, i);
}

/* The same store with the non-temporal hint.  The search kernel stores
 * p_new and Ap this way: with ordinary stores part of their write-back
 * falls into the update kernel that follows (measured at 256^3: search
 * 137 us + update 177 us with ordinary stores, 150 + 150 us with these). */
template <typename V>
__device__ inline void st32_nt(V *__restrict__ a, unsigned i, V v)
{
  __builtin_nontemporal_store(v, (V *) ((char *) a + i * (unsigned) sizeof(V)));
}

/* The same loads without a branch: every thread issues all of them, an
 * entry or plane that holds zeros reading a valid dummy cell instead (the
 * residual is already zero in particle cells).  With no load inside a
 * branch, and none waiting on loaded data, the compiler can keep the loads
 * of the next planes in flight across the loop; a conditional load makes
 * it wait for every outstanding load. */
template <typename V> struct PlaneRawM {
  V r0, p0, r1, p1;
};

template <bool FIRST, typename V, bool NTP = true>
__device__ inline PlaneRawM<V> plane_load_nb(const DevDom &D,
  const PlaneEntry &E0, const PlaneEntry &E1, const V *__restrict__ r,
  const V *__restrict__ pold, int kk)
{
  PlaneRawM<V> R;
  int ks = plane_src(D, kk);
  unsigned off = (unsigned) (ks ? ks : 1) * (unsigned) D.cc.sz;
  unsigned c0 = (unsigned) (E0.cxy >= 0 ? E0.cxy : 0) + off;
  unsigned c1 = (unsigned) (E1.cxy >= 0 ? E1.cxy : 0) + off;
  R.r0 = ld32(r, c0);
  R.r1 = ld32(r, c1);
  R.p0 = FIRST ? (V) 0 : (NTP ? ld32_pold(pold, c0) : ld32(pold, c0));
  R.p1 = FIRST ? (V) 0 : (NTP ? ld32_pold(pold, c1) : ld32(pold, c1));
  return R;
}

template <bool FIRST, typename V>
__device__ inline void plane_store_nb(const DevDom &D, const V *tab,
  const PlaneEntry &E0, const PlaneEntry &E1, const PlaneRawM<V> &R, V beta,
  V *__restrict__ buf, int kk)
{
  int ks = plane_src(D, kk);
  int wz = 9 * wcount(ks, D.nz, D.perz);
  V z0 = R.r0*tab[E0.wxy + wz], z1 = R.r1*tab[E1.wxy + wz];
  V v0 = FIRST ? z0 : z0 + beta*R.p0;
  V v1 = FIRST ? z1 : z1 + beta*R.p1;
  v0 = (ks && E0.cxy >= 0) ? v0 : (V) 0;
  v1 = (ks && E1.cxy >= 0) ? v1 : (V) 0;
  buf[threadIdx.x] = v0;
  if (threadIdx.x + TILE_THREADS < LPLANE) buf[threadIdx.x + TILE_THREADS] = v1;
}


/* (-A) p from the centre value, four in-plane neighbours and the two
 * z neighbours (src/solver_kernel.cu:715-836); the w* say which faces are
 * walls (weight 0). */
template <typename V>
__device__ inline V neg_lap7(const CoefT<V> &K, V pC, V pW, V pE, V pS, V pN,
  V pB, V pT, bool wW, bool wE, bool wS, bool wN, bool wB, bool wT)
{
  return -K.idx2 * (wmul(wE, pE - pC) - wmul(wW, pC - pW))
         -K.idy2 * (wmul(wN, pN - pC) - wmul(wS, pC - pS))
         -K.idz2 * (wmul(wT, pT - pC) - wmul(wB, pC - pB));
}

/* The same pass over a column of `tzb` tiles stacked in z: one workgroup
 * streams all their planes through the LDS ring, so the pipeline start
 * (ring fill, the reduction that yields beta) and the two halo planes are
 * paid once per column instead of once per tile, and each tile's partial
 * is written exactly as above (same per-column z order, same tree).  The
 * ring of raw planes rotates through DEPTH register sets: set d holds
 * plane k + 1 + d at step k (set 0 is stored into LDS at the top of the
 * step and reloaded with plane k + 1 + DEPTH), so DEPTH planes of loads
 * are in flight behind the compute; the LDS planes rotate through three
 * buffers with one barrier per plane. */
template <bool FIRST, typename V, bool MIXED, int DEPTH, bool NTP,
          bool NTS = true>
__global__ void __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(8, 8)))
search_spmv_group_kernel(DevDom D, CoefT<V> K, InvMT<V> T, CGState *st,
  int q, const double *__restrict__ rpart, long ntiles, int tzb,
  const V *__restrict__ r, const V *__restrict__ pold,
  V *__restrict__ pnew, V *__restrict__ Ap, double *__restrict__ gpart)
{
  __shared__ V lds[3 * LPLANE];
  __shared__ V s_invm[27];
  if (st->done | st->inner_done) return;
  if (threadIdx.x < 27) s_invm[threadIdx.x] = T.v[threadIdx.x];
  const Grid &C = D.cc;
  const int ntx_ = (D.nx + TILE_X - 1) / TILE_X;
  const int nty_ = (D.ny + TILE_Y - 1) / TILE_Y;
  const int ntz_ = (D.nz + TILE_Z - 1) / TILE_Z;
  const int ngz = (ntz_ + tzb - 1) / tzb;
  long g = tile_of_block((long) ntx_ * nty_ * ngz);
  int lane = threadIdx.x & (WAVE - 1);
  int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tx = (int) (g % ntx_), ty = (int) ((g / ntx_) % nty_);
  int tz0 = (int) (g / ((long) ntx_ * nty_)) * tzb;
  int xb = tx * TILE_X, jb = ty * TILE_Y;
  int k0 = tz0 * TILE_Z + 1;
  int kend = k0 + tzb * TILE_Z - 1 < D.nz ? k0 + tzb * TILE_Z - 1 : D.nz;
  int x = xb + lane + 1, j = jb + wv + 1;
  bool colok = (x <= D.nx) && (j <= D.ny);
  PlaneEntry E0 = plane_entry(D, xb, jb, threadIdx.x);
  PlaneEntry E1 = plane_entry(D, xb, jb, threadIdx.x + TILE_THREADS);
  PlaneRawM<V> Rp = plane_load_nb<FIRST, V, NTP>(D, E0, E1, r, pold, k0 - 1);
  PlaneRawM<V> Rc = plane_load_nb<FIRST, V, NTP>(D, E0, E1, r, pold, k0);
  /* loads are unconditional (plane index clamped to kend + 1): a load
   * issued on one path only would make the compiler wait for all */
  PlaneRawM<V> R[DEPTH];
#pragma unroll
  for (int d = 0; d < DEPTH; d++)
    R[d] = plane_load_nb<FIRST, V, NTP>(D, E0, E1, r, pold,
                                   min(k0 + 1 + d, kend + 1));
  double beta64 = 0.;
  if (!FIRST && !cg_continue<MIXED>(st, rpart, ntiles, q, &beta64)) return;
  V beta = (V) beta64;
  __syncthreads();                      /* s_invm */
  int xl = lane + 1, yl = wv + 1;
  bool wW = wall_lo(x, D.perx), wE = wall_hi(x, D.nx, D.perx);
  bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);

  /* three LDS planes (k - 1, k, k + 1) rotate: the buffer that receives
   * plane k + 1 held plane k - 2, whose last reads came before the barrier
   * of step k - 1, so one barrier per plane suffices */
  V *bprev = lds, *bcur = lds + LPLANE, *bnext = lds + 2 * LPLANE;
  plane_store_nb<FIRST, V>(D, s_invm, E0, E1, Rp, beta, bprev, k0 - 1);
  plane_store_nb<FIRST, V>(D, s_invm, E0, E1, Rc, beta, bcur, k0);
  __syncthreads();
  V prev = bprev[yl * LX + xl];
  V cur = bcur[yl * LX + xl];
  double acc = 0.0;
  /* Tile partials of the column: each wave's tree sum goes to LDS at the
   * end of each tile (no memory operation in that branch) and the row
   * combination is stored after the loop; the same order as tile_partial. */
  __shared__ double s_tp[8][TILE_Y];
  int ti = 0;
  /* one plane of the column: store plane k + 1 (raw values in Rs, loaded
   * two steps earlier), refill Rs with plane k + 3, compute plane k.
   * Loads and stores are unconditional (lanes outside the domain store to
   * the never-read corner ghost cell 0 and add nothing to the partial). */
  auto step = [&](int k, PlaneRawM<V> &Rs) {
    plane_store_nb<FIRST, V>(D, s_invm, E0, E1, Rs, beta, bnext, k + 1);
    Rs = plane_load_nb<FIRST, V, NTP>(D, E0, E1, r, pold,
                                 min(k + 1 + DEPTH, kend + 1));
    __syncthreads();
    const V *P = bcur;
    V nxt = bnext[yl * LX + xl];
    V a = neg_lap7<V>(K, cur, P[yl * LX + xl - 1], P[yl * LX + xl + 1],
                      P[(yl - 1) * LX + xl], P[(yl + 1) * LX + xl], prev,
                      nxt, wW, wE, wS, wN, wall_lo(k, D.perz),
                      wall_hi(k, D.nz, D.perz));
    unsigned c = colok ? (unsigned) gidx(C, x, j, k) : 0u;
    if (NTS) {
      st32_nt(pnew, c, cur);
      st32_nt(Ap, c, a);
    } else {
      st32(pnew, c, cur);
      st32(Ap, c, a);
    }
    acc += colok ? (double) cur * (double) a : 0.0;
    prev = cur;
    cur = nxt;
    V *t3 = bprev;
    bprev = bcur;
    bcur = bnext;
    bnext = t3;
    if (((k - k0 + 1) & (TILE_Z - 1)) == 0 || k == kend) {
      double v = wave_tree_sum(acc);
      if (lane == 0) s_tp[ti][wv] = v;
      acc = 0.0;
      ti++;
    }
  };
  /* DEPTH planes per trip, so that the register sets take their turns
   * without register copies (a copy of a register with a load in flight
   * waits for that load) and no step inside the loop is conditional; the
   * last planes of the column after it */
  int k = k0;
  for (; k + DEPTH - 1 <= kend; k += DEPTH) {
#pragma unroll
    for (int d = 0; d < DEPTH; d++) step(k + d, R[d]);
  }
#pragma unroll
  for (int d = 0; d < DEPTH - 1; d++)
    if (k + d <= kend) step(k + d, R[d]);
  __syncthreads();
  if (threadIdx.x < ti) {
    const double *w = s_tp[threadIdx.x];
    gpart[tx + (long) ntx_ * (ty + (long) nty_ * (tz0 + (int) threadIdx.x))] =
      ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
  }
}


/* ------------------- search kernel, one wavefront per row ----------------- */
/* The same pass with one wavefront per x row of the domain: lane l holds the
 * CPR cells x = CPR l + 1 .. CPR l + CPR of its row, read with one vector
 * load per array and plane, so x neighbours come from the lane itself and,
 * at the ends of its run, from lanes l - 1 and l + 1 (cross-lane moves, no
 * memory access).  A workgroup of RY wavefronts takes RY consecutive rows
 * and streams `zc` planes of them: y neighbours from an LDS plane of RY + 2
 * rows (its own rows and a halo row on either side, shared out over the
 * wavefronts), z neighbours in registers.  Per plane it loads RY + 2 rows for RY
 * rows and no x halo, where the tile plane of the z-column kernel takes
 * 66 x 10 entries for 64 x 8 cells, 20 of them single elements on cache
 * lines of their own.  Arithmetic per cell and the tile partials (64 x 8 x 8
 * tiles: per cell over the tile's planes in z order, the 64-term tree of
 * wave_tree_sum over the tile's x, then the rows) are those of the z-column
 * kernel, so the two give the same bits.  Needs nx <= 64 CPR, nx a multiple
 * of CPR, and point x = 1 of every row aligned to CPR elements (the grid
 * layout gives 128 bytes); zc a multiple of TILE_Z, at most ROW_ZCMAX. */
#define ROW_ZCMAX 64

template <typename V, int N> struct Run { V v[N]; };

/* N consecutive elements from element i of a grid array (one vector load;
 * NT: with the non-temporal hint). */
template <typename V, int N, bool NT>
__device__ inline Run<V, N> run_load(const V *__restrict__ a, unsigned i)
{
  Run<V, N> o;
  const char *p = (const char *) a + i * (unsigned) sizeof(V);
  if constexpr (N == 1) {
    o.v[0] = NT ? __builtin_nontemporal_load((const V *) p) : *(const V *) p;
  } else {
    typedef V VN __attribute__((ext_vector_type(N)));
    VN t = NT ? __builtin_nontemporal_load((const VN *) p) : *(const VN *) p;
#pragma unroll
    for (int e = 0; e < N; e++) o.v[e] = t[e];
  }
  return o;
}

/* N consecutive elements stored from element i by one vector store; NT:
 * with the non-temporal hint. */
template <typename V, int N, bool NT>
__device__ inline void run_store(V *__restrict__ a, unsigned i,
  const Run<V, N> &o)
{
  char *p = (char *) a + i * (unsigned) sizeof(V);
  if constexpr (N == 1) {
    if (NT) __builtin_nontemporal_store(o.v[0], (V *) p);
    else *(V *) p = o.v[0];
  } else {
    typedef V VN __attribute__((ext_vector_type(N)));
    VN t;
#pragma unroll
    for (int e = 0; e < N; e++) t[e] = o.v[e];
    if (NT) __builtin_nontemporal_store(t, (VN *) p);
    else *(VN *) p = t;
  }
}

template <typename V, int N>
__device__ inline void run_store_nt(V *__restrict__ a, unsigned i,
  const Run<V, N> &o)
{
  char *p = (char *) a + i * (unsigned) sizeof(V);
  if constexpr (N == 1) {
    __builtin_nontemporal_store(o.v[0], (V *) p);
  } else {
    typedef V VN __attribute__((ext_vector_type(N)));
    VN t;
#pragma unroll
    for (int e = 0; e < N; e++) t[e] = o.v[e];
    __builtin_nontemporal_store(t, (VN *) p);
  }
}

/* Lane moves that do not go through the LDS crossbar (ds_bpermute):
 * lane_prev gives lane l the value of lane (l - 1) mod 64 (DPP wave
 * rotate right by one), lane_next that of lane (l + 1) mod 64 (wave rotate
 * left by one), lane_get that of a wavefront-uniform lane (v_readlane).
 * 8-byte values move as two 32-bit halves. */
template <int CTRL>
__device__ inline unsigned dpp32(unsigned x)
{
  return (unsigned) __builtin_amdgcn_mov_dpp((int) x, CTRL, 0xF, 0xF, false);
}

template <int CTRL, typename V>
__device__ inline V dpp_move(V x)
{
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, dpp32<CTRL>(__builtin_bit_cast(unsigned, x)));
  } else {
    typedef unsigned U2 __attribute__((ext_vector_type(2)));
    U2 u = __builtin_bit_cast(U2, x);
    u[0] = dpp32<CTRL>(u[0]);
    u[1] = dpp32<CTRL>(u[1]);
    return __builtin_bit_cast(V, u);
  }
}

#define DPP_WAVE_ROL1 0x134
#define DPP_WAVE_ROR1 0x13C
#define DPP_ROW_SHL(d) (0x100 + (d))

/* The value of lane l + d for d < 16 within l's row of 16 lanes (DPP
 * row shift left); lanes whose source lies past the row keep their own
 * value, as __shfl_down keeps it past the wavefront.  A tree sum whose
 * steps stay inside 16-lane groups reads only valid sources at the
 * group's first lane, so it gives the same result there as with
 * __shfl_down. */
template <int D, typename V>
__device__ inline V lane_down_row(V x)
{
  static_assert(D >= 1 && D < 16, "a DPP row shift moves 1 to 15 lanes");
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, (unsigned) __builtin_amdgcn_update_dpp(
      __builtin_bit_cast(int, x), __builtin_bit_cast(int, x),
      DPP_ROW_SHL(D), 0xF, 0xF, false));
  } else {
    typedef int I2 __attribute__((ext_vector_type(2)));
    I2 u = __builtin_bit_cast(I2, x);
    u[0] = __builtin_amdgcn_update_dpp(u[0], u[0], DPP_ROW_SHL(D), 0xF, 0xF,
                                       false);
    u[1] = __builtin_amdgcn_update_dpp(u[1], u[1], DPP_ROW_SHL(D), 0xF, 0xF,
                                       false);
    return __builtin_bit_cast(V, u);
  }
}

/* One step of the tile tree sum: acc += the value D lanes down (a DPP row
 * shift within 16-lane groups, else a bpermute across them). */
template <int D>
__device__ inline double tree_step(double a)
{
  if constexpr (D < 16) return a + lane_down_row<D>(a);
  else return a + __shfl_down(a, D, WAVE);
}

template <typename V> __device__ inline V lane_prev(V x)
{
  return dpp_move<DPP_WAVE_ROR1>(x);
}

template <typename V> __device__ inline V lane_next(V x)
{
  return dpp_move<DPP_WAVE_ROL1>(x);
}

template <typename V> __device__ inline V lane_get(V x, int src)
{
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, __builtin_amdgcn_readlane(
      __builtin_bit_cast(int, x), src));
  } else {
    typedef int I2 __attribute__((ext_vector_type(2)));
    I2 u = __builtin_bit_cast(I2, x);
    u[0] = __builtin_amdgcn_readlane(u[0], src);
    u[1] = __builtin_amdgcn_readlane(u[1], src);
    return __builtin_bit_cast(V, u);
  }
}

/* A row of WAVE runs of N elements in LDS, laid out as lane-consecutive
 * chunks of at most 16 bytes: chunk c of lane l holds elements
 * c*H .. c*H + H - 1 of the lane's run at (c*WAVE + l)*H.  Each chunk
 * access is then one ds_read_b128 / ds_write_b128 (or narrower) over
 * consecutive addresses, free of bank conflicts; a run stored whole (16
 * bytes of FP32, 32 of FP64) at a run-sized lane stride was split by the
 * compiler into 8-byte accesses that conflict. */
template <typename V, int N> struct LdsRow {
  static constexpr int H = (int) (N * sizeof(V) <= 16 ? N : 16 / sizeof(V));
  static constexpr int NH = N / H;
  static_assert(NH * H == N, "run length must divide into chunks");
  __device__ static void put(V *__restrict__ row, int lane, const Run<V, N> &o)
  {
#pragma unroll
    for (int c = 0; c < NH; c++) {
      /* every row starts on a 16-byte boundary of the aligned LDS array */
      V *q = (V *) __builtin_assume_aligned(row + (c * WAVE + lane) * H,
                                            H * sizeof(V));
      if constexpr (H == 1) {
        *q = o.v[c];
      } else {
        typedef V VH __attribute__((ext_vector_type(H)));
        VH t;
#pragma unroll
        for (int e = 0; e < H; e++) t[e] = o.v[c * H + e];
        *(VH *) q = t;
      }
    }
  }
  __device__ static Run<V, N> get(const V *__restrict__ row, int lane)
  {
    Run<V, N> o;
#pragma unroll
    for (int c = 0; c < NH; c++) {
      const V *q = (const V *) __builtin_assume_aligned(
        row + (c * WAVE + lane) * H, H * sizeof(V));
      if constexpr (H == 1) {
        o.v[c] = *q;
      } else {
        typedef V VH __attribute__((ext_vector_type(H)));
        VH t = *(const VH *) q;
#pragma unroll
        for (int e = 0; e < H; e++) o.v[c * H + e] = t[e];
      }
    }
    return o;
  }
  /* position of element q (0-based x - 1) of the row */
  __device__ static int pos(int q)
  {
    const int l = q / N, e = q - l * N;
    return ((e / H) * WAVE + l) * H + (e % H);
  }
};

/* Source row of LDS row jj (0 and ny + 1 are the halo rows): the row
 * itself, the wrapped row, or -1 for a row that holds zeros (a wall ghost,
 * or a row past the halo in the domain's last workgroup). */
__device__ inline int row_src(const DevDom &D, int jj)
{
  if (jj >= 1 && jj <= D.ny) return jj;
  if (jj == 0) return D.pery ? D.ny : -1;
  if (jj == D.ny + 1) return D.pery ? 1 : -1;
  return -1;
}

template <bool FIRST, typename V, bool MIXED, int CPR, int RY, int DEPTH,
          bool PR = false>
__global__ void __launch_bounds__(RY * WAVE)
search_row_kernel(DevDom D, CoefT<V> K, InvMT<V> T, CGState *st, int q,
  const double *__restrict__ rpart, long ntiles, int zc,
  const V *__restrict__ r, const V *__restrict__ pold,
  V *__restrict__ pnew, V *__restrict__ Ap, double *__restrict__ gpart)
{
  typedef Run<V, CPR> R_t;
  /* FP64 p_old is read with the non-temporal hint, FP32 p_old ordinarily
   * (as in the z-column kernel); p_new and Ap are always stored with the
   * hint.  (A run-time switch between hinted and plain stores let the
   * compiler merge the two store paths and drop the hint; in that build the
   * FP64 update that follows took 193 us per launch at 256^3 instead of
   * 165 us: it paid the search's write-back.) */
  constexpr bool NTP = sizeof(V) == 8;
  typedef LdsRow<V, CPR> LR;
  constexpr int ROWV = WAVE * CPR;        /* elements of one LDS row */
  __shared__ __attribute__((aligned(16))) V lds[3 * (RY + 2) * ROWV];
  __shared__ V s_invm[27];
  __shared__ double s_tp[ROW_ZCMAX / TILE_Z][CPR][RY];
  /* PR (pairs; FP64, four cells per lane): lane m loads r and p_old of
   * cells 2m, 2m + 1 and 128 + 2m, 128 + 2m + 1, forms their p_new and
   * writes it to the LDS row; each lane reads its own run of four back for
   * the stencil; p_new and Ap leave through LDS in the same pairs (s_ap:
   * the wavefront's row of Ap).  Every load and store instruction then
   * moves 1 KB of consecutive cells instead of 16 bytes at a 32-byte lane
   * stride, which wrote each line in halves: at 256^3 the kernel wrote 16 %
   * more than its two arrays and took 124 us per launch; with pairs it
   * writes exactly its two arrays and takes 103 us.  Arithmetic and tile
   * partials are unchanged (same bits). */
  static_assert(!PR || (sizeof(V) == 8 && CPR == 4), "pairs: FP64, 4 cells");
  __shared__ __attribute__((aligned(16))) V s_ap[PR ? RY * ROWV : 2];
  if (st->done | st->inner_done) return;
  if (threadIdx.x < 27) s_invm[threadIdx.x] = T.v[threadIdx.x];
  const Grid &C = D.cc;
  const int nyb = (D.ny + RY - 1) / RY;
  const int nzc = (D.nz + zc - 1) / zc;
  const long g = tile_of_block((long) nyb * nzc);
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int jb = (int) (g % nyb) * RY;
  const int k0 = (int) (g / nyb) * zc + 1;
  const int kend = k0 + zc - 1 < D.nz ? k0 + zc - 1 : D.nz;
  const int j = jb + wv + 1;
  const int x0 = CPR * lane + 1;
  const bool xok = x0 <= D.nx;            /* nx % CPR == 0: the whole run */
  /* pairs: x - 1 of the lane's two pairs, and whether each lies in the row
   * (nx % 4 == 0: a pair lies in it whole or not at all) */
  const int qp0 = 2 * lane, qp1 = 2 * WAVE + 2 * lane;
  const bool pok0 = qp0 < D.nx, pok1 = qp1 < D.nx;
  const bool colok = xok && j <= D.ny;
  /* the wavefront's own row is LDS row wv + 1.  The two halo rows (LDS
   * rows 0 and RY + 1) are shared out evenly, so that every wavefront
   * reaches the plane barrier with the same work: wavefront wv loads one
   * cell per lane of segment wv % (RY / 2) (HC cells) of halo row 0 (wv <
   * RY / 2) or RY + 1.  Every load is unconditional: lanes past the
   * segment or the row read x = 1 and store nothing. */
  constexpr int HSEG = RY / 2, HC = WAVE * CPR / HSEG;
  static_assert(HC <= WAVE, "a halo segment must fit one wavefront");
  const int hs = wv < HSEG ? 0 : RY + 1;
  const int xh = (wv % HSEG) * HC + lane + 1;
  const bool hlane = lane < HC && xh <= D.nx;
  const int so = row_src(D, j), sh = row_src(D, jb + hs);
  const unsigned xo = xok ? (unsigned) x0 : 1u;
  const unsigned co = xo + (unsigned) (so >= 0 ? so : 1) * (unsigned) C.sy;
  const unsigned ch = (hlane ? (unsigned) xh : 1u)
                    + (unsigned) (sh >= 0 ? sh : 1) * (unsigned) C.sy;
  const bool oko = xok && so >= 0, okh = sh >= 0;
  int wx[CPR];
#pragma unroll
  for (int e = 0; e < CPR; e++)
    wx[e] = wcount(PR ? (e < 2 ? qp0 : qp1 - 2) + e + 1 : x0 + e, D.nx,
                   D.perx);
  const int wyo = 3 * wcount(so, D.ny, D.pery);
  const int wxyh = wcount(xh, D.nx, D.perx) + 3 * wcount(sh, D.ny, D.pery);

  struct Raw { R_t r, p; V hr, hp; };
  auto load = [&](int kk) {
    Raw R;
    int ks = plane_src(D, kk);
    unsigned off = (unsigned) (ks ? ks : 1) * (unsigned) C.sz;
    if constexpr (PR) {
      /* the lane's two pairs (x = 1 for a pair past the row) */
      const unsigned rb = 1u + (unsigned) (so >= 0 ? so : 1) * (unsigned) C.sy
                        + off;
      const unsigned a0 = rb + (pok0 ? (unsigned) qp0 : 0u),
                     a1 = rb + (pok1 ? (unsigned) qp1 : 0u);
      const Run<V, 2> r0 = run_load<V, 2, false>(r, a0),
                      r1 = run_load<V, 2, false>(r, a1);
      R.r.v[0] = r0.v[0]; R.r.v[1] = r0.v[1];
      R.r.v[2] = r1.v[0]; R.r.v[3] = r1.v[1];
      if (!FIRST) {
        const Run<V, 2> p0 = run_load<V, 2, NTP>(pold, a0),
                        p1 = run_load<V, 2, NTP>(pold, a1);
        R.p.v[0] = p0.v[0]; R.p.v[1] = p0.v[1];
        R.p.v[2] = p1.v[0]; R.p.v[3] = p1.v[1];
      }
    } else {
      R.r = run_load<V, CPR, false>(r, co + off);
      if (!FIRST) R.p = run_load<V, CPR, NTP>(pold, co + off);
    }
    R.hr = ld32(r, ch + off);
    if (FIRST) {
#pragma unroll
      for (int e = 0; e < CPR; e++) R.p.v[e] = (V) 0;
      R.hp = (V) 0;
    } else {
      R.hp = NTP ? ld32_nt(pold, ch + off) : ld32(pold, ch + off);
    }
    return R;
  };
  Raw Rm = load(k0 - 1), Rc = load(k0);
  Raw Rq[DEPTH];
#pragma unroll
  for (int d = 0; d < DEPTH; d++) Rq[d] = load(min(k0 + 1 + d, kend + 1));
  double beta64 = 0.;
  if (!FIRST && !cg_continue<MIXED, RY * WAVE>(st, rpart, ntiles, q, &beta64))
    return;
  const V beta = (V) beta64;
  __syncthreads();                      /* s_invm */
  /* Jacobi factors of the run's cells in the two z classes a source plane
   * can have: one wall face (planes 1 and nz between walls) or none */
  V mo[2][CPR], mh[2];
#pragma unroll
  for (int c = 0; c < 2; c++) {
#pragma unroll
    for (int e = 0; e < CPR; e++) mo[c][e] = s_invm[wx[e] + wyo + 9 * (c + 1)];
    mh[c] = s_invm[wxyh + 9 * (c + 1)];
  }
  /* p_new of a loaded plane into LDS buffer `buf` (own run and halo cell,
   * the halo cell at its x in the row); returns the own run */
  auto put = [&](const Raw &R, V *buf, int kk) {
    int ks = plane_src(D, kk);
    /* a ghost plane (ks = 0) holds zeros whatever the class */
    const int zc2 = wcount(ks, D.nz, D.perz) == 2;
    R_t o;
#pragma unroll
    for (int e = 0; e < CPR; e++) {
      V z = R.r.v[e]*(zc2 ? mo[1][e] : mo[0][e]);
      V v = FIRST ? z : z + beta*R.p.v[e];
      const bool ok = PR ? so >= 0 && (e < 2 ? pok0 : pok1) : oko;
      o.v[e] = (ks && ok) ? v : (V) 0;
    }
    if constexpr (PR) {
      /* the pairs at LdsRow<V, 4>::pos(qp0) and pos(qp1) */
      typedef V V2 __attribute__((ext_vector_type(2)));
      V *own = buf + (wv + 1) * ROWV;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int ps = ((lane & 1) * WAVE + h * (WAVE / 2) + (lane >> 1)) * 2;
        V2 t;
        t[0] = o.v[2 * h];
        t[1] = o.v[2 * h + 1];
        *(V2 *) __builtin_assume_aligned(own + ps, 16) = t;
      }
    } else {
      LR::put(buf + (wv + 1) * ROWV, lane, o);
    }
    V zh = R.hr*(zc2 ? mh[1] : mh[0]);
    V vh = FIRST ? zh : zh + beta*R.hp;
    vh = (ks && okh) ? vh : (V) 0;
    if (hlane) buf[hs * ROWV + LR::pos(xh - 1)] = vh;
    return o;
  };
  V *bprev = lds, *bcur = lds + (RY + 2) * ROWV,
    *bnext = lds + 2 * (RY + 2) * ROWV;
  R_t prev = put(Rm, bprev, k0 - 1);
  R_t cur = put(Rc, bcur, k0);
  __syncthreads();
  if constexpr (PR) {
    prev = LR::get(bprev + (wv + 1) * ROWV, lane);
    cur = LR::get(bcur + (wv + 1) * ROWV, lane);
  }
  const bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);
  /* x = 1 and x = nx: the first cell of lane 0 and the last of lane
   * nx / CPR - 1, the only cells that can have an x wall */
  const int lend = D.nx / CPR - 1;
  const bool xw_lo = lane == 0 && !D.perx, xw_hi = lane == lend && !D.perx;
  /* (-A) p over the run; YZ: the row or the plane has a wall face in y or
   * z (else those weights are 1 and left out) */
  auto lap = [&](auto yz, const R_t &S, const R_t &N, const R_t &nxt, V wl,
                 V er, V xn, V x1, bool wB, bool wT) {
    constexpr bool YZ = decltype(yz)::value;
    R_t a;
#pragma unroll
    for (int e = 0; e < CPR; e++) {
      V pw = e > 0 ? cur.v[e > 0 ? e - 1 : 0]
                   : (lane == 0 ? (D.perx ? xn : (V) 0) : wl);
      V pe = e < CPR - 1 ? cur.v[e < CPR - 1 ? e + 1 : 0]
                         : (lane == lend ? (D.perx ? x1 : (V) 0) : er);
      const bool ww = e == 0 && xw_lo, we = e == CPR - 1 && xw_hi;
      a.v[e] = neg_lap7<V>(K, cur.v[e], pw, pe, S.v[e], N.v[e], prev.v[e],
                           nxt.v[e], ww, we, YZ && wS, YZ && wN, YZ && wB,
                           YZ && wT);
    }
    return a;
  };
  double acc[CPR];
#pragma unroll
  for (int e = 0; e < CPR; e++) acc[e] = 0.0;
  int ti = 0;
  auto step = [&](int k, Raw &Rs) {
    R_t nxt = put(Rs, bnext, k + 1);
    Rs = load(min(k + 1 + DEPTH, kend + 1));
    __syncthreads();
    if constexpr (PR) nxt = LR::get(bnext + (wv + 1) * ROWV, lane);
    const R_t S = LR::get(bcur + wv * ROWV, lane),
              N = LR::get(bcur + (wv + 2) * ROWV, lane);
    /* run neighbours: the last cell of lane l - 1, the first of l + 1, and
     * the wrapped cells x = nx and x = 1 (0: a wall ghost) */
    V wl = lane_prev(cur.v[CPR - 1]);
    V er = lane_next(cur.v[0]);
    V xn = lane_get(cur.v[CPR - 1], lend);
    V x1 = lane_get(cur.v[0], 0);
    const bool wB = wall_lo(k, D.perz), wT = wall_hi(k, D.nz, D.perz);
    /* wavefront-uniform: interior rows of interior planes skip the y and
     * z weights (a weight of 1 multiplies exactly) */
    R_t a;
    if (wS || wN || wB || wT)
      a = lap(std::true_type(), S, N, nxt, wl, er, xn, x1, wB, wT);
    else
      a = lap(std::false_type(), S, N, nxt, wl, er, xn, x1, wB, wT);
    /* cells outside the domain store to the never-read edge ghost
     * (x = 1, j = 0, k = 0), aligned like every run */
    if constexpr (PR) {
      V *sa = s_ap + wv * ROWV;
      LR::put(sa, lane, a);
      /* the reads below take other lanes' values: LDS operations of one
       * wavefront complete in order; keep the compiler from moving them */
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const V *own = bcur + (wv + 1) * ROWV;
      /* element 1 (x = 1) of this row and plane */
      const unsigned rb = 1u + (unsigned) (so >= 0 ? so : 1) * (unsigned) C.sy
                        + (unsigned) k * (unsigned) C.sz;
      typedef V V2 __attribute__((ext_vector_type(2)));
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int qx = h * (2 * WAVE) + 2 * lane;     /* x - 1 of the pair */
        /* LdsRow<V, 4>::pos(qx), a 16-byte slot */
        const int ps = ((lane & 1) * WAVE + h * (WAVE / 2) + (lane >> 1)) * 2;
        const V2 pv = *(const V2 *) __builtin_assume_aligned(own + ps, 16);
        const V2 av = *(const V2 *) __builtin_assume_aligned(sa + ps, 16);
        if (j <= D.ny && qx < D.nx) {
          __builtin_nontemporal_store(pv, (V2 *) (pnew + rb + qx));
          __builtin_nontemporal_store(av, (V2 *) (Ap + rb + qx));
        }
      }
    } else {
      unsigned c = colok ? co + (unsigned) k * (unsigned) C.sz : 1u;
      run_store<V, CPR, true>(pnew, c, cur);
      run_store<V, CPR, true>(Ap, c, a);
    }
#pragma unroll
    for (int e = 0; e < CPR; e++)
      acc[e] += colok ? (double) cur.v[e] * (double) a.v[e] : 0.0;
    prev = cur;
    cur = nxt;
    V *t3 = bprev;
    bprev = bcur;
    bcur = bnext;
    bnext = t3;
    if (((k - k0 + 1) & (TILE_Z - 1)) == 0 || k == kend) {
      /* wave_tree_sum over each 64-cell tile of the row: steps of CPR cells
       * or more are lane moves within the tile's 64 / CPR lanes, the last
       * log2(CPR) steps add within the lane */
      /* lane distances 32 / CPR, 16 / CPR, ..., 1 (the tile's lanes) */
#pragma unroll
      for (int e = 0; e < CPR; e++) {
        if constexpr (CPR == 1) acc[e] = tree_step<32>(acc[e]);
        if constexpr (CPR <= 2) acc[e] = tree_step<16>(acc[e]);
        acc[e] = tree_step<8>(acc[e]);
        acc[e] = tree_step<4>(acc[e]);
        acc[e] = tree_step<2>(acc[e]);
        acc[e] = tree_step<1>(acc[e]);
      }
#pragma unroll
      for (int off = CPR / 2; off >= 1; off >>= 1)
#pragma unroll
        for (int e = 0; e < off; e++) acc[e] += acc[e + off];
      if ((lane & (WAVE / CPR - 1)) == 0) s_tp[ti][lane / (WAVE / CPR)][wv] = acc[0];
#pragma unroll
      for (int e = 0; e < CPR; e++) acc[e] = 0.0;
      ti++;
    }
  };
  int k = k0;
  for (; k + DEPTH - 1 <= kend; k += DEPTH) {
#pragma unroll
    for (int d = 0; d < DEPTH; d++) step(k + d, Rq[d]);
  }
#pragma unroll
  for (int d = 0; d < DEPTH - 1; d++)
    if (k + d <= kend) step(k + d, Rq[d]);
  __syncthreads();
  /* one partial per (z-tile, x-tile, 8-row group) of the workgroup, rows
   * combined as in tile_partial */
  const int ntx_ = (D.nx + TILE_X - 1) / TILE_X;
  const int nty_ = (D.ny + TILE_Y - 1) / TILE_Y;
  const int nout = ti * CPR * (RY / TILE_Y);
  if ((int) threadIdx.x < nout) {
    int t = threadIdx.x;
    int tx = t % CPR, rg = (t / CPR) % (RY / TILE_Y), tz = t / (CPR * (RY / TILE_Y));
    int ty = jb / TILE_Y + rg;
    if (tx < ntx_ && ty < nty_) {
      const double *w = &s_tp[tz][tx][rg * TILE_Y];
      gpart[tx + (long) ntx_ * (ty + (long) nty_ * ((k0 - 1) / TILE_Z + tz))] =
        ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
    }
  }
}


/* The same pass with every neighbour read from global memory instead of
 * LDS (a second implementation, used to cross-check the tiled one). */
template <bool FIRST>
__global__ void __launch_bounds__(TILE_THREADS)
search_spmv_direct_kernel(DevDom D, Coef K, InvMTab T, CGState *st, int q,
  const double *__restrict__ rpart, long ntiles,
  const double *__restrict__ r, const double *__restrict__ pold,
  const uint8_t *__restrict__ fluid,
  double *__restrict__ pnew, double *__restrict__ Ap,
  double *__restrict__ gpart)
{
  if (st->done) return;
  double beta = 0.;
  if (!FIRST && !cg_continue<false>(st, rpart, ntiles, q, &beta)) return;
  TILE_PROLOGUE
  double acc = 0.0;
  if (colok) {
    int xm = wrap_lo(x, D.nx, D.perx), xp = wrap_hi(x, D.nx, D.perx);
    int jm = wrap_lo(j, D.ny, D.pery), jp = wrap_hi(j, D.ny, D.pery);
    bool wW = wall_lo(x, D.perx), wE = wall_hi(x, D.nx, D.perx);
    bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);
    for (int k = k0; k <= k1; k++) {
      int km = wrap_lo(k, D.nz, D.perz), kp = wrap_hi(k, D.nz, D.perz);
#define PN(xx, jj, kk) p_new<FIRST, double>(r, pold, gidx(C, xx, jj, kk), \
        solid(fluid, gidx(C, xx, jj, kk)) ? 0. : \
        invm_at<double>(T, D, xx, jj, kk), beta)
      double pC = PN(x, j, k), pW = PN(xm, j, k), pE = PN(xp, j, k);
      double pS = PN(x, jm, k), pN = PN(x, jp, k);
      double pB = PN(x, j, km), pT = PN(x, j, kp);
#undef PN
      double a = neg_lap7<double>(K, pC, pW, pE, pS, pN, pB, pT, wW, wE, wS,
                                  wN, wall_lo(k, D.perz),
                                  wall_hi(k, D.nz, D.perz));
      long c = gidx(C, x, j, k);
      pnew[c] = pC;
      Ap[c] = a;
      acc += pC * a;
    }
  }
  tile_partial(acc, gpart, tile);
}

/* -------------------------------- update --------------------------------- */
/* (src/solver_kernel.cu:838-862): alpha = (r, z) / (p, Ap); phi += alpha p,
 * r -= alpha Ap, z = invM r + tile partials of (r, z).  In the mixed form
 * phi is the FP32 correction accumulated since the last FP64 correction. */
/* PHI selects the phi update: PHI_EACH adds alpha p; PHI_SKIP leaves phi
 * alone and keeps alpha in the CG state; PHI_PAIR then adds both terms,
 * (phi + alpha' p') + alpha p with p' = pprev the previous iteration's
 * direction (still in the other p buffer): the same operations in the same
 * order as two PHI_EACH updates, hence the same bits, with phi read and
 * written once per two iterations (8 B per cell per iteration less). */
enum { PHI_EACH = 0, PHI_SKIP = 1, PHI_PAIR = 2 };

template <typename V, bool MASK, int PHI = PHI_EACH>
__global__ void __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(8, 8)))
update_kernel(DevDom D, InvMT<V> T, CGState *st, int q,
  const double *__restrict__ ppart, long ntiles,
  const V *__restrict__ p, const V *__restrict__ Ap,
  const uint8_t *__restrict__ fluid, V *__restrict__ phi,
  V *__restrict__ r, double *__restrict__ gpart,
  const V *__restrict__ pprev = NULL)
{
  __shared__ V s_invm[27];
  if (st->done | st->inner_done) return;
  const V alpha1 = PHI == PHI_PAIR ? (V) st->alpha : (V) 0;
  if (threadIdx.x < 27) s_invm[threadIdx.x] = T.v[threadIdx.x];
  TILE_PROLOGUE
  /* A whole tile (the common case) runs with no load or store behind a
   * branch, so the waits the compiler inserts stay exact; same arithmetic
   * and order as the general path below.  The invM table is read from LDS
   * (tile_block_sum's barriers publish it). */
  if (xb + TILE_X <= D.nx && jb + TILE_Y <= D.ny && k0 + TILE_Z - 1 <= D.nz) {
    const int wxy = wcount(x, D.nx, D.perx) + 3 * wcount(j, D.ny, D.pery);
    unsigned c0 = (unsigned) gidx(C, x, j, k0), sz = (unsigned) C.sz;
    /* planes 0-3 of p and Ap are in flight during the reduction; the tile
     * then goes in two halves of four planes */
    V pv[4], apv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      pv[i] = PHI == PHI_SKIP ? (V) 0 : ld32(p, c0 + i * sz);
      apv[i] = ld32_nt(Ap, c0 + i * sz);
    }
    V alpha = (V) (st->rzs[q & 1]/tile_block_sum(ppart, ntiles));
    if (PHI == PHI_SKIP && blockIdx.x == 0 && threadIdx.x == 0)
      st->alpha = alpha;
    double acc = 0.0;
#pragma unroll
    for (int h = 0; h < TILE_Z; h += 4) {
      V ph[4], rv[4], pp[4];
      uint32_t f[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        unsigned c = c0 + (h + i) * sz;
        if (h > 0) {
          if (PHI != PHI_SKIP) pv[i] = ld32(p, c);
          apv[i] = ld32_nt(Ap, c);
        }
        ph[i] = PHI == PHI_SKIP ? (V) 0 : ld32(phi, c);
        pp[i] = PHI == PHI_PAIR ? ld32(pprev, c) : (V) 0;
        rv[i] = ld32(r, c);
        f[i] = MASK ? (uint32_t) fluid[c] : 1u;
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int kk = h + i;
        unsigned c = c0 + kk * sz;
        if (PHI == PHI_EACH) st32(phi, c, (V) (ph[i] + alpha*pv[i]));
        if (PHI == PHI_PAIR) {
          V t = ph[i] + alpha1*pp[i];
          st32(phi, c, (V) (t + alpha*pv[i]));
        }
        V rr = rv[i] - alpha*apv[i];
        rr = f[i] ? rr : (V) 0;
        st32(r, c, rr);
        V m = s_invm[wxy + 9 * wcount(k0 + kk, D.nz, D.perz)];
        V z = rr*(f[i] ? m : (V) 0);
        acc += (double) rr * (double) z;
      }
      /* the second half's loads are issued after the first half's stores,
       * not hoisted above them (register pressure) */
      __builtin_amdgcn_sched_barrier(0);
    }
    tile_partial_all(acc, gpart, tile);
    return;
  }
  /* general tiles (domain edges): the same halves with every access
   * guarded */
  long cb = gidx(C, x, j, k0);
  V pv[4], apv[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    pv[i] = apv[i] = 0;
    if (colok && k0 + i <= k1) {
      if (PHI != PHI_SKIP) pv[i] = p[cb + (long) i * C.sz];
      apv[i] = Ap[cb + (long) i * C.sz];
    }
  }
  V alpha = (V) (st->rzs[q & 1]/tile_block_sum(ppart, ntiles));
  if (PHI == PHI_SKIP && blockIdx.x == 0 && threadIdx.x == 0)
    st->alpha = alpha;
  double acc = 0.0;
  if (colok) {
#pragma unroll
    for (int h = 0; h < TILE_Z; h += 4) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int kk = h + i;
        if (k0 + kk > k1) continue;
        long c = cb + (long) kk * C.sz;
        if (h > 0) {
          if (PHI != PHI_SKIP) pv[i] = p[c];
          apv[i] = Ap[c];
        }
        if (PHI == PHI_EACH) phi[c] += alpha*pv[i];
        if (PHI == PHI_PAIR) {
          V t = phi[c] + alpha1*pprev[c];
          phi[c] = t + alpha*pv[i];
        }
        V rr = r[c] - alpha*apv[i];
        rr = solid(fluid, c) ? (V) 0 : rr;
        r[c] = rr;
        V z = rr*(solid(fluid, c) ? (V) 0 : invm_at<V>(T, D, x, j, k0 + kk));
        acc += (double) rr * (double) z;
      }
    }
  }
  tile_partial(acc, gpart, tile);
}


/* ----------------- update kernel, one wavefront per row ------------------ */
/* The update pass in the row shape of search_row_kernel: lane l holds the
 * CPR cells x = CPR l + 1 .. CPR l + CPR of its row (vector loads), a
 * workgroup takes RY rows and `zc` planes of them with DEPTH planes of
 * loads in flight.  alpha (the reduction of every tile partial) is formed
 * once per workgroup of RY rows by zc planes instead of once per 64 x 8 x 8
 * tile, each wavefront reads whole rows, and no barrier is needed inside
 * the plane loop.  Arithmetic per cell is that of update_kernel and the
 * tile partials of (r, z) are formed as in search_row_kernel, so the two
 * update kernels give the same bits.  Needs what the search row kernel
 * needs, the fluid mask included (CPR bytes per run, aligned). */
template <typename V, int N>
__device__ inline void run_store(V *__restrict__ a, unsigned i,
  const Run<V, N> &o)
{
  char *p = (char *) a + i * (unsigned) sizeof(V);
  if constexpr (N == 1) {
    *(V *) p = o.v[0];
  } else {
    typedef V VN __attribute__((ext_vector_type(N)));
    VN t;
#pragma unroll
    for (int e = 0; e < N; e++) t[e] = o.v[e];
    *(VN *) p = t;
  }
}

/* N fluid flags from byte i (byte e of the result: cell e). */
template <int N>
__device__ inline uint32_t flags_load(const uint8_t *__restrict__ f,
  unsigned i)
{
  if constexpr (N == 4) return *(const uint32_t *) (f + i);
  else if constexpr (N == 2) return *(const uint16_t *) (f + i);
  else return f[i];
}

template <typename V, bool MASK, int PHI, int CPR, int RY, int DEPTH>
__global__ void __launch_bounds__(RY * WAVE)
update_row_kernel(DevDom D, InvMT<V> T, CGState *st, int q,
  const double *__restrict__ ppart, long ntiles, int zc,
  const V *__restrict__ p, const V *__restrict__ Ap,
  const uint8_t *__restrict__ fluid, V *__restrict__ phi,
  V *__restrict__ r, double *__restrict__ gpart,
  const V *__restrict__ pprev)
{
  typedef Run<V, CPR> R_t;
  __shared__ V s_invm[27];
  __shared__ double s_tp[ROW_ZCMAX / TILE_Z][CPR][RY];
  if (st->done | st->inner_done) return;
  const V alpha1 = PHI == PHI_PAIR ? (V) st->alpha : (V) 0;
  if (threadIdx.x < 27) s_invm[threadIdx.x] = T.v[threadIdx.x];
  const Grid &C = D.cc;
  const int nyb = (D.ny + RY - 1) / RY;
  const int nzc = (D.nz + zc - 1) / zc;
  const long g = tile_of_block((long) nyb * nzc);
  const int lane = threadIdx.x & (WAVE - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int jb = (int) (g % nyb) * RY;
  const int k0 = (int) (g / nyb) * zc + 1;
  const int kend = k0 + zc - 1 < D.nz ? k0 + zc - 1 : D.nz;
  const int j = jb + wv + 1;
  const int x0 = CPR * lane + 1;
  const bool colok = x0 <= D.nx && j <= D.ny;
  /* cells outside the domain read x = 1 of row 1 and store nothing */
  const unsigned co = (colok ? (unsigned) x0 : 1u)
                    + (unsigned) (colok ? j : 1) * (unsigned) C.sy;
  struct Raw { R_t p, ap, ph, pp, r; uint32_t f; };
  auto load = [&](int k) {
    Raw R;
    unsigned c = co + (unsigned) k * (unsigned) C.sz;
    R.ap = run_load<V, CPR, true>(Ap, c);
    R.r = run_load<V, CPR, false>(r, c);
    if (PHI != PHI_SKIP) {
      R.p = run_load<V, CPR, false>(p, c);
      R.ph = run_load<V, CPR, false>(phi, c);
    }
    if (PHI == PHI_PAIR) R.pp = run_load<V, CPR, false>(pprev, c);
    R.f = MASK ? flags_load<CPR>(fluid, c) : 0xffffffffu;
    return R;
  };
  /* the first DEPTH planes are in flight during the reduction */
  Raw R[DEPTH];
#pragma unroll
  for (int d = 0; d < DEPTH; d++) R[d] = load(min(k0 + d, kend));
  const V alpha = (V) (st->rzs[q & 1]/tile_block_sum<RY * WAVE>(ppart, ntiles));
  if (PHI == PHI_SKIP && blockIdx.x == 0 && threadIdx.x == 0)
    st->alpha = alpha;
  /* Jacobi factors by z class (one wall face, none); s_invm was published
   * by tile_block_sum's barriers */
  const int wy = 3 * wcount(j, D.ny, D.pery);
  V m[2][CPR];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int e = 0; e < CPR; e++)
      m[c][e] = s_invm[wcount(x0 + e, D.nx, D.perx) + wy + 9 * (c + 1)];
  double acc[CPR];
#pragma unroll
  for (int e = 0; e < CPR; e++) acc[e] = 0.0;
  int ti = 0;
  auto step = [&](int k, Raw &Rs) {
    const unsigned c = co + (unsigned) k * (unsigned) C.sz;
    const bool zc2 = wcount(k, D.nz, D.perz) == 2;
    R_t ph, rv;
#pragma unroll
    for (int e = 0; e < CPR; e++) {
      if (PHI == PHI_EACH) ph.v[e] = Rs.ph.v[e] + alpha*Rs.p.v[e];
      if (PHI == PHI_PAIR) {
        V t = Rs.ph.v[e] + alpha1*Rs.pp.v[e];
        ph.v[e] = t + alpha*Rs.p.v[e];
      }
      const bool fl = ((Rs.f >> (8 * e)) & 0xffu) != 0;
      V rr = Rs.r.v[e] - alpha*Rs.ap.v[e];
      rr = fl ? rr : (V) 0;
      rv.v[e] = rr;
      V z = rr*(fl ? (zc2 ? m[1][e] : m[0][e]) : (V) 0);
      acc[e] += colok ? (double) rr * (double) z : 0.0;
    }
    if (colok) {
      if (PHI != PHI_SKIP) run_store<V, CPR>(phi, c, ph);
      run_store<V, CPR>(r, c, rv);
    }
    Rs = load(min(k + DEPTH, kend));
    if (((k - k0 + 1) & (TILE_Z - 1)) == 0 || k == kend) {
#pragma unroll
      for (int off = 32; off >= CPR; off >>= 1)
#pragma unroll
        for (int e = 0; e < CPR; e++)
          acc[e] += __shfl_down(acc[e], off / CPR, WAVE);
#pragma unroll
      for (int off = CPR / 2; off >= 1; off >>= 1)
#pragma unroll
        for (int e = 0; e < off; e++) acc[e] += acc[e + off];
      if ((lane & (WAVE / CPR - 1)) == 0)
        s_tp[ti][lane / (WAVE / CPR)][wv] = acc[0];
#pragma unroll
      for (int e = 0; e < CPR; e++) acc[e] = 0.0;
      ti++;
    }
  };
  int k = k0;
  for (; k + DEPTH - 1 <= kend; k += DEPTH) {
#pragma unroll
    for (int d = 0; d < DEPTH; d++) step(k + d, R[d]);
  }
#pragma unroll
  for (int d = 0; d < DEPTH - 1; d++)
    if (k + d <= kend) step(k + d, R[d]);
  __syncthreads();
  const int ntx_ = (D.nx + TILE_X - 1) / TILE_X;
  const int nty_ = (D.ny + TILE_Y - 1) / TILE_Y;
  const int nout = ti * CPR * (RY / TILE_Y);
  if ((int) threadIdx.x < nout) {
    int t = threadIdx.x;
    int tx = t % CPR, rg = (t / CPR) % (RY / TILE_Y), tz = t / (CPR * (RY / TILE_Y));
    int ty = jb / TILE_Y + rg;
    if (tx < ntx_ && ty < nty_) {
      const double *w = &s_tp[tz][tx][rg * TILE_Y];
      gpart[tx + (long) ntx_ * (ty + (long) nty_ * ((k0 - 1) / TILE_Z + tz))] =
        ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
    }
  }
}

/* Every 50th iteration (src/cuda_solver.cu:208-224): phi += alpha p ... */
__global__ void __launch_bounds__(TILE_THREADS)
update_solution_kernel(DevDom D, const CGState *st, int q,
  const double *__restrict__ ppart, long ntiles,
  const double *__restrict__ p, double *__restrict__ phi)
{
  if (st->done) return;
  double alpha = st->rzs[q & 1]/tile_block_sum(ppart, ntiles);
  TILE_PROLOGUE
  if (!colok) return;
  for (int k = k0; k <= k1; k++) {
    long c = gidx(C, x, j, k);
    phi[c] += alpha*p[c];
  }
}

/* phi += alpha p for the term a PHI_SKIP update left when the solve stops
 * right after it; alpha is the one that update kept. */
template <typename V>
__global__ void __launch_bounds__(TILE_THREADS)
phi_finish_kernel(DevDom D, const CGState *st,
  const V *__restrict__ p, V *__restrict__ phi)
{
  V alpha = (V) st->alpha;
  TILE_PROLOGUE
  if (!colok) return;
  for (int k = k0; k <= k1; k++) {
    long c = gidx(C, x, j, k);
    phi[c] += alpha*p[c];
  }
}

/* (-A) phi at (x, j, k) with periodic neighbours by wrapped index. */
__device__ inline double neg_lap_at(const DevDom &D, const Coef &K,
  const double *__restrict__ f, int x, int j, int k)
{
  const Grid &C = D.cc;
  int xm = wrap_lo(x, D.nx, D.perx), xp = wrap_hi(x, D.nx, D.perx);
  int jm = wrap_lo(j, D.ny, D.pery), jp = wrap_hi(j, D.ny, D.pery);
  int km = wrap_lo(k, D.nz, D.perz), kp = wrap_hi(k, D.nz, D.perz);
  return neg_lap7(K, f[gidx(C, x, j, k)], f[gidx(C, xm, j, k)],
                  f[gidx(C, xp, j, k)], f[gidx(C, x, jm, k)],
                  f[gidx(C, x, jp, k)], f[gidx(C, x, j, km)],
                  f[gidx(C, x, j, kp)], wall_lo(x, D.perx),
                  wall_hi(x, D.nx, D.perx), wall_lo(j, D.pery),
                  wall_hi(j, D.ny, D.pery), wall_lo(k, D.perz),
                  wall_hi(k, D.nz, D.perz));
}

/* b - (-A) phi down one thread's column of a whole tile (planes k0 ..
 * k0 + 7): phi of the column's ten planes is loaded first, then each half
 * of four planes loads its x and y neighbours and b before any of it is
 * formed (32-bit offsets, no access behind a branch).  emit(i, c, rr) gets
 * plane k0 + i in order, with the value neg_lap_at gives bit for bit (the
 * same neighbours and wall masks). */
template <bool MASK, typename EMIT>
__device__ inline void defect_column(const DevDom &D, const Coef &K,
  const double *__restrict__ phi, const double *__restrict__ rhs,
  const uint8_t *__restrict__ fluid, int x, int j, int k0, EMIT emit)
{
  const Grid &C = D.cc;
  const int xm = wrap_lo(x, D.nx, D.perx), xp = wrap_hi(x, D.nx, D.perx);
  const int jm = wrap_lo(j, D.ny, D.pery), jp = wrap_hi(j, D.ny, D.pery);
  const bool wW = wall_lo(x, D.perx), wE = wall_hi(x, D.nx, D.perx);
  const bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);
  const unsigned sz = (unsigned) C.sz;
  const unsigned c0 = (unsigned) gidx(C, x, j, k0);
  const unsigned oW = (unsigned) gidx(C, xm, j, k0),
    oE = (unsigned) gidx(C, xp, j, k0), oS = (unsigned) gidx(C, x, jm, k0),
    oN = (unsigned) gidx(C, x, jp, k0);
  const int kb = wrap_lo(k0, D.nz, D.perz);
  const int kt = wrap_hi(k0 + TILE_Z - 1, D.nz, D.perz);
  double pc[TILE_Z + 2];
  pc[0] = ld32(phi, (unsigned) gidx(C, x, j, kb));
#pragma unroll
  for (int i = 0; i < TILE_Z; i++) pc[i + 1] = ld32(phi, c0 + i * sz);
  pc[TILE_Z + 1] = ld32(phi, (unsigned) gidx(C, x, j, kt));
#pragma unroll
  for (int h = 0; h < TILE_Z; h += 4) {
    double w[4], e[4], so[4], no[4], b[4];
    uint32_t f[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const unsigned o = (unsigned) (h + i) * sz;
      w[i] = ld32(phi, oW + o);
      e[i] = ld32(phi, oE + o);
      so[i] = ld32(phi, oS + o);
      no[i] = ld32(phi, oN + o);
      b[i] = ld32(rhs, c0 + o);
      f[i] = MASK ? (uint32_t) fluid[c0 + o] : 1u;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = k0 + h + i;
      const double a = neg_lap7(K, pc[h + i + 1], w[i], e[i], so[i], no[i],
        pc[h + i], pc[h + i + 2], wW, wE, wS, wN, wall_lo(k, D.perz),
        wall_hi(k, D.nz, D.perz));
      emit(h + i, c0 + (unsigned) (h + i) * sz, b[i] - a, f[i] != 0u);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

/* The tile is whole (the fast paths above apply). */
#define WHOLE_TILE \
  (xb + TILE_X <= D.nx && jb + TILE_Y <= D.ny && k0 + TILE_Z - 1 <= D.nz)

/* ... then r = b - (-A) phi, z = invM r + tile partials of (r, z). */
__global__ void __launch_bounds__(TILE_THREADS)
residual_kernel(DevDom D, Coef K, InvMTab T, const CGState *st,
  const double *__restrict__ phi, const double *__restrict__ rhs,
  const uint8_t *__restrict__ fluid, double *__restrict__ r,
  double *__restrict__ gpart)
{
  if (st->done) return;
  TILE_PROLOGUE
  double acc = 0.0;
  if (WHOLE_TILE) {
    const int wxy = wcount(x, D.nx, D.perx) + 3 * wcount(j, D.ny, D.pery);
    auto emit = [&](int i, unsigned c, double rr, bool fl) {
      st32(r, c, fl ? rr : 0.);
      double z = rr*(fl ? T.v[wxy + 9 * wcount(k0 + i, D.nz, D.perz)] : 0.);
      acc += rr * z;
    };
    if (fluid) defect_column<true>(D, K, phi, rhs, fluid, x, j, k0, emit);
    else defect_column<false>(D, K, phi, rhs, fluid, x, j, k0, emit);
  } else if (colok) {
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      double rr = rhs[c] - neg_lap_at(D, K, phi, x, j, k);
      r[c] = solid(fluid, c) ? 0. : rr;
      double z = rr*(solid(fluid, c) ? 0. : invm_at(T, D, x, j, k));
      acc += rr * z;
    }
  }
  tile_partial(acc, gpart, tile);
}

/* Start from a given phi (the learned guess): r = b - (-A) phi and the tile
 * partials of (r, z). */
__global__ void __launch_bounds__(TILE_THREADS)
cg_init_warm_kernel(DevDom D, Coef K, InvMTab T, const double *__restrict__ rhs,
  const uint8_t *__restrict__ fluid, const double *__restrict__ phi,
  double *__restrict__ r, double *__restrict__ gpart)
{
  TILE_PROLOGUE
  double acc = 0.0;
  if (colok) {
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      double rr = rhs[c] - neg_lap_at(D, K, phi, x, j, k);
      r[c] = solid(fluid, c) ? 0. : rr;
      double z = rr*(solid(fluid, c) ? 0. : invm_at(T, D, x, j, k));
      acc += rr * z;
    }
  }
  tile_partial(acc, gpart, tile);
}

/* Tile partials of the Galerkin system on span{psi, q}: (b, psi), (b, q),
 * (psi, -A psi), (psi, -A q), (q, -A q), over fluid cells. */
__global__ void __launch_bounds__(TILE_THREADS)
galerkin_kernel(DevDom D, Coef K, const uint8_t *__restrict__ fluid,
  const double *__restrict__ psi, const double *__restrict__ q,
  const double *__restrict__ rhs, double *__restrict__ part)
{
  TILE_PROLOGUE
  double a[5] = {0., 0., 0., 0., 0.};
  if (colok) {
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      if (solid(fluid, c)) continue;
      double ps = psi[c], qs = q[c], b = rhs[c];
      double Aps = neg_lap_at(D, K, psi, x, j, k);
      double Aq = neg_lap_at(D, K, q, x, j, k);
      a[0] += b * ps;
      a[1] += b * qs;
      a[2] += ps * Aps;
      a[3] += ps * Aq;
      a[4] += qs * Aq;
    }
  }
  long nt = (long) ntx_ * nty_ * ntz_;
  for (int m = 0; m < 5; m++) {
    tile_partial(a[m], part + m * nt, tile);
    __syncthreads();                  /* tile_partial's LDS is reused */
  }
}

/* One damped Jacobi sweep on (-A) phi = b over the fluid cells,
 * out = phi + omega invM (b - (-A) phi); particle cells hold 0 (the value
 * a solve from phi = 0 leaves there, which the fluid rows see as their
 * neighbours). */
__global__ void __launch_bounds__(TILE_THREADS)
jacobi_sweep_kernel(DevDom D, Coef K, InvMTab T, double omega,
  const double *__restrict__ rhs, const uint8_t *__restrict__ fluid,
  const double *__restrict__ phi, double *__restrict__ out)
{
  TILE_PROLOGUE
  if (!colok) return;
  for (int k = k0; k <= k1; k++) {
    long c = gidx(C, x, j, k);
    double rr = rhs[c] - neg_lap_at(D, K, phi, x, j, k);
    out[c] = solid(fluid, c) ? 0.
                             : phi[c] + omega * (invm_at(T, D, x, j, k) * rr);
  }
}

/* ---------------- Galerkin guess on several directions ------------------ */
/* Up to GMAX directions v_0..v_{m-1} (Gcc arrays, zero in particle cells):
 * per workgroup, the partial sums over its tile's fluid cells of the
 * m(m+1)/2 entries (v_a, -A v_b), a <= b, and of the m entries (b, v_a),
 * in the order [(b, v_0) .. (b, v_{m-1}), (v_0, A v_0), (v_0, A v_1) ..,
 * (v_1, A v_1) ..].  These are plain FP64 sums in a fixed order; the
 * guess they give is not part of the bit-exact plain solve. */
#define GMAX 9
#define GTERMS (GMAX + GMAX * (GMAX + 1) / 2)
struct GVecs { const double *v[GMAX]; };

template <int M>
__global__ void __launch_bounds__(TILE_THREADS)
gram_kernel(DevDom D, Coef K, const uint8_t *__restrict__ fluid, GVecs V,
  const double *__restrict__ rhs, double *__restrict__ part)
{
  TILE_PROLOGUE
  constexpr int T = M + M * (M + 1) / 2;
  double a[T];
#pragma unroll
  for (int t = 0; t < T; t++) a[t] = 0.;
  if (colok) {
    const int xm = wrap_lo(x, D.nx, D.perx), xp = wrap_hi(x, D.nx, D.perx);
    const int jm = wrap_lo(j, D.ny, D.pery), jp = wrap_hi(j, D.ny, D.pery);
    const bool wW = wall_lo(x, D.perx), wE = wall_hi(x, D.nx, D.perx);
    const bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);
#pragma unroll 1
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      if (solid(fluid, c)) continue;
      const int km = wrap_lo(k, D.nz, D.perz), kp = wrap_hi(k, D.nz, D.perz);
      const long cW = gidx(C, xm, j, k), cE = gidx(C, xp, j, k);
      const long cS = gidx(C, x, jm, k), cN = gidx(C, x, jp, k);
      const long cB = gidx(C, x, j, km), cT = gidx(C, x, j, kp);
      const bool wB = wall_lo(k, D.perz), wT = wall_hi(k, D.nz, D.perz);
      double vc[M], Av[M];
#pragma unroll
      for (int q = 0; q < M; q++) {
        const double *f = V.v[q];
        vc[q] = f[c];
        Av[q] = neg_lap7(K, vc[q], f[cW], f[cE], f[cS], f[cN], f[cB], f[cT],
                         wW, wE, wS, wN, wB, wT);
      }
      const double b = rhs[c];
      int t = 0;
#pragma unroll
      for (int q = 0; q < M; q++, t++) a[t] += b * vc[q];
#pragma unroll
      for (int p = 0; p < M; p++)
#pragma unroll
        for (int q = p; q < M; q++, t++) a[t] += vc[p] * Av[q];
    }
  }
  /* per-workgroup sums: wavefront trees, then the 8 rows; the terms go
   * to the rows of `part` in the GMAX layout */
  __shared__ double s[TILE_Y][T];
#pragma unroll
  for (int t = 0; t < T; t++) {
    double v = wave_tree_sum(a[t]);
    if (lane == 0) s[wv][t] = v;
  }
  __syncthreads();
  if (threadIdx.x < T) {
    const int t = threadIdx.x;
    int row;
    if (t < M) {
      row = t;
    } else {                            /* pair (p, q), p <= q < M */
      int r = t - M, p = 0;
      while (r >= M - p) { r -= M - p; p++; }
      const int q = p + r;
      row = GMAX + p * GMAX - p * (p - 1) / 2 + (q - p);
    }
    double v = 0.;
    for (int w = 0; w < TILE_Y; w++) v += s[w][t];
    part[(long) row * gridDim.x + blockIdx.x] = v;
  }
}

/* The same sums for up to two new directions v_n1, v_n2 among M (n1 = -1:
 * none; NN = 2: n2 as well): per workgroup, (b, v_a) for every a into rows
 * 0..M-1, (v_a, -A v_n1) into rows GMAX..GMAX+M-1 and (v_a, -A v_n2) into
 * rows 2 GMAX..2 GMAX+M-1.  Used when the other entries of the Gram matrix
 * are still valid (same directions, same operator). */
template <int M, int NN, bool MASK>
__global__ void __launch_bounds__(TILE_THREADS)
gram_new_kernel(DevDom D, Coef K, const uint8_t *__restrict__ fluid,
  GVecs V, int n1, int n2, const double *__restrict__ rhs,
  double *__restrict__ part)
{
  TILE_PROLOGUE
  constexpr int T = (1 + NN) * M;
  double a[T];
#pragma unroll
  for (int t = 0; t < T; t++) a[t] = 0.;
  if (WHOLE_TILE && n1 >= 0) {
    /* whole tile: the new directions' columns are carried in z, every
     * load by a 32-bit offset and none behind a branch (particle cells add
     * nothing: the mask selects) */
    const int xm = wrap_lo(x, D.nx, D.perx), xp = wrap_hi(x, D.nx, D.perx);
    const int jm = wrap_lo(j, D.ny, D.pery), jp = wrap_hi(j, D.ny, D.pery);
    const bool wW = wall_lo(x, D.perx), wE = wall_hi(x, D.nx, D.perx);
    const bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);
    const unsigned sz = (unsigned) C.sz, c0 = (unsigned) gidx(C, x, j, k0);
    const unsigned oW = (unsigned) gidx(C, xm, j, k0),
      oE = (unsigned) gidx(C, xp, j, k0), oS = (unsigned) gidx(C, x, jm, k0),
      oN = (unsigned) gidx(C, x, jp, k0);
    const unsigned cb = (unsigned) gidx(C, x, j, wrap_lo(k0, D.nz, D.perz)),
      ct = (unsigned) gidx(C, x, j, wrap_hi(k0 + TILE_Z - 1, D.nz, D.perz));
    const double *f[2] = {V.v[n1], V.v[NN == 2 && n2 >= 0 ? n2 : n1]};
    double pm[2], pc[2];
#pragma unroll
    for (int d = 0; d < NN; d++) {
      pm[d] = ld32(f[d], cb);
      pc[d] = ld32(f[d], c0);
    }
#pragma unroll
    for (int i = 0; i < TILE_Z; i++) {
      const unsigned o = (unsigned) i * sz, c = c0 + o;
      const int k = k0 + i;
      const bool wB = wall_lo(k, D.perz), wT = wall_hi(k, D.nz, D.perz);
      double An[2];
#pragma unroll
      for (int d = 0; d < NN; d++) {
        const double pt = ld32(f[d], i == TILE_Z - 1 ? ct : c + sz);
        An[d] = neg_lap7(K, pc[d], ld32(f[d], oW + o), ld32(f[d], oE + o),
                         ld32(f[d], oS + o), ld32(f[d], oN + o), pm[d], pt,
                         wW, wE, wS, wN, wB, wT);
        pm[d] = pc[d];
        pc[d] = pt;
      }
      const bool fl = MASK ? fluid[c] != 0 : true;
      const double bl = ld32(rhs, c);
      const double b = fl ? bl : 0.;
#pragma unroll
      for (int q = 0; q < M; q++) {
        const double vl = ld32(V.v[q], c);
        const double vq = fl ? vl : 0.;
        a[q] += b * vq;
        a[M + q] += vq * An[0];
        if (NN == 2) a[2 * M + q] += vq * An[1];
      }
    }
  } else if (colok) {
    const int xm = wrap_lo(x, D.nx, D.perx), xp = wrap_hi(x, D.nx, D.perx);
    const int jm = wrap_lo(j, D.ny, D.pery), jp = wrap_hi(j, D.ny, D.pery);
    const bool wW = wall_lo(x, D.perx), wE = wall_hi(x, D.nx, D.perx);
    const bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);
    const double *f1 = V.v[n1 >= 0 ? n1 : 0];
    const double *f2 = V.v[NN == 2 && n2 >= 0 ? n2 : 0];
#pragma unroll 1
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      if (solid(fluid, c)) continue;
      const int km = wrap_lo(k, D.nz, D.perz), kp = wrap_hi(k, D.nz, D.perz);
      const bool wB = wall_lo(k, D.perz), wT = wall_hi(k, D.nz, D.perz);
      auto nl = [&](const double *f) {
        return neg_lap7(K, f[c], f[gidx(C, xm, j, k)], f[gidx(C, xp, j, k)],
                        f[gidx(C, x, jm, k)], f[gidx(C, x, jp, k)],
                        f[gidx(C, x, j, km)], f[gidx(C, x, j, kp)],
                        wW, wE, wS, wN, wB, wT);
      };
      const double A1 = n1 >= 0 ? nl(f1) : 0.;
      const double A2 = NN == 2 ? nl(f2) : 0.;
      const double b = rhs[c];
#pragma unroll
      for (int q = 0; q < M; q++) {
        const double vq = V.v[q][c];
        a[q] += b * vq;
        a[M + q] += vq * A1;
        if (NN == 2) a[2 * M + q] += vq * A2;
      }
    }
  }
  __shared__ double s[TILE_Y][T];
#pragma unroll
  for (int t = 0; t < T; t++) {
    double v = wave_tree_sum(a[t]);
    if (lane == 0) s[wv][t] = v;
  }
  __syncthreads();
  if (threadIdx.x < T) {
    const int t = threadIdx.x;
    const int row = t < M ? t : (t < 2 * M ? GMAX + (t - M)
                                           : 2 * GMAX + (t - 2 * M));
    double v = 0.;
    for (int w = 0; w < TILE_Y; w++) v += s[w][t];
    part[(long) row * gridDim.x + blockIdx.x] = v;
  }
}

/* Sum of each of the GTERMS rows of n partials (one workgroup per row). */
__global__ void __launch_bounds__(256)
gram_sum_kernel(const double *__restrict__ part, long n,
  double *__restrict__ out)
{
  __shared__ double s[256];
  const double *row = part + (long) blockIdx.x * n;
  double a = 0.;
  for (long i = threadIdx.x; i < n; i += blockDim.x) a += row[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ void zero_rows_kernel(double *__restrict__ a, long n)
{
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0.;
}

static unsigned nblk_zero(long n) { return (unsigned) ((n + 255) / 256); }

/* phi = sum_q c_q v_q over the fluid cells, 0 in particle cells. */
__global__ void __launch_bounds__(TILE_THREADS)
combine_many_kernel(DevDom D, GVecs V, int m, const double *__restrict__ cf,
  const uint8_t *__restrict__ fluid, double *__restrict__ phi)
{
  TILE_PROLOGUE
  if (!colok) return;
  for (int k = k0; k <= k1; k++) {
    long c = gidx(C, x, j, k);
    double s = 0.;
    for (int q = 0; q < m; q++) s += cf[q] * V.v[q][c];
    phi[c] = solid(fluid, c) ? 0. : s;
  }
}

/* Tile partials of ||b - (-A) phi||^2 (diagnostic true residual). */
__global__ void __launch_bounds__(TILE_THREADS)
true_residual_kernel(DevDom D, Coef K, const double *__restrict__ phi,
  const double *__restrict__ rhs, const uint8_t *__restrict__ fluid,
  double *__restrict__ gpart)
{
  TILE_PROLOGUE
  double acc = 0.0;
  if (WHOLE_TILE) {
    auto emit = [&](int, unsigned, double rr, bool fl) {
      acc = fl ? acc + rr * rr : acc;
    };
    if (fluid) defect_column<true>(D, K, phi, rhs, fluid, x, j, k0, emit);
    else defect_column<false>(D, K, phi, rhs, fluid, x, j, k0, emit);
  } else if (colok) {
    for (int k = k0; k <= k1; k++) {
      if (solid(fluid, gidx(C, x, j, k))) continue;
      double rr = rhs[gidx(C, x, j, k)] - neg_lap_at(D, K, phi, x, j, k);
      acc += rr * rr;
    }
  }
  tile_partial(acc, gpart, tile);
}

/* out = b - (-A) x in fluid cells, 0 in particle cells (interior). */
__global__ void __launch_bounds__(TILE_THREADS)
defect_kernel(DevDom D, Coef K, const double *__restrict__ x_in,
  const double *__restrict__ rhs, const uint8_t *__restrict__ fluid,
  double *__restrict__ out)
{
  TILE_PROLOGUE
  if (!colok) return;
  for (int k = k0; k <= k1; k++) {
    long c = gidx(C, x, j, k);
    out[c] = solid(fluid, c) ? 0. : rhs[c] - neg_lap_at(D, K, x_in, x, j, k);
  }
}

/* Stopping test after the last enqueued update of a batch (one workgroup;
 * the next search_spmv repeats it on the same bits). */
template <bool MIXED>
__global__ void __launch_bounds__(TILE_THREADS)
cg_check_kernel(CGState *st, int q, const double *__restrict__ rpart, long n)
{
  if (st->done | st->inner_done) return;
  double beta;
  cg_continue<MIXED>(st, rpart, n, q + 1, &beta);
}

/* --------------------- mixed form: FP64 corrections ---------------------- */
/* Cold start: phi = 0, e = 0, r = b (in FP32) + tile partials of (b, M^-1 b)
 * in FP64. */
__global__ void __launch_bounds__(TILE_THREADS)
mx_init_kernel(DevDom D, InvMTab T, const double *__restrict__ rhs,
  const uint8_t *__restrict__ fluid, double *__restrict__ phi,
  float *__restrict__ e, float *__restrict__ r, double *__restrict__ gpart)
{
  TILE_PROLOGUE
  double acc = 0.0;
  if (colok) {
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      double t = rhs[c];
      phi[c] = 0.;
      e[c] = 0.f;
      r[c] = solid(fluid, c) ? 0.f : (float) t;
      acc += t * (t*(solid(fluid, c) ? 0. : invm_at(T, D, x, j, k)));
    }
  }
  tile_partial(acc, gpart, tile);
}

/* e = 0. */
__global__ void __launch_bounds__(TILE_THREADS)
mx_zero_kernel(DevDom D, float *__restrict__ e)
{
  TILE_PROLOGUE
  if (!colok) return;
  for (int k = k0; k <= k1; k++) e[gidx(C, x, j, k)] = 0.f;
}

/* phi += e (FP64), e = 0. */
__global__ void __launch_bounds__(TILE_THREADS)
mx_fold_kernel(DevDom D, double *__restrict__ phi, float *__restrict__ e)
{
  TILE_PROLOGUE
  if (!colok) return;
  for (int k = k0; k <= k1; k++) {
    long c = gidx(C, x, j, k);
    phi[c] += (double) e[c];
    e[c] = 0.f;
  }
}

/* The FP64 defect r = b - (-A) phi, handed to the inner sweep in FP32, and
 * the tile partials of (r, M^-1 r) in FP64. */
__global__ void __launch_bounds__(TILE_THREADS)
mx_residual_kernel(DevDom D, Coef K, InvMTab T, const double *__restrict__ phi,
  const double *__restrict__ rhs, const uint8_t *__restrict__ fluid,
  float *__restrict__ r, double *__restrict__ gpart)
{
  TILE_PROLOGUE
  double acc = 0.0;
  if (WHOLE_TILE) {
    const int wxy = wcount(x, D.nx, D.perx) + 3 * wcount(j, D.ny, D.pery);
    auto emit = [&](int i, unsigned c, double rr, bool fl) {
      st32(r, c, fl ? (float) rr : 0.f);
      acc += rr * (rr*(fl ? T.v[wxy + 9 * wcount(k0 + i, D.nz, D.perz)] : 0.));
    };
    if (fluid) defect_column<true>(D, K, phi, rhs, fluid, x, j, k0, emit);
    else defect_column<false>(D, K, phi, rhs, fluid, x, j, k0, emit);
  } else if (colok) {
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      double rr = rhs[c] - neg_lap_at(D, K, phi, x, j, k);
      r[c] = solid(fluid, c) ? 0.f : (float) rr;
      acc += rr * (rr*(solid(fluid, c) ? 0. : invm_at(T, D, x, j, k)));
    }
  }
  tile_partial(acc, gpart, tile);
}

/* After a correction: the FP64 (r, M^-1 r) decides convergence; otherwise
 * it becomes the reference for the next inner sweep.  With `restart` the
 * next iteration q + 1 starts a fresh Krylov space (classic defect
 * correction); otherwise the search direction is kept and beta is formed
 * from the FP64 value by the next search launch (a reliable update). */
__global__ void __launch_bounds__(TILE_THREADS)
mx_state_kernel(CGState *st, const double *__restrict__ rpart, long n, int q,
  int restart)
{
  double s = tile_block_sum(rpart, n);
  if (threadIdx.x != 0) return;
  st->rz1 = s;
  if (s <= st->tol2*st->sp_rhs || isnan(s)) {
    st->done = 1;
    st->nan_flag = isnan(s) ? 1 : 0;
    st->q = q;
    return;
  }
  st->rz_ref = s;
  st->inner_done = 0;
  if (restart) st->rzs[(q + 1) & 1] = s;
}

/* ------------------------------- driver --------------------------------- */
static double now_s(void)
{
  struct timeval t;
  gettimeofday(&t, 0);
  return t.tv_sec + t.tv_usec * 1e-6;
}

#define TILES_LAUNCH tiles_grid(ddom), tiles_block(), 0, bb_stream

/* The particle mask of the PCG, NULL without particles. */
static const uint8_t *fluid_mask(void)
{
  return NPARTS > 0 ? _bb_fluid : NULL;
}

static CGState *_cg_state = NULL;
static CGState *cg_host = NULL;
/* The Gram matrix of the last call, kept for its directions (pointer and
 * generation) under the operator of one particle-mask generation: a later
 * call whose directions are those plus at most one new one computes only
 * the new column and the right-hand side terms. */
static const double *gc_ptr[GMAX];
static long gc_gen[GMAX];
static int gc_m = 0;
static long gc_mask = -1;
static double gc_G[GMAX][GMAX];
extern long bb_mask_gen;               /* bb_particle.hip */

/* Gram-system buffers of the multi-direction guess (bb_pp_galerkin_many). */
static double *_gram_part = NULL, *_gram_out = NULL, *_gram_cf = NULL;
static double *gram_host = NULL;
static long gram_blocks = 0;
static double *_p_alt = NULL;          /* second search-direction buffer */
static double *_part_p = NULL;         /* tile partials of (p, Ap) */
static double *_part_r = NULL;         /* tile partials of (r, z) */
static double *_part_w1 = NULL, *_part_w2 = NULL;  /* learned-guess scale */
static long buf_n = 0, part_n = 0;
int bb_pcg_batch = 16;                 /* iterations enqueued per host check */
int bb_pcg_predict = 1;                /* 1: the first check of a sweep comes
                                          after 0.8 of the length the same
                                          sweep (Lamb pass, correction) had
                                          one timestep earlier */
/* Index of the pressure solve within its timestep (the Lamb pass), counted
 * here from the reference's step counter so that it holds under any host
 * loop: set by solve_begin() at the start of every solve. */
static int solve_pass = 0;
static int solve_step = -1;
static void solve_begin(void)
{
  if (stepnum != solve_step) {
    solve_step = stepnum;
    solve_pass = 0;
  } else {
    solve_pass++;
  }
}
/* Length of each inner sweep by Lamb pass and correction (the plain form:
 * correction 0), from the latest timestep that ran it. */
#define SWEEP_PASSES 32
#define SWEEP_OUTER 8
static int sweep_len[2][SWEEP_PASSES][SWEEP_OUTER];

/* Iterations to enqueue before the first check of a sweep. */
static int first_batch(int form, int outer)
{
  const int p = solve_pass;
  int n = bb_pcg_batch;
  if (bb_pcg_predict && p >= 0 && p < SWEEP_PASSES && outer < SWEEP_OUTER) {
    const int m = (int) (0.8 * sweep_len[form][p][outer]);
    if (m > n) n = m;
  }
  return n;
}

static void record_sweep(int form, int outer, int len)
{
  const int p = solve_pass;
  if (p >= 0 && p < SWEEP_PASSES && outer < SWEEP_OUTER)
    sweep_len[form][p][outer] = len;
}
int bb_spmv_variant = 3;               /* 3: search kernel with one
                                          wavefront per x row where the grid
                                          allows it (else 2); 2: over
                                          z-columns of tiles with LDS
                                          planes; 0: direct neighbour loads
                                          (one tile per workgroup, for
                                          cross-checks) */
int bb_row_fp64 = 1;                   /* 1: the FP64 search uses the row
                                          kernel too (256^3 with 500
                                          spheres: 123.8 us against 130.4
                                          for the z-column kernel, with 172
                                          VGPRs, two wavefronts per SIMD);
                                          0: the z-column kernel */
int bb_row_ry = 0;                     /* rows per workgroup of the FP32 row
                                          kernel: 16, else 8 (FP64: 8) */
int bb_spmv_last = -1;                 /* search kernel of the last
                                          launch: 3 row, 2 z-column, 0
                                          direct */
int bb_row_last[4] = {0, 0, 0, 0};     /* the row kernel's last cells per
                                          lane, rows and planes per
                                          workgroup, planes in flight */
int bb_row_depth = 0;                  /* planes of loads in flight in the
                                          FP32 row kernel: 3 or 4, else 2
                                          (FP64: 2) */
int bb_row_pairs = 1;                  /* 1: the FP64 row search with four
                                          cells per lane loads and stores
                                          pairs, 1 KB per instruction (0:
                                          runs of four, 16 B per lane at a
                                          32 B stride) */
int bb_update_variant = 1;             /* 1: update kernel with one
                                          wavefront per x row where the grid
                                          allows it (FP32; FP64 with
                                          bb_urow_fp64), 0: one tile per
                                          workgroup */
int bb_urow_fp64 = 0;                  /* 1: the FP64 update takes the row
                                          kernel too (at 256^3 it measured
                                          slower: pair 177.6 us against
                                          167.2 for the tile kernel) */
int bb_urow_zc = 0;                    /* planes per workgroup of the row
                                          update (multiple of 8 up to 64;
                                          0: automatic) */
int bb_urow_depth = 0;                 /* planes of loads in flight in the
                                          FP32 row update: 4, else 2 (FP64:
                                          2) */
int bb_update_last = -1;               /* update kernel of the last launch:
                                          1 row, 0 tile */
int bb_row_zc = 0;                     /* planes per workgroup of the row
                                          kernel, a multiple of 8 up to 64
                                          (0: automatic) */
int bb_spmv_tzb = 0;                   /* tiles per z-column (0: automatic) */
int bb_spmv_ntp = 1;                   /* 1: FP64 search loads p_old with the
                                          non-temporal hint (0: ordinary
                                          loads, measured slower) */
int bb_spmv_nts = 1;                   /* 1: the FP64 z-column search stores
                                          p_new and Ap with the non-temporal
                                          hint (the row kernels always do) */
int bb_spmv_nts32 = 1;                 /* the same for the FP32 search (with
                                          8-tile columns ordinary stores take
                                          the same time per step, but the
                                          update kernels absorb their
                                          write-back) */
int bb_spmv_tzb32 = 0;                 /* tiles per z-column of the FP32
                                          search (0: automatic) */
int bb_defer_phi = 1;                  /* 1: phi updated every second
                                          iteration with both terms */
int bb_mixed_restart = 0;              /* 1: restart CG at every correction */
long bb_pcg_iters_total = 0;           /* PCG iterations over all solves */
long bb_pcg_solves_total = 0;
/* Iterations of each solve, in order (entry n % BB_SOLVE_LOG of solve n). */
#define BB_SOLVE_LOG 8192
int bb_solve_log[BB_SOLVE_LOG];

static void log_solve(int niter)
{
  bb_solve_log[bb_pcg_solves_total % BB_SOLVE_LOG] = niter;
  bb_pcg_iters_total += niter;
  bb_pcg_solves_total++;
}

/* FP32 vectors of the mixed form's inner sweeps (Gcc layout, grid arrays
 * of bb_grid_alloc: x = 1 of every row on a 128-byte boundary). */
static float *_r32 = NULL, *_p32a = NULL, *_p32b = NULL, *_Ap32 = NULL,
             *_e32 = NULL;
static long buf32_n = 0;

static void cg_buffers32_free(void)
{
  float **bufs[5] = {&_r32, &_p32a, &_p32b, &_Ap32, &_e32};
  for (int i = 0; i < 5; i++) {
    bb_grid_free(*bufs[i]);
    *bufs[i] = NULL;
  }
  buf32_n = 0;
}

static void cg_buffers32(void)
{
  if (buf32_n == ddom.cc.n) return;
  cg_buffers32_free();
  float **bufs[5] = {&_r32, &_p32a, &_p32b, &_Ap32, &_e32};
  for (int i = 0; i < 5; i++)
    *bufs[i] = (float *) bb_grid_alloc(ddom.cc.n, sizeof(float));
  buf32_n = ddom.cc.n;
}

/* Set while cuda_PP_cg_noalloc runs: the CG buffers then point into the
 * static scratch array and must not be (re)allocated. */
static int cg_in_scratch = 0;

static void cg_buffers(void)
{
  if (cg_in_scratch) return;
  if (!_cg_state) {
    BB_CHECK(hipMalloc((void **) &_cg_state, sizeof(CGState)));
    BB_CHECK(hipMemset(_cg_state, 0, sizeof(CGState)));
    BB_CHECK(hipHostMalloc((void **) &cg_host, sizeof(CGState),
      hipHostMallocDefault));
  }
  if (buf_n != ddom.cc.n) {
    bb_grid_free(_p_alt);
    _p_alt = (double *) bb_grid_alloc(ddom.cc.n, sizeof(double));
    buf_n = ddom.cc.n;
  }
  long nt = n_tiles(ddom);
  if (part_n != nt) {
    if (_part_p) BB_CHECK(hipFree(_part_p));
    if (_part_r) BB_CHECK(hipFree(_part_r));
    if (_part_w1) BB_CHECK(hipFree(_part_w1));
    if (_part_w2) BB_CHECK(hipFree(_part_w2));
    BB_CHECK(hipMalloc((void **) &_part_p, nt * sizeof(double)));
    BB_CHECK(hipMalloc((void **) &_part_r, nt * sizeof(double)));
    BB_CHECK(hipMalloc((void **) &_part_w1, 5 * nt * sizeof(double)));
    BB_CHECK(hipMalloc((void **) &_part_w2, nt * sizeof(double)));
    part_n = nt;
  }
}

void bb_solver_free(void)
{
  bb_grid_free(_p_alt);
  if (_part_p) BB_CHECK(hipFree(_part_p));
  if (_part_r) BB_CHECK(hipFree(_part_r));
  if (_part_w1) BB_CHECK(hipFree(_part_w1));
  if (_part_w2) BB_CHECK(hipFree(_part_w2));
  _part_w1 = _part_w2 = NULL;
  if (_cg_state) BB_CHECK(hipFree(_cg_state));
  if (cg_host) BB_CHECK(hipHostFree(cg_host));
  cg_buffers32_free();
  _cg_state = NULL;
  cg_host = NULL;
  _p_alt = _part_p = _part_r = NULL;
  buf_n = part_n = 0;
  if (_gram_part) BB_CHECK(hipFree(_gram_part));
  if (_gram_out) BB_CHECK(hipFree(_gram_out));
  if (_gram_cf) BB_CHECK(hipFree(_gram_cf));
  if (gram_host) BB_CHECK(hipHostFree(gram_host));
  _gram_part = _gram_out = _gram_cf = gram_host = NULL;
  gram_blocks = 0;
  gc_m = 0;
}

static void fetch_state(void)
{
  BB_CHECK(hipMemcpyAsync(cg_host, _cg_state, sizeof(CGState),
    hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
  /* launches past the iteration the solve (or inner sweep) stopped at
   * returned at once: their timing samples are dropped */
  if (bb_time_kernels)
    bb_kt_collect_upto((cg_host->done || cg_host->inner_done)
                       ? (long) cg_host->q : LONG_MAX);
}

/* rhs and the tile partials of (b, b) in _part_p. */
void bb_pp_rhs(void)
{
  cg_buffers();
  hipLaunchKernelGGL(rhs_kernel, TILES_LAUNCH, ddom, 1. / dom[0].dx,
    1. / dom[0].dy, 1. / dom[0].dz, rho_f / dt, _u_star, _v_star, _w_star,
    fluid_mask(), _rhs_p, _part_p);
  BB_LAUNCH_CHECK();
}

/* Diagnostic: sqrt(||b + A phi||^2 / (b, b)). */
double bb_true_residual(double sp_rhs)
{
  hipLaunchKernelGGL(true_residual_kernel, TILES_LAUNCH, ddom, coef(), _phi,
    _rhs_p, fluid_mask(), _part_r);
  BB_LAUNCH_CHECK();
  bb_reduce_rows(_part_r, n_tiles(ddom), _bb_scalars + 8);
  double rr = bb_fetch_scalar(8);
  return sp_rhs > 0 ? sqrt(rr) / sqrt(sp_rhs) : 0.;
}


/* Eigen-decomposition of a symmetric m x m matrix (m <= GMAX) by cyclic
 * Jacobi rotations: on return the diagonal of a holds the eigenvalues and
 * the columns of v the eigenvectors. */
static void sym_eigen(int m, double a[GMAX][GMAX], double v[GMAX][GMAX])
{
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++) v[i][j] = i == j ? 1. : 0.;
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = 0.;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) off += a[p][q] * a[p][q];
    if (off < 1e-300) break;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) {
        if (fabs(a[p][q]) < 1e-300) continue;
        double th = 0.5 * (a[q][q] - a[p][p]) / a[p][q];
        double t = (th >= 0. ? 1. : -1.) / (fabs(th) + sqrt(th * th + 1.));
        double c = 1. / sqrt(t * t + 1.), sn = t * c;
        for (int k = 0; k < m; k++) {
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - sn * akq;
          a[k][q] = sn * akp + c * akq;
        }
        for (int k = 0; k < m; k++) {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - sn * aqk;
          a[q][k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < m; k++) {
          double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - sn * vkq;
          v[k][q] = sn * vkp + c * vkq;
        }
      }
  }
}


/* Coefficients c of the Galerkin system G c = g on the leading m
 * directions, through the eigen-decomposition of G with directions of
 * relative eigenvalue below 1e-12 dropped (nearly parallel history
 * vectors); returns the energy c . g they capture, 0 for no guess (c = 0,
 * also when a coefficient is not finite). */
static double galerkin_solve(int m, const double G0[GMAX][GMAX],
  const double g[GMAX], double c[GMAX])
{
  double G[GMAX][GMAX], E[GMAX][GMAX];
  for (int p = 0; p < GMAX; p++) {
    c[p] = 0.;
    for (int q = 0; q < GMAX; q++) G[p][q] = G0[p][q];
  }
  sym_eigen(m, G, E);
  double lmax = 0.;
  for (int q = 0; q < m; q++) lmax = fmax(lmax, G[q][q]);
  if (!(lmax > 0.) || !isfinite(lmax)) return 0.;
  for (int q = 0; q < m; q++) {
    double lam = G[q][q];
    if (!(lam > 1e-12 * lmax)) continue;
    double proj = 0.;
    for (int k = 0; k < m; k++) proj += E[k][q] * g[k];
    for (int k = 0; k < m; k++) c[k] += E[k][q] * proj / lam;
  }
  double energy = 0.;
  for (int q = 0; q < m; q++) energy += c[q] * g[q];
  for (int q = 0; q < m; q++)
    if (!isfinite(c[q])) {
      for (int k = 0; k < GMAX; k++) c[k] = 0.;
      return 0.;
    }
  return energy > 0. ? energy : 0.;
}

/* The initial guess with the least A-norm error in span{v_0 .. v_{m-1}}
 * (m <= GMAX, each zero in particle cells) into _phi: the Gram system
 * G c = g, G_ab = (v_a, -A v_b), g_a = (b, v_a), from one pass over the
 * vectors.  Returns the energy 2 (b, phi0) - (phi0, -A phi0) = c . g the
 * guess captures (0: no guess, phi = 0).  From the same Gram matrix,
 * *e_first gets the energy of the Galerkin guess on v_0 alone and
 * *e_lead that on v_0 .. v_{m-2} (without the last direction), when the
 * pointers are given. */
double bb_pp_galerkin_many(const double *const *vecs, const long *gens, int m,
  double *e_first, double *e_lead)
{
  cg_buffers();
  if (e_first) *e_first = 0.;
  if (e_lead) *e_lead = 0.;
  if (m <= 0) return 0.;
  if (m > GMAX) m = GMAX;
  const long nt = n_tiles(ddom);
  if (gram_blocks != nt) {
    if (_gram_part) BB_CHECK(hipFree(_gram_part));
    if (!_gram_out) {
      BB_CHECK(hipMalloc((void **) &_gram_out, GTERMS * sizeof(double)));
      BB_CHECK(hipMalloc((void **) &_gram_cf, GMAX * sizeof(double)));
      BB_CHECK(hipHostMalloc((void **) &gram_host,
        (GTERMS + GMAX) * sizeof(double), hipHostMallocDefault));
    }
    BB_CHECK(hipMalloc((void **) &_gram_part, GTERMS * nt * sizeof(double)));
    gram_blocks = nt;
  }
  GVecs V;
  for (int q = 0; q < GMAX; q++) V.v[q] = q < m ? vecs[q] : vecs[0];
  /* which directions the kept Gram matrix already holds */
  int idx[GMAX], nnew = 0, inew[2] = {-1, -1};
  for (int q = 0; q < m; q++) {
    idx[q] = -1;
    if (gens && gc_mask == bb_mask_gen)
      for (int r = 0; r < gc_m; r++)
        if (gc_ptr[r] == vecs[q] && gc_gen[r] == gens[q]) idx[q] = r;
    if (idx[q] < 0) {
      if (nnew < 2) inew[nnew] = q;
      nnew++;
    }
  }
  const bool incr = nnew <= 2;
  const uint8_t *fl = fluid_mask();
  /* rows of the terms a smaller m leaves out are summed too: zero them */
  hipLaunchKernelGGL(zero_rows_kernel, dim3(nblk_zero(GTERMS * nt)),
    dim3(256), 0, bb_stream, _gram_part, (long) GTERMS * nt);
  if (incr && nnew <= 1) {
    switch (m) {
#define GN(M) case M: if (fl) hipLaunchKernelGGL((gram_new_kernel<M, 1, true>), \
    TILES_LAUNCH, ddom, coef(), fl, V, inew[0], -1, _rhs_p, _gram_part);     \
    else hipLaunchKernelGGL((gram_new_kernel<M, 1, false>), TILES_LAUNCH,     \
    ddom, coef(), fl, V, inew[0], -1, _rhs_p, _gram_part); break;
      GN(1) GN(2) GN(3) GN(4) GN(5) GN(6) GN(7) GN(8) GN(9)
#undef GN
    }
  } else if (incr) {
    switch (m) {
#define GN(M) case M: if (fl) hipLaunchKernelGGL((gram_new_kernel<M, 2, true>), \
    TILES_LAUNCH, ddom, coef(), fl, V, inew[0], inew[1], _rhs_p,            \
    _gram_part); else hipLaunchKernelGGL((gram_new_kernel<M, 2, false>),   \
    TILES_LAUNCH, ddom, coef(), fl, V, inew[0], inew[1], _rhs_p,            \
    _gram_part); break;
      GN(2) GN(3) GN(4) GN(5) GN(6) GN(7) GN(8) GN(9)
#undef GN
    }
  } else {
    switch (m) {
#define GK(M) case M: hipLaunchKernelGGL(gram_kernel<M>, TILES_LAUNCH, ddom, \
    coef(), fluid_mask(), V, _rhs_p, _gram_part); break;
      GK(1) GK(2) GK(3) GK(4) GK(5) GK(6) GK(7) GK(8) GK(9)
#undef GK
    }
  }
  hipLaunchKernelGGL(gram_sum_kernel, dim3(GTERMS), dim3(256), 0, bb_stream,
    _gram_part, nt, _gram_out);
  BB_CHECK(hipMemcpyAsync(gram_host, _gram_out, GTERMS * sizeof(double),
    hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
  double g[GMAX], G[GMAX][GMAX];
  for (int q = 0; q < GMAX; q++) g[q] = gram_host[q];
  if (incr) {
    for (int p = 0; p < m; p++)
      for (int q = 0; q < m; q++)
        if (idx[p] >= 0 && idx[q] >= 0) G[p][q] = gc_G[idx[p]][idx[q]];
    for (int d = 0; d < 2; d++)
      if (inew[d] >= 0)
        for (int a = 0; a < m; a++)
          G[a][inew[d]] = G[inew[d]][a] = gram_host[(1 + d) * GMAX + a];
  } else {
    int t = GMAX;
    for (int p = 0; p < GMAX; p++)
      for (int q = p; q < GMAX; q++, t++) G[p][q] = G[q][p] = gram_host[t];
  }
  /* keep this call's matrix for the next */
  for (int p = 0; p < m; p++) {
    gc_ptr[p] = vecs[p];
    gc_gen[p] = gens ? gens[p] : -1;
    for (int q = 0; q < m; q++) gc_G[p][q] = G[p][q];
  }
  gc_m = gens ? m : 0;
  gc_mask = bb_mask_gen;
  /* The nested Galerkin solutions on v_0 alone, on v_0 .. v_{m-2} and on
   * all m directions.  In exact arithmetic each captures at least the
   * energy of the one before; the eigen-solve's dropping of near-null
   * directions of nearly parallel vectors can break that by rounding, so
   * the guess takes whichever captures the most, and is never worse in
   * the A-norm than the smaller spans it reports. */
  double c[GMAX], cl[GMAX];
  const double ef = G[0][0] > 0. ? g[0] * g[0] / G[0][0] : 0.;
  const double el = m > 1 ? galerkin_solve(m - 1, G, g, cl) : 0.;
  double energy = galerkin_solve(m, G, g, c);
  if (e_first) *e_first = ef;
  if (e_lead) *e_lead = el;
  if (m > 1 && el > energy) {
    for (int q = 0; q < GMAX; q++) c[q] = q < m - 1 ? cl[q] : 0.;
    energy = el;
  }
  if (ef > energy) {
    for (int q = 0; q < GMAX; q++) c[q] = 0.;
    c[0] = g[0] / G[0][0];
    energy = ef;
  }
  if (energy <= 0.) return 0.;
  double *ch = gram_host + GTERMS;
  for (int q = 0; q < m; q++) ch[q] = c[q];
  BB_CHECK(hipMemcpyAsync(_gram_cf, ch, m * sizeof(double),
    hipMemcpyHostToDevice, bb_stream));
  hipLaunchKernelGGL(combine_many_kernel, TILES_LAUNCH, ddom, V, m, _gram_cf,
    fluid_mask(), _phi);
  BB_LAUNCH_CHECK();
  return energy;
}

/* nsweeps damped Jacobi sweeps (rounded up to even) on the initial guess
 * in _phi for the right-hand side in _rhs_p.  The iteration matrix
 * I - omega invM (-A) is self-adjoint in the A inner product with
 * eigenvalues in (-1, 1) for 0 < omega < 1, so each sweep lowers the
 * guess's A-norm error, damping its high-frequency part fastest.  The
 * second search-direction buffer serves as the other half of the
 * ping-pong (the solve that follows overwrites it). */
void bb_pp_smooth(int nsweeps, double omega)
{
  if (nsweeps <= 0) return;
  cg_buffers();
  Coef K = coef();
  InvMTab T = invm_table(K);
  for (int s = 0; s < nsweeps; s += 2) {
    hipLaunchKernelGGL(jacobi_sweep_kernel, TILES_LAUNCH, ddom, K, T, omega,
      _rhs_p, fluid_mask(), _phi, _p_alt);
    hipLaunchKernelGGL(jacobi_sweep_kernel, TILES_LAUNCH, ddom, K, T, omega,
      _rhs_p, fluid_mask(), _p_alt, _phi);
  }
  BB_LAUNCH_CHECK();
}

/* (b, x) over the fluid cells of the current right-hand side. */
double bb_pp_energy(const double *x)
{
  cg_buffers();
  const long nt = n_tiles(ddom);
  hipLaunchKernelGGL(galerkin_kernel, TILES_LAUNCH, ddom, coef(), fluid_mask(),
    x, x, _rhs_p, _part_w1);
  bb_reduce_rows(_part_w1, nt, _bb_scalars + 16);
  return bb_fetch_scalar(16);
}

/* Tiles per workgroup of the z-column search kernel: as many as keep at
 * least four workgroups per CU (1024 on MI355X), at most 4. */
static int spmv_tzb(void)
{
  /* the z-column kernels keep at most 8 tile partials per column */
  if (bb_spmv_tzb > 0) return bb_spmv_tzb < 8 ? bb_spmv_tzb : 8;
  long nt = n_tiles(ddom);
  long t = nt / 1024;
  return t < 1 ? 1 : (t > 4 ? 4 : (int) t);
}

/* The FP32 search moves half the bytes per cell: longer columns (up to 8
 * tiles, two workgroups per CU at 256^3) measured faster for it, and
 * slower for the FP64 one. */
static int spmv_tzb32(void)
{
  if (bb_spmv_tzb32 > 0) return bb_spmv_tzb32 < 8 ? bb_spmv_tzb32 : 8;
  if (bb_spmv_tzb > 0) return spmv_tzb();
  long t = n_tiles(ddom) / 512;
  return t < 1 ? 1 : (t > 8 ? 8 : (int) t);
}

static dim3 spmv_group_grid(int tzb)
{
  long ntx = (ddom.nx + TILE_X - 1) / TILE_X;
  long nty = (ddom.ny + TILE_Y - 1) / TILE_Y;
  long ntz = (ddom.nz + TILE_Z - 1) / TILE_Z;
  return dim3((unsigned) (ntx * nty * ((ntz + tzb - 1) / tzb)));
}

/* The defect b - (-A) x of the current right-hand side (ghosts of x must
 * be current). */
void bb_pp_defect(const double *x, double *out)
{
  hipLaunchKernelGGL(defect_kernel, TILES_LAUNCH, ddom, coef(), x, _rhs_p,
    fluid_mask(), out);
  BB_LAUNCH_CHECK();
}

/* Cells per lane of the row kernel on this grid (a whole x row in one
 * wavefront, nx a multiple of it), 0 where it does not apply. */
static int row_cpr(void)
{
  const int nx = ddom.nx;
  if (nx <= WAVE) return 1;
  if (nx <= 2 * WAVE && nx % 2 == 0) return 2;
  if (nx <= 4 * WAVE && nx % 4 == 0) return 4;
  return 0;
}

/* Point x = 1 of every row of `a` aligned to `bytes`. */
static bool rows_aligned(const void *a, size_t esize, size_t bytes)
{
  const Grid &C = ddom.cc;
  uintptr_t x1 = (uintptr_t) a + esize;
  return x1 % bytes == 0 && (C.sy * esize) % bytes == 0;
}

template <bool FIRST, typename V, bool MIXED, int CPR, int RY, int DEPTH>
static void launch_row_t(const CoefT<V> &K, const InvMT<V> &T, int q,
  long nrp, int zc, const V *r, const V *pold, V *pnew, V *Ap)
{
  const long nyb = (ddom.ny + RY - 1) / RY, nzc = (ddom.nz + zc - 1) / zc;
  if constexpr (sizeof(V) == 8 && CPR == 4) {
    if (bb_row_pairs) {
      BB_TIMED_LAUNCH((search_row_kernel<FIRST, V, MIXED, CPR, RY, DEPTH,
                                         true>),
        dim3((unsigned) (nyb * nzc)), dim3(RY * WAVE), 0, bb_stream, ddom, K,
        T, _cg_state, q, _part_r, nrp, zc, r, pold, pnew, Ap, _part_p);
      return;
    }
  }
  BB_TIMED_LAUNCH((search_row_kernel<FIRST, V, MIXED, CPR, RY, DEPTH>),
    dim3((unsigned) (nyb * nzc)), dim3(RY * WAVE), 0, bb_stream, ddom, K, T,
    _cg_state, q, _part_r, nrp, zc, r, pold, pnew, Ap, _part_p);
}

template <bool FIRST, typename V, bool MIXED, int RY, int DEPTH>
static void launch_row_ry(int cpr, const CoefT<V> &K, const InvMT<V> &T,
  int q, long nrp, int zc, const V *r, const V *pold, V *pnew, V *Ap)
{
  if (cpr == 4)
    launch_row_t<FIRST, V, MIXED, 4, RY, DEPTH>(K, T, q, nrp, zc, r, pold,
                                                pnew, Ap);
  else if (cpr == 2)
    launch_row_t<FIRST, V, MIXED, 2, RY, DEPTH>(K, T, q, nrp, zc, r, pold,
                                                pnew, Ap);
  else
    launch_row_t<FIRST, V, MIXED, 1, RY, DEPTH>(K, T, q, nrp, zc, r, pold,
                                                pnew, Ap);
}

/* The row kernel where it applies; false: the caller launches another. */
template <bool FIRST, typename V, bool MIXED>
static bool launch_row(const CoefT<V> &K, const InvMT<V> &T, int q,
  long nrp, const V *r, const V *pold, V *pnew, V *Ap)
{
  if (sizeof(V) == 8 && !bb_row_fp64) return false;
  const int cpr = row_cpr();
  if (!cpr) return false;
  const size_t run = (size_t) cpr * sizeof(V);
  if (!rows_aligned(r, sizeof(V), run) || !rows_aligned(pold, sizeof(V), run)
      || !rows_aligned(pnew, sizeof(V), run) || !rows_aligned(Ap, sizeof(V), run))
    return false;
  /* 8 rows per workgroup (FP32 at 256^3 inside the solver: 61.5 us with 8
   * rows and 32 planes, 64.1 with 16 rows and 16 planes); FP64 always 8
   * (with 16 the 1024-thread workgroup leaves 128 VGPRs per lane and the
   * FP64 kernel spills) */
  int ry = sizeof(V) == 4 && bb_row_ry == 16 ? 16 : 8;
  /* planes per workgroup: about 256 workgroups of RY rows over the grid
   * (one per CU of MI355X), in whole z-tiles, at most 8 of them */
  int zc = bb_row_zc;
  if (zc <= 0 || zc % TILE_Z || zc > ROW_ZCMAX) {
    long nyb = (ddom.ny + ry - 1) / ry;
    long ntz = (ddom.nz + TILE_Z - 1) / TILE_Z;
    long t = nyb * ntz / 256;
    zc = TILE_Z * (int) (t < 1 ? 1 : (t > ROW_ZCMAX / TILE_Z ? ROW_ZCMAX / TILE_Z : t));
  }
  bb_spmv_last = 3;
  bb_row_last[0] = cpr;
  bb_row_last[1] = ry;
  bb_row_last[2] = zc;
  bb_row_last[3] = 2;
  if constexpr (sizeof(V) == 4) {
    /* planes of loads in flight: 2 (as the z-column kernel), 3 or 4 */
    const int depth = bb_row_depth == 3 || bb_row_depth == 4 ? bb_row_depth : 2;
    bb_row_last[3] = depth;
#define RL(RY_, D_) launch_row_ry<FIRST, V, MIXED, RY_, D_>(cpr, K, T, q, nrp, \
    zc, r, pold, pnew, Ap)
    if (ry == 16) {
      if (depth == 4) RL(16, 4); else if (depth == 3) RL(16, 3); else RL(16, 2);
    } else {
      if (depth == 4) RL(8, 4); else if (depth == 3) RL(8, 3); else RL(8, 2);
    }
#undef RL
    return true;
  }
  launch_row_ry<FIRST, V, MIXED, 8, 2>(cpr, K, T, q, nrp, zc, r, pold, pnew,
                                       Ap);
  return true;
}

/* The z-column search kernel; the two hint switches select the template
 * (non-temporal p_old loads: FP64 only, the FP32 kernel loads p_old
 * ordinarily either way). */
template <bool FIRST, typename V, bool MIXED>
static void launch_group(const CoefT<V> &K, const InvMT<V> &T, int q,
  long nrp, const V *r, const V *pold, V *pnew, V *Ap)
{
  if (bb_spmv_variant == 3 &&
      launch_row<FIRST, V, MIXED>(K, T, q, nrp, r, pold, pnew, Ap))
    return;
  bb_spmv_last = 2;
  const bool f32 = sizeof(V) == 4;
  int tzb = f32 ? spmv_tzb32() : spmv_tzb();
  bool nts = (f32 ? bb_spmv_nts32 : bb_spmv_nts) != 0;
  bool ntp = bb_spmv_ntp != 0;
  dim3 grid = spmv_group_grid(tzb);
  /* two planes in flight: three (FP32: 72.4 against 71.5 us at 256^3) and
   * four did not help, and past two the FP64 kernel spills at 8 waves */
#define SG(NTP, NTS) BB_TIMED_LAUNCH((search_spmv_group_kernel<FIRST, V,    \
    MIXED, 2, NTP, NTS>), grid, tiles_block(), 0, bb_stream, ddom, K, T,    \
    _cg_state, q, _part_r, nrp, tzb, r, pold, pnew, Ap, _part_p)
  if (ntp && nts) SG(true, true);
  else if (ntp) SG(true, false);
  else if (nts) SG(false, true);
  else SG(false, false);
#undef SG
}

/* Paired phi updates (bb_defer_phi, plain form): odd iterations leave phi
 * to the next one, except before an iteration that recomputes the true
 * residual (every 50th), whose phi update adds its own term only. */
static bool phi_skips(int q)
{
  return bb_defer_phi && (q & 1) && (q + 1) % 50 != 0;
}

/* The mixed form's inner sweep pairs its updates of the correction e the
 * same way, counting from the first iteration after each FP64 correction
 * (where e is folded into phi and must be whole). */
static int pair_base32 = 1;

static bool phi_skips32(int q)
{
  return bb_defer_phi && ((q - pair_base32) & 1) == 0;
}

template <typename V, bool MASK, int PHI, int CPR, int DEPTH>
static void urow_t(const InvMT<V> &T, int q, long nt, int zc, const V *p,
  const V *Ap, const uint8_t *fm, V *phi, V *r, const V *pprev)
{
  const long nyb = (ddom.ny + 7) / 8, nzc = (ddom.nz + zc - 1) / zc;
  BB_TIMED_LAUNCH((update_row_kernel<V, MASK, PHI, CPR, 8, DEPTH>),
    dim3((unsigned) (nyb * nzc)), dim3(8 * WAVE), 0, bb_stream, ddom, T,
    _cg_state, q, _part_p, nt, zc, p, Ap, fm, phi, r, _part_r, pprev);
}

template <typename V, bool MASK, int PHI, int DEPTH>
static void urow_cpr(int cpr, const InvMT<V> &T, int q, long nt, int zc,
  const V *p, const V *Ap, const uint8_t *fm, V *phi, V *r, const V *pprev)
{
  if (cpr == 4)
    urow_t<V, MASK, PHI, 4, DEPTH>(T, q, nt, zc, p, Ap, fm, phi, r, pprev);
  else if (cpr == 2)
    urow_t<V, MASK, PHI, 2, DEPTH>(T, q, nt, zc, p, Ap, fm, phi, r, pprev);
  else
    urow_t<V, MASK, PHI, 1, DEPTH>(T, q, nt, zc, p, Ap, fm, phi, r, pprev);
}

/* The row update where it applies (phi update `mode`); false: the caller
 * launches update_kernel. */
template <typename V>
static bool launch_update_row(int mode, const InvMT<V> &T, int q, long nt,
  const V *p, const V *Ap, const uint8_t *fm, V *phi, V *r, const V *pprev)
{
  if (!bb_update_variant || (sizeof(V) == 8 && !bb_urow_fp64)) return false;
  const int cpr = row_cpr();
  if (!cpr) return false;
  const size_t run = (size_t) cpr * sizeof(V);
  if (!rows_aligned(p, sizeof(V), run) || !rows_aligned(Ap, sizeof(V), run)
      || !rows_aligned(phi, sizeof(V), run) || !rows_aligned(r, sizeof(V), run)
      || !rows_aligned(pprev, sizeof(V), run)
      || (fm && !rows_aligned(fm, 1, (size_t) cpr)))
    return false;
  /* planes per workgroup: about 256 workgroups of 8 rows (one per CU of
   * MI355X), in whole z-tiles, at most 8 of them (FP32 at 256^3: skip
   * update 38.0 us with 32 planes, 40.1 with 16; tile kernel 44.8) */
  int zc = bb_urow_zc;
  if (zc <= 0 || zc % TILE_Z || zc > ROW_ZCMAX) {
    long nyb = (ddom.ny + 7) / 8;
    long ntz = (ddom.nz + TILE_Z - 1) / TILE_Z;
    long t = nyb * ntz / 256;
    zc = TILE_Z * (int) (t < 1 ? 1 : (t > ROW_ZCMAX / TILE_Z ? ROW_ZCMAX / TILE_Z : t));
  }
  /* two planes in flight (FP32 pair update at 256^3: 76.2 us, 77.6 with
   * four, which take 184 VGPRs; FP64 with four spills) */
  const int depth = sizeof(V) == 4 && bb_urow_depth == 4 ? 4 : 2;
  bb_update_last = 1;
#define UR(M, P, D) urow_cpr<V, M, P, D>(cpr, T, q, nt, zc, p, Ap, fm, phi, r, \
    pprev)
#define UR_MODE(M, D) do {                                                    \
    if (mode == PHI_SKIP) UR(M, PHI_SKIP, D);                                 \
    else if (mode == PHI_PAIR) UR(M, PHI_PAIR, D);                            \
    else UR(M, PHI_EACH, D);                                                  \
  } while (0)
  if constexpr (sizeof(V) == 4) {
    if (depth == 4) {
      if (fm) UR_MODE(true, 4); else UR_MODE(false, 4);
      return true;
    }
  }
  if (fm) UR_MODE(true, 2); else UR_MODE(false, 2);
#undef UR_MODE
#undef UR
  return true;
}

/* Enqueue iteration q; p buffers alternate by parity. */
static void enqueue_iteration(int q, const Coef &K, const InvMTab &T)
{
  const long nt = n_tiles(ddom);
  double *pnew = (q & 1) ? _p_q : _p_alt;
  double *pold = (q & 1) ? _p_alt : _p_q;
  bb_kt_q = q;
  bb_kt_begin(KT_SPMV);
  if (bb_spmv_variant != 0) {
    if (q == 1)
      launch_group<true, double, false>(K, T, q, nt, _r_q, pold, pnew, _Apb_q);
    else
      launch_group<false, double, false>(K, T, q, nt, _r_q, pold, pnew,
                                         _Apb_q);
  } else {
    bb_spmv_last = 0;
    if (q == 1)
      BB_TIMED_LAUNCH(search_spmv_direct_kernel<true>, TILES_LAUNCH, ddom,
        K, T, _cg_state, q, _part_r, nt, _r_q, pold, fluid_mask(), pnew,
        _Apb_q, _part_p);
    else
      BB_TIMED_LAUNCH(search_spmv_direct_kernel<false>, TILES_LAUNCH, ddom,
        K, T, _cg_state, q, _part_r, nt, _r_q, pold, fluid_mask(), pnew,
        _Apb_q, _part_p);
  }
  bb_kt_end(KT_SPMV);
  if (q % 50 == 0) {
    hipLaunchKernelGGL(update_solution_kernel, TILES_LAUNCH, ddom, _cg_state,
      q, _part_p, nt, pnew, _phi);
    hipLaunchKernelGGL(residual_kernel, TILES_LAUNCH, ddom, K, T, _cg_state,
      _phi, _rhs_p, fluid_mask(), _r_q, _part_r);
  } else {
    int mode = phi_skips(q) ? PHI_SKIP
             : (q > 1 && phi_skips(q - 1)) ? PHI_PAIR : PHI_EACH;
    int kt = mode == PHI_SKIP ? KT_UPDATE_SKIP
           : mode == PHI_PAIR ? KT_UPDATE_PAIR : KT_UPDATE;
    bb_kt_begin(kt);
    const uint8_t *fm = fluid_mask();
#define UPD(M, P) BB_TIMED_LAUNCH((update_kernel<double, M, P>),             \
      TILES_LAUNCH, ddom, T, _cg_state, q, _part_p, nt, pnew, _Apb_q, fm,   \
      _phi, _r_q, _part_r, pold)
    if (!launch_update_row<double>(mode, T, q, nt, pnew, _Apb_q, fm, _phi,
                                   _r_q, pold)) {
      bb_update_last = 0;
      if (fm) {
        if (mode == PHI_SKIP) UPD(true, PHI_SKIP);
        else if (mode == PHI_PAIR) UPD(true, PHI_PAIR);
        else UPD(true, PHI_EACH);
      } else {
        if (mode == PHI_SKIP) UPD(false, PHI_SKIP);
        else if (mode == PHI_PAIR) UPD(false, PHI_PAIR);
        else UPD(false, PHI_EACH);
      }
    }
#undef UPD
    bb_kt_end(kt);
  }
  BB_LAUNCH_CHECK();
}

/* The plain FP64 Jacobi-PCG (src/cuda_solver.cu:574-761) on the RHS already
 * in _rhs_p with the tile partials of (b, b) in _part_p. */
static void pcg_plain(double ts, int warm)
{
  solve_begin();
  Coef K = coef();
  InvMTab T = invm_table(K);
  const char *rname = "solver_expd.rec";
  const long nt = n_tiles(ddom);
  cg_buffers();

  if (warm)
    hipLaunchKernelGGL(cg_init_warm_kernel, TILES_LAUNCH, ddom, K, T, _rhs_p,
      fluid_mask(), _phi, _r_q, _part_r);
  else
    hipLaunchKernelGGL(cg_init_kernel, TILES_LAUNCH, ddom, T, _rhs_p,
      fluid_mask(), _r_q, _phi, _part_r);
  hipLaunchKernelGGL(cg_state_init_kernel, dim3(1), dim3(TILE_THREADS), 0,
    bb_stream, _part_p, _part_r, nt, _cg_state, pp_residual*pp_residual, 0.);
  BB_LAUNCH_CHECK();
  fetch_state();
  real sp_rhs = cg_host->sp_rhs;
  real RHS_TOL = 1.e-8;

  pp_last.sp_rhs = sp_rhs;
  pp_last.nouter = 1;
  if (sp_rhs < RHS_TOL*RHS_TOL) {
    double etime = now_s() - ts;
    recorder_PP(rname, 0, 0., etime);
    if (rank == 0)
      printf("N%d >> Norm of the rhs is less than %.1e, exiting solver\n",
        rank, RHS_TOL);
    pp_last.niter = 0; pp_last.resid = 0.; pp_last.true_resid = 0.;
    pp_last.etime = etime;
    log_solve(0);
    return;
  }
  int q = 0;
  while (q <= pp_max_iter) {
    int qend = q + (q == 0 ? first_batch(0, 0) : bb_pcg_batch);
    if (qend > pp_max_iter + 1) qend = pp_max_iter + 1;
    for (int qq = q + 1; qq <= qend; qq++) enqueue_iteration(qq, K, T);
    hipLaunchKernelGGL(cg_check_kernel<false>, dim3(1), dim3(TILE_THREADS),
      0, bb_stream, _cg_state, qend, _part_r, nt);
    q = qend;
    fetch_state();
    if (cg_host->done) break;
  }
  if (cg_host->done) record_sweep(0, 0, cg_host->q);
  /* a solve that stops right after a PHI_SKIP update: add its term */
  if (cg_host->done && cg_host->q >= 1 && phi_skips(cg_host->q)) {
    hipLaunchKernelGGL(phi_finish_kernel<double>, TILES_LAUNCH, ddom,
      _cg_state, (cg_host->q & 1) ? _p_q : _p_alt, _phi);
    BB_LAUNCH_CHECK();
  }
  real sp_rq1 = cg_host->rz1;
  if (cg_host->done && cg_host->nan_flag) {
    if (rank == 0) {
      printf("N%d >> The PP equation did not converge.\n", rank);
      printf("N%d >> The residual at iteration %d is nan (%lf).\n", rank,
        cg_host->q, sqrt(sp_rq1));
    }
    exit(EXIT_FAILURE);
  }
  if (!cg_host->done) {
    printf("N%d >> The pressure-Poisson equation did not converge.\n", rank);
    printf("N%d >> (rhs, rhs) is %e\n", rank, sp_rhs);
    printf("N%d >> Residual at iteration %d is %lf\n", rank, q,
      sqrt(sp_rq1)/sqrt(sp_rhs));
    exit(EXIT_FAILURE);
  }
  double etime = now_s() - ts;
  recorder_PP(rname, cg_host->q, sqrt(sp_rq1)/sqrt(sp_rhs), etime);
  pp_last.etime = etime;
  pp_last.niter = cg_host->q;
  pp_last.resid = sqrt(sp_rq1)/sqrt(sp_rhs);
  log_solve(cg_host->q);
}

int bb_log_true_residual = 1;

extern "C" int cuda_PP_warmstart_infer(void);
extern "C" void cuda_PP_warmstart_train(void);

extern "C" void cuda_PP_cg_noparts(void)
{
  double ts = now_s();
  bb_pp_rhs();
  int warm = pp_warmstart == PP_WARM_LEARNED ? cuda_PP_warmstart_infer() : 0;
  pcg_plain(ts, warm);
  if (pp_warmstart == PP_WARM_LEARNED) cuda_PP_warmstart_train();
  pp_last.true_resid = (bb_log_true_residual && pp_last.sp_rhs > 0)
    ? bb_true_residual(pp_last.sp_rhs) : 0.;
}

extern "C" void cuda_PP_cg(void)
{
  /* The particle form (src/cuda_solver.cu:39-300): the same iteration with
   * the particle cells masked out (fluid_mask). */
  cuda_PP_cg_noparts();
}

/* ---------------------- allocation-free plain solve ---------------------- */
/* Scratch of cuda_PP_cg_noalloc: the CG state, the second search-direction
 * buffer and the two tile-partial vectors of the plain solve, carved from
 * one static device array (each device holds its own copy until the
 * process ends).  Host code never passes `scratch` itself to a kernel: its
 * device address comes from hipGetSymbolAddress, once per device. */
#define BB_SCRATCH_BYTES ((size_t) 64 << 20)
#define BB_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[BB_SCRATCH_BYTES];
static unsigned char *scratch_dev[BB_MAX_DEVICES];
static CGState cg_host_plain;          /* pageable host copy of the state */
extern "C" int recorder_enabled;       /* host/recorder.c */

__global__ void scratch_zero_kernel(unsigned long long *p, long n)
{
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  for (; t < n; t += (long) gridDim.x * blockDim.x) p[t] = 0ull;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t) 255; }

/* Bytes of scratch the plain solve needs on the current grid. */
static size_t noalloc_bytes(void)
{
  const long nt = n_tiles(ddom);
  return align256(sizeof(CGState))
       + align256((ddom.cc.n + GRID_ORIGIN + 1) * sizeof(double))
       + 2 * align256(nt * sizeof(double));
}

extern "C" int cuda_PP_cg_noalloc(void)
{
  int dev = 0;
  BB_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= BB_MAX_DEVICES) {
    fprintf(stderr, "cuda_PP_cg_noalloc: device %d beyond the %d supported\n",
      dev, BB_MAX_DEVICES);
    return 1;
  }
  if (!scratch_dev[dev]) {
    void *ptr = NULL;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == NULL) {
      fprintf(stderr, "cuda_PP_cg_noalloc: no scratch address on device %d "
        "(%s)\n", dev, hipGetErrorString(e));
      return 1;
    }
    scratch_dev[dev] = (unsigned char *) ptr;
  }
  const size_t need = noalloc_bytes();
  if (need > BB_SCRATCH_BYTES) {
    fprintf(stderr, "cuda_PP_cg_noalloc: the grid needs %zu bytes of scratch, "
      "more than the %zu available (at most about %zu cells)\n", need,
      BB_SCRATCH_BYTES, BB_SCRATCH_BYTES / sizeof(double) * 15 / 16);
    return 2;
  }
  if (pp_warmstart == PP_WARM_LEARNED) {
    fprintf(stderr, "cuda_PP_cg_noalloc: the learned initial guess "
      "allocates its network; use cuda_PP_cg\n");
    return 3;
  }
  /* carve the scratch and clear it (a fill kernel, no memset node) */
  unsigned char *base = scratch_dev[dev];
  const long nt = n_tiles(ddom);
  CGState *st = (CGState *) base;
  base += align256(sizeof(CGState));
  double *palt = (double *) base + GRID_ORIGIN;   /* as bb_grid_alloc */
  base += align256((ddom.cc.n + GRID_ORIGIN + 1) * sizeof(double));
  double *pp = (double *) base;
  base += align256(nt * sizeof(double));
  double *pr = (double *) base;
  hipLaunchKernelGGL(scratch_zero_kernel, dim3(1024), dim3(256), 0, bb_stream,
    (unsigned long long *) scratch_dev[dev], (long) (need / 8));
  BB_LAUNCH_CHECK();

  /* run the plain solve on the scratch buffers, then put the library's
   * own buffers back */
  CGState *s_state = _cg_state, *s_host = cg_host;
  double *s_alt = _p_alt, *s_pp = _part_p, *s_pr = _part_r;
  const int s_timers = bb_time_kernels, s_rec = recorder_enabled;
  _cg_state = st;
  cg_host = &cg_host_plain;
  _p_alt = palt;
  _part_p = pp;
  _part_r = pr;
  bb_time_kernels = 0;                 /* its events are created lazily */
  recorder_enabled = 0;
  cg_in_scratch = 1;
  double ts = now_s();
  bb_pp_rhs();
  pcg_plain(ts, 0);
  pp_last.true_resid = (bb_log_true_residual && pp_last.sp_rhs > 0)
    ? bb_true_residual(pp_last.sp_rhs) : 0.;
  BB_CHECK(hipStreamSynchronize(bb_stream));
  cg_in_scratch = 0;
  _cg_state = s_state;
  cg_host = s_host;
  _p_alt = s_alt;
  _part_p = s_pp;
  _part_r = s_pr;
  bb_time_kernels = s_timers;
  recorder_enabled = s_rec;
  return 0;
}

/* Enqueue inner iteration q of the mixed form (FP32 vectors); `first`
 * starts a new Krylov space (p = z). */
static void enqueue_iteration32(int q, bool first, const CoefT<float> &K,
  const InvMT<float> &T)
{
  const long nt = n_tiles(ddom);
  float *pnew = (q & 1) ? _p32a : _p32b;
  float *pold = (q & 1) ? _p32b : _p32a;
  bb_kt_q = q;
  bb_kt_begin(KT_SPMV32);
  if (first)
    launch_group<true, float, true>(K, T, q, nt, _r32, pold, pnew, _Ap32);
  else
    launch_group<false, float, true>(K, T, q, nt, _r32, pold, pnew, _Ap32);
  bb_kt_end(KT_SPMV32);
  int mode = phi_skips32(q) ? PHI_SKIP
           : (q > pair_base32 && phi_skips32(q - 1)) ? PHI_PAIR : PHI_EACH;
  int kt = mode == PHI_SKIP ? KT_UPDATE32_SKIP
         : mode == PHI_PAIR ? KT_UPDATE32_PAIR : KT_UPDATE32;
  bb_kt_begin(kt);
  const uint8_t *fm = fluid_mask();
#define UPD(M, P) BB_TIMED_LAUNCH((update_kernel<float, M, P>),              \
    TILES_LAUNCH, ddom, T, _cg_state, q, _part_p, nt, pnew, _Ap32, fm,      \
    _e32, _r32, _part_r, pold)
  if (!launch_update_row<float>(mode, T, q, nt, pnew, _Ap32, fm, _e32, _r32,
                                pold)) {
    bb_update_last = 0;
    if (fm) {
      if (mode == PHI_SKIP) UPD(true, PHI_SKIP);
      else if (mode == PHI_PAIR) UPD(true, PHI_PAIR);
      else UPD(true, PHI_EACH);
    } else {
      if (mode == PHI_SKIP) UPD(false, PHI_SKIP);
      else if (mode == PHI_PAIR) UPD(false, PHI_PAIR);
      else UPD(false, PHI_EACH);
    }
  }
#undef UPD
  bb_kt_end(kt);
  BB_LAUNCH_CHECK();
}

/* FP32-PCG inside an FP64 defect-correction loop.  The inner sweep is the
 * plain form's Jacobi-PCG with FP32 vectors (search direction, residual,
 * A p and the accumulated correction e) and FP64 dot products; it runs
 * until its (r, z) has fallen by pp_inner_tol^2 since the last correction
 * or reaches the target.  A correction then folds e into phi (FP64), forms
 * the FP64 defect b - (-A) phi, and applies the stopping test of the plain
 * form, (r, M^-1 r) <= pp_residual^2 (b, b), to it in FP64; the solve ends
 * only there.  The inner sweep continues with the defect as its residual:
 * by default keeping its search direction (a reliable update, so the
 * Krylov space survives the correction), or with bb_mixed_restart = 1
 * restarting from p = z (classic defect correction).  No counterpart in the
 * reference; it replaces src/cuda_solver.cu:574-761 when selected. */
static void pcg_mixed(double ts, int warm)
{
  solve_begin();
  Coef K = coef();
  InvMTab T = invm_table(K);
  CoefT<float> K32 = coef32(K);
  InvMT<float> T32 = invm_table32(T);
  const char *rname = "solver_expd.rec";
  const long nt = n_tiles(ddom);
  if (!(pp_inner_tol > 0. && pp_inner_tol < 1.)) {
    fprintf(stderr, "pp_inner_tol must lie in (0, 1), got %g\n", pp_inner_tol);
    exit(EXIT_FAILURE);
  }
  cg_buffers();
  cg_buffers32();

  if (warm) {
    hipLaunchKernelGGL(mx_zero_kernel, TILES_LAUNCH, ddom, _e32);
    hipLaunchKernelGGL(mx_residual_kernel, TILES_LAUNCH, ddom, K, T, _phi,
      _rhs_p, fluid_mask(), _r32, _part_r);
  } else {
    hipLaunchKernelGGL(mx_init_kernel, TILES_LAUNCH, ddom, T, _rhs_p,
      fluid_mask(), _phi, _e32, _r32, _part_r);
  }
  hipLaunchKernelGGL(cg_state_init_kernel, dim3(1), dim3(TILE_THREADS), 0,
    bb_stream, _part_p, _part_r, nt, _cg_state, pp_residual*pp_residual,
    pp_inner_tol*pp_inner_tol);
  BB_LAUNCH_CHECK();
  fetch_state();
  real sp_rhs = cg_host->sp_rhs;
  real RHS_TOL = 1.e-8;
  pp_last.sp_rhs = sp_rhs;
  pp_last.nouter = 1;
  if (sp_rhs < RHS_TOL*RHS_TOL) {
    double etime = now_s() - ts;
    recorder_PP(rname, 0, 0., etime);
    if (rank == 0)
      printf("N%d >> Norm of the rhs is less than %.1e, exiting solver\n",
        rank, RHS_TOL);
    pp_last.niter = 0; pp_last.resid = 0.; pp_last.true_resid = 0.;
    pp_last.etime = etime;
    log_solve(0);
    return;
  }
  int q = 0, first_q = 1, nouter = 1;
  pair_base32 = 1;
  int done = cg_host->rz1 <= pp_residual*pp_residual*sp_rhs;
  int sweep0 = 0;                       /* q at the start of the sweep */
  while (!done && q <= pp_max_iter && nouter <= pp_max_iter) {
    int qend = q + (q == sweep0 ? first_batch(1, nouter - 1) : bb_pcg_batch);
    if (qend > pp_max_iter + 1) qend = pp_max_iter + 1;
    for (int qq = q + 1; qq <= qend; qq++)
      enqueue_iteration32(qq, qq == first_q, K32, T32);
    hipLaunchKernelGGL(cg_check_kernel<true>, dim3(1), dim3(TILE_THREADS),
      0, bb_stream, _cg_state, qend, _part_r, nt);
    fetch_state();
    if (!cg_host->inner_done) {
      q = qend;
      continue;
    }
    q = cg_host->q;
    record_sweep(1, nouter - 1, q - sweep0);
    sweep0 = q;
    /* e must be whole before the fold: add a skipped update's term */
    if (q >= pair_base32 && phi_skips32(q))
      hipLaunchKernelGGL(phi_finish_kernel<float>, TILES_LAUNCH, ddom,
        _cg_state, (q & 1) ? _p32a : _p32b, _e32);
    pair_base32 = q + 1;
    hipLaunchKernelGGL(mx_fold_kernel, TILES_LAUNCH, ddom, _phi, _e32);
    hipLaunchKernelGGL(mx_residual_kernel, TILES_LAUNCH, ddom, K, T, _phi,
      _rhs_p, fluid_mask(), _r32, _part_r);
    hipLaunchKernelGGL(mx_state_kernel, dim3(1), dim3(TILE_THREADS), 0,
      bb_stream, _cg_state, _part_r, nt, q, bb_mixed_restart);
    BB_LAUNCH_CHECK();
    nouter++;
    fetch_state();
    done = cg_host->done;
    if (bb_mixed_restart) first_q = q + 1;
  }
  real sp_rq1 = cg_host->rz1;
  if (done && (cg_host->nan_flag || isnan(sp_rq1))) {
    if (rank == 0) {
      printf("N%d >> The PP equation did not converge.\n", rank);
      printf("N%d >> The residual at iteration %d is nan (%lf).\n", rank, q,
        sqrt(sp_rq1));
    }
    exit(EXIT_FAILURE);
  }
  if (!done) {
    printf("N%d >> The pressure-Poisson equation did not converge.\n", rank);
    printf("N%d >> (rhs, rhs) is %e\n", rank, sp_rhs);
    printf("N%d >> Residual at iteration %d is %lf\n", rank, q,
      sqrt(sp_rq1)/sqrt(sp_rhs));
    exit(EXIT_FAILURE);
  }
  double etime = now_s() - ts;
  recorder_PP(rname, q, sqrt(sp_rq1)/sqrt(sp_rhs), etime);
  pp_last.etime = etime;
  pp_last.niter = q;
  pp_last.nouter = nouter;
  pp_last.resid = sqrt(sp_rq1)/sqrt(sp_rhs);
  log_solve(q);
}

extern "C" void cuda_PP_cg_mixed(void)
{
  double ts = now_s();
  bb_pp_rhs();
  int warm = pp_warmstart == PP_WARM_LEARNED ? cuda_PP_warmstart_infer() : 0;
  pcg_mixed(ts, warm);
  if (pp_warmstart == PP_WARM_LEARNED) cuda_PP_warmstart_train();
  pp_last.true_resid = (bb_log_true_residual && pp_last.sp_rhs > 0)
    ? bb_true_residual(pp_last.sp_rhs) : 0.;
}
