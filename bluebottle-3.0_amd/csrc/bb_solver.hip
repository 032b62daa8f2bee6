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
template <bool MIXED>
__device__ inline bool cg_continue(CGState *st, const double *__restrict__ rpart,
  long n, int q, double *beta)
{
  double s = tile_block_sum(rpart, n);
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

/* Raw r and p_old of a thread's two entries of one plane. */
template <typename V> struct PlaneRaw {
  V r0, p0, r1, p1;
};

/* Element `i` of a Gcc array by a 32-bit byte offset from the (uniform)
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

template <bool FIRST, typename V>
__device__ inline PlaneRaw<V> plane_load(const DevDom &D,
  const PlaneEntry &E0, const PlaneEntry &E1, const V *__restrict__ r,
  const V *__restrict__ pold, const uint8_t *__restrict__ fluid, int kk)
{
  PlaneRaw<V> R;
  R.r0 = R.p0 = R.r1 = R.p1 = 0;
  int ks = plane_src(D, kk);
  unsigned off = (unsigned) ks * (unsigned) D.cc.sz;
  if (ks && E0.cxy >= 0) {
    R.r0 = ld32(r, E0.cxy + off);
    if (!FIRST) R.p0 = ld32(pold, E0.cxy + off);
    if (fluid != NULL && fluid[E0.cxy + off] == 0) R.r0 = 0;
  }
  if (ks && E1.cxy >= 0) {
    R.r1 = ld32(r, E1.cxy + off);
    if (!FIRST) R.p1 = ld32(pold, E1.cxy + off);
    if (fluid != NULL && fluid[E1.cxy + off] == 0) R.r1 = 0;
  }
  return R;
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

/* `tab`: the 27-entry invM table (kernel argument or a workgroup's LDS
 * copy). */
template <bool FIRST, typename V>
__device__ inline void plane_store(const DevDom &D, const V *tab,
  const PlaneEntry &E0, const PlaneEntry &E1, const PlaneRaw<V> &R, V beta,
  V *__restrict__ buf, int kk)
{
  int ks = plane_src(D, kk);
  int wz = 9 * wcount(ks, D.nz, D.perz);
  V v0 = 0, v1 = 0;
  if (ks && E0.cxy >= 0) {
    V z = R.r0*tab[E0.wxy + wz];
    v0 = FIRST ? z : z + beta*R.p0;
  }
  if (ks && E1.cxy >= 0) {
    V z = R.r1*tab[E1.wxy + wz];
    v1 = FIRST ? z : z + beta*R.p1;
  }
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

/* The tile's planes stream through two LDS buffers; the raw r and p_old of
 * each plane are loaded one (FP64) or two (FP32) planes ahead of their use,
 * the first three before the reduction that yields beta, so memory latency
 * overlaps the work of the planes in between. */
template <bool FIRST, typename V, bool MIXED>
__global__ void __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(8, 8)))
search_spmv_kernel(DevDom D, CoefT<V> K, InvMT<V> T, CGState *st, int q,
  const double *__restrict__ rpart, long ntiles,
  const V *__restrict__ r, const V *__restrict__ pold,
  const uint8_t *__restrict__ fluid,
  V *__restrict__ pnew, V *__restrict__ Ap, double *__restrict__ gpart)
{
  __shared__ V lds[2 * LPLANE];
  if (st->done | st->inner_done) return;
  TILE_PROLOGUE
  PlaneEntry E0 = plane_entry(D, xb, jb, threadIdx.x);
  PlaneEntry E1 = plane_entry(D, xb, jb, threadIdx.x + TILE_THREADS);
  /* Ring of raw planes in registers: slot (p - k0 + 1) mod RING holds plane
   * p; plane k + RING - 1 is loaded while plane k is computed. */
  constexpr int RING = sizeof(V) == 8 ? 3 : 4;
  PlaneRaw<V> R[RING];
  R[0] = plane_load<FIRST, V>(D, E0, E1, r, pold, fluid, k0 - 1);
  R[1] = plane_load<FIRST, V>(D, E0, E1, r, pold, fluid, k0);
  R[2] = plane_load<FIRST, V>(D, E0, E1, r, pold, fluid, k0 + 1);
  double beta64 = 0.;
  if (!FIRST && !cg_continue<MIXED>(st, rpart, ntiles, q, &beta64)) return;
  if (RING > 3 && k0 + 2 <= k1 + 1)
    R[3 % RING] = plane_load<FIRST, V>(D, E0, E1, r, pold, fluid, k0 + 2);
  V beta = (V) beta64;
  int xl = lane + 1, yl = wv + 1;
  bool wW = wall_lo(x, D.perx), wE = wall_hi(x, D.nx, D.perx);
  bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);

  plane_store<FIRST, V>(D, T.v, E0, E1, R[0], beta,
                        lds + ((k0 - 1) & 1) * LPLANE, k0 - 1);
  plane_store<FIRST, V>(D, T.v, E0, E1, R[1], beta, lds + (k0 & 1) * LPLANE,
                        k0);
  __syncthreads();
  V prev = lds[((k0 - 1) & 1) * LPLANE + yl * LX + xl];
  V cur = lds[(k0 & 1) * LPLANE + yl * LX + xl];
  double acc = 0.0;
#pragma unroll
  for (int kk = 0; kk < TILE_Z; kk++) {
    const int k = k0 + kk;
    if (k <= k1) {                    /* uniform over the workgroup */
      __syncthreads();
      plane_store<FIRST, V>(D, T.v, E0, E1, R[(kk + 2) % RING], beta,
                            lds + ((k + 1) & 1) * LPLANE, k + 1);
      if (k + RING - 1 <= k1 + 1)
        R[kk % RING] = plane_load<FIRST, V>(D, E0, E1, r, pold, fluid, k + RING - 1);
      __syncthreads();
      const V *P = lds + (k & 1) * LPLANE;
      V nxt = lds[((k + 1) & 1) * LPLANE + yl * LX + xl];
      if (colok) {
        V a = neg_lap7<V>(K, cur, P[yl * LX + xl - 1], P[yl * LX + xl + 1],
                          P[(yl - 1) * LX + xl], P[(yl + 1) * LX + xl], prev,
                          nxt, wW, wE, wS, wN, wall_lo(k, D.perz),
                          wall_hi(k, D.nz, D.perz));
        long c = gidx(C, x, j, k);
        pnew[c] = cur;
        Ap[c] = a;
        acc += (double) cur * (double) a;
      }
      prev = cur;
      cur = nxt;
    }
  }
  tile_partial(acc, gpart, tile);
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

/* ----------------- search + SpMV without LDS planes ---------------------- */
/* Each wavefront owns one row (64 cells along x) of a z-column of tiles and
 * streams it through z on its own: no LDS planes and no barrier per plane.
 * For plane k it forms p_new of its own row at plane k + 1 (the z
 * neighbour) and of rows j - 1 and j + 1 at plane k (loaded again by the
 * neighbouring wavefronts; the cache serves the repeats); the x neighbours
 * come from the lanes on either side (cross-lane shuffles), and the two
 * cells past the row's ends from one extra load.  Every p_new is formed
 * with the same arithmetic as in search_spmv_group_kernel, so the results
 * are the same bits.  Raw values are loaded two planes ahead into two
 * alternating register sets. */

/* Source offset and invM x/y index of cell (xg, jj) after the periodic
 * wrap (cxy -1: a wall ghost or outside, value 0). */
__device__ inline PlaneEntry cell_entry(const DevDom &D, int xg, int jj)
{
  PlaneEntry E;
  int js = jj, xs = xg;
  if (jj < 1) js = D.pery ? D.ny : 0;
  else if (jj > D.ny) js = (D.pery && jj == D.ny + 1) ? 1 : 0;
  if (xg < 1) xs = D.perx ? D.nx : 0;
  else if (xg > D.nx) xs = (D.perx && xg == D.nx + 1) ? 1 : 0;
  bool ok = js && xs;
  E.cxy = ok ? xs + js * D.cc.sy : -1;
  E.wxy = ok ? wcount(xs, D.nx, D.perx) + 3 * wcount(js, D.ny, D.pery) : 0;
  return E;
}

/* Raw r and p_old: own row at plane kc (c), rows j - 1 (s), j + 1 (n) and
 * the end-of-row cell (h) at plane ks. */
template <typename V> struct RowsRaw {
  V rc, pc, rs, ps, rn, pn, rh, ph;
};

template <bool FIRST, typename V>
__device__ inline void rows_ld(const PlaneEntry &E, unsigned off,
  const V *__restrict__ r, const V *__restrict__ pold, V *rv, V *pv)
{
  unsigned c = (unsigned) (E.cxy >= 0 ? E.cxy : 0) + off;
  *rv = ld32(r, c);
  *pv = FIRST ? (V) 0 : ld32(pold, c);
}

template <bool FIRST, typename V>
__device__ inline RowsRaw<V> rows_load(const DevDom &D, const PlaneEntry &EC,
  const PlaneEntry &ES, const PlaneEntry &EN, const PlaneEntry &EH, int kc,
  int ks, const V *__restrict__ r, const V *__restrict__ pold)
{
  RowsRaw<V> R;
  int pc = plane_src(D, kc), psd = plane_src(D, ks);
  unsigned oc = (unsigned) (pc ? pc : 1) * (unsigned) D.cc.sz;
  unsigned os = (unsigned) (psd ? psd : 1) * (unsigned) D.cc.sz;
  rows_ld<FIRST, V>(EC, oc, r, pold, &R.rc, &R.pc);
  rows_ld<FIRST, V>(ES, os, r, pold, &R.rs, &R.ps);
  rows_ld<FIRST, V>(EN, os, r, pold, &R.rn, &R.pn);
  rows_ld<FIRST, V>(EH, os, r, pold, &R.rh, &R.ph);
  return R;
}

/* p_new = z + beta p_old of one entry at plane kk (0 for a zero entry or a
 * wall plane) */
template <bool FIRST, typename V>
__device__ inline V pnew_of(const DevDom &D, const V *tab,
  const PlaneEntry &E, int kk, V rv, V pv, V beta)
{
  int ks = plane_src(D, kk);
  int wz = 9 * wcount(ks, D.nz, D.perz);
  V z = rv*tab[E.wxy + wz];
  V v = FIRST ? z : z + beta*pv;
  return (ks && E.cxy >= 0) ? v : (V) 0;
}

template <typename V>
__device__ inline V shfl_up1(V v)
{
  return __shfl_up(v, 1, WAVE);
}

template <typename V>
__device__ inline V shfl_down1(V v)
{
  return __shfl_down(v, 1, WAVE);
}

template <bool FIRST, typename V, bool MIXED>
__global__ void __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(6, 8)))
search_spmv_rows_kernel(DevDom D, CoefT<V> K, InvMT<V> T, CGState *st,
  int q, const double *__restrict__ rpart, long ntiles, int tzb,
  const V *__restrict__ r, const V *__restrict__ pold,
  V *__restrict__ pnew, V *__restrict__ Ap, double *__restrict__ gpart)
{
  __shared__ V s_invm[27];
  __shared__ double s_tp[8][TILE_Y];
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
  PlaneEntry EC = cell_entry(D, x, j);
  PlaneEntry ES = cell_entry(D, x, j - 1);
  PlaneEntry EN = cell_entry(D, x, j + 1);
  /* the cells past the row's ends: lane 0 needs x - 1, lane 63 x + 1; the
   * other lanes repeat one of the two addresses (two cache lines a wave) */
  PlaneEntry EH = cell_entry(D, lane < WAVE / 2 ? xb : xb + WAVE + 1, j);
  /* raw values: centre planes k0 - 1 and k0; then sets A (centre k0 + 1,
   * rows k0) and B (centre k0 + 2, rows k0 + 1) */
  V rp0, pp0, rc0, pc0;
  {
    int pa = plane_src(D, k0 - 1);
    unsigned oa = (unsigned) (pa ? pa : 1) * (unsigned) D.cc.sz;
    rows_ld<FIRST, V>(EC, oa, r, pold, &rp0, &pp0);
    unsigned ob = (unsigned) k0 * (unsigned) D.cc.sz;
    rows_ld<FIRST, V>(EC, ob, r, pold, &rc0, &pc0);
  }
  RowsRaw<V> Ra = rows_load<FIRST, V>(D, EC, ES, EN, EH, k0 + 1, k0, r, pold);
  RowsRaw<V> Rb = rows_load<FIRST, V>(D, EC, ES, EN, EH, min(k0 + 2, kend + 1),
                                      min(k0 + 1, kend), r, pold);
  double beta64 = 0.;
  if (!FIRST && !cg_continue<MIXED>(st, rpart, ntiles, q, &beta64)) return;
  V beta = (V) beta64;
  __syncthreads();                      /* s_invm */
  bool wW = wall_lo(x, D.perx), wE = wall_hi(x, D.nx, D.perx);
  bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);
  V prev = pnew_of<FIRST, V>(D, s_invm, EC, k0 - 1, rp0, pp0, beta);
  V cur = pnew_of<FIRST, V>(D, s_invm, EC, k0, rc0, pc0, beta);
  double acc = 0.0;
  int ti = 0;
  auto step = [&](int k, RowsRaw<V> &X) {
    V nxt = pnew_of<FIRST, V>(D, s_invm, EC, k + 1, X.rc, X.pc, beta);
    V Sv = pnew_of<FIRST, V>(D, s_invm, ES, k, X.rs, X.ps, beta);
    V Nv = pnew_of<FIRST, V>(D, s_invm, EN, k, X.rn, X.pn, beta);
    V Hv = pnew_of<FIRST, V>(D, s_invm, EH, k, X.rh, X.ph, beta);
    X = rows_load<FIRST, V>(D, EC, ES, EN, EH, min(k + 3, kend + 1),
                            min(k + 2, kend), r, pold);
    V up = shfl_up1(cur), dn = shfl_down1(cur);
    V Wv = lane == 0 ? Hv : up;
    V Ev = lane == WAVE - 1 ? Hv : dn;
    V a = neg_lap7<V>(K, cur, Wv, Ev, Sv, Nv, prev, nxt, wW, wE, wS, wN,
                      wall_lo(k, D.perz), wall_hi(k, D.nz, D.perz));
    unsigned c = colok ? (unsigned) gidx(C, x, j, k) : 0u;
    st32(pnew, c, cur);
    st32(Ap, c, a);
    acc += colok ? (double) cur * (double) a : 0.0;
    prev = cur;
    cur = nxt;
    if (((k - k0 + 1) & (TILE_Z - 1)) == 0 || k == kend) {
      double v = wave_tree_sum(acc);
      if (lane == 0) s_tp[ti][wv] = v;
      acc = 0.0;
      ti++;
    }
  };
  int k = k0;
  for (; k + 1 <= kend; k += 2) {
    step(k, Ra);
    step(k + 1, Rb);
  }
  if (k <= kend) step(k, Ra);
  __syncthreads();
  if (threadIdx.x < ti) {
    const double *w = s_tp[threadIdx.x];
    gpart[tx + (long) ntx_ * (ty + (long) nty_ * (tz0 + (int) threadIdx.x))] =
      ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
  }
}

/* ------------- FP32 inner sweep: two cells per lane (x pairs) ------------- */
/* The mixed form's FP32 vectors halve the bytes per cell, so one cell per
 * lane would leave the FP32 kernels issue-bound.  These kernels give each
 * lane the x pair (x, x + 1), x odd, moved as one float2 (the FP32
 * vectors are offset so that the pair is 8-byte aligned; needs even nx).
 * Tiles are 128 x 8 x 8 cells (64 lanes x 2, eight rows, eight planes)
 * grouped into z-columns of tzb tiles like the FP64 kernels; each tile's
 * partial is the per-lane sum over its planes of (c0 + c1) products,
 * combined by the tile tree.  The LDS row holds [pad, x = xb, x = xb + 1
 * .. xb + 128, x = xb + 129, pad], so interior pairs are 8-byte aligned;
 * raw planes are loaded one plane ahead (Ra holds plane k + 1). */
#define TX2 128
#define LX2 (TX2 + 4)
#define LPLANE2 (LX2 * LY)

__device__ inline int n_tiles_x2(const DevDom &D)
{
  return (D.nx + TX2 - 1) / TX2;
}

/* One plane's raw values held by a thread: pair entries e = t and
 * t + TILE_THREADS (of 64 pairs x LY rows) and halo entry t (of 2 x LY). */
struct Plane2 {
  float2 r0, p0, r1, p1;
  float rh, ph;
};

/* Source cell (x) of local column xl (0: x = xb, 129: x = xb + 129) after
 * the periodic wrap; 0 if the entry is a wall ghost or outside. */
__device__ inline int x2_src(const DevDom &D, int xb, int xl)
{
  int xg = xb + xl;
  if (xg < 1) return D.perx ? D.nx : 0;
  if (xg > D.nx) return (D.perx && xg == D.nx + 1) ? 1 : 0;
  return xg;
}

__device__ inline int y2_src(const DevDom &D, int jj)
{
  if (jj < 1) return D.pery ? D.ny : 0;
  if (jj > D.ny) return (D.pery && jj == D.ny + 1) ? 1 : 0;
  return jj;
}

/* Where a thread's entry of every plane comes from, fixed per workgroup:
 * cxy = x + y*sy of the source (first) cell, -1 for an all-zero entry;
 * info = w0 | w1 << 4 | single << 8 with w0, w1 the x and y part of the
 * invM table index of the two cells and `single` for an entry whose only
 * nonzero cell is the first (the wrapped or wall neighbour x = nx + 1). */
struct Entry2 {
  int cxy, info;
};

__device__ inline Entry2 pair_entry(const DevDom &D, int xb, int jb, int e)
{
  Entry2 E;
  E.cxy = -1;
  E.info = 0;
  int rr = e >> 6, pp = e & 63;
  int js = y2_src(D, jb + rr);
  if (e >= 64 * LY || !js) return E;
  int x0 = xb + 1 + 2 * pp;
  int wy = 3 * wcount(js, D.ny, D.pery);
  if (x0 + 1 <= D.nx) {
    E.cxy = x0 + js * D.cc.sy;
    E.info = (wcount(x0, D.nx, D.perx) + wy) |
             ((wcount(x0 + 1, D.nx, D.perx) + wy) << 4);
  } else {
    int xs = x2_src(D, xb, 1 + 2 * pp);
    if (xs) {
      E.cxy = xs + js * D.cc.sy;
      E.info = (wcount(xs, D.nx, D.perx) + wy) | (1 << 8);
    }
  }
  return E;
}

__device__ inline Entry2 halo_entry(const DevDom &D, int xb, int jb, int t)
{
  Entry2 E;
  E.cxy = -1;
  E.info = 0;
  if (t >= 2 * LY) return E;
  int rr = t >> 1, xl = (t & 1) ? TX2 + 1 : 0;
  int js = y2_src(D, jb + rr), xs = x2_src(D, xb, xl);
  if (js && xs) {
    E.cxy = xs + js * D.cc.sy;
    E.info = wcount(xs, D.nx, D.perx) + 3 * wcount(js, D.ny, D.pery);
  }
  return E;
}

/* Every load is issued unconditionally (a zero entry or plane reads the
 * aligned dummy pair at x = 1 of row 0 instead and is zeroed when stored;
 * a `single` entry reads its pair and keeps the first value), so the
 * compiler can keep loads of the next plane in flight across the loop.
 * The residual is already zero in particle cells. */
template <bool FIRST>
__device__ inline void pair_load(const Entry2 &E, unsigned off,
  const float *__restrict__ r, const float *__restrict__ pold, float2 *rv,
  float2 *pv)
{
  unsigned c = (unsigned) (E.cxy >= 0 ? E.cxy : 1) + off;
  *rv = *(const float2 *) (r + c);
  *pv = FIRST ? make_float2(0.f, 0.f) : *(const float2 *) (pold + c);
}

template <bool FIRST>
__device__ inline Plane2 plane2_load(const DevDom &D, const Entry2 &E0,
  const Entry2 &E1, const Entry2 &EH, int kk, const float *__restrict__ r,
  const float *__restrict__ pold)
{
  Plane2 P;
  int ks = plane_src(D, kk);
  unsigned off = (unsigned) (ks ? ks : 1) * (unsigned) D.cc.sz;
  pair_load<FIRST>(E0, off, r, pold, &P.r0, &P.p0);
  pair_load<FIRST>(E1, off, r, pold, &P.r1, &P.p1);
  unsigned c = (unsigned) (EH.cxy >= 0 ? EH.cxy : 1) + off;
  P.rh = r[c];
  P.ph = FIRST ? 0.f : pold[c];
  return P;
}

/* invM of the entry's cells from the workgroup's LDS copy of the table
 * (wz: z part of the index, -1 for a zero plane) */
template <bool FIRST>
__device__ inline float2 pair_value(const float *s_invm, const Entry2 &E,
  int wz, float2 rv, float2 pv, float beta)
{
  int w = wz < 0 ? 0 : wz;
  float z0 = rv.x * s_invm[(E.info & 15) + w];
  float z1 = rv.y * s_invm[((E.info >> 4) & 15) + w];
  float2 v = FIRST ? make_float2(z0, z1)
                   : make_float2(z0 + beta * pv.x, z1 + beta * pv.y);
  if (E.info & (1 << 8)) v.y = 0.f;
  return (E.cxy < 0 || wz < 0) ? make_float2(0.f, 0.f) : v;
}

template <bool FIRST>
__device__ inline void plane2_store(const DevDom &D, const float *s_invm,
  const Entry2 &E0, const Entry2 &E1, const Entry2 &EH, int kk,
  const Plane2 &P, float beta, float *__restrict__ buf)
{
  int ks = plane_src(D, kk);
  int wz = ks ? 9 * wcount(ks, D.nz, D.perz) : -1;
  int t = threadIdx.x;
  {
    float2 v = pair_value<FIRST>(s_invm, E0, wz, P.r0, P.p0, beta);
    int rr = t >> 6, pp = t & 63;
    *(float2 *) &buf[rr * LX2 + 2 + 2 * pp] = v;
  }
  int e1 = t + TILE_THREADS;
  if (e1 < 64 * LY) {
    float2 v = pair_value<FIRST>(s_invm, E1, wz, P.r1, P.p1, beta);
    int rr = e1 >> 6, pp = e1 & 63;
    *(float2 *) &buf[rr * LX2 + 2 + 2 * pp] = v;
  }
  if (t < 2 * LY) {
    int rr = t >> 1, xl = (t & 1) ? TX2 + 1 : 0;
    float z = P.rh * s_invm[EH.info + (wz < 0 ? 0 : wz)];
    float v = FIRST ? z : z + beta * P.ph;
    buf[rr * LX2 + 1 + xl] = (EH.cxy >= 0 && wz >= 0) ? v : 0.f;
  }
}

template <bool FIRST>
__global__ void __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(8, 8)))
search_spmv_x2_kernel(DevDom D, CoefT<float> K, InvMT<float> T, CGState *st,
  int q, const double *__restrict__ rpart, long nrpart, int tzb,
  const float *__restrict__ r, const float *__restrict__ pold,
  float *__restrict__ pnew, float *__restrict__ Ap, double *__restrict__ gpart)
{
  __shared__ __attribute__((aligned(16))) float lds[3 * LPLANE2];
  __shared__ float s_invm[27];
  if (st->done | st->inner_done) return;
  if (threadIdx.x < 27) s_invm[threadIdx.x] = T.v[threadIdx.x];
  const Grid &C = D.cc;
  const int ntx_ = n_tiles_x2(D);
  const int nty_ = (D.ny + TILE_Y - 1) / TILE_Y;
  const int ntz_ = (D.nz + TILE_Z - 1) / TILE_Z;
  const int ngz = (ntz_ + tzb - 1) / tzb;
  long g = tile_of_block((long) ntx_ * nty_ * ngz);
  int lane = threadIdx.x & (WAVE - 1);
  int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tx = (int) (g % ntx_), ty = (int) ((g / ntx_) % nty_);
  int tz0 = (int) (g / ((long) ntx_ * nty_)) * tzb;
  int xb = tx * TX2, jb = ty * TILE_Y;
  int k0 = tz0 * TILE_Z + 1;
  int kend = k0 + tzb * TILE_Z - 1 < D.nz ? k0 + tzb * TILE_Z - 1 : D.nz;
  int x0 = xb + 2 * lane + 1, j = jb + wv + 1;
  bool colok = (x0 <= D.nx) && (j <= D.ny);
  Entry2 E0 = pair_entry(D, xb, jb, threadIdx.x);
  Entry2 E1 = pair_entry(D, xb, jb, threadIdx.x + TILE_THREADS);
  Entry2 EH = halo_entry(D, xb, jb, threadIdx.x);
  Plane2 Rp = plane2_load<FIRST>(D, E0, E1, EH, k0 - 1, r, pold);
  Plane2 Rc = plane2_load<FIRST>(D, E0, E1, EH, k0, r, pold);
  Plane2 Ra = plane2_load<FIRST>(D, E0, E1, EH, k0 + 1, r, pold);
  double beta64 = 0.;
  if (!FIRST && !cg_continue<true>(st, rpart, nrpart, q, &beta64)) return;
  float beta = (float) beta64;
  __syncthreads();                      /* s_invm */
  int yl = wv + 1, o = yl * LX2 + 2 + 2 * lane;
  bool wW = wall_lo(x0, D.perx), wE = wall_hi(x0 + 1, D.nx, D.perx);
  bool wS = wall_lo(j, D.pery), wN = wall_hi(j, D.ny, D.pery);
  float *bprev = lds, *bcur = lds + LPLANE2, *bnext = lds + 2 * LPLANE2;
  plane2_store<FIRST>(D, s_invm, E0, E1, EH, k0 - 1, Rp, beta, bprev);
  plane2_store<FIRST>(D, s_invm, E0, E1, EH, k0, Rc, beta, bcur);
  __syncthreads();
  float2 prev = *(const float2 *) &bprev[o];
  float2 cur = *(const float2 *) &bcur[o];
  double acc = 0.0;
  long tile = tx + (long) ntx_ * (ty + (long) nty_ * tz0);
  const long tstride = (long) ntx_ * nty_;
  for (int k = k0; k <= kend; k++) {
    plane2_store<FIRST>(D, s_invm, E0, E1, EH, k + 1, Ra, beta, bnext);
    Ra = plane2_load<FIRST>(D, E0, E1, EH, min(k + 2, kend + 1), r, pold);
    __syncthreads();
    float2 nxt = *(const float2 *) &bnext[o];
    if (colok) {
      float2 S = *(const float2 *) &bcur[o - LX2];
      float2 N = *(const float2 *) &bcur[o + LX2];
      float Wv = bcur[o - 1], Ev = bcur[o + 2];
      bool wB = wall_lo(k, D.perz), wT = wall_hi(k, D.nz, D.perz);
      float a0 = neg_lap7<float>(K, cur.x, Wv, cur.y, S.x, N.x, prev.x, nxt.x,
                                 wW, false, wS, wN, wB, wT);
      float a1 = neg_lap7<float>(K, cur.y, cur.x, Ev, S.y, N.y, prev.y, nxt.y,
                                 false, wE, wS, wN, wB, wT);
      unsigned c = (unsigned) gidx(C, x0, j, k);
      *(float2 *) (pnew + c) = cur;
      *(float2 *) (Ap + c) = make_float2(a0, a1);
      acc += (double) cur.x * (double) a0 + (double) cur.y * (double) a1;
    }
    prev = cur;
    cur = nxt;
    float *t3 = bprev;
    bprev = bcur;
    bcur = bnext;
    bnext = t3;
    if (((k - k0 + 1) & (TILE_Z - 1)) == 0 || k == kend) {
      tile_partial(acc, gpart, tile);   /* uniform over the workgroup */
      acc = 0.0;
      tile += tstride;
    }
  }
}

/* alpha = (r, z) / (p, Ap); e += alpha p, r -= alpha Ap, z = invM r and the
 * tile partials of (r, z), on x pairs over z-columns of tiles. */
__global__ void __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(8, 8)))
update_x2_kernel(DevDom D, InvMT<float> T, const CGState *st, int q,
  const double *__restrict__ ppart, long nppart, int tzb,
  const float *__restrict__ p, const float *__restrict__ Ap,
  const uint8_t *__restrict__ fluid, float *__restrict__ e,
  float *__restrict__ r, double *__restrict__ gpart)
{
  __shared__ float s_invm[27];
  if (st->done | st->inner_done) return;
  if (threadIdx.x < 27) s_invm[threadIdx.x] = T.v[threadIdx.x];
  const Grid &C = D.cc;
  const int ntx_ = n_tiles_x2(D);
  const int nty_ = (D.ny + TILE_Y - 1) / TILE_Y;
  const int ntz_ = (D.nz + TILE_Z - 1) / TILE_Z;
  const int ngz = (ntz_ + tzb - 1) / tzb;
  long g = tile_of_block((long) ntx_ * nty_ * ngz);
  int lane = threadIdx.x & (WAVE - 1);
  int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int tx = (int) (g % ntx_), ty = (int) ((g / ntx_) % nty_);
  int tz0 = (int) (g / ((long) ntx_ * nty_)) * tzb;
  int x0 = tx * TX2 + 2 * lane + 1, j = ty * TILE_Y + wv + 1;
  bool colok = (x0 <= D.nx) && (j <= D.ny);
  int tz1 = tz0 + tzb < ntz_ ? tz0 + tzb : ntz_;
  long tile = tx + (long) ntx_ * (ty + (long) nty_ * tz0);
  const long tstride = (long) ntx_ * nty_;
  /* the first tile's p and Ap are in flight during the reduction */
  int kf = tz0 * TILE_Z + 1;
  int kf1 = kf + TILE_Z - 1 < D.nz ? kf + TILE_Z - 1 : D.nz;
  unsigned sz = (unsigned) C.sz;
  float2 pv[TILE_Z], apv[TILE_Z];
  {
    unsigned c0 = (unsigned) gidx(C, x0, j, kf);
#pragma unroll
    for (int kk = 0; kk < TILE_Z; kk++) {
      pv[kk] = apv[kk] = make_float2(0.f, 0.f);
      if (colok && kf + kk <= kf1) {
        pv[kk] = *(const float2 *) (p + c0 + kk * sz);
        apv[kk] = *(const float2 *) (Ap + c0 + kk * sz);
      }
    }
  }
  /* tile_block_sum's barriers also publish s_invm */
  float alpha = (float) (st->rzs[q & 1]/tile_block_sum(ppart, nppart));
  int wj = 3 * wcount(j, D.ny, D.pery);
  int w0 = wcount(x0, D.nx, D.perx) + wj, w1 = wcount(x0 + 1, D.nx, D.perx) + wj;
  for (int tz = tz0; tz < tz1; tz++, tile += tstride) {
    int k0 = tz * TILE_Z + 1;
    int k1 = k0 + TILE_Z - 1 < D.nz ? k0 + TILE_Z - 1 : D.nz;
    unsigned c0 = (unsigned) gidx(C, x0, j, k0);
    if (tz > tz0) {
#pragma unroll
      for (int kk = 0; kk < TILE_Z; kk++) {
        pv[kk] = apv[kk] = make_float2(0.f, 0.f);
        if (colok && k0 + kk <= k1) {
          pv[kk] = *(const float2 *) (p + c0 + kk * sz);
          apv[kk] = *(const float2 *) (Ap + c0 + kk * sz);
        }
      }
    }
    double acc = 0.0;
    if (colok) {
#pragma unroll
      for (int kk = 0; kk < TILE_Z; kk++) {
        if (k0 + kk <= k1) {
          unsigned c = c0 + kk * sz;
          float2 ev = *(const float2 *) (e + c);
          float2 rv = *(const float2 *) (r + c);
          ev.x += alpha*pv[kk].x;
          ev.y += alpha*pv[kk].y;
          rv.x = rv.x - alpha*apv[kk].x;
          rv.y = rv.y - alpha*apv[kk].y;
          rv.x = solid(fluid, c) ? 0.f : rv.x;
          rv.y = solid(fluid, c + 1) ? 0.f : rv.y;
          *(float2 *) (e + c) = ev;
          *(float2 *) (r + c) = rv;
          int wz = 9 * wcount(k0 + kk, D.nz, D.perz);
          float m0 = solid(fluid, c) ? 0.f : s_invm[w0 + wz];
          float m1 = solid(fluid, c + 1) ? 0.f : s_invm[w1 + wz];
          float z0 = rv.x*m0, z1 = rv.y*m1;
          acc += (double) rv.x * (double) z0 + (double) rv.y * (double) z1;
        }
      }
    }
    tile_partial(acc, gpart, tile);
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

/* The same update over a column of `tzb` tiles stacked in z (the grouping
 * of search_spmv_group_kernel): alpha's reduction once per column, each
 * tile's planes of p and Ap loaded together, each tile's partial written
 * as above. */
template <typename V>
__global__ void __launch_bounds__(TILE_THREADS)
__attribute__((amdgpu_waves_per_eu(8, 8)))
update_group_kernel(DevDom D, InvMT<V> T, const CGState *st, int q,
  const double *__restrict__ ppart, long ntiles, int tzb,
  const V *__restrict__ p, const V *__restrict__ Ap,
  const uint8_t *__restrict__ fluid, V *__restrict__ phi,
  V *__restrict__ r, double *__restrict__ gpart)
{
  if (st->done | st->inner_done) return;
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
  int x = tx * TILE_X + lane + 1, j = ty * TILE_Y + wv + 1;
  bool colok = (x <= D.nx) && (j <= D.ny);
  int tz1 = tz0 + tzb < ntz_ ? tz0 + tzb : ntz_;
  long tile = tx + (long) ntx_ * (ty + (long) nty_ * tz0);
  const long tstride = (long) ntx_ * nty_;
  /* the reduction runs once per column (the single-tile kernel overlaps
   * it with its loads; here it is amortised over the column instead) */
  V alpha = (V) (st->rzs[q & 1]/tile_block_sum(ppart, ntiles));
  for (int tz = tz0; tz < tz1; tz++, tile += tstride) {
    int k0 = tz * TILE_Z + 1;
    int k1 = k0 + TILE_Z - 1 < D.nz ? k0 + TILE_Z - 1 : D.nz;
    /* 32-bit element offsets: one vector register per address */
    unsigned c0 = (unsigned) gidx(C, x, j, k0), sz = (unsigned) C.sz;
    V pv[TILE_Z], apv[TILE_Z];
#pragma unroll
    for (int kk = 0; kk < TILE_Z; kk++) {
      pv[kk] = apv[kk] = 0;
      if (colok && k0 + kk <= k1) {
        pv[kk] = ld32(p, c0 + kk * sz);
        apv[kk] = ld32(Ap, c0 + kk * sz);
      }
    }
    double acc = 0.0;
    if (colok) {
#pragma unroll
      for (int kk = 0; kk < TILE_Z; kk++) {
        if (k0 + kk <= k1) {
          unsigned c = c0 + kk * sz;
          V *ph = (V *) ((char *) phi + c * (unsigned) sizeof(V));
          V *rp = (V *) ((char *) r + c * (unsigned) sizeof(V));
          *ph += alpha*pv[kk];
          V rr = *rp - alpha*apv[kk];
          rr = solid(fluid, c) ? (V) 0 : rr;
          *rp = rr;
          V z = rr*(solid(fluid, c) ? (V) 0 : invm_at<V>(T, D, x, j, k0 + kk));
          acc += (double) rr * (double) z;
        }
      }
    }
    tile_partial(acc, gpart, tile);
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

/* Tile partials of the least-squares system for the Jacobi-weighted
 * residual of phi0 = alpha psi + beta q: (b, W A psi), (b, W A q),
 * (A psi, W A psi), (A psi, W A q), (A q, W A q) with A = -Laplacian and
 * W = invM, over fluid cells. */
__global__ void __launch_bounds__(TILE_THREADS)
galerkin_mr_kernel(DevDom D, Coef K, InvMTab T,
  const uint8_t *__restrict__ fluid, const double *__restrict__ psi,
  const double *__restrict__ q, const double *__restrict__ rhs,
  double *__restrict__ part)
{
  TILE_PROLOGUE
  double a[5] = {0., 0., 0., 0., 0.};
  if (colok) {
    for (int k = k0; k <= k1; k++) {
      long c = gidx(C, x, j, k);
      if (solid(fluid, c)) continue;
      double w = invm_at(T, D, x, j, k), b = rhs[c];
      double Aps = neg_lap_at(D, K, psi, x, j, k);
      double Aq = neg_lap_at(D, K, q, x, j, k);
      a[0] += w * b * Aps;
      a[1] += w * b * Aq;
      a[2] += w * Aps * Aps;
      a[3] += w * Aps * Aq;
      a[4] += w * Aq * Aq;
    }
  }
  long nt = (long) ntx_ * nty_ * ntz_;
  for (int m = 0; m < 5; m++) {
    tile_partial(a[m], part + m * nt, tile);
    __syncthreads();                  /* tile_partial's LDS is reused */
  }
}

/* phi = alpha psi + beta q on interior cells. */
__global__ void __launch_bounds__(TILE_THREADS)
combine_kernel(DevDom D, double al, const double *__restrict__ psi, double be,
  const double *__restrict__ q, double *__restrict__ phi)
{
  TILE_PROLOGUE
  if (!colok) return;
  for (int k = k0; k <= k1; k++) {
    long c = gidx(C, x, j, k);
    phi[c] = al * psi[c] + be * q[c];
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
  if (colok) {
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
  if (colok) {
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
static double *_p_alt = NULL;          /* second search-direction buffer */
static double *_part_p = NULL;         /* tile partials of (p, Ap) */
static double *_part_r = NULL;         /* tile partials of (r, z) */
static double *_part_w1 = NULL, *_part_w2 = NULL;  /* learned-guess scale */
static long buf_n = 0, part_n = 0;
int bb_pcg_batch = 16;                 /* iterations enqueued per host check */
int bb_spmv_variant = 2;               /* 2: LDS kernel over z-columns of
                                          tiles, 3: the same columns with one
                                          wavefront per row and no LDS
                                          planes, 1: one tile per workgroup,
                                          0: direct loads */
int bb_spmv_tzb = 0;                   /* tiles per z-column (0: automatic) */
int bb_spmv_ntp = 1;                   /* 1: FP64 search loads p_old with the
                                          non-temporal hint (0: ordinary
                                          loads, measured slower) */
int bb_spmv_nts = 1;                   /* 1: the FP64 search stores p_new and
                                          Ap with the non-temporal hint */
int bb_spmv_nts32 = 1;                 /* the same for the FP32 search (with
                                          8-tile columns ordinary stores take
                                          the same time per step, but the
                                          update kernels absorb their
                                          write-back) */
int bb_spmv_tzb32 = 0;                 /* tiles per z-column of the FP32
                                          search (0: automatic) */
int bb_defer_phi = 1;                  /* 1: phi updated every second
                                          iteration with both terms */
int bb_update_group = 0;               /* 1: update kernel over z-columns too */
int bb_mixed_x2 = 0;                   /* 1: FP32 kernels on x pairs (even nx);
                                          0 (default): the one-cell z-column
                                          kernels, measured as fast */
int bb_spmv_kchunk = 8;                /* kept for the C ABI; tiles fix 8 */
int bb_mixed_restart = 0;              /* 1: restart CG at every correction */
long bb_pcg_iters_total = 0;           /* PCG iterations over all solves */
long bb_pcg_solves_total = 0;

/* FP32 vectors of the mixed form's inner sweeps (Gcc layout).  Each
 * vector starts one float into its allocation, so that element 1 (the
 * first interior x) and, for even nx, every odd x of every row sits on an
 * 8-byte boundary: the two-cells-per-lane kernels move (x, x + 1) pairs
 * with odd x as one float2. */
static float *_r32 = NULL, *_p32a = NULL, *_p32b = NULL, *_Ap32 = NULL,
             *_e32 = NULL;
static long buf32_n = 0;

static void cg_buffers32_free(void)
{
  float **bufs[5] = {&_r32, &_p32a, &_p32b, &_Ap32, &_e32};
  for (int i = 0; i < 5; i++) {
    if (*bufs[i]) BB_CHECK(hipFree(*bufs[i] - 1));
    *bufs[i] = NULL;
  }
  buf32_n = 0;
}

static void cg_buffers32(void)
{
  if (buf32_n == ddom.cc.n) return;
  cg_buffers32_free();
  float **bufs[5] = {&_r32, &_p32a, &_p32b, &_Ap32, &_e32};
  for (int i = 0; i < 5; i++) {
    float *a = NULL;
    BB_CHECK(hipMalloc((void **) &a, (ddom.cc.n + 2) * sizeof(float)));
    BB_CHECK(hipMemset(a, 0, (ddom.cc.n + 2) * sizeof(float)));
    *bufs[i] = a + 1;
  }
  buf32_n = ddom.cc.n;
}

static void cg_buffers(void)
{
  if (!_cg_state) {
    BB_CHECK(hipMalloc((void **) &_cg_state, sizeof(CGState)));
    BB_CHECK(hipMemset(_cg_state, 0, sizeof(CGState)));
    BB_CHECK(hipHostMalloc((void **) &cg_host, sizeof(CGState),
      hipHostMallocDefault));
  }
  if (buf_n != ddom.cc.n) {
    if (_p_alt) BB_CHECK(hipFree(_p_alt));
    BB_CHECK(hipMalloc((void **) &_p_alt, ddom.cc.n * sizeof(double)));
    BB_CHECK(hipMemset(_p_alt, 0, ddom.cc.n * sizeof(double)));
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
  if (_p_alt) BB_CHECK(hipFree(_p_alt));
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
}

static void fetch_state(void)
{
  BB_CHECK(hipMemcpyAsync(cg_host, _cg_state, sizeof(CGState),
    hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
  if (bb_time_kernels) bb_kt_collect();
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

/* The initial guess with the least A-norm error in span{psi, q}:
 * phi = alpha psi + beta q solving the 2 x 2 Galerkin system
 *   [(psi,A psi) (psi,A q); (q,A psi) (q,A q)] [alpha; beta]
 *     = [(b, psi); (b, q)]
 * (A = -Laplacian, symmetric positive semidefinite).  use_psi / use_q
 * select the directions; a singular or non-positive system drops the
 * direction at fault.  Returns 1 if phi was set (0: start from phi = 0). */
/* Energy 2 (b, phi0) - (phi0, A phi0) captured by the last guess, and by
 * the Galerkin guess on q alone; with the solution phi* the A-norm errors
 * squared are (b, phi*) minus these. */
double bb_galerkin_e_guess = 0., bb_galerkin_e_q = 0.;
/* 0: the guess minimises the A-norm error (Galerkin); 1: it minimises the
 * Jacobi-weighted residual (b - A phi0, invM (b - A phi0)), the norm of the
 * PCG stopping test. */
int bb_galerkin_norm = 0;

int bb_pp_galerkin_guess(const double *psi, int use_psi, const double *q,
  int use_q, double *alpha, double *beta)
{
  cg_buffers();
  bb_galerkin_e_guess = bb_galerkin_e_q = 0.;
  const long nt = n_tiles(ddom);
  *alpha = *beta = 0.;
  if (!use_psi && !use_q) return 0;
  const double *ps = use_psi ? psi : q, *qq = use_q ? q : psi;
  hipLaunchKernelGGL(galerkin_kernel, TILES_LAUNCH, ddom, coef(), fluid_mask(),
    ps, qq, _rhs_p, _part_w1);
  for (int m = 0; m < 5; m++)
    bb_reduce_rows(_part_w1 + m * nt, nt, _bb_scalars + 16 + m);
  bb_fetch_scalars(21);
  double *S = bb_host_scalars + 16;
  double bp = S[0], bq = S[1], app = S[2], apq = S[3], aqq = S[4];
  /* the normal equations of the chosen norm */
  double Bp = bp, Bq = bq, Ap = app, Apq = apq, Aq = aqq;
  if (bb_galerkin_norm == 1) {
    hipLaunchKernelGGL(galerkin_mr_kernel, TILES_LAUNCH, ddom, coef(),
      invm_table(coef()), fluid_mask(), ps, qq, _rhs_p, _part_w1);
    for (int m = 0; m < 5; m++)
      bb_reduce_rows(_part_w1 + m * nt, nt, _bb_scalars + 16 + m);
    bb_fetch_scalars(21);
    Bp = S[0]; Bq = S[1]; Ap = S[2]; Apq = S[3]; Aq = S[4];
  }
  double a = 0., b = 0.;
  double det = Ap * Aq - Apq * Apq;
  if (use_psi && use_q && det > 1e-12 * Ap * Aq && Ap > 0. && Aq > 0.) {
    a = (Bp * Aq - Bq * Apq) / det;
    b = (Bq * Ap - Bp * Apq) / det;
  } else if (use_psi && Ap > 0.) {
    a = Bp / Ap;
  } else if (use_q && Aq > 0.) {
    b = Bq / Aq;
  }
  if (!isfinite(a) || !isfinite(b) || (a == 0. && b == 0.)) return 0;
  bb_galerkin_e_guess = 2. * (a * bp + b * bq) -
                        (a * a * app + 2. * a * b * apq + b * b * aqq);
  if (use_q && aqq > 0.) bb_galerkin_e_q = bq * bq / aqq;
  hipLaunchKernelGGL(combine_kernel, TILES_LAUNCH, ddom, a, ps, b, qq, _phi);
  BB_LAUNCH_CHECK();
  *alpha = a;
  *beta = b;
  return 1;
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

/* The z-column search kernel. */
template <bool FIRST, typename V, bool MIXED>
static void launch_group(const CoefT<V> &K, const InvMT<V> &T, int q,
  long nrp, const V *r, const V *pold, V *pnew, V *Ap)
{
  const bool f32 = sizeof(V) == 4;
  int tzb = f32 ? spmv_tzb32() : spmv_tzb();
  int nts = f32 ? bb_spmv_nts32 : bb_spmv_nts;
  if (bb_spmv_variant == 3) {
    hipLaunchKernelGGL((search_spmv_rows_kernel<FIRST, V, MIXED>),
      spmv_group_grid(tzb), tiles_block(), 0, bb_stream, ddom, K, T,
      _cg_state, q, _part_r, nrp, tzb, r, pold, pnew, Ap, _part_p);
    return;
  }
  /* two planes in flight: three (FP32: 72.4 against 71.5 us at 256^3) and
   * four did not help, and past two the FP64 kernel spills at 8 waves */
  if (!nts)
    hipLaunchKernelGGL((search_spmv_group_kernel<FIRST, V, MIXED, 2, true,
      false>), spmv_group_grid(tzb), tiles_block(), 0, bb_stream, ddom, K, T,
      _cg_state, q, _part_r, nrp, tzb, r, pold, pnew, Ap, _part_p);
  else if (bb_spmv_ntp)
    hipLaunchKernelGGL((search_spmv_group_kernel<FIRST, V, MIXED, 2, true>),
      spmv_group_grid(tzb), tiles_block(), 0, bb_stream, ddom, K, T,
      _cg_state, q, _part_r, nrp, tzb, r, pold, pnew, Ap, _part_p);
  else
    hipLaunchKernelGGL((search_spmv_group_kernel<FIRST, V, MIXED, 2, false>),
      spmv_group_grid(tzb), tiles_block(), 0, bb_stream, ddom, K, T,
      _cg_state, q, _part_r, nrp, tzb, r, pold, pnew, Ap, _part_p);
}

/* Paired phi updates (bb_defer_phi, plain form, one-tile update kernel):
 * odd iterations leave phi to the next one, except before an iteration
 * that recomputes the true residual (every 50th), whose phi update adds
 * its own term only. */
static bool update_is_group(void)
{
  return bb_spmv_variant == 2 && bb_update_group && spmv_tzb() > 1;
}

static bool phi_skips(int q)
{
  return bb_defer_phi && !update_is_group() && (q & 1) && (q + 1) % 50 != 0;
}

/* The mixed form's inner sweep pairs its updates of the correction e the
 * same way, counting from the first iteration after each FP64 correction
 * (where e is folded into phi and must be whole). */
static int pair_base32 = 1;

static bool use_x2(void);

static bool phi_skips32(int q)
{
  return bb_defer_phi && !update_is_group() && !use_x2() &&
         ((q - pair_base32) & 1) == 0;
}

/* Enqueue iteration q; p buffers alternate by parity. */
static void enqueue_iteration(int q, const Coef &K, const InvMTab &T)
{
  const long nt = n_tiles(ddom);
  double *pnew = (q & 1) ? _p_q : _p_alt;
  double *pold = (q & 1) ? _p_alt : _p_q;
  bb_kt_begin(KT_SPMV);
  if (bb_spmv_variant >= 2) {
    if (q == 1)
      launch_group<true, double, false>(K, T, q, nt, _r_q, pold, pnew, _Apb_q);
    else
      launch_group<false, double, false>(K, T, q, nt, _r_q, pold, pnew,
                                         _Apb_q);
  } else if (bb_spmv_variant == 1) {
    if (q == 1)
      hipLaunchKernelGGL((search_spmv_kernel<true, double, false>),
        TILES_LAUNCH, ddom, K, T, _cg_state, q, _part_r, nt, _r_q, pold, fluid_mask(), pnew,
        _Apb_q, _part_p);
    else
      hipLaunchKernelGGL((search_spmv_kernel<false, double, false>),
        TILES_LAUNCH, ddom, K, T, _cg_state, q, _part_r, nt, _r_q, pold, fluid_mask(), pnew,
        _Apb_q, _part_p);
  } else {
    if (q == 1)
      hipLaunchKernelGGL(search_spmv_direct_kernel<true>, TILES_LAUNCH, ddom,
        K, T, _cg_state, q, _part_r, nt, _r_q, pold, fluid_mask(), pnew, _Apb_q, _part_p);
    else
      hipLaunchKernelGGL(search_spmv_direct_kernel<false>, TILES_LAUNCH, ddom,
        K, T, _cg_state, q, _part_r, nt, _r_q, pold, fluid_mask(), pnew, _Apb_q, _part_p);
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
    int tzb = spmv_tzb();
    if (bb_spmv_variant == 2 && bb_update_group && tzb > 1) {
      hipLaunchKernelGGL(update_group_kernel<double>, spmv_group_grid(tzb),
        tiles_block(), 0, bb_stream, ddom, T, _cg_state, q, _part_p, nt, tzb,
        pnew, _Apb_q, fluid_mask(), _phi, _r_q, _part_r);
    } else {
      const uint8_t *fm = fluid_mask();
#define UPD(M, P) hipLaunchKernelGGL((update_kernel<double, M, P>),          \
        TILES_LAUNCH, ddom, T, _cg_state, q, _part_p, nt, pnew, _Apb_q, fm, \
        _phi, _r_q, _part_r, pold)
      if (fm) {
        if (mode == PHI_SKIP) UPD(true, PHI_SKIP);
        else if (mode == PHI_PAIR) UPD(true, PHI_PAIR);
        else UPD(true, PHI_EACH);
      } else {
        if (mode == PHI_SKIP) UPD(false, PHI_SKIP);
        else if (mode == PHI_PAIR) UPD(false, PHI_PAIR);
        else UPD(false, PHI_EACH);
      }
#undef UPD
    }
    bb_kt_end(kt);
  }
  BB_LAUNCH_CHECK();
}

/* The plain FP64 Jacobi-PCG (src/cuda_solver.cu:574-761) on the RHS already
 * in _rhs_p with the tile partials of (b, b) in _part_p. */
static void pcg_plain(double ts, int warm)
{
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
    return;
  }
  int q = 0;
  while (q <= pp_max_iter) {
    int qend = q + bb_pcg_batch;
    if (qend > pp_max_iter + 1) qend = pp_max_iter + 1;
    for (int qq = q + 1; qq <= qend; qq++) enqueue_iteration(qq, K, T);
    hipLaunchKernelGGL(cg_check_kernel<false>, dim3(1), dim3(TILE_THREADS),
      0, bb_stream, _cg_state, qend, _part_r, nt);
    q = qend;
    fetch_state();
    if (cg_host->done) break;
  }
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
  bb_pcg_iters_total += cg_host->q;
  bb_pcg_solves_total++;
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

/* Number of tile partials now in _part_r (the FP64 tile kernels and the
 * x-pair FP32 kernels use different tile counts). */
static long rp_n = 0;

static bool use_x2(void) { return bb_mixed_x2 && ddom.nx % 2 == 0; }

static long n_tiles_x2_host(void)
{
  long ntx = (ddom.nx + TX2 - 1) / TX2;
  long nty = (ddom.ny + TILE_Y - 1) / TILE_Y;
  long ntz = (ddom.nz + TILE_Z - 1) / TILE_Z;
  return ntx * nty * ntz;
}

static dim3 x2_group_grid(int tzb)
{
  long ntx = (ddom.nx + TX2 - 1) / TX2;
  long nty = (ddom.ny + TILE_Y - 1) / TILE_Y;
  long ntz = (ddom.nz + TILE_Z - 1) / TILE_Z;
  return dim3((unsigned) (ntx * nty * ((ntz + tzb - 1) / tzb)));
}

/* z-column length of the x-pair kernels: at least four workgroups per CU */
static int x2_tzb(void)
{
  if (bb_spmv_tzb > 0) return bb_spmv_tzb;
  long t = n_tiles_x2_host() / 1024;
  return t < 1 ? 1 : (t > 4 ? 4 : (int) t);
}

/* Enqueue inner iteration q of the mixed form (FP32 vectors); `first`
 * starts a new Krylov space (p = z). */
static void enqueue_iteration32(int q, bool first, const CoefT<float> &K,
  const InvMT<float> &T)
{
  const long nt = n_tiles(ddom);
  float *pnew = (q & 1) ? _p32a : _p32b;
  float *pold = (q & 1) ? _p32b : _p32a;
  if (use_x2()) {
    const long nt2 = n_tiles_x2_host();
    int tzb = x2_tzb();
    bb_kt_begin(KT_SPMV32);
    if (first)
      hipLaunchKernelGGL(search_spmv_x2_kernel<true>, x2_group_grid(tzb),
        tiles_block(), 0, bb_stream, ddom, K, T, _cg_state, q, _part_r, rp_n,
        tzb, _r32, pold, pnew, _Ap32, _part_p);
    else
      hipLaunchKernelGGL(search_spmv_x2_kernel<false>, x2_group_grid(tzb),
        tiles_block(), 0, bb_stream, ddom, K, T, _cg_state, q, _part_r, rp_n,
        tzb, _r32, pold, pnew, _Ap32, _part_p);
    bb_kt_end(KT_SPMV32);
    bb_kt_begin(KT_UPDATE32);
    hipLaunchKernelGGL(update_x2_kernel, x2_group_grid(tzb), tiles_block(), 0,
      bb_stream, ddom, T, _cg_state, q, _part_p, nt2, tzb, pnew, _Ap32,
      fluid_mask(), _e32, _r32, _part_r);
    bb_kt_end(KT_UPDATE32);
    rp_n = nt2;
    BB_LAUNCH_CHECK();
    return;
  }
  bb_kt_begin(KT_SPMV32);
  if (bb_spmv_variant >= 2) {
    if (first)
      launch_group<true, float, true>(K, T, q, nt, _r32, pold, pnew, _Ap32);
    else
      launch_group<false, float, true>(K, T, q, nt, _r32, pold, pnew, _Ap32);
  } else if (first) {
    hipLaunchKernelGGL((search_spmv_kernel<true, float, true>), TILES_LAUNCH,
      ddom, K, T, _cg_state, q, _part_r, nt, _r32, pold, fluid_mask(), pnew,
      _Ap32, _part_p);
  } else {
    hipLaunchKernelGGL((search_spmv_kernel<false, float, true>), TILES_LAUNCH,
      ddom, K, T, _cg_state, q, _part_r, nt, _r32, pold, fluid_mask(), pnew,
      _Ap32, _part_p);
  }
  bb_kt_end(KT_SPMV32);
  int mode = phi_skips32(q) ? PHI_SKIP
           : (q > pair_base32 && phi_skips32(q - 1)) ? PHI_PAIR : PHI_EACH;
  int kt = mode == PHI_SKIP ? KT_UPDATE32_SKIP
         : mode == PHI_PAIR ? KT_UPDATE32_PAIR : KT_UPDATE32;
  bb_kt_begin(kt);
  int tzb = spmv_tzb();
  if (bb_spmv_variant == 2 && bb_update_group && tzb > 1) {
    hipLaunchKernelGGL(update_group_kernel<float>, spmv_group_grid(tzb),
      tiles_block(), 0, bb_stream, ddom, T, _cg_state, q, _part_p, nt, tzb,
      pnew, _Ap32, fluid_mask(), _e32, _r32, _part_r);
  } else {
    const uint8_t *fm = fluid_mask();
#define UPD(M, P) hipLaunchKernelGGL((update_kernel<float, M, P>),           \
      TILES_LAUNCH, ddom, T, _cg_state, q, _part_p, nt, pnew, _Ap32, fm,    \
      _e32, _r32, _part_r, pold)
    if (fm) {
      if (mode == PHI_SKIP) UPD(true, PHI_SKIP);
      else if (mode == PHI_PAIR) UPD(true, PHI_PAIR);
      else UPD(true, PHI_EACH);
    } else {
      if (mode == PHI_SKIP) UPD(false, PHI_SKIP);
      else if (mode == PHI_PAIR) UPD(false, PHI_PAIR);
      else UPD(false, PHI_EACH);
    }
#undef UPD
  }
  bb_kt_end(kt);
  rp_n = nt;
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
  rp_n = nt;
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
    return;
  }
  int q = 0, first_q = 1, nouter = 1;
  pair_base32 = 1;
  int done = cg_host->rz1 <= pp_residual*pp_residual*sp_rhs;
  while (!done && q <= pp_max_iter && nouter <= pp_max_iter) {
    int qend = q + bb_pcg_batch;
    if (qend > pp_max_iter + 1) qend = pp_max_iter + 1;
    for (int qq = q + 1; qq <= qend; qq++)
      enqueue_iteration32(qq, qq == first_q, K32, T32);
    hipLaunchKernelGGL(cg_check_kernel<true>, dim3(1), dim3(TILE_THREADS),
      0, bb_stream, _cg_state, qend, _part_r, rp_n);
    fetch_state();
    if (!cg_host->inner_done) {
      q = qend;
      continue;
    }
    q = cg_host->q;
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
    rp_n = nt;
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
  bb_pcg_iters_total += q;
  bb_pcg_solves_total++;
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
