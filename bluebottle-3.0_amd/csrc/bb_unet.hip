/*
 * Learned warm start for the pressure solve: a 3-level 3-D convolutional
 * U-Net (16/32/64 channels, 3x3x3 kernels, zero padding, ReLU, 2x2x2
 * average pooling, nearest-neighbour upsampling, skip connections by
 * channel concatenation, a final 1x1x1 convolution to one channel).
 *
 * Weights and activations are BF16 on the MFMA units with FP32
 * accumulation (v_mfma_f32_16x16x16_bf16); master weights, gradients and
 * the Adam moments are FP32.  Tensors are channels-last: element (v, c) of
 * a tensor with row length ld sits at v*ld + c, v = x + nx*(y + ny*z), so a
 * lane's four consecutive channels are one 8-byte load.
 *
 * A 3x3x3 convolution is an implicit GEMM: M = voxels, N = output channels,
 * K = 27 taps x input channels.  One wavefront owns a 16-voxel x 16-channel
 * output tile and runs one MFMA per (tap, 16-channel input chunk).  The
 * backward data pass is the same kernel on the output gradient with the
 * weights flipped and transposed; the weight gradient is a GEMM over voxels
 * (M = output channel, N = input channel, K = voxels) split into voxel
 * chunks whose partial sums are added in a fixed order (deterministic).
 *
 * No counterpart in the reference (it solves every Poisson problem from
 * phi = 0, src/solver_kernel.cu:258-283).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "bb_common.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

/* BF16 <-> FP32, round to nearest even. */
__host__ __device__ inline uint16_t f2bf(float f)
{
  union { float f; uint32_t u; } x;
  x.f = f;
  uint32_t u = x.u;
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t) (u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t) (u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b)
{
  union { float f; uint32_t u; } x;
  x.u = ((uint32_t) b) << 16;
  return x.f;
}

/* A tensor view: base, row length (channels per voxel in the buffer),
 * channel offset, volume dims. */
struct TV {
  void *p;
  int ld, c0;
  int nx, ny, nz;
};

__device__ inline bool vox_xyz(const TV &T, long v, int *x, int *y, int *z)
{
  long n = (long) T.nx * T.ny * T.nz;
  if (v >= n) return false;
  *x = (int) (v % T.nx);
  long r = v / T.nx;
  *y = (int) (r % T.ny);
  *z = (int) (r / T.ny);
  return true;
}

/* Four consecutive channels (c .. c+3) of voxel (x, y, z) as BF16, zero
 * outside the volume or past cmax. */
template <typename IN>
__device__ inline bf16x4 load4(const TV &T, int x, int y, int z, int c,
  int cmax)
{
  bf16x4 r = {0, 0, 0, 0};
  if (x < 0 || y < 0 || z < 0 || x >= T.nx || y >= T.ny || z >= T.nz)
    return r;
  long base = ((long) x + (long) T.nx * (y + (long) T.ny * z)) * T.ld + T.c0;
  const IN *p = (const IN *) T.p + base;
  for (int e = 0; e < 4; e++) {
    if (c + e < cmax) {
      if (sizeof(IN) == 2) r[e] = (short) ((const uint16_t *) p)[c + e];
      else r[e] = (short) f2bf(((const float *) p)[c + e]);
    }
  }
  return r;
}

/* ------------------------------- convolution ----------------------------- */
/* Y[v, co] = act(sum_tap sum_ci X[v + off(tap), ci] W[tap][co][ci] + b[co]).
 * W is BF16 [taps][Cout_pad][Cin_pad]; Cin_pad, Cout_pad multiples of 16.
 *
 * MFMA operands: A = W (m = output channel, k = input channel), B = X
 * (k = input channel, n = voxel), D = Y (m = output channel, n = voxel).
 * Fragment layouts of v_mfma_f32_16x16x{16,32}_bf16 (KPL = K/4 elements
 * per lane): A[m][k] in lane m + 16*(k/KPL), element k%KPL; B[k][n] in lane
 * n + 16*(k/KPL); D[m][n] in lane n + 16*(m/4), element m%4.  A lane thus
 * loads KPL consecutive channels of one voxel (one 8- or 16-byte load of
 * the channels-last row) and KPL consecutive input channels of one weight
 * row, and holds four consecutive output channels of one voxel.
 *
 * One wavefront: NT tiles of 16 consecutive voxels x NCO tiles of 16
 * output channels (each B fragment feeds NCO MFMAs, each A fragment NT).
 * Block: 4 wavefronts; grid (ceil(nvox / (64 NT)), Cout_pad / (16 NCO)).
 * `vec`: rows aligned for vector loads of KPL channels (else per element). */
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int KS> struct Frag;
template <> struct Frag<16> { typedef bf16x4 T; };
template <> struct Frag<32> { typedef bf16x8 T; };

template <int KS>
__device__ inline f32x4 mfma_bf16(typename Frag<KS>::T a,
  typename Frag<KS>::T b, f32x4 c);
template <> __device__ inline f32x4 mfma_bf16<16>(bf16x4 a, bf16x4 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
template <> __device__ inline f32x4 mfma_bf16<32>(bf16x8 a, bf16x8 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
template <int KPL> struct U16V;
template <> struct U16V<4> { typedef u16x4 T; };
template <> struct U16V<8> { typedef u16x8 T; };

/* KPL channels c .. c+KPL-1 (zero at c >= cmax) of row `row` of T as BF16. */
template <typename IN, int KS>
__device__ inline typename Frag<KS>::T load_row(const TV &T, long row, int c,
  int cmax, bool vec)
{
  constexpr int KPL = KS / 4;
  typedef typename Frag<KS>::T F;
  typedef typename U16V<KPL>::T U;
  const IN *p = (const IN *) T.p + row * T.ld + T.c0 + c;
  U h;
  if (vec && c + KPL <= cmax) {
    if (sizeof(IN) == 2) {
      h = *(const U *) p;
    } else {
#pragma unroll
      for (int e4 = 0; e4 < KPL / 4; e4++) {
        float4 f = ((const float4 *) p)[e4];
        h[4 * e4] = f2bf(f.x); h[4 * e4 + 1] = f2bf(f.y);
        h[4 * e4 + 2] = f2bf(f.z); h[4 * e4 + 3] = f2bf(f.w);
      }
    }
    return __builtin_bit_cast(F, h);
  }
#pragma unroll
  for (int e = 0; e < KPL; e++) {
    h[e] = 0;
    if (c + e < cmax)
      h[e] = sizeof(IN) == 2 ? ((const uint16_t *) p)[e]
                             : f2bf(((const float *) p)[e]);
  }
  return __builtin_bit_cast(F, h);
}

template <int KS>
__device__ inline typename Frag<KS>::T load_w(const uint16_t *__restrict__ w)
{
  typedef typename U16V<KS / 4>::T U;
  return __builtin_bit_cast(typename Frag<KS>::T, *(const U *) w);
}

template <typename IN, int KS, int NCO>
__global__ void __launch_bounds__(256)
conv_fwd_kernel(TV X, int cin, int cin_pad, const uint16_t *__restrict__ W,
  int taps, const float *__restrict__ bias, int cout, int cout_pad,
  TV Yf, TV Yb, int relu, int vec)
{
  constexpr int NT = 2, KPL = KS / 4;
  typedef typename Frag<KS>::T F;
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  long vbase = ((long) blockIdx.x * 4 + wv) * (16 * NT);
  int co0 = blockIdx.y * 16 * NCO;
  int q = lane >> 4, n = lane & 15;
  long nvox = (long) X.nx * X.ny * X.nz;
  /* this lane's voxel of each voxel tile (B operand and D column) */
  int vx[NT], vy[NT], vz[NT];
  bool vok[NT];
#pragma unroll
  for (int t = 0; t < NT; t++)
    vok[t] = vox_xyz(X, vbase + 16 * t + n, &vx[t], &vy[t], &vz[t]);
  f32x4 acc[NT][NCO];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int c = 0; c < NCO; c++) acc[t][c] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < taps; tap++) {
    int dx = 0, dy = 0, dz = 0;
    if (taps == 27) { dx = tap % 3 - 1; dy = (tap / 3) % 3 - 1; dz = tap / 9 - 1; }
    long row[NT];
    bool rok[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
      int xs = vx[t] + dx, ys = vy[t] + dy, zs = vz[t] + dz;
      rok[t] = vok[t] && xs >= 0 && ys >= 0 && zs >= 0 && xs < X.nx &&
               ys < X.ny && zs < X.nz;
      row[t] = rok[t] ? (long) xs + (long) X.nx * (ys + (long) X.ny * zs) : 0;
    }
    const uint16_t *Wt = W + ((long) tap * cout_pad + co0 + n) * cin_pad + KPL * q;
    for (int ch = 0; ch < cin_pad; ch += KS) {
      F b[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) {
        b[t] = __builtin_bit_cast(F, (typename U16V<KPL>::T) 0);
        if (rok[t]) b[t] = load_row<IN, KS>(X, row[t], ch + KPL * q, cin, vec);
      }
#pragma unroll
      for (int c = 0; c < NCO; c++) {
        F a = load_w<KS>(Wt + (long) c * 16 * cin_pad + ch);
#pragma unroll
        for (int t = 0; t < NT; t++)
          acc[t][c] = mfma_bf16<KS>(a, b[t], acc[t][c]);
      }
    }
  }
  (void) nvox;
  /* D[m][n]: lane holds output channels co0 + 16 c + 4 q + e of voxel n */
  const TV &O = Yf.p ? Yf : Yb;
#pragma unroll
  for (int t = 0; t < NT; t++) {
    if (!vok[t]) continue;
    long v = vbase + 16 * t + n;
#pragma unroll
    for (int c = 0; c < NCO; c++) {
      int cb = co0 + 16 * c + 4 * q;
      if (cb >= cout) continue;
      float val[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        float bb = (bias && cb + e < cout) ? bias[cb + e] : 0.f;
        val[e] = acc[t][c][e] + bb;
        if (relu) val[e] = val[e] > 0.f ? val[e] : 0.f;
      }
      bool full = cb + 4 <= cout && (O.ld & 3) == 0 && (O.c0 & 3) == 0;
      if (Yf.p) {
        float *o = (float *) Yf.p + v * Yf.ld + Yf.c0 + cb;
        if (full) *(float4 *) o = make_float4(val[0], val[1], val[2], val[3]);
        else for (int e = 0; e < 4; e++) if (cb + e < cout) o[e] = val[e];
      }
      if (Yb.p) {
        uint16_t *o = (uint16_t *) Yb.p + v * Yb.ld + Yb.c0 + cb;
        if (full) {
          uint2 u;
          u.x = (uint32_t) f2bf(val[0]) | ((uint32_t) f2bf(val[1]) << 16);
          u.y = (uint32_t) f2bf(val[2]) | ((uint32_t) f2bf(val[3]) << 16);
          *(uint2 *) o = u;
        } else {
          for (int e = 0; e < 4; e++) if (cb + e < cout) o[e] = f2bf(val[e]);
        }
      }
    }
  }
}

/* Four consecutive channels of one voxel as loaded by the staging loop
 * of conv_lds_kernel: BF16 inputs stay packed (two registers), FP32 inputs
 * are converted to BF16 when they are written to LDS. */
template <typename IN> struct RawUnit;
template <> struct RawUnit<uint16_t> {
  typedef u16x4 T;
  __device__ static T zero() { return (T) 0; }
  __device__ static void set(T &r, int e, uint16_t x) { r[e] = x; }
  __device__ static u16x4 bf16(const T &r) { return r; }
};
template <> struct RawUnit<float> {
  typedef float4 T;
  __device__ static T zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
  __device__ static void set(T &r, int e, float x)
  {
    if (e == 0) r.x = x; else if (e == 1) r.y = x; else if (e == 2) r.z = x;
    else r.w = x;
  }
  __device__ static u16x4 bf16(const T &r)
  {
    u16x4 h;
    h[0] = f2bf(r.x); h[1] = f2bf(r.y); h[2] = f2bf(r.z); h[3] = f2bf(r.w);
    return h;
  }
};

/* The same convolution with the input staged through LDS.  A workgroup
 * owns a 16 x 4 x 4 block of output voxels (16 rows of 16 voxels along x,
 * four rows per wavefront); for each chunk of KS input channels it stages
 * the 18 x 6 x 6 input voxels around the block (zero outside the volume)
 * in LDS, then runs the 27 taps from LDS, so each input element is read
 * from memory about 1.7 times per chunk instead of once per tap.  Output
 * channels: NCO tiles of 16 per wavefront, Cout_pad / (16 NCO) workgroups
 * along grid.y.  Same fragment layouts as conv_fwd_kernel. */
#define CB_X 16
#define CB_Y 4
#define CB_Z 4
#define CB_V ((CB_X + 2) * (CB_Y + 2) * (CB_Z + 2))

template <typename IN, int KS, int NCO>
__global__ void __launch_bounds__(256)
conv_lds_kernel(TV X, int cin, int cin_pad, const uint16_t *__restrict__ W,
  const float *__restrict__ bias, int cout, int cout_pad, TV Yf, TV Yb,
  int relu, int vec)
{
  constexpr int NT = 4, KPL = KS / 4;
  constexpr int KCP = KS + 8;           /* LDS row: KS channels + padding */
  /* taps per weight group: a z-plane of taps (9) where their fragments
   * take at most 72 registers (KPL / 2 per fragment), else a row (3);
   * all 27 at once cost the 16-channel kernel an occupancy step */
  constexpr int TG = NCO * KPL * 9 <= 144 ? 9 : 3;
  typedef typename Frag<KS>::T F;
  typedef typename U16V<KPL>::T U;
  __shared__ __attribute__((aligned(16))) uint16_t tile[CB_V * KCP];
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int q = lane >> 4, n = lane & 15;
  int nbx = (X.nx + CB_X - 1) / CB_X, nby = (X.ny + CB_Y - 1) / CB_Y;
  int b = blockIdx.x;
  int x0 = (b % nbx) * CB_X, y0 = ((b / nbx) % nby) * CB_Y,
      z0 = (b / (nbx * nby)) * CB_Z;
  int co0 = blockIdx.y * 16 * NCO;
  f32x4 acc[NT][NCO];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int c = 0; c < NCO; c++) acc[t][c] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const uint16_t *Wl = W + (long) (co0 + n) * cin_pad + KPL * q;
  for (int ch = 0; ch < cin_pad; ch += KS) {
    /* stage: units of 4 channels of one voxel, NU per thread.  Every load
     * of the chunk is issued before the first LDS store, so the
     * workgroup waits for its loads once and not once per unit. */
    constexpr int NUNIT = CB_V * (KS / 4), NU = (NUNIT + 255) / 256;
    /* BF16 input: the four channels packed as loaded; FP32: four floats */
    typedef typename RawUnit<IN>::T RU;
    RU rv[NU];
#pragma unroll
    for (int i = 0; i < NU; i++) {
      const int u = threadIdx.x + 256 * i;
      rv[i] = RawUnit<IN>::zero();
      const int v = u / (KS / 4), c4 = 4 * (u % (KS / 4));
      const int vx = v % (CB_X + 2), vy = (v / (CB_X + 2)) % (CB_Y + 2),
                vz = v / ((CB_X + 2) * (CB_Y + 2));
      const int gx = x0 + vx - 1, gy = y0 + vy - 1, gz = z0 + vz - 1;
      if (u < NUNIT && gx >= 0 && gy >= 0 && gz >= 0 && gx < X.nx &&
          gy < X.ny && gz < X.nz) {
        long row = (long) gx + (long) X.nx * (gy + (long) X.ny * gz);
        int c = ch + c4;
        const IN *p = (const IN *) X.p + row * X.ld + X.c0 + c;
        if (vec && c + 4 <= cin) {
          rv[i] = *(const RU *) p;
        } else {
#pragma unroll
          for (int e = 0; e < 4; e++)
            if (c + e < cin) RawUnit<IN>::set(rv[i], e, p[e]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NU; i++) {
      const int u = threadIdx.x + 256 * i;
      if (u < NUNIT) {
        const int v = u / (KS / 4), c4 = 4 * (u % (KS / 4));
        *(u16x4 *) &tile[v * KCP + c4] = RawUnit<IN>::bf16(rv[i]);
      }
    }
    __syncthreads();
    /* taps in groups of TG: the group's weight fragments are loaded
     * together before its MFMAs, so their latency is paid once per group
     * and not once per tap (same MFMAs in the same order) */
    for (int g0 = 0; g0 < 27; g0 += TG) {
      F aw[TG][NCO];
#pragma unroll
      for (int tg = 0; tg < TG; tg++)
#pragma unroll
        for (int c = 0; c < NCO; c++)
          aw[tg][c] = load_w<KS>(Wl + (long) (g0 + tg) * cout_pad * cin_pad +
                                 ch + (long) c * 16 * cin_pad);
#pragma unroll
      for (int tg = 0; tg < TG; tg++) {
        const int tap = g0 + tg;
        int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;   /* offset + 1 */
        F bfr[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
          int r = wv * NT + t;
          int yy = r % CB_Y, zz = r / CB_Y;
          int v = (n + dx) + (CB_X + 2) * ((yy + dy) + (CB_Y + 2) * (zz + dz));
          bfr[t] = __builtin_bit_cast(F, *(const U *) &tile[v * KCP + KPL * q]);
        }
#pragma unroll
        for (int c = 0; c < NCO; c++)
#pragma unroll
          for (int t = 0; t < NT; t++)
            acc[t][c] = mfma_bf16<KS>(aw[tg][c], bfr[t], acc[t][c]);
      }
    }
    __syncthreads();
  }
  /* D[m][n]: lane holds output channels co0 + 16 c + 4 q + e of voxel n of
   * row wv * NT + t */
  const TV &O = Yf.p ? Yf : Yb;
#pragma unroll
  for (int t = 0; t < NT; t++) {
    int r = wv * NT + t;
    int gx = x0 + n, gy = y0 + r % CB_Y, gz = z0 + r / CB_Y;
    if (gx >= O.nx || gy >= O.ny || gz >= O.nz) continue;
    long v = (long) gx + (long) O.nx * (gy + (long) O.ny * gz);
#pragma unroll
    for (int c = 0; c < NCO; c++) {
      int cb = co0 + 16 * c + 4 * q;
      if (cb >= cout) continue;
      float val[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        float bb = (bias && cb + e < cout) ? bias[cb + e] : 0.f;
        val[e] = acc[t][c][e] + bb;
        if (relu) val[e] = val[e] > 0.f ? val[e] : 0.f;
      }
      bool full = cb + 4 <= cout && (O.ld & 3) == 0 && (O.c0 & 3) == 0;
      if (Yf.p) {
        float *o = (float *) Yf.p + v * Yf.ld + Yf.c0 + cb;
        if (full) *(float4 *) o = make_float4(val[0], val[1], val[2], val[3]);
        else for (int e = 0; e < 4; e++) if (cb + e < cout) o[e] = val[e];
      }
      if (Yb.p) {
        uint16_t *o = (uint16_t *) Yb.p + v * Yb.ld + Yb.c0 + cb;
        if (full) {
          uint2 u;
          u.x = (uint32_t) f2bf(val[0]) | ((uint32_t) f2bf(val[1]) << 16);
          u.y = (uint32_t) f2bf(val[2]) | ((uint32_t) f2bf(val[3]) << 16);
          *(uint2 *) o = u;
        } else {
          for (int e = 0; e < 4; e++) if (cb + e < cout) o[e] = f2bf(val[e]);
        }
      }
    }
  }
}

/* Weight gradient partials: P[chunk][tap][co][ci] = sum over the chunk's
 * voxels v of G[v, co] X[v + off(tap), ci].  Grid: (nchunks, taps *
 * cout_pad/16 * cin_pad/16 / 4); 4 wavefronts, each one (tap, co-tile,
 * ci-tile).  G is FP32, X is BF16. */
__global__ void __launch_bounds__(256)
conv_wgrad_kernel(TV G, int cout, TV X, int cin, int taps, int cout_pad,
  int cin_pad, long chunk, float *__restrict__ P)
{
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int nco = cout_pad / 16, nci = cin_pad / 16;
  int job = blockIdx.y * 4 + wv;
  if (job >= taps * nco * nci) return;
  int tap = job / (nco * nci), rem = job % (nco * nci);
  int co0 = (rem / nci) * 16, ci0 = (rem % nci) * 16;
  int dx = 0, dy = 0, dz = 0;
  if (taps == 27) { dx = tap % 3 - 1; dy = (tap / 3) % 3 - 1; dz = tap / 9 - 1; }
  long nv = (long) G.nx * G.ny * G.nz;
  long vb = (long) blockIdx.x * chunk, ve = vb + chunk < nv ? vb + chunk : nv;
  int m = lane & 15, kq = 4 * (lane >> 4);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (long v0 = vb; v0 < ve; v0 += 16) {
    bf16x4 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
    for (int e = 0; e < 4; e++) {
      long v = v0 + kq + e;
      int x, y, z;
      if (v >= ve || !vox_xyz(G, v, &x, &y, &z)) continue;
      if (co0 + m < cout)
        a[e] = (short) f2bf(((const float *) G.p)[v * G.ld + G.c0 + co0 + m]);
      int xs = x + dx, ys = y + dy, zs = z + dz;
      if (ci0 + m < cin && xs >= 0 && ys >= 0 && zs >= 0 && xs < X.nx &&
          ys < X.ny && zs < X.nz) {
        long vs = (long) xs + (long) X.nx * (ys + (long) X.ny * zs);
        b[e] = (short) ((const uint16_t *) X.p)[vs * X.ld + X.c0 + ci0 + m];
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  /* D[m = co][n = ci] */
  int nn = lane & 15;
  float *Pc = P + (long) blockIdx.x * taps * cout_pad * cin_pad;
  for (int e = 0; e < 4; e++) {
    int co = co0 + 4 * (lane >> 4) + e;
    Pc[((long) tap * cout_pad + co) * cin_pad + ci0 + nn] = acc[e];
  }
}

/* dW = sum of the chunk partials in chunk order; db[co] likewise from the
 * bias partials. */
__global__ void sum_chunks_kernel(const float *__restrict__ P, long n,
  int nchunks, float *__restrict__ out)
{
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int c = 0; c < nchunks; c++) s += P[(long) c * n + i];
  out[i] = s;
}

/* Bias-gradient partials: Pb[chunk][co] = sum over the chunk of G[v, co];
 * zero for the padding channels co >= cout. */
__global__ void bias_grad_kernel(TV G, int cout, long chunk, float *Pb)
{
  int co = threadIdx.x;
  if (co >= cout) {
    Pb[(long) blockIdx.x * blockDim.x + co] = 0.f;
    return;
  }
  long nv = (long) G.nx * G.ny * G.nz;
  long vb = (long) blockIdx.x * chunk, ve = vb + chunk < nv ? vb + chunk : nv;
  float s = 0.f;
  for (long v = vb; v < ve; v++) s += ((const float *) G.p)[v * G.ld + G.c0 + co];
  Pb[(long) blockIdx.x * blockDim.x + co] = s;
}

/* Backward-data weights: Wt[tap][ci][co] = W[26 - tap][co][ci] (flipped,
 * transposed), BF16. */
__global__ void flip_weights_kernel(const uint16_t *__restrict__ W, int taps,
  int cout_pad, int cin_pad, uint16_t *__restrict__ Wt)
{
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  long n = (long) taps * cout_pad * cin_pad;
  if (i >= n) return;
  int tap = (int) (i / ((long) cin_pad * cout_pad));
  int rem = (int) (i % ((long) cin_pad * cout_pad));
  int ci = rem / cout_pad, co = rem % cout_pad;
  int ft = taps - 1 - tap;
  Wt[i] = W[((long) ft * cout_pad + co) * cin_pad + ci];
}

/* ------------------------------ elementwise ------------------------------ */
/* G[v, c] *= (Y[v, c] > 0) over c < C (ReLU backward). */
__global__ void relu_bwd_kernel(TV G, TV Y, int C)
{
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  long nv = (long) G.nx * G.ny * G.nz;
  if (t >= nv * C) return;
  long v = t / C;
  int c = (int) (t % C);
  uint16_t y = ((const uint16_t *) Y.p)[v * Y.ld + Y.c0 + c];
  if (!(bf2f(y) > 0.f)) ((float *) G.p)[v * G.ld + G.c0 + c] = 0.f;
}

/* 2x2x2 average pool, BF16 in and out (coarse dims from C): one thread per
 * four channels of a coarse voxel (C, both row lengths and channel offsets
 * multiples of 4, as every pooled view of the network has), 32-bit voxel
 * indices; the eight children are summed in the same order for every
 * channel. */
__global__ void pool_fwd_kernel(TV F, TV Cr, int C)
{
  /* blockDim.x is a multiple of C / 4: a workgroup covers whole voxels */
  const unsigned nv = (unsigned) Cr.nx * Cr.ny * Cr.nz, C4 = (unsigned) C / 4;
  const unsigned v = blockIdx.x * (blockDim.x / C4) + threadIdx.x / C4;
  const unsigned c = 4 * (threadIdx.x % C4);
  if (v >= nv) return;
  const unsigned x = v % Cr.nx, yz = v / Cr.nx;
  const unsigned y = yz % Cr.ny, z = yz / Cr.ny;
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  for (int q = 0; q < 8; q++) {
    unsigned xf = 2 * x + (q & 1), yf = 2 * y + ((q >> 1) & 1), zf = 2 * z + (q >> 2);
    unsigned vf = xf + F.nx * (yf + F.ny * zf);
    u16x4 h = *(const u16x4 *) ((const uint16_t *) F.p + (size_t) vf * F.ld + F.c0 + c);
#pragma unroll
    for (int e = 0; e < 4; e++) s[e] += bf2f(h[e]);
  }
  u16x4 o;
#pragma unroll
  for (int e = 0; e < 4; e++) o[e] = f2bf(0.125f * s[e]);
  *(u16x4 *) ((uint16_t *) Cr.p + (size_t) v * Cr.ld + Cr.c0 + c) = o;
}

/* Pool backward: G_fine[v, c] += 0.125 G_coarse[parent(v), c] (FP32). */
__global__ void pool_bwd_kernel(TV Gf, TV Gc, int C)
{
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  long nv = (long) Gf.nx * Gf.ny * Gf.nz;
  if (t >= nv * C) return;
  long v = t / C;
  int c = (int) (t % C), x, y, z;
  vox_xyz(Gf, v, &x, &y, &z);
  long vc = (long) (x / 2) + (long) Gc.nx * ((y / 2) + (long) Gc.ny * (z / 2));
  ((float *) Gf.p)[v * Gf.ld + Gf.c0 + c] +=
    0.125f * ((const float *) Gc.p)[vc * Gc.ld + Gc.c0 + c];
}

/* Nearest upsample x2, BF16: one thread per four channels of a fine voxel
 * (the same alignment as the pool), 32-bit voxel indices. */
__global__ void up_fwd_kernel(TV Cr, TV F, int C)
{
  /* blockDim.x is a multiple of C / 4: a workgroup covers whole voxels */
  const unsigned nv = (unsigned) F.nx * F.ny * F.nz, C4 = (unsigned) C / 4;
  const unsigned v = blockIdx.x * (blockDim.x / C4) + threadIdx.x / C4;
  const unsigned c = 4 * (threadIdx.x % C4);
  if (v >= nv) return;
  const unsigned x = v % F.nx, yz = v / F.nx;
  const unsigned y = yz % F.ny, z = yz / F.ny;
  const unsigned vc = x / 2 + Cr.nx * (y / 2 + Cr.ny * (z / 2));
  *(u16x4 *) ((uint16_t *) F.p + (size_t) v * F.ld + F.c0 + c) =
    *(const u16x4 *) ((const uint16_t *) Cr.p + (size_t) vc * Cr.ld + Cr.c0 + c);
}

/* Upsample backward: G_coarse[v, c] = sum of the 8 children (FP32). */
__global__ void up_bwd_kernel(TV Gf, TV Gc, int C)
{
  long t = (long) blockIdx.x * blockDim.x + threadIdx.x;
  long nv = (long) Gc.nx * Gc.ny * Gc.nz;
  if (t >= nv * C) return;
  long v = t / C;
  int c = (int) (t % C), x, y, z;
  vox_xyz(Gc, v, &x, &y, &z);
  float s = 0.f;
  for (int q = 0; q < 8; q++) {
    int xf = 2 * x + (q & 1), yf = 2 * y + ((q >> 1) & 1), zf = 2 * z + (q >> 2);
    long vf = (long) xf + (long) Gf.nx * (yf + (long) Gf.ny * zf);
    s += ((const float *) Gf.p)[vf * Gf.ld + Gf.c0 + c];
  }
  ((float *) Gc.p)[v * Gc.ld + Gc.c0 + c] = s;
}

__global__ void fill_f32_kernel(float *a, long n, float v)
{
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

/* Loss 0.5 * mean (y - t)^2 over the volume: per-block partials, and the
 * output gradient dy = (y - t) / n. */
__global__ void __launch_bounds__(256)
mse_kernel(const float *__restrict__ y, const float *__restrict__ t, long n,
  float *__restrict__ dy, float *__restrict__ part)
{
  __shared__ float s[256];
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  float d = 0.f;
  if (i < n) {
    d = y[i] - t[i];
    dy[i] = d / (float) n;
  }
  s[threadIdx.x] = 0.5f * d * d;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

/* Adam on FP32 master weights with BF16 copies (biases: wb == NULL). */
__global__ void adam_kernel(float *w, const float *__restrict__ g, float *m,
  float *v, long n, float lr, float b1, float b2, float eps, float c1,
  float c2, uint16_t *wb)
{
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void to_bf16_kernel(const float *__restrict__ a, long n,
  uint16_t *__restrict__ b)
{
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

/* ================================ network ================================ */
#define NLAYER 11
/* layer table: cin, cout, taps, level, relu */
static const int LC[NLAYER][5] = {
  { 2, 16, 27, 0, 1},   /* 0 enc0a */
  {16, 16, 27, 0, 1},   /* 1 enc0b -> cat0[32..47] */
  {16, 32, 27, 1, 1},   /* 2 enc1a */
  {32, 32, 27, 1, 1},   /* 3 enc1b -> cat1[64..95] */
  {32, 64, 27, 2, 1},   /* 4 bot_a */
  {64, 64, 27, 2, 1},   /* 5 bot_b */
  {96, 32, 27, 1, 1},   /* 6 dec1a (cat1) */
  {32, 32, 27, 1, 1},   /* 7 dec1b */
  {48, 16, 27, 0, 1},   /* 8 dec0a (cat0) */
  {16, 16, 27, 0, 1},   /* 9 dec0b */
  {16,  1,  1, 0, 0},   /* 10 out 1x1 */
};

static inline int pad16(int c) { return (c + 15) / 16 * 16; }

struct Layer {
  int cin, cout, taps, level, relu, cinp, coutp;
  long nw;                      /* taps * coutp * cinp */
  float *w, *b, *gw, *gb, *mw, *vw, *mb, *vb;
  uint16_t *wb, *wt;            /* BF16 weights, flipped/transposed */
};

struct Acts {
  int n[3][3];                  /* dims per level */
  long nv[3];
  /* BF16 activations */
  uint16_t *in0, *e0a, *cat0, *p0, *e1a, *cat1, *p1, *b0, *b1, *d1a, *d1,
           *d0a, *d0;
  float *y;                     /* output */
  /* FP32 gradients */
  float *g_e0a, *g_cat0, *g_p0, *g_e1a, *g_cat1, *g_p1, *g_b0, *g_b1, *g_d1a,
        *g_d1, *g_d0a, *g_d0, *g_y, *g_in0;
  float *part, *loss_part;
  long part_n, loss_n;
};

/* Two activation sets: the training crop and the full domain. */
struct UNet {
  Layer L[NLAYER];
  Acts acts[2];
  int have[2];
  int next;
  Acts *A;
  long step;
};

static UNet *net = NULL;

static unsigned long long urng = 0x9e3779b97f4a7c15ULL;
static double urand(void)
{
  urng ^= urng >> 12; urng ^= urng << 25; urng ^= urng >> 27;
  return (double) ((urng * 2685821657736338717ULL) >> 11) * (1.0 / 9007199254740992.0);
}

static void *dzalloc(size_t bytes)
{
  void *p = NULL;
  BB_CHECK(hipMalloc(&p, bytes > 0 ? bytes : 1));
  BB_CHECK(hipMemset(p, 0, bytes > 0 ? bytes : 1));
  return p;
}

static inline unsigned blocks(long n, int t) { return (unsigned) ((n + t - 1) / t); }

static void upload_weights(Layer &Y)
{
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(Y.nw, 256)), dim3(256), 0,
    bb_stream, Y.w, Y.nw, Y.wb);
}

/* Create the network with He-initialised weights from `seed`. */
extern "C" int bb_unet_create(unsigned long long seed)
{
  if (net) return 0;
  net = (UNet *) calloc(1, sizeof(UNet));
  urng = seed ? seed : 0x9e3779b97f4a7c15ULL;
  for (int l = 0; l < NLAYER; l++) {
    Layer &Y = net->L[l];
    Y.cin = LC[l][0]; Y.cout = LC[l][1]; Y.taps = LC[l][2];
    Y.level = LC[l][3]; Y.relu = LC[l][4];
    Y.cinp = pad16(Y.cin); Y.coutp = pad16(Y.cout);
    Y.nw = (long) Y.taps * Y.coutp * Y.cinp;
    float *hw = (float *) calloc(Y.nw, sizeof(float));
    double sd = sqrt(2.0 / (Y.taps * Y.cin));
    for (int t = 0; t < Y.taps; t++)
      for (int co = 0; co < Y.cout; co++)
        for (int ci = 0; ci < Y.cin; ci++) {
          /* Box-Muller */
          double u1 = urand() + 1e-300, u2 = urand();
          double g = sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
          hw[((long) t * Y.coutp + co) * Y.cinp + ci] = (float) (g * sd);
        }
    Y.w = (float *) dzalloc(Y.nw * sizeof(float));
    BB_CHECK(hipMemcpy(Y.w, hw, Y.nw * sizeof(float), hipMemcpyHostToDevice));
    free(hw);
    Y.gw = (float *) dzalloc(Y.nw * sizeof(float));
    Y.mw = (float *) dzalloc(Y.nw * sizeof(float));
    Y.vw = (float *) dzalloc(Y.nw * sizeof(float));
    Y.b = (float *) dzalloc(Y.coutp * sizeof(float));
    Y.gb = (float *) dzalloc(Y.coutp * sizeof(float));
    Y.mb = (float *) dzalloc(Y.coutp * sizeof(float));
    Y.vb = (float *) dzalloc(Y.coutp * sizeof(float));
    Y.wb = (uint16_t *) dzalloc(Y.nw * sizeof(uint16_t));
    Y.wt = (uint16_t *) dzalloc(Y.nw * sizeof(uint16_t));
    upload_weights(Y);
  }
  BB_CHECK(hipStreamSynchronize(bb_stream));
  return 0;
}

static void free_acts(Acts &A)
{
  void *p[] = {A.in0, A.e0a, A.cat0, A.p0, A.e1a, A.cat1, A.p1, A.b0, A.b1,
    A.d1a, A.d1, A.d0a, A.d0, A.y, A.g_e0a, A.g_cat0, A.g_p0, A.g_e1a,
    A.g_cat1, A.g_p1, A.g_b0, A.g_b1, A.g_d1a, A.g_d1, A.g_d0a, A.g_d0, A.g_y,
    A.g_in0, A.part, A.loss_part};
  for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++)
    if (p[i]) BB_CHECK(hipFree(p[i]));
  memset(&A, 0, sizeof(A));
}

/* Voxel chunk of the weight- and bias-gradient partials: at most
 * MAX_CHUNKS chunks of at least 256 voxels (a multiple of 16), so that a
 * 64^3 training crop spreads over 512 workgroups per tile job. */
#define MAX_CHUNKS 512L
static long chunk_of(long nv)
{
  long c = (nv + MAX_CHUNKS - 1) / MAX_CHUNKS;
  c = c < 256 ? 256 : c;
  return (c + 15) / 16 * 16;
}

/* Activation buffers for a volume nx x ny x nz (each divisible by 4). */
extern "C" int bb_unet_volume(int nx, int ny, int nz)
{
  if (!net) return -1;
  if (nx % 4 || ny % 4 || nz % 4 || nx < 4 || ny < 4 || nz < 4) {
    fprintf(stderr, "bb_unet_volume: dims must be positive multiples of 4 "
      "(%d %d %d)\n", nx, ny, nz);
    return -2;
  }
  for (int s = 0; s < 2; s++)
    if (net->have[s] && net->acts[s].n[0][0] == nx &&
        net->acts[s].n[0][1] == ny && net->acts[s].n[0][2] == nz) {
      net->A = &net->acts[s];
      return 0;
    }
  int slot = net->next;
  net->next = 1 - net->next;
  if (net->have[slot]) free_acts(net->acts[slot]);
  net->A = &net->acts[slot];
  Acts &A = *net->A;
  for (int l = 0; l < 3; l++) {
    A.n[l][0] = nx >> l; A.n[l][1] = ny >> l; A.n[l][2] = nz >> l;
    A.nv[l] = (long) A.n[l][0] * A.n[l][1] * A.n[l][2];
  }
  long n0 = A.nv[0], n1 = A.nv[1], n2 = A.nv[2];
  A.in0 = (uint16_t *) dzalloc(n0 * 16 * 2);
  A.e0a = (uint16_t *) dzalloc(n0 * 16 * 2);
  A.cat0 = (uint16_t *) dzalloc(n0 * 48 * 2);
  A.p0 = (uint16_t *) dzalloc(n1 * 16 * 2);
  A.e1a = (uint16_t *) dzalloc(n1 * 32 * 2);
  A.cat1 = (uint16_t *) dzalloc(n1 * 96 * 2);
  A.p1 = (uint16_t *) dzalloc(n2 * 32 * 2);
  A.b0 = (uint16_t *) dzalloc(n2 * 64 * 2);
  A.b1 = (uint16_t *) dzalloc(n2 * 64 * 2);
  A.d1a = (uint16_t *) dzalloc(n1 * 32 * 2);
  A.d1 = (uint16_t *) dzalloc(n1 * 32 * 2);
  A.d0a = (uint16_t *) dzalloc(n0 * 16 * 2);
  A.d0 = (uint16_t *) dzalloc(n0 * 16 * 2);
  A.y = (float *) dzalloc(n0 * 4);
  A.g_e0a = (float *) dzalloc(n0 * 16 * 4);
  A.g_cat0 = (float *) dzalloc(n0 * 48 * 4);
  A.g_p0 = (float *) dzalloc(n1 * 16 * 4);
  A.g_e1a = (float *) dzalloc(n1 * 32 * 4);
  A.g_cat1 = (float *) dzalloc(n1 * 96 * 4);
  A.g_p1 = (float *) dzalloc(n2 * 32 * 4);
  A.g_b0 = (float *) dzalloc(n2 * 64 * 4);
  A.g_b1 = (float *) dzalloc(n2 * 64 * 4);
  A.g_d1a = (float *) dzalloc(n1 * 32 * 4);
  A.g_d1 = (float *) dzalloc(n1 * 32 * 4);
  A.g_d0a = (float *) dzalloc(n0 * 16 * 4);
  A.g_d0 = (float *) dzalloc(n0 * 16 * 4);
  A.g_y = (float *) dzalloc(n0 * 4);
  A.g_in0 = (float *) dzalloc(n0 * 16 * 4);
  A.part_n = MAX_CHUNKS * 27L * 64 * 96;
  A.part = (float *) dzalloc(A.part_n * 4);
  A.loss_n = blocks(n0, 256);
  A.loss_part = (float *) dzalloc(A.loss_n * 4);
  net->have[slot] = 1;
  return 0;
}

static TV tv(void *p, int ld, int c0, const Acts &A, int level)
{
  TV T;
  T.p = p; T.ld = ld; T.c0 = c0;
  T.nx = A.n[level][0]; T.ny = A.n[level][1]; T.nz = A.n[level][2];
  return T;
}

static TV none(void)
{
  TV T;
  memset(&T, 0, sizeof(T));
  return T;
}

/* Launch of conv_fwd_kernel: K step 32 where the input channels allow,
 * up to four output-channel tiles per wavefront. */
int bb_unet_conv_lds = 1;    /* 3x3x3 convolutions: 1 LDS-staged kernel,
                                0 direct loads */

template <typename IN>
static void conv_launch(TV X, int cin, int cinp, const uint16_t *W, int taps,
  const float *bias, int cout, int coutp, TV Yf, TV Yb, int relu)
{
  int ks = cinp % 32 == 0 ? 32 : 16, kpl = ks / 4;
  int tiles = coutp / 16;
  int nco = tiles % 4 == 0 ? 4 : (tiles % 3 == 0 ? 3 : (tiles % 2 == 0 ? 2 : 1));
  long nv = (long) X.nx * X.ny * X.nz;
  if (taps == 27 && bb_unet_conv_lds) {
    int vec4 = X.ld % 4 == 0 && X.c0 % 4 == 0;
    long nb = (long) ((X.nx + CB_X - 1) / CB_X) * ((X.ny + CB_Y - 1) / CB_Y) *
              ((X.nz + CB_Z - 1) / CB_Z);
    dim3 g((unsigned) nb, tiles / nco);
#define CONV_S(KS, NCO) hipLaunchKernelGGL((conv_lds_kernel<IN, KS, NCO>), \
    g, dim3(256), 0, bb_stream, X, cin, cinp, W, bias, cout, coutp, Yf, Yb, \
    relu, vec4)
    if (ks == 32) {
      if (nco == 4) CONV_S(32, 4);
      else if (nco == 3) CONV_S(32, 3);
      else if (nco == 2) CONV_S(32, 2);
      else CONV_S(32, 1);
    } else {
      if (nco == 4) CONV_S(16, 4);
      else if (nco == 3) CONV_S(16, 3);
      else if (nco == 2) CONV_S(16, 2);
      else CONV_S(16, 1);
    }
#undef CONV_S
    BB_LAUNCH_CHECK();
    return;
  }
  int vec = X.ld % kpl == 0 && X.c0 % kpl == 0;
  dim3 grid(blocks(nv, 128), tiles / nco);
#define CONV_L(KS, NCO) hipLaunchKernelGGL((conv_fwd_kernel<IN, KS, NCO>), \
    grid, dim3(256), 0, bb_stream, X, cin, cinp, W, taps, bias, cout, coutp, \
    Yf, Yb, relu, vec)
  if (ks == 32) {
    if (nco == 4) CONV_L(32, 4);
    else if (nco == 3) CONV_L(32, 3);
    else if (nco == 2) CONV_L(32, 2);
    else CONV_L(32, 1);
  } else {
    if (nco == 4) CONV_L(16, 4);
    else if (nco == 3) CONV_L(16, 3);
    else if (nco == 2) CONV_L(16, 2);
    else CONV_L(16, 1);
  }
#undef CONV_L
  BB_LAUNCH_CHECK();
}

static void conv(int l, TV X, bool xf32, TV Yf, TV Yb)
{
  Layer &Y = net->L[l];
  if (xf32)
    conv_launch<float>(X, Y.cin, Y.cinp, Y.wb, Y.taps, Y.b, Y.cout, Y.coutp,
                       Yf, Yb, Y.relu);
  else
    conv_launch<uint16_t>(X, Y.cin, Y.cinp, Y.wb, Y.taps, Y.b, Y.cout,
                          Y.coutp, Yf, Yb, Y.relu);
}

template <typename K, typename... Args>
static void ew(K kern, long n, Args... args)
{
  hipLaunchKernelGGL(kern, dim3(blocks(n, 256)), dim3(256), 0, bb_stream,
    args...);
}

/* Launch of a kernel with one thread per four channels of each of nv
 * voxels (C channels, C / 4 dividing the 256-thread workgroup). */
template <typename K, typename... Args>
static void vox4(K kern, long nv, int C, Args... args)
{
  const long vpb = 256 / (C / 4);
  if (nv > 0xffffffffL) {
    fprintf(stderr, "U-Net volume too large (%ld voxels)\n", nv);
    exit(EXIT_FAILURE);
  }
  hipLaunchKernelGGL(kern, dim3((unsigned) ((nv + vpb - 1) / vpb)), dim3(256),
    0, bb_stream, args...);
}

/* Forward pass on the BF16 2-channel input already in A.in0 (16-channel
 * rows, channels 0 and 1); output FP32 in A.y. */
static void forward(void)
{
  Acts &A = *net->A;
  conv(0, tv(A.in0, 16, 0, A, 0), false, none(), tv(A.e0a, 16, 0, A, 0));
  conv(1, tv(A.e0a, 16, 0, A, 0), false, none(), tv(A.cat0, 48, 32, A, 0));
  vox4(pool_fwd_kernel, A.nv[1], 16, tv(A.cat0, 48, 32, A, 0),
       tv(A.p0, 16, 0, A, 1), 16);
  conv(2, tv(A.p0, 16, 0, A, 1), false, none(), tv(A.e1a, 32, 0, A, 1));
  conv(3, tv(A.e1a, 32, 0, A, 1), false, none(), tv(A.cat1, 96, 64, A, 1));
  vox4(pool_fwd_kernel, A.nv[2], 32, tv(A.cat1, 96, 64, A, 1),
       tv(A.p1, 32, 0, A, 2), 32);
  conv(4, tv(A.p1, 32, 0, A, 2), false, none(), tv(A.b0, 64, 0, A, 2));
  conv(5, tv(A.b0, 64, 0, A, 2), false, none(), tv(A.b1, 64, 0, A, 2));
  vox4(up_fwd_kernel, A.nv[1], 64, tv(A.b1, 64, 0, A, 2),
       tv(A.cat1, 96, 0, A, 1), 64);
  conv(6, tv(A.cat1, 96, 0, A, 1), false, none(), tv(A.d1a, 32, 0, A, 1));
  conv(7, tv(A.d1a, 32, 0, A, 1), false, none(), tv(A.d1, 32, 0, A, 1));
  vox4(up_fwd_kernel, A.nv[0], 32, tv(A.d1, 32, 0, A, 1),
       tv(A.cat0, 48, 0, A, 0), 32);
  conv(8, tv(A.cat0, 48, 0, A, 0), false, none(), tv(A.d0a, 16, 0, A, 0));
  conv(9, tv(A.d0a, 16, 0, A, 0), false, none(), tv(A.d0, 16, 0, A, 0));
  conv(10, tv(A.d0, 16, 0, A, 0), false, tv(A.y, 1, 0, A, 0), none());
}

/* Backward of layer l: G (FP32, the gradient of the layer's output after
 * its ReLU mask has been applied), X (its BF16 input); accumulates dW, db
 * and writes the input gradient into GX (FP32) when GX.p != NULL. */
static void layer_bwd(int l, TV G, TV X, TV GX)
{
  Layer &Y = net->L[l];
  Acts &A = *net->A;
  long nv = (long) G.nx * G.ny * G.nz;
  long chunk = chunk_of(nv);
  long nchunks = (nv + chunk - 1) / chunk;
  int jobs = Y.taps * (Y.coutp / 16) * (Y.cinp / 16);
  hipLaunchKernelGGL(conv_wgrad_kernel, dim3((unsigned) nchunks,
    (unsigned) ((jobs + 3) / 4)), dim3(256), 0, bb_stream, G, Y.cout, X,
    Y.cin, Y.taps, Y.coutp, Y.cinp, chunk, A.part);
  ew(sum_chunks_kernel, Y.nw, A.part, Y.nw, (int) nchunks, Y.gw);
  hipLaunchKernelGGL(bias_grad_kernel, dim3((unsigned) nchunks),
    dim3(Y.coutp), 0, bb_stream, G, Y.cout, chunk, A.part);
  ew(sum_chunks_kernel, (long) Y.coutp, A.part, (long) Y.coutp,
     (int) nchunks, Y.gb);
  if (GX.p) {
    ew(flip_weights_kernel, Y.nw, Y.wb, Y.taps, Y.coutp, Y.cinp, Y.wt);
    conv_launch<float>(G, Y.cout, Y.coutp, Y.wt, Y.taps, (const float *) NULL,
                       Y.cin, Y.cinp, GX, none(), 0);
  }
  BB_LAUNCH_CHECK();
}

static void relu_bwd(TV G, TV Y, int C)
{
  ew(relu_bwd_kernel, (long) G.nx * G.ny * G.nz * C, G, Y, C);
}

/* Backward pass from the output gradient in A.g_y. */
static void backward(void)
{
  Acts &A = *net->A;
  /* out (1x1, no ReLU): d0 <- */
  layer_bwd(10, tv(A.g_y, 1, 0, A, 0), tv(A.d0, 16, 0, A, 0),
            tv(A.g_d0, 16, 0, A, 0));
  relu_bwd(tv(A.g_d0, 16, 0, A, 0), tv(A.d0, 16, 0, A, 0), 16);
  layer_bwd(9, tv(A.g_d0, 16, 0, A, 0), tv(A.d0a, 16, 0, A, 0),
            tv(A.g_d0a, 16, 0, A, 0));
  relu_bwd(tv(A.g_d0a, 16, 0, A, 0), tv(A.d0a, 16, 0, A, 0), 16);
  layer_bwd(8, tv(A.g_d0a, 16, 0, A, 0), tv(A.cat0, 48, 0, A, 0),
            tv(A.g_cat0, 48, 0, A, 0));
  /* cat0 = [up(d1) 0..31 | e0 32..47] */
  ew(up_bwd_kernel, A.nv[1] * 32, tv(A.g_cat0, 48, 0, A, 0),
     tv(A.g_d1, 32, 0, A, 1), 32);
  relu_bwd(tv(A.g_d1, 32, 0, A, 1), tv(A.d1, 32, 0, A, 1), 32);
  layer_bwd(7, tv(A.g_d1, 32, 0, A, 1), tv(A.d1a, 32, 0, A, 1),
            tv(A.g_d1a, 32, 0, A, 1));
  relu_bwd(tv(A.g_d1a, 32, 0, A, 1), tv(A.d1a, 32, 0, A, 1), 32);
  layer_bwd(6, tv(A.g_d1a, 32, 0, A, 1), tv(A.cat1, 96, 0, A, 1),
            tv(A.g_cat1, 96, 0, A, 1));
  /* cat1 = [up(b1) 0..63 | e1 64..95] */
  ew(up_bwd_kernel, A.nv[2] * 64, tv(A.g_cat1, 96, 0, A, 1),
     tv(A.g_b1, 64, 0, A, 2), 64);
  relu_bwd(tv(A.g_b1, 64, 0, A, 2), tv(A.b1, 64, 0, A, 2), 64);
  layer_bwd(5, tv(A.g_b1, 64, 0, A, 2), tv(A.b0, 64, 0, A, 2),
            tv(A.g_b0, 64, 0, A, 2));
  relu_bwd(tv(A.g_b0, 64, 0, A, 2), tv(A.b0, 64, 0, A, 2), 64);
  layer_bwd(4, tv(A.g_b0, 64, 0, A, 2), tv(A.p1, 32, 0, A, 2),
            tv(A.g_p1, 32, 0, A, 2));
  /* e1 gradient: skip path (already in g_cat1[64..95]) + pool path */
  ew(pool_bwd_kernel, A.nv[1] * 32, tv(A.g_cat1, 96, 64, A, 1),
     tv(A.g_p1, 32, 0, A, 2), 32);
  relu_bwd(tv(A.g_cat1, 96, 64, A, 1), tv(A.cat1, 96, 64, A, 1), 32);
  layer_bwd(3, tv(A.g_cat1, 96, 64, A, 1), tv(A.e1a, 32, 0, A, 1),
            tv(A.g_e1a, 32, 0, A, 1));
  relu_bwd(tv(A.g_e1a, 32, 0, A, 1), tv(A.e1a, 32, 0, A, 1), 32);
  layer_bwd(2, tv(A.g_e1a, 32, 0, A, 1), tv(A.p0, 16, 0, A, 1),
            tv(A.g_p0, 16, 0, A, 1));
  ew(pool_bwd_kernel, A.nv[0] * 16, tv(A.g_cat0, 48, 32, A, 0),
     tv(A.g_p0, 16, 0, A, 1), 16);
  relu_bwd(tv(A.g_cat0, 48, 32, A, 0), tv(A.cat0, 48, 32, A, 0), 16);
  layer_bwd(1, tv(A.g_cat0, 48, 32, A, 0), tv(A.e0a, 16, 0, A, 0),
            tv(A.g_e0a, 16, 0, A, 0));
  relu_bwd(tv(A.g_e0a, 16, 0, A, 0), tv(A.e0a, 16, 0, A, 0), 16);
  layer_bwd(0, tv(A.g_e0a, 16, 0, A, 0), tv(A.in0, 16, 0, A, 0), none());
}

/* ---------------------------- C entry points ----------------------------- */
/* Input: two FP32 channels-last fields (v*2 + c) on the device, copied to
 * the network's BF16 input. */
__global__ void pack_input_kernel(const float *__restrict__ in, long nv,
  uint16_t *__restrict__ x)
{
  long v = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const float2 f = ((const float2 *) in)[v];
  *(uint32_t *) &x[v * 16] = (uint32_t) f2bf(f.x) | ((uint32_t) f2bf(f.y) << 16);
}

/* Forward on a device input (FP32, 2 channels, channels-last) of the
 * current volume; the FP32 output (one channel) goes to `out` (device). */
extern "C" int bb_unet_forward(const float *in, float *out)
{
  if (!net || !net->A) return -1;
  Acts &A = *net->A;
  ew(pack_input_kernel, A.nv[0], in, A.nv[0], A.in0);
  forward();
  BB_CHECK(hipMemcpyAsync(out, A.y, A.nv[0] * sizeof(float),
    hipMemcpyDeviceToDevice, bb_stream));
  BB_LAUNCH_CHECK();
  return 0;
}

/* Loss 0.5 mean (y - t)^2 of the last forward against `target` (device,
 * FP32), its backward pass, and the gradients in the layers; returns the
 * loss (host). */
extern "C" double bb_unet_backward(const float *target)
{
  Acts &A = *net->A;
  hipLaunchKernelGGL(mse_kernel, dim3((unsigned) A.loss_n), dim3(256), 0,
    bb_stream, A.y, target, A.nv[0], A.g_y, A.loss_part);
  backward();
  float *h = (float *) malloc(A.loss_n * sizeof(float));
  BB_CHECK(hipMemcpyAsync(h, A.loss_part, A.loss_n * sizeof(float),
    hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
  double s = 0.;
  for (long i = 0; i < A.loss_n; i++) s += h[i];
  free(h);
  return s / (double) A.nv[0];
}

extern "C" void bb_unet_adam(float lr)
{
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  net->step++;
  float c1 = 1.f - powf(b1, (float) net->step);
  float c2 = 1.f - powf(b2, (float) net->step);
  for (int l = 0; l < NLAYER; l++) {
    Layer &Y = net->L[l];
    ew(adam_kernel, Y.nw, Y.w, Y.gw, Y.mw, Y.vw, Y.nw, lr, b1, b2, eps, c1,
       c2, Y.wb);
    ew(adam_kernel, (long) Y.coutp, Y.b, Y.gb, Y.mb, Y.vb, (long) Y.coutp,
       lr, b1, b2, eps, c1, c2, (uint16_t *) NULL);
  }
  BB_LAUNCH_CHECK();
}

/* Parameter access for tests: layer count and sizes, FP32 master weights
 * [taps][cout_pad][cin_pad], biases [cout_pad], gradients likewise. */
extern "C" int bb_unet_layer_info(int l, int *info)
{
  if (!net || l < 0 || l >= NLAYER) return -1;
  Layer &Y = net->L[l];
  info[0] = Y.cin; info[1] = Y.cout; info[2] = Y.taps; info[3] = Y.level;
  info[4] = Y.relu; info[5] = Y.cinp; info[6] = Y.coutp;
  return NLAYER;
}

extern "C" int bb_unet_get(int l, int what, float *host)
{
  if (!net || l < 0 || l >= NLAYER) return -1;
  Layer &Y = net->L[l];
  const float *src = what == 0 ? Y.w : what == 1 ? Y.b : what == 2 ? Y.gw : Y.gb;
  long n = (what == 0 || what == 2) ? Y.nw : Y.coutp;
  BB_CHECK(hipStreamSynchronize(bb_stream));
  BB_CHECK(hipMemcpy(host, src, n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int bb_unet_set(int l, int what, const float *host)
{
  if (!net || l < 0 || l >= NLAYER || (what != 0 && what != 1)) return -1;
  Layer &Y = net->L[l];
  float *dst = what == 0 ? Y.w : Y.b;
  long n = what == 0 ? Y.nw : Y.coutp;
  BB_CHECK(hipMemcpy(dst, host, n * sizeof(float), hipMemcpyHostToDevice));
  if (what == 0) upload_weights(Y);
  BB_CHECK(hipStreamSynchronize(bb_stream));
  return 0;
}

extern "C" void bb_unet_destroy(void)
{
  if (!net) return;
  for (int s = 0; s < 2; s++)
    if (net->have[s]) free_acts(net->acts[s]);
  for (int l = 0; l < NLAYER; l++) {
    Layer &Y = net->L[l];
    void *p[] = {Y.w, Y.b, Y.gw, Y.gb, Y.mw, Y.vw, Y.mb, Y.vb, Y.wb, Y.wt};
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++)
      if (p[i]) BB_CHECK(hipFree(p[i]));
  }
  free(net);
  net = NULL;
}
