/*
 * Learned initial guess for the pressure solve (pp_warmstart LEARNED).
 *
 * Each solve of -A phi = -b starts from the guess with the least A-norm
 * error in span{phi_1 .. phi_h, psi}: phi_1 .. phi_h are the solutions of
 * the last h solves (most recent first; h <= bb_warm_history), and
 * psi = U-Net(r0 / |r0|_rms, fluid mask) is the network's prediction of
 * the correction phi - phi_1 from the defect r0 = b - (-A) phi_1 of the
 * coming solve (for a linear operator the correction is a fixed function
 * of the defect, which is what the network learns).  The Galerkin system
 * fixes every weight, so the guess is never worse in the A-norm than
 * phi0 = 0 (the reference's start) or than phi_1 alone (the fallback).
 * Successive solves of a timestep (the Lamb passes) change the right-hand
 * side by amounts that shrink as the particle boundary values converge, so
 * the earlier solutions span most of the next correction; the first solve
 * of a timestep is where they span little, and where the network runs.
 * By default (bb_warm_coarse) the network works on the grid coarsened 2x
 * per direction: its input is the defect averaged over 2x2x2 cells, and
 * its output is interpolated back trilinearly, which leaves psi smooth
 * (the part of the correction the PCG reduces slowest) at an eighth of
 * the full grid's cost.  A guess that came from the network can then be
 * smoothed by bb_warm_sweeps damped Jacobi sweeps.  Every solve still ends
 * at the same test (r, M^-1 r) <= pp_residual^2 (b, b).
 *
 * The network is trained online: the first bb_warm_train_solves solves
 * each contribute one r0 -> (phi - phi_1) pair (both scaled to unit
 * RMS, averaged like the input on the coarse grid), on which
 * bb_warm_steps Adam steps are taken over random cubic crops of
 * bb_warm_crop voxels; afterwards the network only infers, by default for
 * the first solve of each timestep.
 *
 * No counterpart in the reference, which always starts from phi = 0
 * (src/solver_kernel.cu:258-283).
 */
#include <math.h>
#include <stdint.h>

#include "bb_common.h"

extern "C" int bb_unet_create(unsigned long long seed);
extern "C" int bb_unet_volume(int nx, int ny, int nz);
extern "C" int bb_unet_forward(const float *in, float *out);
extern "C" double bb_unet_backward(const float *target);
extern "C" void bb_unet_adam(float lr);
extern "C" void bb_unet_destroy(void);

int bb_warm_train_solves = 64;         /* solves that feed the training */
int bb_warm_steps = 32;                /* Adam steps per training pair */
int bb_warm_crop = 64;                 /* training crop edge (cells) */
float bb_warm_lr = 1e-3f;
unsigned long long bb_warm_seed = 20260104ULL;
/* Statistics of the last solve. */
int bb_warm_used = 0;                  /* 1: phi0 came from the network */
double bb_warm_alpha = 0.;
double bb_warm_loss = 0.;              /* last training loss */
#define LOSS_LOG 4096
double bb_warm_loss_log[LOSS_LOG];     /* loss of every Adam step, in order
                                          (the first LOSS_LOG) */
int bb_warm_loss_n = 0;
int bb_warm_solves = 0;                /* solves seen */

int bb_warm_use_net = 1;               /* 0: span{phi_last} only */
static int net_step = -1;              /* timestep of the last inference */
int bb_warm_net_all_passes = 0;        /* 0: the network runs for the first
                                          solve of a timestep only; the
                                          later Lamb passes of the step use
                                          span{phi_last} */
int bb_warm_use_last = 1;              /* 0: span{psi} only */
int bb_warm_history = 6;               /* earlier solutions in the span
                                          (1: the last one only; at 256^3
                                          with 500 spheres 6 measured 1.6 %
                                          faster per step than 4 and than 8) */
int bb_warm_sweeps = 0;                /* damped Jacobi sweeps on a guess
                                          that came from the network (the
                                          coarse network's interpolated
                                          output needs none: 2 sweeps saved
                                          3 iterations per step at 256^3
                                          and cost more than that) */
double bb_warm_omega = 0.8;            /* their damping factor */
/* What the network gained on the last solve that used it: A-norm errors
 * squared of the guess from span{phi_last} alone and of the guess used,
 * and their running sums over all such solves. */
double bb_warm_err_last = 0., bb_warm_err_guess = 0.;
double bb_warm_err_last_sum = 0., bb_warm_err_guess_sum = 0.;
/* the same for the guess from the earlier solutions without psi */
double bb_warm_err_hist = 0., bb_warm_err_hist_sum = 0.;
int bb_warm_net_solves = 0;
static int net_this_solve = 0;
double bb_warm_beta = 0.;
#define HIST_MAX 8
static double *_hist[HIST_MAX];        /* ring of earlier solutions */
static int hist_n = 0, hist_head = 0;  /* count, next slot */
/* Generation of each direction's content: a new value for every write
 * (never reused, also across set-ups), so the solver can keep the Gram
 * entries of directions that did not change. */
static long gen_counter = 0;
static long hist_gen[HIST_MAX];
static long psi_gen = 0;
/* Particle-mask generation each earlier solution is zero in the solid
 * cells of (a solution is, for the mask it was solved with). */
extern long bb_mask_gen;               /* bb_particle.hip */
static long hist_mask[HIST_MAX];
/* Solutions by Lamb pass: set cur holds this timestep's, 1 - cur the
 * previous timestep's (pass j of a step tends to differ from pass j - 1 as
 * it did one step earlier).  Allocated on first use. */
#define PASS_MAX 16
#define SPAN_MAX 9                     /* directions of the Galerkin system */
int bb_warm_pass_span = 3;             /* the previous timestep's solutions
                                          in the span: 1 of this Lamb pass,
                                          2 also of the pass before, 3 also
                                          of the pass after */
static double *_pass[2][PASS_MAX];
static long pass_gen[2][PASS_MAX], pass_mask[2][PASS_MAX];
static int pass_n[2] = {0, 0}, pass_cur = 0, pass_idx = 0;
static long pass_step = -1;
static double *_psi = NULL, *_wres = NULL;
static float *_win = NULL, *_wtgt = NULL, *_wout = NULL;
static float *_cin = NULL, *_ctgt = NULL;
static double *_wpart = NULL, *wpart_host = NULL;
static long wnv = 0;                   /* interior cells of the grid */
static int crop_edge = 0;
int bb_warm_coarse = 1;                /* 1: the network works on the grid
                                          coarsened 2x in each direction
                                          (inputs and targets averaged over
                                          2x2x2 cells, its output
                                          interpolated back trilinearly);
                                          0: on the grid itself */
/* The network's volume: the grid's interior divided by nf = 1 or 2. */
struct NetVol { int nf, nx, ny, nz; };
static NetVol nvol = {1, 0, 0, 0};
static long nnv = 0;                   /* voxels of the network's volume */

/* Interior cell v = x + nx*(y + ny*z) of the Gcc grid. */
__device__ inline long cell_of_v(const DevDom &D, long v)
{
  int x = (int) (v % D.nx);
  long r = v / D.nx;
  int y = (int) (r % D.ny), z = (int) (r / D.ny);
  return gidx(D.cc, x + 1, y + 1, z + 1);
}

/* Per-block partial sums of {d, d^2, p, p^2} with d = b - w p over
 * interior cells. */
__global__ void __launch_bounds__(256)
moments_kernel(DevDom D, const double *__restrict__ b,
  const double *__restrict__ p, double w, double *__restrict__ part)
{
  __shared__ double s[4][256];
  long nv = (long) D.nx * D.ny * D.nz;
  double a[4] = {0., 0., 0., 0.};
  for (long v = (long) blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += (long) gridDim.x * blockDim.x) {
    long c = cell_of_v(D, v);
    double y = p[c], x = b[c] - w * y;
    a[0] += x; a[1] += x * x; a[2] += y; a[3] += y * y;
  }
  for (int q = 0; q < 4; q++) s[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int q = 0; q < 4; q++) s[q][threadIdx.x] += s[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; q++) part[blockIdx.x * 4 + q] = s[q][0];
}

#define MBLK 512

static void moments(const double *b, const double *p, double w, double m[4])
{
  hipLaunchKernelGGL(moments_kernel, dim3(MBLK), dim3(256), 0, bb_stream,
    ddom, b, p, w, _wpart);
  BB_CHECK(hipMemcpyAsync(wpart_host, _wpart, MBLK * 4 * sizeof(double),
    hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
  for (int q = 0; q < 4; q++) m[q] = 0.;
  for (int i = 0; i < MBLK; i++)
    for (int q = 0; q < 4; q++) m[q] += wpart_host[i * 4 + q];
}

/* Network input per voxel of the network's volume: [(r0 - rm) * ir,
 * fluid] averaged over the voxel's nf^3 cells (nf = 1: the cell itself;
 * fluid = 1 in fluid cells, 0 in particle cells, where r0 is 0 too). */
__device__ inline long cell_of_child(const DevDom &D, const NetVol &N, long v,
  int q)
{
  int x = (int) (v % N.nx);
  long r = v / N.nx;
  int y = (int) (r % N.ny), z = (int) (r / N.ny);
  x = x * N.nf + (q % N.nf);
  y = y * N.nf + (q / N.nf) % N.nf;
  z = z * N.nf + q / (N.nf * N.nf);
  return gidx(D.cc, x + 1, y + 1, z + 1);
}

__global__ void warm_input_kernel(DevDom D, NetVol N, const double *__restrict__ r0,
  const uint8_t *__restrict__ fluid, double rm, double ir,
  float *__restrict__ in)
{
  long v = (long) blockIdx.x * blockDim.x + threadIdx.x;
  long nv = (long) N.nx * N.ny * N.nz;
  if (v >= nv) return;
  const int nc = N.nf * N.nf * N.nf;
  double a = 0., fr = 0.;
  for (int q = 0; q < nc; q++) {
    long c = cell_of_child(D, N, v, q);
    bool f = fluid == NULL || fluid[c];
    a += f ? (r0[c] - rm) * ir : 0.;
    fr += f ? 1. : 0.;
  }
  in[2 * v] = (float) (a / nc);
  in[2 * v + 1] = (float) (fr / nc);
}

/* Training target ((phi - phi_last) - m) * s, averaged over the nf^3
 * cells of a network voxel (0 in particle cells). */
__global__ void warm_target_kernel(DevDom D, NetVol N,
  const double *__restrict__ phi, const double *__restrict__ last,
  const uint8_t *__restrict__ fluid, double m, double s, float *__restrict__ t)
{
  long v = (long) blockIdx.x * blockDim.x + threadIdx.x;
  long nv = (long) N.nx * N.ny * N.nz;
  if (v >= nv) return;
  const int nc = N.nf * N.nf * N.nf;
  double a = 0.;
  for (int q = 0; q < nc; q++) {
    long c = cell_of_child(D, N, v, q);
    bool f = fluid == NULL || fluid[c];
    a += f ? ((phi[c] - last[c]) - m) * s : 0.;
  }
  t[v] = (float) (a / nc);
}

/* Random cubic crop (edge e, origin ox, oy, oz) of the input and target. */
__global__ void crop_kernel(NetVol N, const float *__restrict__ in,
  const float *__restrict__ t, int e, int ox, int oy, int oz,
  float *__restrict__ cin, float *__restrict__ ct)
{
  long v = (long) blockIdx.x * blockDim.x + threadIdx.x;
  long n = (long) e * e * e;
  if (v >= n) return;
  int x = (int) (v % e), y = (int) ((v / e) % e), z = (int) (v / ((long) e * e));
  long src = (long) (x + ox) + (long) N.nx * ((y + oy) + (long) N.ny * (z + oz));
  cin[2 * v] = in[2 * src];
  cin[2 * v + 1] = in[2 * src + 1];
  ct[v] = t[src];
}

/* The two network voxels along one direction that interpolate fine cell
 * x (nf = 2): its own coarse voxel with weight 3/4 and the neighbour on
 * the side of x with 1/4 (cell-centred trilinear weights), wrapped in a
 * periodic direction and clamped to the voxel itself at a wall. */
__device__ inline void prolong_1d(int x, int n, int per, int *i0, int *i1)
{
  int I = x >> 1, J = (x & 1) ? I + 1 : I - 1;
  if (J < 0) J = per ? n - 1 : I;
  if (J >= n) J = per ? 0 : I;
  *i0 = I;
  *i1 = J;
}

/* psi (FP32 network output) into the Gcc FP64 array phi, zero in particle
 * cells and ghosts untouched: the voxel's value (nf = 1) or the trilinear
 * interpolation of the coarse voxels around the cell (nf = 2). */
__global__ void warm_place_kernel(DevDom D, NetVol N, const float *__restrict__ y,
  const uint8_t *__restrict__ fluid, double *__restrict__ phi)
{
  long v = (long) blockIdx.x * blockDim.x + threadIdx.x;
  long nv = (long) D.nx * D.ny * D.nz;
  if (v >= nv) return;
  long c = cell_of_v(D, v);
  bool f = fluid == NULL || fluid[c];
  if (!f) { phi[c] = 0.; return; }
  if (N.nf == 1) { phi[c] = (double) y[v]; return; }
  int x = (int) (v % D.nx);
  long r = v / D.nx;
  int yy = (int) (r % D.ny), z = (int) (r / D.ny);
  int ix[2], iy[2], iz[2];
  prolong_1d(x, N.nx, D.perx, &ix[0], &ix[1]);
  prolong_1d(yy, N.ny, D.pery, &iy[0], &iy[1]);
  prolong_1d(z, N.nz, D.perz, &iz[0], &iz[1]);
  const double w[2] = {0.75, 0.25};
  double a = 0.;
  for (int cz = 0; cz < 2; cz++)
    for (int cy = 0; cy < 2; cy++)
      for (int cx = 0; cx < 2; cx++)
        a += w[cx] * w[cy] * w[cz] *
             (double) y[ix[cx] + (long) N.nx * (iy[cy] + (long) N.ny * iz[cz])];
  phi[c] = a;
}

/* Zero in the particle cells of the current mask (interior cells).  A
 * solution kept from an earlier solve holds values in cells that particles
 * have entered since; the fluid rows of the coming solve would see them
 * as boundary data, where a solve from phi = 0 sees zeros. */
__global__ void mask_solid_kernel(DevDom D, const uint8_t *__restrict__ fluid,
  double *__restrict__ a)
{
  long v = (long) blockIdx.x * blockDim.x + threadIdx.x;
  long nv = (long) D.nx * D.ny * D.nz;
  if (v >= nv) return;
  long c = cell_of_v(D, v);
  if (!fluid[c]) a[c] = 0.;
}

static inline unsigned nb(long n) { return (unsigned) ((n + 255) / 256); }

static bool dims_ok(int nf)
{
  const int m = 4 * nf;
  return ddom.nx % m == 0 && ddom.ny % m == 0 && ddom.nz % m == 0;
}

extern "C" void bb_warm_setup(void)
{
  if (pp_warmstart != PP_WARM_LEARNED) return;
  const int nf = bb_warm_coarse ? 2 : 1;
  if (!dims_ok(nf)) {
    fprintf(stderr, "pp_warmstart LEARNED needs grid dimensions divisible by "
      "%d (got %d %d %d)\n", 4 * nf, ddom.nx, ddom.ny, ddom.nz);
    exit(EXIT_FAILURE);
  }
  bb_unet_create(bb_warm_seed);
  wnv = (long) ddom.nx * ddom.ny * ddom.nz;
  nvol.nf = nf;
  nvol.nx = ddom.nx / nf; nvol.ny = ddom.ny / nf; nvol.nz = ddom.nz / nf;
  nnv = (long) nvol.nx * nvol.ny * nvol.nz;
  BB_CHECK(hipMalloc((void **) &_win, 2 * nnv * sizeof(float)));
  BB_CHECK(hipMalloc((void **) &_wtgt, nnv * sizeof(float)));
  BB_CHECK(hipMalloc((void **) &_wout, nnv * sizeof(float)));
  int e = bb_warm_crop;
  int m = nvol.nx < nvol.ny ? nvol.nx : nvol.ny;
  m = m < nvol.nz ? m : nvol.nz;
  if (e > m) e = m;
  e -= e % 4;
  crop_edge = e;
  long ne = (long) e * e * e;
  BB_CHECK(hipMalloc((void **) &_cin, 2 * ne * sizeof(float)));
  BB_CHECK(hipMalloc((void **) &_ctgt, ne * sizeof(float)));
  BB_CHECK(hipMalloc((void **) &_wpart, MBLK * 4 * sizeof(double)));
  _psi = (double *) bb_grid_alloc(ddom.cc.n, sizeof(double));
  for (int h = 0; h < HIST_MAX; h++)
    _hist[h] = (double *) bb_grid_alloc(ddom.cc.n, sizeof(double));
  _wres = (double *) bb_grid_alloc(ddom.cc.n, sizeof(double));
  hist_n = hist_head = 0;
  for (int h = 0; h < HIST_MAX; h++) hist_mask[h] = -1;
  pass_n[0] = pass_n[1] = 0;
  pass_cur = pass_idx = 0;
  pass_step = -1;
  net_step = -1;
  net_this_solve = 0;
  bb_warm_err_last = bb_warm_err_guess = bb_warm_err_hist = 0.;
  bb_warm_err_last_sum = bb_warm_err_guess_sum = bb_warm_err_hist_sum = 0.;
  bb_warm_net_solves = 0;
  bb_warm_loss_n = 0;
  BB_CHECK(hipHostMalloc((void **) &wpart_host, MBLK * 4 * sizeof(double),
    hipHostMallocDefault));
  bb_warm_solves = 0;
}

extern "C" void bb_warm_teardown(void)
{
  if (!_win) return;
  float *f[] = {_win, _wtgt, _wout, _cin, _ctgt};
  for (int i = 0; i < 5; i++) BB_CHECK(hipFree(f[i]));
  BB_CHECK(hipFree(_wpart));
  BB_CHECK(hipHostFree(wpart_host));
  bb_grid_free(_psi);
  for (int h = 0; h < HIST_MAX; h++) {
    bb_grid_free(_hist[h]);
    _hist[h] = NULL;
  }
  bb_grid_free(_wres);
  for (int c = 0; c < 2; c++)
    for (int j = 0; j < PASS_MAX; j++) {
      if (_pass[c][j]) bb_grid_free(_pass[c][j]);
      _pass[c][j] = NULL;
    }
  _psi = _wres = NULL;
  _win = _wtgt = _wout = _cin = _ctgt = NULL;
  _wpart = NULL;
  wpart_host = NULL;
  bb_unet_destroy();
}

static double rms_scale(double sum, double sum2, long n, double *mean)
{
  *mean = sum / n;
  double var = sum2 / n - (*mean) * (*mean);
  return var > 1e-300 ? 1. / sqrt(var) : 0.;
}

/* The defect b - (-A) x, bb_solver.hip. */
void bb_pp_defect(const double *x, double *out);

/* The most recent earlier solution (a zero array before the first). */
static double *phi_last(void)
{
  return _hist[(hist_head + HIST_MAX - 1) % HIST_MAX];
}

/* Network input from the defect of phi_1 for the current right-hand
 * side (phi_1 = 0 before the first solve: the defect is b). */
static void build_input(const uint8_t *fluid)
{
  bb_pp_defect(phi_last(), _wres);
  double m[4];
  moments(_wres, _wres, 0., m);
  double rm;
  double ir = rms_scale(m[0], m[1], wnv, &rm);
  hipLaunchKernelGGL(warm_input_kernel, dim3(nb(nnv)), dim3(256), 0,
    bb_stream, ddom, nvol, _wres, fluid, rm, ir, _win);
}

/* Galerkin guess on several directions and (b, x), bb_solver.hip. */
double bb_pp_galerkin_many(const double *const *vecs, const long *gens, int m,
  double *e_first, double *e_lead);
double bb_pp_energy(const double *x);
void bb_pp_smooth(int nsweeps, double omega);
static double e_guess = 0., e_last = 0., e_hist = 0.;

/* C ABI: the guess for the coming solve into _phi.  Returns 1 if it was
 * placed, 0 if the solve starts from phi = 0. */
extern "C" int cuda_PP_warmstart_infer(void)
{
  bb_warm_used = 0;
  bb_warm_alpha = bb_warm_beta = 0.;
  if (pp_warmstart != PP_WARM_LEARNED || !_win) return 0;
  const uint8_t *fluid = NPARTS > 0 ? _bb_fluid : NULL;
  int net_ready = bb_warm_use_net && bb_warm_solves >= bb_warm_train_solves &&
                  (bb_warm_net_all_passes || stepnum != net_step);
  if (net_ready) net_step = stepnum;
  int hmax = bb_warm_history < 1 ? 1 : (bb_warm_history > HIST_MAX
                                          ? HIST_MAX : bb_warm_history);
  int nh = bb_warm_use_last ? (hist_n < hmax ? hist_n : hmax) : 0;
  if (stepnum != pass_step) {            /* a new timestep: pass 0 */
    pass_cur = 1 - pass_cur;
    pass_n[pass_cur] = 0;
    pass_idx = 0;
    pass_step = stepnum;
  }
  /* the previous timestep's solutions of this pass and the one before
   * join the span; the span holds at most SPAN_MAX directions, psi
   * included (the solver's Gram system) */
  const int np = bb_warm_pass_span < 0 ? 0 : (bb_warm_pass_span > 3 ? 3
                                              : bb_warm_pass_span);
  int npass = 0;
  for (int d = 0; d < np; d++) {
    const int jj = pass_idx + (d == 2 ? 1 : -d);
    npass += jj >= 0 && jj < pass_n[1 - pass_cur];
  }
  if (nh + npass + net_ready > SPAN_MAX) nh = SPAN_MAX - npass - net_ready;
  /* the earlier solutions, most recent first; particles may have moved
   * since they were solved, and a value kept in a cell a particle has
   * entered would act as boundary data of the fluid rows */
  const double *vecs[SPAN_MAX];
  long gens[SPAN_MAX];
  for (int h = 0; h < nh; h++) {
    const int slot = (hist_head + HIST_MAX - 1 - h) % HIST_MAX;
    double *v = _hist[slot];
    /* zero in the cells of the current mask, once per mask generation
     * (the particles moved; the solver's kept Gram entries lapse then
     * anyway) */
    if (fluid && hist_mask[slot] != bb_mask_gen) {
      hipLaunchKernelGGL(mask_solid_kernel, dim3(nb(wnv)), dim3(256), 0,
        bb_stream, ddom, fluid, v);
      hist_mask[slot] = bb_mask_gen;
    }
    vecs[h] = v;
    gens[h] = hist_gen[slot];
  }
  {
    const int pv = 1 - pass_cur;
    for (int d = 0; d < np; d++) {
      const int jj = pass_idx + (d == 2 ? 1 : -d);
      if (jj < 0 || jj >= pass_n[pv]) continue;
      double *v = _pass[pv][jj];
      if (fluid && pass_mask[pv][jj] != bb_mask_gen) {
        hipLaunchKernelGGL(mask_solid_kernel, dim3(nb(wnv)), dim3(256), 0,
          bb_stream, ddom, fluid, v);
        pass_mask[pv][jj] = bb_mask_gen;
      }
      vecs[nh] = v;
      gens[nh++] = pass_gen[pv][jj];
    }
  }
  if (net_ready) {
    build_input(fluid);
    bb_unet_volume(nvol.nx, nvol.ny, nvol.nz);
    bb_unet_forward(_win, _wout);
    hipLaunchKernelGGL(warm_place_kernel, dim3(nb(wnv)), dim3(256), 0,
      bb_stream, ddom, nvol, _wout, fluid, _psi);
    BB_LAUNCH_CHECK();
    psi_gen = ++gen_counter;
  }
  int m = nh;
  if (net_ready) {
    gens[m] = psi_gen;
    vecs[m++] = _psi;
  }
  e_guess = e_last = e_hist = 0.;
  if (m == 0) return 0;
  /* one Gram pass; the energies of the span without psi and of phi_1
   * alone come from the same matrix, for the record of what psi adds */
  double ef = 0., el = 0.;
  e_guess = bb_pp_galerkin_many(vecs, gens, m, &ef, &el);
  e_last = nh > 0 ? ef : 0.;
  e_hist = (net_ready && nh > 0) ? el : e_guess;
  bb_warm_used = e_guess > 0.;
  net_this_solve = net_ready && nh > 0 && bb_warm_used;
  /* smoothing sweeps on the network's guess */
  if (net_ready && bb_warm_used) bb_pp_smooth(bb_warm_sweeps, bb_warm_omega);
  return bb_warm_used;
}

static unsigned long long crng = 0x2545F4914F6CDD1DULL;
static int crand(int n)
{
  crng ^= crng << 13; crng ^= crng >> 7; crng ^= crng << 17;
  return n > 0 ? (int) (crng % (unsigned long long) n) : 0;
}

/* Adam steps on random crops of the pair (defect of phi_last -> phi -
 * phi_last) of the solve just finished. */
static void train_pair(void)
{
  const uint8_t *fluid = NPARTS > 0 ? _bb_fluid : NULL;
  build_input(fluid);
  double m[4];
  moments(_phi, phi_last(), 1., m);
  double em;
  double is = rms_scale(m[0], m[1], wnv, &em);
  if (is == 0.) return;
  hipLaunchKernelGGL(warm_target_kernel, dim3(nb(nnv)), dim3(256), 0,
    bb_stream, ddom, nvol, _phi, phi_last(), fluid, em, is, _wtgt);
  int e = crop_edge;
  bb_unet_volume(e, e, e);
  for (int s = 0; s < bb_warm_steps; s++) {
    int ox = crand(nvol.nx - e + 1), oy = crand(nvol.ny - e + 1),
        oz = crand(nvol.nz - e + 1);
    hipLaunchKernelGGL(crop_kernel, dim3(nb((long) e * e * e)), dim3(256), 0,
      bb_stream, nvol, _win, _wtgt, e, ox, oy, oz, _cin, _ctgt);
    bb_unet_forward(_cin, _wout);
    bb_warm_loss = bb_unet_backward(_ctgt);
    if (bb_warm_loss_n < LOSS_LOG) bb_warm_loss_log[bb_warm_loss_n++] = bb_warm_loss;
    bb_unet_adam(bb_warm_lr);
  }
  BB_LAUNCH_CHECK();
}

/* C ABI: one training round on the solve just finished (input: the defect
 * of phi_1; target: phi - phi_1), then phi joins the earlier solutions. */
extern "C" void cuda_PP_warmstart_train(void)
{
  if (pp_warmstart != PP_WARM_LEARNED || !_win) return;
  int train = bb_warm_use_net && bb_warm_solves < bb_warm_train_solves;
  bb_warm_solves++;
  if (net_this_solve) {
    /* (b, phi*) - E is the A-norm error squared of a Galerkin guess that
     * captures the energy E, against the converged phi* */
    double es = bb_pp_energy(_phi);
    bb_warm_err_last = es - e_last;
    bb_warm_err_hist = es - e_hist;
    bb_warm_err_guess = es - e_guess;
    bb_warm_err_last_sum += bb_warm_err_last;
    bb_warm_err_hist_sum += bb_warm_err_hist;
    bb_warm_err_guess_sum += bb_warm_err_guess;
    bb_warm_net_solves++;
    net_this_solve = 0;
  }
  if (train) train_pair();
  BB_CHECK(hipMemcpyAsync(_hist[hist_head], _phi,
    ddom.cc.n * sizeof(double), hipMemcpyDeviceToDevice, bb_stream));
  hist_gen[hist_head] = ++gen_counter;
  hist_mask[hist_head] = bb_mask_gen;
  hist_head = (hist_head + 1) % HIST_MAX;
  if (hist_n < HIST_MAX) hist_n++;
  if (bb_warm_pass_span && pass_step == stepnum && pass_idx < PASS_MAX) {
    double *&v = _pass[pass_cur][pass_idx];
    if (!v) v = (double *) bb_grid_alloc(ddom.cc.n, sizeof(double));
    BB_CHECK(hipMemcpyAsync(v, _phi, ddom.cc.n * sizeof(double),
      hipMemcpyDeviceToDevice, bb_stream));
    pass_gen[pass_cur][pass_idx] = ++gen_counter;
    pass_mask[pass_cur][pass_idx] = bb_mask_gen;
    pass_n[pass_cur] = pass_idx + 1;
  }
  pass_idx++;
}
