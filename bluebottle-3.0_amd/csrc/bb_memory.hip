/*
 * Device selection, allocation and host <-> device transfer.
 *
 * Follows src/cuda_bluebottle.cu:44-509: pinned host arrays in the reference
 * layouts, device arrays x-fastest (see bb_common.h).  cuda_dom_push and
 * cuda_dom_pull convert between the two layouts on the host.  The device
 * forcing is a per-component constant (src/cuda_bluebottle.cu:1811-1982
 * only ever adds constants in this path), so no device f_x/f_y/f_z arrays
 * exist; the host arrays are filled with the constant on pull.
 */
#include <string.h>

#include "bb_common.h"

DevDom ddom;
hipStream_t bb_stream = 0;
double *_bb_rows = NULL;
double *_bb_rows2 = NULL;
double *_bb_scalars = NULL;
double *bb_host_scalars = NULL;

/* Forcing constants, set by cuda_compute_forcing. */
double bb_force[3] = {0., 0., 0.};
extern int bb_history_rotated;

#define NSCALARS 64

extern "C" int cuda_device_count(void)
{
  int n = 0;
  BB_CHECK(hipGetDeviceCount(&n));
  return n;
}

extern "C" void cuda_device_init(int device)
{
  BB_CHECK(hipSetDevice(device));
}

extern "C" void cuda_block(void)
{
  BB_CHECK(hipStreamSynchronize(bb_stream));
}

extern "C" void cuda_check_errors(int line)
{
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "HIP error at line %d: %s\n", line, hipGetErrorString(e));
    exit(EXIT_FAILURE);
  }
}

extern "C" void cuda_blocks_init(void)
{
  /* Launch shapes are computed per call from ddom; nothing to cache. */
}

extern "C" void cuda_update_bc(void)
{
  /* The BC values travel to the kernels as arguments on every launch
   * (src/cuda_bluebottle.cu:435-440 keeps a device copy for the particle
   * kernels only). */
}

static Grid make_grid(const grid_info &G)
{
  Grid g;
  g.ni = G.in; g.nj = G.jn; g.nk = G.kn;
  g.ib = G.inb; g.jb = G.jnb; g.kb = G.knb;
  g.sy = (G.inb + GRID_ROW_ALIGN - 1) / GRID_ROW_ALIGN * GRID_ROW_ALIGN;
  g.sz = g.sy * G.jnb;
  g.n = (long) g.sz * G.knb;
  return g;
}

/* Device arrays on a grid: n elements of esize bytes starting GRID_ORIGIN
 * elements into a zeroed allocation (the allocation holds GRID_ORIGIN + 1
 * spare elements, so index n is still inside).  The table maps the
 * returned pointer back to the allocation for bb_grid_free. */
#define GRID_ALLOCS 512
static struct { void *user, *raw; } grid_tab[GRID_ALLOCS];

void *bb_grid_alloc(long n, size_t esize)
{
  size_t bytes = (size_t) (n + GRID_ORIGIN + 1) * esize;
  char *raw = NULL;
  BB_CHECK(hipMalloc((void **) &raw, bytes));
  BB_CHECK(hipMemset(raw, 0, bytes));
  gpumem += bytes;
  void *user = raw + (size_t) GRID_ORIGIN * esize;
  for (int i = 0; i < GRID_ALLOCS; i++)
    if (grid_tab[i].user == NULL) {
      grid_tab[i].user = user;
      grid_tab[i].raw = raw;
      return user;
    }
  fprintf(stderr, "bb_grid_alloc: more than %d grid arrays\n", GRID_ALLOCS);
  exit(EXIT_FAILURE);
}

void bb_grid_free(void *p)
{
  if (p == NULL) return;
  for (int i = 0; i < GRID_ALLOCS; i++)
    if (grid_tab[i].user == p) {
      BB_CHECK(hipFree(grid_tab[i].raw));
      grid_tab[i].user = grid_tab[i].raw = NULL;
      return;
    }
  fprintf(stderr, "bb_grid_free: %p is not a grid array\n", p);
  exit(EXIT_FAILURE);
}

static void *host_alloc(size_t bytes)
{
  void *ptr = NULL;
  BB_CHECK(hipHostMalloc(&ptr, bytes, hipHostMallocDefault));
  memset(ptr, 0, bytes);
  cpumem += bytes;
  return ptr;
}

extern "C" void cuda_dom_malloc_host(void)
{
  /* src/cuda_bluebottle.cu:76-141 */
  dom_struct *d = &dom[0];
  size_t cc = d->Gcc.s3b * sizeof(real);
  size_t fx = d->Gfx.s3b * sizeof(real);
  size_t fy = d->Gfy.s3b * sizeof(real);
  size_t fz = d->Gfz.s3b * sizeof(real);
  p = (real *) host_alloc(cc); p0 = (real *) host_alloc(cc);
  phi = (real *) host_alloc(cc);
  u = (real *) host_alloc(fx); u0 = (real *) host_alloc(fx);
  v = (real *) host_alloc(fy); v0 = (real *) host_alloc(fy);
  w = (real *) host_alloc(fz); w0 = (real *) host_alloc(fz);
  u_star = (real *) host_alloc(fx); v_star = (real *) host_alloc(fy);
  w_star = (real *) host_alloc(fz);
  f_x = (real *) host_alloc(fx); f_y = (real *) host_alloc(fy);
  f_z = (real *) host_alloc(fz);
  conv_u = (real *) host_alloc(fx); conv_v = (real *) host_alloc(fy);
  conv_w = (real *) host_alloc(fz);
  conv0_u = (real *) host_alloc(fx); conv0_v = (real *) host_alloc(fy);
  conv0_w = (real *) host_alloc(fz);
  diff_u = (real *) host_alloc(fx); diff_v = (real *) host_alloc(fy);
  diff_w = (real *) host_alloc(fz);
  diff0_u = (real *) host_alloc(fx); diff0_v = (real *) host_alloc(fy);
  diff0_w = (real *) host_alloc(fz);
  flag_u = (int *) host_alloc(d->Gfx.s3b * sizeof(int));
  flag_v = (int *) host_alloc(d->Gfy.s3b * sizeof(int));
  flag_w = (int *) host_alloc(d->Gfz.s3b * sizeof(int));
}

extern "C" void cuda_part_malloc_host(void)
{
  /* src/cuda_particle.cu:47-55: the host phase arrays, needed with or
   * without particles (called once after parts_read_input) */
  size_t cc = dom[0].Gcc.s3b * sizeof(int);
  if (phase == NULL) phase = (int *) host_alloc(cc);
  if (phase_shell == NULL) phase_shell = (int *) host_alloc(cc);
}

static void *dev_alloc(size_t bytes)
{
  void *ptr = NULL;
  BB_CHECK(hipMalloc(&ptr, bytes));
  BB_CHECK(hipMemset(ptr, 0, bytes));
  gpumem += bytes;
  return ptr;
}

extern "C" void cuda_dom_malloc_dev(void)
{
  /* src/cuda_bluebottle.cu:144-432 */
  dom_struct *d = &dom[0];
  bb_history_rotated = 0;
  bb_force[0] = bb_force[1] = bb_force[2] = 0.;
  ddom.cc = make_grid(d->Gcc);
  ddom.fx = make_grid(d->Gfx);
  ddom.fy = make_grid(d->Gfy);
  ddom.fz = make_grid(d->Gfz);
  ddom.nx = d->xn; ddom.ny = d->yn; ddom.nz = d->zn;
  ddom.dx = d->dx; ddom.dy = d->dy; ddom.dz = d->dz;
  ddom.perx = (bc.pW == PERIODIC);
  ddom.pery = (bc.pS == PERIODIC);
  ddom.perz = (bc.pB == PERIODIC);

  const long cc = ddom.cc.n, fx = ddom.fx.n, fy = ddom.fy.n, fz = ddom.fz.n;
  /* the stencil and PCG kernels address a grid array by a 32-bit byte
   * offset from its base: every FP64 grid array must stay below 4 GiB
   * (about 810^3 cells) */
  {
    long big = cc > fx ? cc : fx;
    big = big > fy ? big : fy;
    big = big > fz ? big : fz;
    if ((unsigned long) (big + GRID_ORIGIN) * sizeof(real) >= (1UL << 32)) {
      fprintf(stderr, "grid arrays of %ld elements exceed the 4 GiB that "
        "32-bit byte offsets address\n", big);
      exit(EXIT_FAILURE);
    }
  }
#define GA(n) ((real *) bb_grid_alloc(n, sizeof(real)))
  _p = GA(cc); _p0 = GA(cc);
  _phi = GA(cc); _invM = GA(cc);
  _rhs_p = GA(cc); _r_q = GA(cc);
  _z_q = GA(cc); _p_q = GA(cc);
  _pb_q = _p_q;                       /* one ghosted search direction */
  _Apb_q = GA(cc);
  _u = GA(fx); _u0 = GA(fx);
  _v = GA(fy); _v0 = GA(fy);
  _w = GA(fz); _w0 = GA(fz);
  _u_star = GA(fx); _v_star = GA(fy);
  _w_star = GA(fz);
  _conv_u = GA(fx); _conv_v = GA(fy);
  _conv_w = GA(fz);
  _conv0_u = GA(fx); _conv0_v = GA(fy);
  _conv0_w = GA(fz);
  _diff_u = GA(fx); _diff_v = GA(fy);
  _diff_w = GA(fz);
  _diff0_u = GA(fx); _diff0_v = GA(fy);
  _diff0_w = GA(fz);
#undef GA
  _flag_u = (int *) bb_grid_alloc(fx, sizeof(int));
  _flag_v = (int *) bb_grid_alloc(fy, sizeof(int));
  _flag_w = (int *) bb_grid_alloc(fz, sizeof(int));
  _phase = (int *) bb_grid_alloc(cc, sizeof(int));
  _phase_shell = (int *) bb_grid_alloc(cc, sizeof(int));

  long rows = (long) ddom.ny * ddom.nz + n_tiles(ddom);
  long faces = (long) (ddom.nx + 1) * (ddom.ny + 1);
  long cap = rows > faces ? rows : faces;
  faces = (long) (ddom.ny + 1) * (ddom.nz + 1);
  cap = cap > faces ? cap : faces;
  faces = (long) (ddom.nx + 1) * (ddom.nz + 1);
  cap = cap > faces ? cap : faces;
  _bb_rows = (double *) dev_alloc(cap * sizeof(double));
  _bb_rows2 = (double *) dev_alloc(cap * sizeof(double));
  _bb_scalars = (double *) dev_alloc(NSCALARS * sizeof(double));
  BB_CHECK(hipHostMalloc((void **) &bb_host_scalars,
    NSCALARS * sizeof(double), hipHostMallocDefault));
  /* All fluid until particle cages are built (src/particle.c:556-598). */
  bb_phase_fluid();
  bb_flags_reset();
}

/* ---- layout conversion between reference host order and device order ---- */
enum { L_GCC = 0, L_GFX, L_GFY, L_GFZ };

static long host_loc(const grid_info &G, int kind, int i, int j, int k)
{
  switch (kind) {
    case L_GFX: return GFX_LOC(i, j, k, G.s1b, G.s2b);
    case L_GFY: return GFY_LOC(i, j, k, G.s1b, G.s2b);
    case L_GFZ: return GFZ_LOC(i, j, k, G.s1b, G.s2b);
    default:    return GCC_LOC(i, j, k, G.s1b, G.s2b);
  }
}

static const grid_info &ginfo(int kind)
{
  switch (kind) {
    case L_GFX: return dom[0].Gfx;
    case L_GFY: return dom[0].Gfy;
    case L_GFZ: return dom[0].Gfz;
    default:    return dom[0].Gcc;
  }
}

static const Grid &dgrid(int kind)
{
  switch (kind) {
    case L_GFX: return ddom.fx;
    case L_GFY: return ddom.fy;
    case L_GFZ: return ddom.fz;
    default:    return ddom.cc;
  }
}

/* Host (reference layout, dense) to device (x-fastest, padded rows) and
 * back; the padding is zero on the device. */
template <typename T>
static void push_field(const T *host, T *dev, int kind)
{
  const grid_info &G = ginfo(kind);
  const Grid &D = dgrid(kind);
  size_t n = (size_t) D.n;
  T *stage = (T *) calloc(n, sizeof(T));
  for (int k = 0; k < G.knb; k++)
    for (int j = 0; j < G.jnb; j++)
      for (int i = 0; i < G.inb; i++)
        stage[gidx(D, i, j, k)] = host[host_loc(G, kind, i, j, k)];
  BB_CHECK(hipMemcpy(dev, stage, n * sizeof(T), hipMemcpyHostToDevice));
  free(stage);
}

template <typename T>
static void pull_field(T *host, const T *dev, int kind)
{
  const grid_info &G = ginfo(kind);
  const Grid &D = dgrid(kind);
  size_t n = (size_t) D.n;
  T *stage = (T *) malloc(n * sizeof(T));
  BB_CHECK(hipMemcpy(stage, dev, n * sizeof(T), hipMemcpyDeviceToHost));
  for (int k = 0; k < G.knb; k++)
    for (int j = 0; j < G.jnb; j++)
      for (int i = 0; i < G.inb; i++)
        host[host_loc(G, kind, i, j, k)] = stage[gidx(D, i, j, k)];
  free(stage);
}

extern "C" void cuda_dom_push(void)
{
  /* src/cuda_bluebottle.cu:443-490 */
  BB_CHECK(hipDeviceSynchronize());
  push_field(p, _p, L_GCC); push_field(p0, _p0, L_GCC);
  push_field(phi, _phi, L_GCC);
  push_field(u, _u, L_GFX); push_field(u0, _u0, L_GFX);
  push_field(v, _v, L_GFY); push_field(v0, _v0, L_GFY);
  push_field(w, _w, L_GFZ); push_field(w0, _w0, L_GFZ);
  push_field(u_star, _u_star, L_GFX); push_field(v_star, _v_star, L_GFY);
  push_field(w_star, _w_star, L_GFZ);
  push_field(conv_u, _conv_u, L_GFX); push_field(conv_v, _conv_v, L_GFY);
  push_field(conv_w, _conv_w, L_GFZ);
  push_field(conv0_u, _conv0_u, L_GFX); push_field(conv0_v, _conv0_v, L_GFY);
  push_field(conv0_w, _conv0_w, L_GFZ);
  push_field(diff_u, _diff_u, L_GFX); push_field(diff_v, _diff_v, L_GFY);
  push_field(diff_w, _diff_w, L_GFZ);
  push_field(diff0_u, _diff0_u, L_GFX); push_field(diff0_v, _diff0_v, L_GFY);
  push_field(diff0_w, _diff0_w, L_GFZ);
}

extern "C" void cuda_dom_pull(void)
{
  /* src/cuda_bluebottle.cu:492-509 */
  BB_CHECK(hipDeviceSynchronize());
  pull_field(p, _p, L_GCC); pull_field(p0, _p0, L_GCC);
  pull_field(phi, _phi, L_GCC);
  pull_field(u, _u, L_GFX); pull_field(u0, _u0, L_GFX);
  pull_field(v, _v, L_GFY); pull_field(v0, _v0, L_GFY);
  pull_field(w, _w, L_GFZ); pull_field(w0, _w0, L_GFZ);
  pull_field(u_star, _u_star, L_GFX); pull_field(v_star, _v_star, L_GFY);
  pull_field(w_star, _w_star, L_GFZ);
  /* After cuda_store_u the conv/diff histories are rotated by pointer: the
   * reference's conv == conv0 at that point, so both come from conv0. */
  pull_field(conv_u, bb_history_rotated ? _conv0_u : _conv_u, L_GFX);
  pull_field(conv_v, bb_history_rotated ? _conv0_v : _conv_v, L_GFY);
  pull_field(conv_w, bb_history_rotated ? _conv0_w : _conv_w, L_GFZ);
  pull_field(conv0_u, _conv0_u, L_GFX); pull_field(conv0_v, _conv0_v, L_GFY);
  pull_field(conv0_w, _conv0_w, L_GFZ);
  pull_field(diff_u, bb_history_rotated ? _diff0_u : _diff_u, L_GFX);
  pull_field(diff_v, bb_history_rotated ? _diff0_v : _diff_v, L_GFY);
  pull_field(diff_w, bb_history_rotated ? _diff0_w : _diff_w, L_GFZ);
  pull_field(diff0_u, _diff0_u, L_GFX); pull_field(diff0_v, _diff0_v, L_GFY);
  pull_field(diff0_w, _diff0_w, L_GFZ);
  pull_field(flag_u, _flag_u, L_GFX); pull_field(flag_v, _flag_v, L_GFY);
  pull_field(flag_w, _flag_w, L_GFZ);
  if (phase != NULL) pull_field(phase, _phase, L_GCC);
  if (phase_shell != NULL) pull_field(phase_shell, _phase_shell, L_GCC);
  dom_struct *d = &dom[0];
  for (int i = 0; i < d->Gfx.s3b; i++) f_x[i] = bb_force[0];
  for (int i = 0; i < d->Gfy.s3b; i++) f_y[i] = bb_force[1];
  for (int i = 0; i < d->Gfz.s3b; i++) f_z[i] = bb_force[2];
}

extern "C" void cuda_dom_pull_phase(void)
{
  /* src/cuda_bluebottle.cu:1455 */
  BB_CHECK(hipDeviceSynchronize());
  cuda_part_malloc_host();
  pull_field(phase, _phase, L_GCC);
  pull_field(phase_shell, _phase_shell, L_GCC);
}

extern "C" void cuda_dom_pull_debug(void)
{
  /* src/cuda_bluebottle.cu:1464-1505: phi, the intermediate velocity, the
   * convection and diffusion terms of this step and the face flags */
  BB_CHECK(hipDeviceSynchronize());
  pull_field(phi, _phi, L_GCC);
  pull_field(u_star, _u_star, L_GFX); pull_field(v_star, _v_star, L_GFY);
  pull_field(w_star, _w_star, L_GFZ);
  pull_field(conv_u, bb_history_rotated ? _conv0_u : _conv_u, L_GFX);
  pull_field(conv_v, bb_history_rotated ? _conv0_v : _conv_v, L_GFY);
  pull_field(conv_w, bb_history_rotated ? _conv0_w : _conv_w, L_GFZ);
  pull_field(diff_u, bb_history_rotated ? _diff0_u : _diff_u, L_GFX);
  pull_field(diff_v, bb_history_rotated ? _diff0_v : _diff_v, L_GFY);
  pull_field(diff_w, bb_history_rotated ? _diff0_w : _diff_w, L_GFZ);
  pull_field(flag_u, _flag_u, L_GFX); pull_field(flag_v, _flag_v, L_GFY);
  pull_field(flag_w, _flag_w, L_GFZ);
}

extern "C" void cuda_dom_pull_restart(void)
{
  /* src/cuda_bluebottle.cu:1507-1530: the fields of the previous step that
   * a restart needs besides those of cuda_dom_pull */
  BB_CHECK(hipDeviceSynchronize());
  pull_field(p0, _p0, L_GCC);
  pull_field(u0, _u0, L_GFX); pull_field(v0, _v0, L_GFY);
  pull_field(w0, _w0, L_GFZ);
  pull_field(conv0_u, _conv0_u, L_GFX); pull_field(conv0_v, _conv0_v, L_GFY);
  pull_field(conv0_w, _conv0_w, L_GFZ);
  pull_field(diff0_u, _diff0_u, L_GFX); pull_field(diff0_v, _diff0_v, L_GFY);
  pull_field(diff0_w, _diff0_w, L_GFZ);
}

extern "C" void printMemInfo(void)
{
  /* src/cuda_scalar.cu:972-986: device memory in use.  One GPU holds every
   * particle, so the subdomain count is nparts. */
  size_t free_byte = 0, total_byte = 0;
  hipError_t e = hipMemGetInfo(&free_byte, &total_byte);
  if (e != hipSuccess) {
    printf("Error: hipMemGetInfo fails, %s\n", hipGetErrorString(e));
    exit(1);
  }
  printf("N%d >> nparts = %d, nparts_subdom = %d, GPU memory usage: used = "
    "%zu B, free = %zu B, total = %zu B\n", rank, nparts, nparts,
    total_byte - free_byte, free_byte, total_byte);
  fflush(stdout);
}

/* Pull one Gcc device array (x-fastest, ghosted) into a host buffer of the
 * same order; used by tests and the solver statistics. */
extern "C" void bb_pull_raw(const void *dev, void *host, size_t bytes)
{
  BB_CHECK(hipDeviceSynchronize());
  BB_CHECK(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
}

extern "C" void bb_push_raw(void *dev, const void *host, size_t bytes)
{
  BB_CHECK(hipDeviceSynchronize());
  BB_CHECK(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
}

static void dfree(void *ptr)
{
  if (ptr) BB_CHECK(hipFree(ptr));
}

static void hfree(void *ptr)
{
  if (ptr) BB_CHECK(hipHostFree(ptr));
}

void bb_solver_free(void);

extern "C" void cuda_dom_free(void)
{
  /* src/cuda_bluebottle.cu:2619 */
  BB_CHECK(hipDeviceSynchronize());
  bb_solver_free();
  real *dd[] = {_p, _p0, _phi, _invM, _rhs_p, _r_q, _z_q, _p_q, _Apb_q,
    _u, _u0, _v, _v0, _w, _w0, _u_star, _v_star, _w_star,
    _conv_u, _conv_v, _conv_w, _conv0_u, _conv0_v, _conv0_w,
    _diff_u, _diff_v, _diff_w, _diff0_u, _diff0_v, _diff0_w};
  for (size_t i = 0; i < sizeof(dd) / sizeof(dd[0]); i++) bb_grid_free(dd[i]);
  bb_grid_free(_flag_u); bb_grid_free(_flag_v); bb_grid_free(_flag_w);
  bb_grid_free(_phase); bb_grid_free(_phase_shell);
  dfree(_bb_rows); dfree(_bb_rows2); dfree(_bb_scalars);
  hfree(bb_host_scalars);
  _p = _p0 = _phi = _invM = _rhs_p = _r_q = _z_q = _p_q = _pb_q = _Apb_q = NULL;
  _u = _u0 = _v = _v0 = _w = _w0 = _u_star = _v_star = _w_star = NULL;
  _conv_u = _conv_v = _conv_w = _conv0_u = _conv0_v = _conv0_w = NULL;
  _diff_u = _diff_v = _diff_w = _diff0_u = _diff0_v = _diff0_w = NULL;
  _flag_u = _flag_v = _flag_w = _phase = _phase_shell = NULL;
  _bb_rows = _bb_rows2 = _bb_scalars = bb_host_scalars = NULL;

  real *hh[] = {p, p0, phi, u, u0, v, v0, w, w0, u_star, v_star, w_star,
    f_x, f_y, f_z, conv_u, conv_v, conv_w, conv0_u, conv0_v, conv0_w,
    diff_u, diff_v, diff_w, diff0_u, diff0_v, diff0_w};
  for (size_t i = 0; i < sizeof(hh) / sizeof(hh[0]); i++) hfree(hh[i]);
  hfree(flag_u); hfree(flag_v); hfree(flag_w); hfree(phase);
  hfree(phase_shell);
  p = p0 = phi = u = u0 = v = v0 = w = w0 = NULL;
  u_star = v_star = w_star = f_x = f_y = f_z = NULL;
  conv_u = conv_v = conv_w = conv0_u = conv0_v = conv0_w = NULL;
  diff_u = diff_v = diff_w = diff0_u = diff0_v = diff0_w = NULL;
  flag_u = flag_v = flag_w = phase = phase_shell = NULL;
}

/* ------------------------- scalar fields -------------------------------- */
/* src/scalar.c:41-53; src/cuda_scalar.cu:18-146.  Gcc arrays, host in the
 * reference layout, device x-fastest. */
real *s = NULL, *s0 = NULL, *s_conv = NULL, *s_conv0 = NULL, *s_diff = NULL,
     *s_diff0 = NULL;
real *_s = NULL, *_s0 = NULL, *_s_conv = NULL, *_s_conv0 = NULL,
     *_s_diff = NULL, *_s_diff0 = NULL;

extern "C" void cuda_scalar_malloc_host(void)
{
  size_t cc = dom[0].Gcc.s3b * sizeof(real);
  s = (real *) host_alloc(cc); s0 = (real *) host_alloc(cc);
  s_conv = (real *) host_alloc(cc); s_conv0 = (real *) host_alloc(cc);
  s_diff = (real *) host_alloc(cc); s_diff0 = (real *) host_alloc(cc);
}

extern "C" void cuda_scalar_malloc_dev(void)
{
  const long cc = ddom.cc.n;
  _s = (real *) bb_grid_alloc(cc, sizeof(real));
  _s0 = (real *) bb_grid_alloc(cc, sizeof(real));
  _s_conv = (real *) bb_grid_alloc(cc, sizeof(real));
  _s_conv0 = (real *) bb_grid_alloc(cc, sizeof(real));
  _s_diff = (real *) bb_grid_alloc(cc, sizeof(real));
  _s_diff0 = (real *) bb_grid_alloc(cc, sizeof(real));
}

extern "C" void cuda_scalar_push(void)
{
  BB_CHECK(hipDeviceSynchronize());
  push_field(s, _s, L_GCC); push_field(s0, _s0, L_GCC);
  push_field(s_conv, _s_conv, L_GCC); push_field(s_conv0, _s_conv0, L_GCC);
  push_field(s_diff, _s_diff, L_GCC); push_field(s_diff0, _s_diff0, L_GCC);
}

extern "C" void cuda_scalar_pull(void)
{
  BB_CHECK(hipDeviceSynchronize());
  pull_field(s, _s, L_GCC); pull_field(s0, _s0, L_GCC);
  pull_field(s_conv, _s_conv, L_GCC); pull_field(s_conv0, _s_conv0, L_GCC);
  pull_field(s_diff, _s_diff, L_GCC); pull_field(s_diff0, _s_diff0, L_GCC);
}

extern "C" void cuda_scalar_pull_debug(void)
{
  /* src/cuda_scalar.cu:108-115 */
  BB_CHECK(hipDeviceSynchronize());
  pull_field(s_conv, _s_conv, L_GCC); pull_field(s_diff, _s_diff, L_GCC);
}

extern "C" void cuda_scalar_pull_restart(void)
{
  /* src/cuda_scalar.cu:117-126 */
  BB_CHECK(hipDeviceSynchronize());
  pull_field(s0, _s0, L_GCC); pull_field(s_conv0, _s_conv0, L_GCC);
  pull_field(s_diff0, _s_diff0, L_GCC);
}

extern "C" void cuda_scalar_free(void)
{
  BB_CHECK(hipDeviceSynchronize());
  real *dd[] = {_s, _s0, _s_conv, _s_conv0, _s_diff, _s_diff0};
  for (size_t i = 0; i < sizeof(dd) / sizeof(dd[0]); i++) bb_grid_free(dd[i]);
  real *hh[] = {s, s0, s_conv, s_conv0, s_diff, s_diff0};
  for (size_t i = 0; i < sizeof(hh) / sizeof(hh[0]); i++) hfree(hh[i]);
  _s = _s0 = _s_conv = _s_conv0 = _s_diff = _s_diff0 = NULL;
  s = s0 = s_conv = s_conv0 = s_diff = s_diff0 = NULL;
}
