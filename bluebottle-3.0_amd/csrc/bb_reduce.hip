/*
 * Second stage of the canonical reduction (see bb_common.h): one 256-thread
 * workgroup folds the row partials into a single double in a device slot.
 * Replaces the thrust::reduce / thrust::inner_product calls of
 * src/cuda_solver.cu:151-231 and src/cuda_bluebottle.cu:2341-2526.
 */
#include <limits.h>

#include "bb_common.h"

__global__ void __launch_bounds__(REDUCE_THREADS)
reduce_rows_kernel(const double *__restrict__ rows, long nrows,
  double *__restrict__ out)
{
  double v = block_canonical_sum(rows, nrows);
  if (threadIdx.x == 0) *out = v;
}

void bb_reduce_rows(const double *rows, long nrows, double *out_slot)
{
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(1), dim3(REDUCE_THREADS), 0,
    bb_stream, rows, nrows, out_slot);
  BB_LAUNCH_CHECK();
}

/* Copy slots [0, n) to the pinned host mirror and wait for them. */
void bb_fetch_scalars(int n)
{
  BB_CHECK(hipMemcpyAsync(bb_host_scalars, _bb_scalars, n * sizeof(double),
    hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
  if (bb_time_kernels) bb_kt_collect();
}

double bb_fetch_scalar(int slot)
{
  BB_CHECK(hipMemcpyAsync(bb_host_scalars + slot, _bb_scalars + slot,
    sizeof(double), hipMemcpyDeviceToHost, bb_stream));
  BB_CHECK(hipStreamSynchronize(bb_stream));
  return bb_host_scalars[slot];
}

/* ------------------------------ kernel timing ---------------------------- */
/* One launch in bb_time_every of each timed kernel carries a pair of HIP
 * events, recorded by the runtime at the kernel's start and end
 * (BB_TIMED_LAUNCH).  A sample is tagged with the PCG
 * iteration (bb_kt_q) it belongs to and resolved when the solver has
 * fetched its state: the launches a batch enqueues past the iteration at
 * which the solve stopped return at once, and their samples are dropped
 * (bb_kt_collect_upto) so that averages cover full launches only. */
int bb_time_kernels = 0;
int bb_time_every = 8;                 /* time one launch in this many */
long bb_kt_q = 0;                      /* iteration of the next launch */
#define KT_SLOTS 32
static hipEvent_t kt_a[KT_COUNT][KT_SLOTS], kt_b[KT_COUNT][KT_SLOTS];
static long kt_tag[KT_COUNT][KT_SLOTS];
static int kt_pending[KT_COUNT], kt_init = 0, kt_active[KT_COUNT];
static long kt_seen[KT_COUNT];
static double kt_ms[KT_COUNT];
static long kt_n[KT_COUNT], kt_dropped[KT_COUNT];

hipEvent_t bb_kt_start = NULL, bb_kt_stop = NULL;

void bb_kt_begin(int id)
{
  kt_active[id] = 0;
  bb_kt_start = bb_kt_stop = NULL;
  if (!bb_time_kernels) return;
  if (bb_time_every > 1 && (kt_seen[id]++ % bb_time_every) != 0) return;
  kt_active[id] = 1;
  if (!kt_init) {
    for (int i = 0; i < KT_COUNT; i++)
      for (int j = 0; j < KT_SLOTS; j++) {
        BB_CHECK(hipEventCreate(&kt_a[i][j]));
        BB_CHECK(hipEventCreate(&kt_b[i][j]));
      }
    kt_init = 1;
  }
  if (kt_pending[id] == KT_SLOTS) bb_kt_collect();
  kt_tag[id][kt_pending[id]] = bb_kt_q;
  bb_kt_start = kt_a[id][kt_pending[id]];
  bb_kt_stop = kt_b[id][kt_pending[id]];
}

/* Every timed region launches exactly one kernel through BB_TIMED_LAUNCH. */
void bb_kt_end(int id)
{
  bb_kt_start = bb_kt_stop = NULL;
  if (!bb_time_kernels || !kt_active[id]) return;
  kt_pending[id]++;
}

void bb_kt_collect_upto(long qmax)
{
  if (!kt_init) return;
  for (int i = 0; i < KT_COUNT; i++) {
    for (int j = 0; j < kt_pending[i]; j++) {
      BB_CHECK(hipEventSynchronize(kt_b[i][j]));
      if (kt_tag[i][j] > qmax) {
        kt_dropped[i]++;
        continue;
      }
      float ms = 0.f;
      BB_CHECK(hipEventElapsedTime(&ms, kt_a[i][j], kt_b[i][j]));
      kt_ms[i] += ms;
      kt_n[i] += 1;
    }
    kt_pending[i] = 0;
  }
}

void bb_kt_collect(void)
{
  bb_kt_collect_upto(LONG_MAX);
}

/* Samples dropped as launches past the end of their solve. */
extern "C" long bb_ktimer_dropped(int id)
{
  return (id >= 0 && id < KT_COUNT) ? kt_dropped[id] : 0;
}

/* Accumulated milliseconds and launch count of timed kernel `id`. */
extern "C" void bb_ktimer_read(int id, double *ms, long *n)
{
  bb_kt_collect();
  *ms = (id >= 0 && id < KT_COUNT) ? kt_ms[id] : 0.;
  *n = (id >= 0 && id < KT_COUNT) ? kt_n[id] : 0;
}

extern "C" void bb_ktimer_reset(void)
{
  bb_kt_collect();
  for (int i = 0; i < KT_COUNT; i++) {
    kt_ms[i] = 0.; kt_n[i] = 0; kt_seen[i] = 0; kt_dropped[i] = 0;
  }
}

extern "C" void bb_ktimer_enable(int on)
{
  bb_kt_collect();
  bb_time_kernels = on;
}
