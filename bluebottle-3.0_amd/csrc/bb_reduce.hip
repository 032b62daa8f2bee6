/*
 * Second stage of the canonical reduction (see bb_common.h): one 256-thread
 * workgroup folds the row partials into a single double in a device slot.
 * Replaces the thrust::reduce / thrust::inner_product calls of
 * src/cuda_solver.cu:151-231 and src/cuda_bluebottle.cu:2341-2526.
 */
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
int bb_time_kernels = 0;
int bb_time_every = 8;                 /* time one launch in this many */
static hipEvent_t kt_a[KT_COUNT], kt_b[KT_COUNT];
static int kt_pending[KT_COUNT], kt_init = 0, kt_active[KT_COUNT];
static long kt_seen[KT_COUNT];
static double kt_ms[KT_COUNT];
static long kt_n[KT_COUNT];

void bb_kt_begin(int id)
{
  kt_active[id] = 0;
  if (!bb_time_kernels) return;
  if (bb_time_every > 1 && (kt_seen[id]++ % bb_time_every) != 0) return;
  kt_active[id] = 1;
  if (!kt_init) {
    for (int i = 0; i < KT_COUNT; i++) {
      BB_CHECK(hipEventCreate(&kt_a[i]));
      BB_CHECK(hipEventCreate(&kt_b[i]));
    }
    kt_init = 1;
  }
  if (kt_pending[id]) bb_kt_collect();
  BB_CHECK(hipEventRecord(kt_a[id], bb_stream));
}

void bb_kt_end(int id)
{
  if (!bb_time_kernels || !kt_active[id]) return;
  BB_CHECK(hipEventRecord(kt_b[id], bb_stream));
  kt_pending[id] = 1;
}

void bb_kt_collect(void)
{
  if (!kt_init) return;
  for (int i = 0; i < KT_COUNT; i++) {
    if (!kt_pending[i]) continue;
    BB_CHECK(hipEventSynchronize(kt_b[i]));
    float ms = 0.f;
    BB_CHECK(hipEventElapsedTime(&ms, kt_a[i], kt_b[i]));
    kt_ms[i] += ms;
    kt_n[i] += 1;
    kt_pending[i] = 0;
  }
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
  for (int i = 0; i < KT_COUNT; i++) { kt_ms[i] = 0.; kt_n[i] = 0; kt_seen[i] = 0; }
}

extern "C" void bb_ktimer_enable(int on)
{
  bb_kt_collect();
  bb_time_kernels = on;
}
