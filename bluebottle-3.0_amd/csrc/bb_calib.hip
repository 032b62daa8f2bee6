/*
 * Stream-copy calibration: b[i] = a[i] over n doubles, timed with HIP events
 * on the library stream.  Gives the measured HBM copy bandwidth of the box
 * (16 bytes moved per element) and a kernel of known traffic against which
 * the rocprofv3 FETCH_SIZE / WRITE_SIZE counters are scaled.
 */
#include "bb_common.h"

__global__ void __launch_bounds__(256)
stream_copy_kernel(const double *__restrict__ a, double *__restrict__ b, long n)
{
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

/* Returns the best-of-`reps` copy bandwidth in GB/s (1e9 bytes/s). */
extern "C" double bb_stream_copy_gbs(long n, int reps)
{
  double *a = NULL, *b = NULL;
  BB_CHECK(hipMalloc((void **) &a, n * sizeof(double)));
  BB_CHECK(hipMalloc((void **) &b, n * sizeof(double)));
  BB_CHECK(hipMemset(a, 0, n * sizeof(double)));
  BB_CHECK(hipMemset(b, 0, n * sizeof(double)));
  hipEvent_t e0, e1;
  BB_CHECK(hipEventCreate(&e0));
  BB_CHECK(hipEventCreate(&e1));
  long blocks = (n + 255) / 256;
  double best = 0.;
  for (int r = 0; r < reps + 1; r++) {
    BB_CHECK(hipEventRecord(e0, bb_stream));
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned) blocks), dim3(256), 0,
      bb_stream, a, b, n);
    BB_CHECK(hipEventRecord(e1, bb_stream));
    BB_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    BB_CHECK(hipEventElapsedTime(&ms, e0, e1));
    double gbs = 16.0 * (double) n / (ms * 1e-3) / 1e9;
    if (r > 0 && gbs > best) best = gbs;
  }
  BB_CHECK(hipEventDestroy(e0));
  BB_CHECK(hipEventDestroy(e1));
  BB_CHECK(hipFree(a));
  BB_CHECK(hipFree(b));
  return best;
}

/* The access pattern of the FP32 search kernel without its stencil: two
 * FP32 arrays read and two written at the same index (16 bytes per
 * element), one element per thread, the stores non-temporal as the search
 * kernel's.  Over the workload's cell count this is the simplest kernel
 * that moves the search kernel's algorithmic bytes in one launch: a
 * measured ceiling for it at that size. */
__global__ void __launch_bounds__(256)
stream4_kernel(const float *__restrict__ r, const float *__restrict__ p,
  float *__restrict__ q, float *__restrict__ a, long n, float w, float b)
{
  long i = (long) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float v = r[i] * w + b * p[i];
    __builtin_nontemporal_store(v, &q[i]);
    __builtin_nontemporal_store(v * w, &a[i]);
  }
}

/* Best-of-`reps` GB/s (1e9 bytes/s) of stream4_kernel over n elements. */
extern "C" double bb_stream4_gbs(long n, int reps)
{
  float *buf[4];
  for (int k = 0; k < 4; k++) {
    BB_CHECK(hipMalloc((void **) &buf[k], n * sizeof(float)));
    BB_CHECK(hipMemset(buf[k], 0, n * sizeof(float)));
  }
  hipEvent_t e0, e1;
  BB_CHECK(hipEventCreate(&e0));
  BB_CHECK(hipEventCreate(&e1));
  long blocks = (n + 255) / 256;
  double best = 0.;
  for (int r = 0; r < reps + 1; r++) {
    BB_CHECK(hipEventRecord(e0, bb_stream));
    hipLaunchKernelGGL(stream4_kernel, dim3((unsigned) blocks), dim3(256), 0,
      bb_stream, buf[0], buf[1], buf[2], buf[3], n, 0.5f, 0.25f);
    BB_CHECK(hipEventRecord(e1, bb_stream));
    BB_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    BB_CHECK(hipEventElapsedTime(&ms, e0, e1));
    double gbs = 16.0 * (double) n / (ms * 1e-3) / 1e9;
    if (r > 0 && gbs > best) best = gbs;
  }
  BB_CHECK(hipEventDestroy(e0));
  BB_CHECK(hipEventDestroy(e1));
  for (int k = 0; k < 4; k++) BB_CHECK(hipFree(buf[k]));
  return best;
}
