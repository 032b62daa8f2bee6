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
