// Residual of the step loop's convergence check (src/main/cruna_hpc.f90:185-186):
//   residual = maxval(abs((qn - q1) / dt))
// over every point and variable of two FP64 states.  Division by dt > 0 is
// monotone and sign-symmetric, so max|qn - q1| / dt is the same number.
// Two passes: k_absdiff_max leaves one maximum per work-group in `part`
// (grid-stride over 16-byte pairs, wavefront then work-group reduction);
// k_max_final reduces the partial maxima and divides by dt.  A maximum does
// not depend on the order of its operands, so the result is deterministic.
#pragma once

#include <hip/hip_runtime.h>

namespace cruna {

constexpr int RES_THREADS = 256;
constexpr int RES_MAX_PARTS = 1024;  // partial maxima of the first pass

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ double block_max(double v) {
  __shared__ double w[RES_THREADS / 64];
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) w[wid] = v;
  __syncthreads();
  v = threadIdx.x < RES_THREADS / 64 ? w[threadIdx.x] : 0.0;
  if (wid == 0) v = wave_max(v);
  return v;
}

// M values of a and b (M even: 5 fields of an even number of points or not,
// the tail is handled element-wise)
__global__ __launch_bounds__(RES_THREADS) void k_absdiff_max(const double *__restrict__ a,
                                                              const double *__restrict__ b, long M,
                                                              double *__restrict__ part) {
  double m = 0.0;
  const long M2 = M / 2;
  const double2 *a2 = reinterpret_cast<const double2 *>(a);
  const double2 *b2 = reinterpret_cast<const double2 *>(b);
  for (long i = (long)blockIdx.x * RES_THREADS + threadIdx.x; i < M2; i += (long)gridDim.x * RES_THREADS) {
    const double2 x = a2[i], y = b2[i];
    m = fmax(m, fmax(fabs(x.x - y.x), fabs(x.y - y.y)));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && (M & 1)) m = fmax(m, fabs(a[M - 1] - b[M - 1]));
  m = block_max(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(RES_THREADS) void k_max_final(const double *__restrict__ part, int nparts,
                                                            double dt, double *__restrict__ out) {
  double m = 0.0;
  for (int i = threadIdx.x; i < nparts; i += RES_THREADS) m = fmax(m, part[i]);
  m = block_max(m);
  if (threadIdx.x == 0) *out = m / dt;
}

}  // namespace cruna
