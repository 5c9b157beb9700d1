// rkdg_batch.hip — device-side time control of batched ensembles, so that a
// whole step (stages, limiter, dt update) runs without the host and can be
// captured in a HIP graph. Per instance it reproduces the scalar time loop of
// main.cpp: first dt from face wave speeds (main.cpp:467-484), the step taken
// with dt_real = min(dt, t_end - t) (:512), the next dt = cfl*hmin/lambda/(2p+1)
// from the last stage's wave speed (:538-553), t += dt (:701) and the stop test
// t >= t_end - 1e-8 dt (:733). Contraction is off so every instance's dt and t
// are the same IEEE doubles a host loop computes.
#include "rkdg_kernels.hpp"

namespace rkdg {

namespace {

__device__ __forceinline__ double next_dt(double cfl, double h, double lam, int p) {
#pragma clang fp contract(off)
  return cfl * h / lam / (double)(2 * p + 1);
}

// t already holds the start times
__global__ __launch_bounds__(256) void batch_init_kernel(BatchTime b) {
#pragma clang fp contract(off)
  __shared__ int s_n;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  int n = 0;
  for (int i = threadIdx.x; i < b.count; i += blockDim.x) {
    b.dt[i] = next_dt(b.cfl, b.hmin[i], b.lam[i], b.p);
    b.dt_real[i] = b.dt[i];
    b.steps[i] = 0;
    b.active[i] = 1;
    n++;
  }
  atomicAdd(&s_n, n);
  __syncthreads();
  if (threadIdx.x == 0) b.n_active[0] = s_n;
}

__global__ __launch_bounds__(256) void batch_pre_kernel(BatchTime b) {
#pragma clang fp contract(off)
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < b.count; i += gridDim.x * blockDim.x) {
    if (b.active[i]) {
      const double r = b.t_final[i] - b.t[i];
      const double d = b.dt[i];
      b.dt_real[i] = r < d ? r : d;
    }
  }
}

__global__ __launch_bounds__(256) void batch_post_kernel(BatchTime b) {
#pragma clang fp contract(off)
  __shared__ int s_n;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  int n = 0;
  for (int i = threadIdx.x; i < b.count; i += blockDim.x) {
    if (!b.active[i]) continue;
    const double dt = next_dt(b.cfl, b.hmin[i], b.lam[i], b.p);
    const double t = b.t[i] + dt;
    b.dt[i] = dt;
    b.t[i] = t;
    b.steps[i] += 1;
    const double stop = b.t_final[i] - 1e-8 * dt;
    const bool done = t >= stop || (b.max_steps > 0 && b.steps[i] >= b.max_steps);
    if (done) b.active[i] = 0;
    else n++;
  }
  atomicAdd(&s_n, n);
  __syncthreads();
  if (threadIdx.x == 0) b.n_active[0] = s_n;
}

}  // namespace

hipError_t launch_batch_init(const BatchTime& b, hipStream_t s) {
  batch_init_kernel<<<1, 256, 0, s>>>(b);
  return hipGetLastError();
}

hipError_t launch_batch_pre(const BatchTime& b, hipStream_t s) {
  const int grid = (b.count + 255) / 256;
  batch_pre_kernel<<<grid, 256, 0, s>>>(b);
  return hipGetLastError();
}

hipError_t launch_batch_post(const BatchTime& b, hipStream_t s) {
  batch_post_kernel<<<1, 256, 0, s>>>(b);
  return hipGetLastError();
}

}  // namespace rkdg
