// rkdg_io.hip — kernels of the post-step diagnostics (no hot-path role):
// conservation totals (src/postprocessing/check_total_energy.cpp:3-63, done
// with the element means so that every order and element type integrates
// exactly), the primitive fields and their time integrals of main.cpp:388-451,
// 705-730, and evaluation of the state at given reference points (output).
#include "rkdg_kernels.hpp"

namespace rkdg {

// sum over a fixed range of elements of means[v*NE + e] * vol[e]; one partial
// per workgroup and variable, reduced in LDS in a fixed order (deterministic)
__global__ __launch_bounds__(256) void totals_kernel(const double* __restrict__ means,
                                                     const double* __restrict__ vol, int64_t NE, int neq,
                                                     int64_t per_block, double* __restrict__ partial) {
  __shared__ double s[256];
  const int64_t e0 = (int64_t)blockIdx.x * per_block;
  const int64_t e1 = e0 + per_block < NE ? e0 + per_block : NE;
  for (int v = 0; v < neq; v++) {
    double acc = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) acc += means[v * NE + e] * vol[e];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * neq + v] = s[0];
    __syncthreads();
  }
}

// rho, velocity, pressure (covolume EoS, physics.cpp:101-112) and temperature
// (physics.cpp:115-126, without the covolume as there) at every node:
// prim[q*ndofs + i] for q = rho, u_0..u_{dim-1}, p, T; acc += w * prim
__global__ __launch_bounds__(256) void primitives_kernel(const double* __restrict__ x, int64_t ndofs, int dim,
                                                         double gamma, double R, double b, double* prim,
                                                         double* acc, double w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndofs) return;
  const double rho = x[i];
  double m[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < dim; d++) m[d] = x[(1 + d) * ndofs + i];
  const double E = x[(1 + dim) * ndofs + i];
  double m2 = 0.0;
  for (int d = 0; d < dim; d++) m2 += m[d] * m[d];
  const double ei = E - 0.5 * m2 / rho;
  const double p = (gamma - 1.0) * ei / (1.0 - rho * b);
  const double T = (gamma - 1.0) * ei / rho / R;
  double q[6];
  q[0] = rho;
  for (int d = 0; d < dim; d++) q[1 + d] = m[d] / rho;
  q[1 + dim] = p;
  q[2 + dim] = T;
  for (int k = 0; k < dim + 3; k++) {
    if (prim) prim[k * ndofs + i] = q[k];
    if (acc) acc[k * ndofs + i] += w * q[k];
  }
}

// out[(v*NE + e)*npts + q] = sum_k phi[q*nd + k] x[v*NE*nd + e*nd + k]
__global__ __launch_bounds__(256) void interp_kernel(const double* __restrict__ x, int64_t NE, int nd, int neq,
                                                     const double* __restrict__ phi, int npts,
                                                     double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NE * npts) return;
  const int64_t e = t / npts;
  const int q = (int)(t - e * npts);
  const int64_t ndofs = NE * nd;
  for (int v = 0; v < neq; v++) {
    const double* xe = x + v * ndofs + e * nd;
    double s = 0.0;
    for (int k = 0; k < nd; k++) s += phi[q * nd + k] * xe[k];
    out[(v * NE + e) * npts + q] = s;
  }
}

hipError_t launch_totals(const double* means, const double* vol, int64_t NE, int neq, double* partial, int nblocks,
                         hipStream_t s) {
  const int64_t per = (NE + nblocks - 1) / nblocks;
  totals_kernel<<<nblocks, 256, 0, s>>>(means, vol, NE, neq, per, partial);
  return hipGetLastError();
}

hipError_t launch_primitives(const double* x, int64_t ndofs, int dim, double gamma, double R, double b, double* prim,
                             double* acc, double w, hipStream_t s) {
  if (ndofs <= 0) return hipSuccess;
  primitives_kernel<<<(unsigned)((ndofs + 255) / 256), 256, 0, s>>>(x, ndofs, dim, gamma, R, b, prim, acc, w);
  return hipGetLastError();
}

hipError_t launch_interp(const double* x, int64_t NE, int nd, int neq, const double* phi, int npts, double* out,
                         hipStream_t s) {
  const int64_t n = NE * npts;
  if (n <= 0) return hipSuccess;
  interp_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(x, NE, nd, neq, phi, npts, out);
  return hipGetLastError();
}

}  // namespace rkdg
