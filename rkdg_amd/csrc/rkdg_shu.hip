// rkdg_shu.hip — Shu's troubled-cell indicator on neighbour-extrapolated
// means (src/indicator/indicator_shu.cpp:13-74, with the averages of
// src/averager/averager.cpp:152-273) and the FinDiff limiter it pairs with
// (src/limiter/limiter_findiff.cpp:3-33). A selectable option, off the
// benchmark's path: one thread per element.
//
// For element t and every cell k of its stencil (t, then its face
// neighbours), the extrapolated mean of variable v is coef_k . u_k[v], the
// host-built weights of k's nodal basis over t (the rule, the weight read
// without a point and the division by |k| of assembleShiftedElementMatrix
// and computeStencilExtrapAverages). The indicator of the last variable
// decides (limiter_findiff.cpp:7-8):
//   value = 1 if |mean_t| < eps and S < eps, else min(1, Ck |mean_t| / (S + eps)),
//   S = sum_k |mean_t - extrap_k|, Ck = 0.03, eps = 1e-6;
// maxFabsPj is the troubled cell's own |mean| (indicator_shu.cpp:40-48).
// The flags come first and the means replace the flagged elements in a second
// kernel, so every indicator reads the state before limiting (the reference
// writes into a copy, limiter.cpp:30,85,91).
#include "rkdg_kernels.hpp"

namespace rkdg {

__global__ __launch_bounds__(256) void shu_indicator_kernel(const double* __restrict__ x,
                                                            const double* __restrict__ means,
                                                            const double* __restrict__ coef,
                                                            const int* __restrict__ nbr, int64_t NE, int nd, int neq,
                                                            int nf, int8_t* __restrict__ flag,
                                                            int8_t* __restrict__ troubled) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= NE) return;
  const int64_t ndofs = NE * nd;
  const int v = neq - 1;
  const double mt = means[v * NE + e];
  const double* xv = x + v * ndofs;
  double S = 0.0;
  for (int s = 0; s <= nf; s++) {
    const int64_t k = s == 0 ? e : (int64_t)nbr[(s - 1) * NE + e];
    if (k < 0) continue;  // boundary faces are not in the stencil
    double ex = 0.0;
    for (int j = 0; j < nd; j++) ex += coef[((int64_t)s * nd + j) * NE + e] * xv[k * nd + j];
    S += fabs(mt - ex);
  }
  const double Ck = 0.03, eps = 1e-6;
  const double mx = fabs(mt);
  const double val = (mx < eps && S < eps) ? 1.0 : fmin(1.0, Ck * mx / (S + eps));
  const int8_t f = val < 0.999 ? 1 : 0;
  flag[e] = f;
  if (troubled) troubled[e] = f;
}

__global__ __launch_bounds__(256) void shu_apply_kernel(double* __restrict__ x, const double* __restrict__ means,
                                                        const int8_t* __restrict__ flag, int64_t NE, int nd, int neq) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t ndofs = NE * nd;
  if (i >= ndofs) return;
  const int64_t e = i / nd;
  if (!flag[e]) return;
  for (int v = 0; v < neq; v++) x[v * ndofs + i] = means[v * NE + e];
}

hipError_t launch_shu_indicator(const double* x, const double* means, const double* coef, const int* nbr, int64_t NE,
                                int nd, int neq, int nf, int8_t* flag, int8_t* troubled, hipStream_t s) {
  if (NE <= 0) return hipSuccess;
  shu_indicator_kernel<<<(unsigned)((NE + 255) / 256), 256, 0, s>>>(x, means, coef, nbr, NE, nd, neq, nf, flag,
                                                                      troubled);
  return hipGetLastError();
}

hipError_t launch_shu_apply(double* x, const double* means, const int8_t* flag, int64_t NE, int nd, int neq,
                            hipStream_t s) {
  const int64_t n = NE * nd;
  if (n <= 0) return hipSuccess;
  shu_apply_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(x, means, flag, NE, nd, neq);
  return hipGetLastError();
}

}  // namespace rkdg
