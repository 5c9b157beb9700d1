// rkdg_dense.hpp — compile-time layout of the dense small-element kernels
// (affine triangles P_p, bilinear quadrilaterals Q_p): work decomposition, the
// per-element geometry record and the reference tables, shared by the forward
// kernels (rkdg_dense.hip) and their adjoints (rkdg_adjoint.hip).
#pragma once

#include "rkdg_kernels.hpp"

// element types and orders instantiated
#define RKDG_DENSE_CASES(X) X(RKDG_GEOM_TRI, 1) X(RKDG_GEOM_TRI, 2) X(RKDG_GEOM_TRI, 3) \
  X(RKDG_GEOM_QUAD, 1) X(RKDG_GEOM_QUAD, 2) X(RKDG_GEOM_QUAD, 3)

namespace rkdg {

template <int GEOM, int P>
struct DC {
  static constexpr bool TRI = GEOM == RKDG_GEOM_TRI;
  static constexpr int DIM = 2;
  static constexpr int NEQ = 4;
  static constexpr int ND = TRI ? (P + 1) * (P + 2) / 2 : (P + 1) * (P + 1);
  static constexpr int NF = TRI ? 3 : 4;
  static constexpr int NV = TRI ? 3 : 4;
  // face rule order OrderW + 2p (+1 for P_k): triangle 2p+1, bilinear quad 2p+1
  static constexpr int NQ = TRI ? P + 1 : P + 1;
  static constexpr int NPT = NF * NQ;
  static constexpr int m1 = ND > NPT ? ND : NPT;
  static constexpr int T = m1 <= 4 ? 4 : (m1 <= 8 ? 8 : (m1 <= 16 ? 16 : 32));
  static constexpr int BLOCK = 64;
  static constexpr int EPB = BLOCK / T;
  // geometry record per element: affine triangle adj(J)[4], det J, area normals[2*NF];
  // bilinear quad: adj(J) at nodes [4*ND], Minv[ND*ND], node mass weights[ND], area normals[2*NF]
  static constexpr int G_ADJ = 0;
  static constexpr int G_DET = TRI ? 4 : -1;
  static constexpr int G_NRM = TRI ? 5 : 4 * ND + ND * ND + ND;
  static constexpr int G_MINV = TRI ? -1 : 4 * ND;
  static constexpr int G_MW = TRI ? -1 : 4 * ND + ND * ND;
  static constexpr int GSIZE = G_NRM + 2 * NF;
};

template <int GEOM, int P>
struct DenseTab {
  using C = DC<GEOM, P>;
  double K[2][C::ND][C::ND];        // volume: K[r][j][k]
  double Phi[C::NF][C::NQ][C::ND];  // face traces
  double wq[C::NQ];
  double Minv[C::ND][C::ND];        // reference inverse mass (triangles)
  double mhat[C::ND];               // integral of the reference basis (triangles)
  double PhiI[C::NF][2][C::ND];     // indicator points (IntRules order 2 on edges)
  double PhiV[C::NV][C::ND];        // reference vertices (last hope)
};

template <int GEOM, int P>
__device__ __forceinline__ void stage_dense(DenseTab<GEOM, P>& d, const double* s) {
  constexpr int n = sizeof(DenseTab<GEOM, P>) / sizeof(double);
  double* o = reinterpret_cast<double*>(&d);
  for (int i = threadIdx.x; i < n; i += blockDim.x) o[i] = s[i];
}

__device__ __forceinline__ int64_t xcd_remap_d(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t x = b % 8, l = b / 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + l;
}

}  // namespace rkdg
