// rkdg_nc.hip — gfx950 kernel for the non-conforming faces of tensor meshes
// (MFEM NCMesh slave faces, as the reference's face loop sees them:
// face_integrator.cpp:11-276 over an interior face whose Elem1 is the fine
// (slave) element and whose Elem2 is the coarser (master) element, Loc2 being
// the slave's point matrix in the master).
//
// One workgroup (one wavefront) per element that has such faces. For each
// slave face touching the element, the lanes take the face points (the slave
// face's Gauss-Legendre rule, the order of the conforming faces): both states
// by tensor Lagrange interpolation at the point's reference coordinates, the
// Riemann flux with the arguments ordered (slave state, master state, slave
// normal), so the two elements of the face see bitwise identical fluxes, and
// the element's signed share flux * w * |n|. Node lanes then lift the shares
// with their basis at the element's own reference points and write the sum,
// before the inverse mass; the element kernels add it to their face terms.
// Non-conforming faces are few (the interfaces between refinement levels), so
// the kernel favours simplicity: states staged through LDS, basis values formed
// from the points' coordinates rather than tabulated.
#include "rkdg_kernels.hpp"

namespace rkdg {

namespace {

// 1D Lagrange basis on the N Gauss-Legendre nodes xs at t
template <int N>
__device__ __forceinline__ void lagrange(const double* xs, double t, double* L) {
#pragma unroll
  for (int a = 0; a < N; a++) {
    double v = 1.0;
#pragma unroll
    for (int b = 0; b < N; b++)
      if (b != a) v *= (t - xs[b]) / (xs[a] - xs[b]);
    L[a] = v;
  }
}

}  // namespace

template <int DIM, int P>
__global__ __launch_bounds__(64) void nc_face_kernel(const NcArgs A, const SchemeParams g,
                                                     const TensorTables* __restrict__ gtb) {
  constexpr int N = P + 1;
  constexpr int ND = DIM == 3 ? N * N * N : N * N;
  constexpr int NEQ = DIM + 2;
  constexpr int NQ = DIM == 3 ? P + 2 : P + 1;
  constexpr int NP = DIM == 3 ? NQ * NQ : NQ;
  constexpr int PR = 3 * DIM + 1;  // per point: slave xi, master xi, w |n|, unit normal
  constexpr int RG = NP * PR;
  static_assert(ND <= 64 && NP <= 64, "one lane per node and per face point");

  __shared__ double s_x[N];
  __shared__ double s_u[2][NEQ][ND];   // slave and master states
  __shared__ double s_c[NP][NEQ];      // the element's share of flux * w * |n| per point
  __shared__ double s_xi[NP][DIM];     // the element's own reference coordinates of the points

  const int tid = threadIdx.x;
  const int64_t slot = blockIdx.x;
  const int64_t NDOFS = A.NE * ND;
  if (tid < N) s_x[tid] = gtb->x[tid];
  double z[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; v++) z[v] = 0.0;
  double lmax = 0.0;
  const int i0 = A.ptr[slot], i1 = A.ptr[slot + 1];
  for (int it = i0; it < i1; it++) {
    const int item = A.item[it];
    const int r = item >> 1, side = item & 1;
    const int es = A.rec[3 * r], em = A.rec[3 * r + 1];
    __syncthreads();  // the previous face's LDS reads are done
    for (int k = tid; k < NEQ * ND; k += 64) {
      const int v = k / ND, i = k - v * ND;
      s_u[0][v][i] = A.x[v * NDOFS + (int64_t)es * ND + i];
      s_u[1][v][i] = A.x[v * NDOFS + (int64_t)em * ND + i];
    }
    __syncthreads();
    if (tid < NP) {
      const double* G = A.geo + (int64_t)r * RG;
      const double* gp = G + tid * PR;
      double us[NEQ], um[NEQ];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        double L[3][N];
#pragma unroll
        for (int d = 0; d < DIM; d++) lagrange<N>(s_x, gp[h * DIM + d], L[d]);
        double* u = h == 0 ? us : um;
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < ND; i++) {
            const int a = i % N, b = (i / N) % N, c = i / (N * N);
            double phi = L[0][a] * L[1][b];
            if (DIM == 3) phi *= L[2][c];
            acc += s_u[h][v][i] * phi;
          }
          u[v] = acc;
        }
      }
      const double* nrm = gp + 2 * DIM + 1;
      double fl[NEQ];
      bool ok = true;
      const double lam = riemann<DIM>(g, us, um, nrm, fl, ok);
      if (!ok || lam < 0.0) {
        const bool nf = nonfinite_nodes(A.x, NDOFS, es, ND, NEQ) || nonfinite_nodes(A.x, NDOFS, em, ND, NEQ);
        flag_error(A.err, es, em, A.rec[3 * r + 2], nf ? kErrNonFinite : kErrUnphysical);
#pragma unroll
        for (int v = 0; v < NEQ; v++) fl[v] = 0.0;
      } else {
        lmax = lam > lmax ? lam : lmax;
      }
      const double wn = gp[2 * DIM] * (side == 0 ? -1.0 : 1.0);
#pragma unroll
      for (int v = 0; v < NEQ; v++) s_c[tid][v] = fl[v] * wn;
#pragma unroll
      for (int d = 0; d < DIM; d++) s_xi[tid][d] = gp[side * DIM + d];
    }
    __syncthreads();
    if (tid < ND) {
      const int a = tid % N, b = (tid / N) % N, c = tid / (N * N);
      for (int q = 0; q < NP; q++) {
        double L[3][N];
#pragma unroll
        for (int d = 0; d < DIM; d++) lagrange<N>(s_x, s_xi[q][d], L[d]);
        double phi = L[0][a] * L[1][b];
        if (DIM == 3) phi *= L[2][c];
#pragma unroll
        for (int v = 0; v < NEQ; v++) z[v] += s_c[q][v] * phi;
      }
    }
  }
  if (tid < ND) {
#pragma unroll
    for (int v = 0; v < NEQ; v++) A.out[(slot * NEQ + v) * ND + tid] = z[v];
  }
  if (!A.no_lambda) {
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(lmax, off, 64);
      lmax = o > lmax ? o : lmax;
    }
    if (tid == 0) atomic_max_pos(A.lam, lmax);
  }
}

hipError_t launch_nc_faces(int dim, int p, const NcArgs& a, const SchemeParams& g, const TensorTables* t,
                           hipStream_t s) {
  if (a.n_elem <= 0) return hipSuccess;
#define X(D, PP)                                                                 \
  if (dim == D && p == PP) {                                                     \
    nc_face_kernel<D, PP><<<(unsigned)a.n_elem, 64, 0, s>>>(a, g, t);            \
    return hipGetLastError();                                                    \
  }
  X(2, 1) X(2, 2) X(2, 3) X(2, 4) X(2, 5) X(3, 1) X(3, 2) X(3, 3)
#undef X
  return hipErrorInvalidValue;
}

}  // namespace rkdg
