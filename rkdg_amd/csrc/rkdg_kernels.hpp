// rkdg_kernels.hpp — argument blocks and launchers shared by the host code and
// the HIP kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rkdg_physics.hpp"

namespace rkdg {

constexpr int kMaxN = 6;  // up to p = 5 nodes per direction
constexpr int kMaxQ = 8;  // face quadrature points per direction
constexpr int kMaxPacked = 3 * kMaxN + kMaxN * kMaxN + 2 * kMaxN + kMaxQ * kMaxN + kMaxQ + 2 * kMaxN;

// 1D tables of a tensor-product element of order p (N = p+1 Gauss-Legendre
// nodes) and its face rule (NQ Gauss-Legendre points).
struct TensorTables {
  double x[kMaxN], w[kMaxN], invw[kMaxN];
  double D[kMaxN][kMaxN];   // D[a][b] = psi_b'(x_a)
  double E[2][kMaxN];       // E[c][b] = psi_b(c), c in {0, 1}
  double B[kMaxQ][kMaxN];   // B[q][b] = psi_b(s_q), face points s_q
  double wq[kMaxQ];         // face weights
  double BI[2][kMaxN];      // psi_b at the two indicator points (IntRules order 2)
  // the tables of this order packed as the kernels keep them in LDS: w[N],
  // invw[N], D[N][N], E[2][N], B[NQ][N], wq[NQ], BI[2][N]
  double packed[kMaxPacked];
};

// Butcher combination (wi == 0): out = (base ? base : stage input) + coef * k;
// Shu-Osher combination (wi != 0): out = wb * base + wi * (stage input + coef * k)
struct StageOut {
  double* out;
  const double* base;
  double coef;
  double wb, wi;
};

struct RhsArgs {
  const double* x;   // stage input (limited)
  double* k;         // optional: store the RHS itself
  int nout;
  StageOut outs[4];
  double* means;     // optional: element means of outs[means_of] ([v*NE + e])
  int means_of;      // the tensor kernel takes 0 (rhs_launch reorders the outputs)
  const double* geo;   // affine tensor geometry [g][NE]: adj(J) row-major, then det(J)
  const int* nbr;      // [lf][NE] neighbour element, or -1-bc_slot for boundary faces
  const int* fcode;    // [lf][NE] bits 0-2 neighbour local face, 3-5 orientation, 6 = Elem1
  int64_t NE;
  double* lam;         // device max wave speed (non-negative doubles compared as int64)
  long long* err;      // device error word [4]: flag, e1, e2, local face
  int face_only;       // first-dt path: face terms and wave speeds only, no outputs
  // batched ensembles (all null for a single instance): element -> instance,
  // per-instance step size multiplying every stage coefficient, per-instance
  // boundary parameters [inst][RKDG_MAX_BC][8], per-instance active flag; lam
  // then holds one slot per instance
  const int* inst;
  const double* dt_inst;
  const double* bcp;
  const int* active;
  int grid_cap;  // dense kernel: persistent grid of at most this many workgroups (0: one per work block)
  int no_lambda; // 1: the wave speed of this launch is not needed (tensor kernel skips the nodal one)
  // tensor kernel (nullable): work block processed by each workgroup (a block is
  // the EPB elements from block * EPB), in an order that keeps face neighbours
  // close in time so their gathered states are still in L2
  const int* order;
  // Two-colour (red-black) form of the tensor kernel, for meshes whose face
  // graph is bipartite and has no boundary faces: every face joins a "solver"
  // element and a "receiver" element. role 1 runs over the solver elements,
  // solves each of their faces once and writes the receiver's share of the
  // lifted face flux into fbuf; role 2 runs over the receivers, which add it
  // instead of solving faces. role 0: every element solves its own faces.
  int role;
  const int* elist;    // roles 1, 2: the elements of work block b are elist[b*EPB .. b*EPB+EPB-1]
  int64_t nlist;       // number of elements in elist
  const int* fslot;    // [NE]: receiver slot of each element (its index in the receivers' elist)
  double* fbuf;        // [receivers][NF][NEQ][face nodes]: lifted face fluxes in the receiver's face frame
  int aligned;         // tensor kernel: every interior face joins faces (r, c) and (r, 1-c), identity orientation
  // Limiter applied on load (tensor kernel, role 0; both null otherwise): x
  // holds the unlimited stage value and the kernel works on the limited one,
  // mean + f (u - mean) per element and variable with f = lim_f[e] and
  // mean = lim_mean[v*NE + e] (from the limiter kernel's factor pass). The
  // stage input of the epilogue's Shu-Osher combinations is then the
  // unlimited value, read again from x.
  const double* lim_f;
  const double* lim_mean;
  // non-conforming faces (tensor kernel, nullable): the element's share of
  // the fluxes over its non-conforming faces, before the inverse mass, is
  // ncbuf[(ncslot[e] * NEQ + v) * ND + node] where ncslot[e] >= 0
  const int* ncslot;
  const double* ncbuf;
  // curved (non-affine) hexahedra (tensor kernel, CV instantiation): the
  // per-element record of CurvedLayout, AoS [e][CurvedLayout::size]
  const double* cgeo;
};

struct LimitArgs {
  double* x;           // state limited in place
  const double* means; // [v*NE + e] of the pre-limit state
  const double* geo;
  const int* nbr;
  const int* fcode;
  int64_t NE;
  int8_t* troubled;
  const int* inst;    // batched ensembles: element -> instance (nullable)
  const int* active;  // per-instance active flag (nullable)
  double* out;        // dense kernels: write the limited state here instead of in place (nullable)
  // tensor kernel, factor pass (nullable): write each element's limiter factor
  // f here and leave x as it is; the limited state is mean + f (u - mean)
  // (f = y^nd or y, 0 after a reset to the mean, 1 where nothing is limited)
  double* factor;
  // non-conforming faces (tensor kernel, nullable): a face coded
  // kNcCode0 - k in nbr has the neighbours ncnb[4 k .. 4 k + 3] (-1 padded)
  const int* ncnb;
};

struct MeansArgs {
  const double* x;
  double* means;
  const double* geo;
  int64_t NE;
  const double* cgeo;  // curved hexahedra (nullable): the mean weights of CurvedLayout
};

// Geometry record of a curved (non-affine, trilinear) hexahedron of order P,
// AoS per element: adj(J) at every node [ND][3][3] (row-major) for the
// collocated volume term; the inverse of the exact mass matrix [ND][ND],
// transposed (MassIntegrator's rule, order 2p + OrderW); the mean weights [ND] (the
// integrals of the basis, Averager's rule); per local face (MFEM numbering)
// and face point (qi + NQ qj in the face frame of the kernels) the element's
// outward unit normal, |n| and 0.5 / |unit normal| [NF][NFP][5]; the volume.
template <int P>
struct CurvedLayout {
  static constexpr int N = P + 1, ND = N * N * N, NQ = P + 2, NFP = NQ * NQ, NF = 6;
  static constexpr int adj = 0;
  static constexpr int minv = adj + ND * 9;
  static constexpr int mw = minv + ND * ND;
  static constexpr int fn = mw + ND;
  static constexpr int vol = fn + NF * NFP * 5;
  static constexpr int size = vol + 1;
};

// LDS staging of a kernel's setup data in two steps, so that a workgroup issues
// every global load of its setup before the first wait: load() fills
// registers, store() writes them to LDS. (A strided copy loop that the
// compiler keeps as a loop waits for each load before its store.)
//
// n doubles from s, copied by the BLOCK lanes of the workgroup
template <int n, int BLOCK>
struct CopyRegs {
  static constexpr int passes = (n + BLOCK - 1) / BLOCK;
  double v[passes];
  __device__ __forceinline__ void load(const double* s, int tid) {
#pragma unroll
    for (int p = 0; p < passes; p++) v[p] = tid + p * BLOCK < n ? s[tid + p * BLOCK] : 0.0;
  }
  __device__ __forceinline__ void store(double* o, int tid) const {
#pragma unroll
    for (int p = 0; p < passes; p++)
      if (tid + p * BLOCK < n) o[tid + p * BLOCK] = v[p];
  }
};

// a wave-uniform address kept in scalar registers: passing it through
// readfirstlane stops the compiler from re-associating the uniform part of an
// address into per-lane 64-bit arithmetic
template <typename Tp>
__device__ __forceinline__ Tp* sbase(Tp* p) {
  const uint64_t u = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return reinterpret_cast<Tp*>(((uint64_t)hi << 32) | lo);
}

// the NEQ-variable states of the EPB elements from e0 (byNODES layout) in one
// pass: lane tid holds node tid % ND of element e0 + tid / ND; past the last
// element the value is `fill`
template <int EPB, int NEQ, int ND, int BLOCK>
struct StateRegs {
  static_assert(EPB * ND <= BLOCK, "one pass of the workgroup covers its elements' nodes");
  double v[NEQ];
  __device__ __forceinline__ void load(const double* x, int64_t NDOFS, int64_t e0, int64_t NE, int tid,
                                       double fill = 1.0) {
    const bool ok = tid < EPB * ND && e0 + tid / ND < NE;
#pragma unroll
    for (int q = 0; q < NEQ; q++) v[q] = ok ? sbase(x + (q * NDOFS + e0 * ND))[(unsigned)tid] : fill;
  }
  // the same for the elements elist[li0 .. li0+EPB-1] (entries from nlist on:
  // `fill`); the list entries are read at wave-uniform addresses
  __device__ __forceinline__ void load_list(const double* x, int64_t NDOFS, const int* elist, int64_t li0,
                                            int64_t nlist, int tid, double fill = 1.0) {
    const int te = tid / ND, tn = tid - te * ND;
    int el[EPB];
#pragma unroll
    for (int t = 0; t < EPB; t++) el[t] = li0 + t < nlist ? elist[li0 + t] : -1;
    int me = el[0];
#pragma unroll
    for (int t = 1; t < EPB; t++) me = te == t ? el[t] : me;
    const bool ok = tid < EPB * ND && me >= 0;
#pragma unroll
    for (int q = 0; q < NEQ; q++)
      v[q] = ok ? sbase(x + q * NDOFS)[(unsigned)me * (unsigned)ND + (unsigned)tn] : fill;
  }
  __device__ __forceinline__ void store(double (&su)[EPB][NEQ][ND], int tid) const {
    if (tid < EPB * ND) {
      const int te = tid / ND, tn = tid - te * ND;
#pragma unroll
      for (int q = 0; q < NEQ; q++) su[te][q][tn] = v[q];
    }
  }
};

// Max-accumulate a non-negative double into *addr (non-negative doubles order
// like their bit patterns). The current value is read first and only a larger
// value issues the atomic: one same-address atomic per workgroup over a whole
// grid would serialise in L2 and cost milliseconds.
__device__ __forceinline__ void atomic_max_pos(double* addr, double v) {
  if (!(v > 0.0)) return;
  unsigned long long* p = reinterpret_cast<unsigned long long*>(addr);
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned long long cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (bits > cur) atomicMax(p, bits);
}

// Device error word [4]: kind, first element, second element (-1 at a
// boundary), local face. The first face to fail records itself; later ones
// leave the word alone until the host reads and clears it.
constexpr unsigned long long kErrUnphysical = 1;  // a Riemann solver saw an unphysical state
constexpr unsigned long long kErrNonFinite = 2;   // ... and one of its inputs was NaN or infinite
__device__ __forceinline__ void flag_error(long long* err, int64_t e1, int64_t e2, int lf,
                                           unsigned long long kind) {
  if (atomicCAS(reinterpret_cast<unsigned long long*>(err), 0ull, kind) == 0ull) {
    err[1] = e1;
    err[2] = e2;
    err[3] = lf;
  }
}
// Error path only: true if a nodal value of element e (of a byNODES state with
// ndofs = NE * nd scalar nodes) is a NaN or an infinity (x - x is not 0).
// Reading the nodes from memory here, instead of testing the traces the
// Riemann solver saw, keeps the check out of the hot path's registers.
__device__ __attribute__((noinline)) bool nonfinite_nodes(const double* x, int64_t ndofs, int64_t e, int nd,
                                                           int neq) {
  if (e < 0) return false;
  bool r = false;
  for (int v = 0; v < neq; v++)
    for (int i = 0; i < nd; i++) {
      const double a = x[v * ndofs + e * nd + i];
      r = r || !(a - a == 0.0);
    }
  return r;
}

// LimiterMultiplier::linearize (limiter_multiplier.cpp:101-178) for p = 1
// tensor elements: the node values of the plane through the mean whose
// reference gradient is the element's at the centre. With two Gauss-Legendre
// nodes per axis, xi_i - 1/2 = s (x1 - x0)/2 and
// d phi_j / d xi_d (1/2) = s_j (1/2)^(dim-1) / (x1 - x0), with s = -1 at the
// lower node and +1 at the upper, so u_i = mean + sum_d s_d(i) D_d / 2, D_d the
// mean over the 2^(dim-1) node pairs along axis d of their differences. u holds
// the element's 2^DIM node values, lexicographic with x fastest.
template <int DIM>
__device__ __forceinline__ double linearize_q1(const double* u, double mean, int i) {
  constexpr int ND = 1 << DIM;
  double acc = mean;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    double D = 0.0;
#pragma unroll
    for (int j = 0; j < ND; j++) D += ((j >> d) & 1) ? u[j] : -u[j];
    D *= 1.0 / (double)(ND / 2);
    acc += (((i >> d) & 1) ? 0.5 : -0.5) * D;
  }
  return acc;
}

// launchers (rkdg_tensor.hip)
// t: device copy of the tables, th: the host copy (its packed tables also go to
// the kernel as an argument, for reads at wave-uniform indices from scalar
// registers); has_bdr selects the kernel with boundary faces
hipError_t launch_tensor_rhs(int dim, int p, const RhsArgs& a, const SchemeParams& g,
                             const TensorTables* t, const TensorTables& th, bool has_bdr, hipStream_t s);
hipError_t launch_tensor_limit(int dim, int p, const LimitArgs& a, const SchemeParams& g,
                               const TensorTables* t, hipStream_t s);
hipError_t launch_tensor_means(int dim, int p, const MeansArgs& a, const TensorTables* t,
                               hipStream_t s);
hipError_t launch_fill(double* p, double v, int64_t n, hipStream_t s);
// diagnostics (rkdg_io.hip): partial[b*neq + v] over nblocks fixed element
// ranges; primitive fields (and their weighted accumulation); evaluation at
// reference points with a basis table phi[npts][nd]
hipError_t launch_totals(const double* means, const double* vol, int64_t NE, int neq, double* partial, int nblocks,
                         hipStream_t s);
hipError_t launch_primitives(const double* x, int64_t ndofs, int dim, double gamma, double R, double b, double* prim,
                             double* acc, double w, hipStream_t s);
hipError_t launch_interp(const double* x, int64_t NE, int nd, int neq, const double* phi, int npts, double* out,
                         hipStream_t s);
// Shu's indicator with the FinDiff limiter (rkdg_shu.hip): flag[e] = 1 where
// the last variable's indicator is below 0.999, from the element means and the
// neighbour-extrapolated means coef . u_k; then troubled elements take their
// means. coef[(s*nd + j)*NE + e]: s = 0 the element itself, s = 1 + lf its
// neighbour across local face lf.
hipError_t launch_shu_indicator(const double* x, const double* means, const double* coef, const int* nbr, int64_t NE,
                                int nd, int neq, int nf, int8_t* flag, int8_t* troubled, hipStream_t s);
hipError_t launch_shu_apply(double* x, const double* means, const int8_t* flag, int64_t NE, int nd, int neq,
                            hipStream_t s);
// Non-conforming faces (MFEM NCMesh master/slave faces) of tensor meshes. In
// nbr such a face is coded kNcCode0 - k (k >= 0): below every boundary code,
// so the element kernels give it no flux, as a face without an integrator.
// Its flux comes from nc_face_kernel: for every element with such faces, the
// slave faces touching it are integrated at the slave face's points (Elem1 =
// the slave, Elem2 = the master, at the master's reference points given by
// the slave's point matrix) and the element's share is lifted to its nodes.
constexpr int kNcCode0 = -(RKDG_MAX_BC + 2);
struct NcArgs {
  const double* x;        // stage input
  int64_t NE;
  int64_t n_elem;         // elements with non-conforming faces
  const int* elem;        // [n_elem] their ids
  const int* ptr;         // [n_elem + 1] ranges in item
  const int* item;        // slave-face record * 2 + side (0: the element is the slave, 1: the master)
  const int* rec;         // [records][3]: slave element, master element, slave local face
  const double* geo;      // [records][NP][3 DIM + 1]: per point the slave and master reference
                          // coordinates, weight * |n| and the slave's outward unit normal
  double* out;            // [n_elem][NEQ][ND]
  double* lam;            // wave speed slot
  long long* err;
  int no_lambda;
};
hipError_t launch_nc_faces(int dim, int p, const NcArgs& a, const SchemeParams& g, const TensorTables* t,
                           hipStream_t s);
// the static scratch array of rkdg_step_noalloc (rkdg_noalloc.hip): its
// address on the current device, looked up once per device
constexpr int64_t kScratchDoubles = int64_t(1) << 24;
constexpr int kMaxDevices = 64;
hipError_t noalloc_scratch(double** out);
hipError_t launch_copy(double* dst, const double* src, int64_t n, hipStream_t s);
bool tensor_supported(int dim, int p);
int tensor_elems_per_block(int dim, int p);  // elements per workgroup of the tensor RHS

// launchers (rkdg_dense.hip): affine triangles and bilinear quadrilaterals;
// t points to the device copy of the dense reference tables
bool dense_supported(int geom, int p);
int dense_table_size(int geom, int p);  // doubles
int dense_geo_size(int geom, int p);    // doubles per element
hipError_t launch_dense_rhs(int geom, int p, const RhsArgs& a, const SchemeParams& g, const double* t,
                            bool has_bdr, hipStream_t s);
hipError_t launch_dense_limit(int geom, int p, const LimitArgs& a, const SchemeParams& g, const double* t,
                              hipStream_t s);
hipError_t launch_dense_means(int geom, int p, const MeansArgs& a, const double* t, hipStream_t s);

// adjoint kernels of the dense 2D path (rkdg_adjoint.hip)
struct RhsAdjArgs {
  const double* x;       // state the right-hand side was evaluated at
  const double* kbar;    // adjoint of the right-hand side
  double* xbar;          // out: addend + (dR/dx)^T kbar
  const double* addend;  // nullable
  double* pbar;          // [RKDG_MAX_BC][8] boundary-parameter adjoint, accumulated (nullable)
  const double* geo;
  const int* nbr;
  const int* fcode;
  int64_t NE;
};
// y_k = (set_k ? 0 : y_k) + c_k v for up to four targets (adjoint of the
// stage combinations)
struct AxpyTargets {
  int n;
  double* y[4];
  double c[4];
  int set[4];
};
struct LimAdjArgs {
  const double* x;      // pre-limit state
  const double* means;  // its element means [v*NE + e]
  const double* ubar;   // adjoint of the limited state
  double* vbar;         // out: adjoint of the pre-limit state
  double* meanbar;      // scratch [NEQ*NE]
  int64_t* send_to;     // scratch [NE]
  double* send_val;     // scratch [NE]
  const double* geo;
  const int* nbr;
  const int* fcode;
  int64_t NE;
  int obj_slot;         // >= 0: add obj_coef * d(objective)/du of the limited state to ubar first
  double obj_coef;
  AxpyTargets ax;       // applied with v = the final vbar by pass B (ax.n = 0: none)
};
struct ObjArgs {
  const double* x;
  double* J;     // device accumulator (nullable)
  double* xbar;  // nullable: += coef * dJ/dx
  double coef;
  int slot;      // boundary slot whose faces carry the objective
  const double* geo;
  const int* nbr;
  int64_t NE;
  const int* elems;  // the elements with a face on the slot, ascending
  int nelems;
};
hipError_t launch_dense_rhs_adj(int geom, int p, const RhsAdjArgs& a, const SchemeParams& g, const double* t,
                                bool has_bdr, hipStream_t s);
hipError_t launch_dense_limit_adj(int geom, int p, const LimAdjArgs& a, const SchemeParams& g, const double* t,
                                  hipStream_t s);
hipError_t launch_dense_objective(int geom, int p, const ObjArgs& a, const SchemeParams& g, const double* t,
                                  hipStream_t s);
hipError_t launch_axpy_multi(const double* v, int64_t n, const AxpyTargets& t, hipStream_t s);

// per-instance time control of a batched ensemble (rkdg_batch.hip)
struct BatchTime {
  int count;
  double* t;
  double* dt;           // next step size (from the last stage's wave speed)
  double* dt_real;      // size of the step being taken: min(dt, t_final - t)
  const double* t_final;
  const double* hmin;
  const double* lam;    // wave-speed slot of each instance (first dt: face speeds; after a step: last stage)
  int* active;
  long long* steps;
  int* n_active;        // [1]
  double cfl;
  int p;
  long long max_steps;  // 0: no step limit
};
hipError_t launch_batch_init(const BatchTime& b, hipStream_t s);  // t holds the start times
hipError_t launch_batch_pre(const BatchTime& b, hipStream_t s);
hipError_t launch_batch_post(const BatchTime& b, hipStream_t s);

}  // namespace rkdg
