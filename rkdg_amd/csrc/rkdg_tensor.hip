// rkdg_tensor.hip — gfx950 kernels for tensor-product elements (Q_p quads and
// hexes on affine geometry): the fused right-hand side FE_Evolution::Mult
// (src/fe_evolution/fe_evolution.cpp:102-216) with its RK-stage epilogue
// (src/ode_solver/rk_explicit_limited.cpp:51-76), element means
// (src/averager/averager.cpp:84-109) and the limiter
// (src/limiter/limiter.cpp:22-94, src/indicator/indicator_bj.cpp:15-200,
// src/limiter/limiter_multiplier.cpp:12-99, src/limiter/limiter_findiff.cpp:3-33).
//
// Work decomposition: one team of T lanes per element (T = next power of two of
// max(nodes, 2 * face points)); a workgroup is one 64-lane wavefront holding
// EPB = 64/T elements, so every barrier is a single-wave barrier. The 1D basis
// tables are staged in LDS, as is the element's state. The volume term is the
// collocated weak derivative (Gauss-Legendre nodes are the volume quadrature
// points, so Aflux reduces to a 1D derivative matrix per axis); faces are
// processed one axis at a time (the -r and +r faces together): normal-direction
// interpolation to face nodes, tangential interpolation to the face Gauss
// points, Riemann solve, lift back. Every face flux is evaluated with the
// arguments ordered as (Elem1 state, Elem2 state, Elem1 normal), so both
// elements of a face see bitwise identical fluxes and the scheme is
// conservative to round-off. Kernels are specialised on the Riemann solver and
// on the presence of boundary faces.
#include <cstring>

#include "rkdg_kernels.hpp"

namespace rkdg {

template <int DIM, int P>
struct TC {
  static constexpr int N = P + 1;
  static constexpr int ND = DIM == 3 ? N * N * N : N * N;
  static constexpr int NEQ = DIM + 2;
  static constexpr int NF = 2 * DIM;
  // face rule: IntRules order OrderW + 2p (OrderW = 2 trilinear hex, 1 bilinear quad)
  static constexpr int NQ = DIM == 3 ? P + 2 : P + 1;
  static constexpr int NFP = DIM == 3 ? NQ * NQ : NQ;
  static constexpr int NFN = DIM == 3 ? N * N : N;
  static constexpr int NVERT = DIM == 3 ? 8 : 4;
  static constexpr int NIP = DIM == 3 ? 4 : 2;  // indicator points per face (order 2 rule)
  static constexpr int m1 = ND > 2 * NFP ? ND : 2 * NFP;
  static constexpr int T = m1 <= 4 ? 4 : (m1 <= 8 ? 8 : (m1 <= 16 ? 16 : (m1 <= 32 ? 32 : 64)));
  static constexpr int BLOCK = 64;
  static constexpr int EPB = BLOCK / T;
};

// MFEM local face numbering <-> (normal axis r, side c). 3D: (r, c) = (0,0)
// -> 4, (0,1) -> 2, (1,0) -> 1, (1,1) -> 3, (2,0) -> 0, (2,1) -> 5; 2D: (0,0) ->
// 3, (0,1) -> 1, (1,0) -> 0, (1,1) -> 2. The tables are packed into integer
// constants (3 or 2 bits per entry) so that a run-time index is a shift, not a
// memory lookup.
template <int DIM>
__host__ __device__ __forceinline__ int lf_of(int r, int c) {
  if (DIM == 3) return (0x28654 >> (3 * (2 * r + c))) & 7;  // 4,2,1,3,0,5
  return (0x87 >> (2 * (2 * r + c))) & 3;                  // 3,1,0,2
}
template <int DIM>
__host__ __device__ __forceinline__ void rc_of(int lf, int& r, int& c) {
  if (DIM == 3) {
    r = (0x846 >> (2 * lf)) & 3;  // 2,1,0,1,0,2
    c = (0x2c >> lf) & 1;         // 0,0,1,1,0,1
  } else {
    r = (0x5 >> lf) & 1;  // 1,0,1,0
    c = (0x6 >> lf) & 1;  // 0,1,1,0
  }
}

// volume node with index k along axis r and tangential indices (a, b)
template <int DIM, int N>
__device__ __forceinline__ int vnode(int r, int k, int a, int b) {
  if (DIM == 2) return r == 0 ? k + N * a : a + N * k;
  if (r == 0) return k + N * (a + N * b);
  if (r == 1) return a + N * (k + N * b);
  return a + N * (b + N * k);
}

// vnode for an axis known only at run time, without branches: the node index
// is a * sa + b * sb + k * sk
template <int DIM, int N>
__device__ __forceinline__ void vstrides(int r, int& sk, int& sa, int& sb) {
  if (DIM == 2) {
    sk = r == 0 ? 1 : N;
    sa = r == 0 ? N : 1;
    sb = 0;
  } else {
    sk = r == 0 ? 1 : (r == 1 ? N : N * N);
    sa = r == 0 ? N : 1;
    sb = r == 2 ? N : N * N;
  }
}

// orientation: my face-point indices (i, j) -> neighbour frame (i', j')
template <int DIM, int NQ>
__device__ __forceinline__ int map_point(int code, int qi, int qj) {
  int o = (code >> 3) & 7;
  int i = qi, j = qj;
  if (DIM == 3 && (o & 1)) {
    int t = i;
    i = j;
    j = t;
  }
  if (o & 2) i = NQ - 1 - i;
  if (DIM == 3 && (o & 4)) j = NQ - 1 - j;
  return DIM == 3 ? i + NQ * j : i;
}

__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nb) {
  // workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
  // contiguous range of elements so face neighbours share its L2.
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t x = b % 8, l = b / 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + l;
}


// the part of TensorTables an element of order N-1 uses, packed for LDS
template <int N, int NQ>
struct TT {
  double w[N], invw[N];
  double D[N][N];
  double E[2][N];
  double B[NQ][N];
  double wq[NQ];
  double BI[2][N];
};

// tables read at wave-uniform indices, passed by value as a kernel argument
// (scalar registers): the 1D tables and the tensor-product quadrature weights
// w_k = w_a w_b (w_c) of the nodes with their sum
template <int DIM, int N, int NQ>
struct TArg {
  static constexpr int ND = DIM == 3 ? N * N * N : N * N;
  TT<N, NQ> t;
  double wk[ND];
  double wsum;
};

// ---------------------------------------------------------------------------
// ROLE 0: every element solves its faces. ROLE 1 (solver elements of a
// two-coloured mesh): as ROLE 0, and the receiver's share of each lifted face
// flux is written to A.fbuf. ROLE 2 (receivers): the face fluxes are read from
// A.fbuf instead of being solved; no traces, Riemann solves or lifts.
// AL: every interior face joins local face (r, c) to the neighbour's (r, 1-c)
// with the identity orientation (structured meshes), so the neighbour's face,
// node strides and point map are known at compile time.
// CV: curved (trilinear) hexahedra, role 0 only: adj(J) per node in the volume
// term, the normal and |n| per face point, the dense inverse mass and the mass
// weights of the means, all from the element's CurvedLayout record.
template <int DIM, int P, int RS, bool BC, int ROLE, bool AL, bool CV = false>
__global__ __launch_bounds__(64, ROLE == 1 ? 4 : 1) void tensor_rhs_kernel(const RhsArgs A, const SchemeParams g,
                                                        const TensorTables* __restrict__ gtb,
                                                        const TArg<DIM, P + 1, TC<DIM, P>::NQ> ka) {
  const TT<P + 1, TC<DIM, P>::NQ>& kt = ka.t;
  using C = TC<DIM, P>;
  constexpr int N = C::N, ND = C::ND, NEQ = C::NEQ, NQ = C::NQ, NFP = C::NFP, NFN = C::NFN;
  constexpr int NF = C::NF, T = C::T, EPB = C::EPB, BLOCK = C::BLOCK;
  constexpr int GD = DIM * DIM + 1;
  // face-point fluxes are stored in rows per (face, variable), padded by one
  // entry when NFP is a power of two so that the lift's lanes, which read ten
  // rows at once, hit different LDS banks
  constexpr int RW = (NFP & (NFP - 1)) == 0 ? NFP + 1 : NFP;
  constexpr int WORK = (NEQ + 1) * ND > 2 * NEQ * RW ? (NEQ + 1) * ND : 2 * NEQ * RW;
  constexpr int NB = DIM == 3 ? N : 1;  // face nodes per row (second tangential index b)
  // The lifted fluxes may share the face-point flux rows when every phase is a
  // single pass of the team (one wavefront per workgroup executes in order, so a
  // phase's reads of the rows come before its writes of the lifted values, and
  // the node accumulation reads them before the next points phase writes rows).
  constexpr bool LIFT_IN_WORK = 2 * NEQ * N <= T && 2 * NFP <= T;
  constexpr int LIFT = 2 * NEQ * NFN;

  static_assert(!CV || (DIM == 3 && ROLE == 0 && !AL), "curved elements: 3D, one-kernel form, any orientation");
  using CL = CurvedLayout<P>;
  constexpr bool SOLVE = ROLE != 2;    // this kernel solves face fluxes
  constexpr int ES = SOLVE ? EPB : 1;  // extent of the arrays only the face solves use
  constexpr int FB = NF * NEQ * NFN;   // lifted face fluxes of one receiver element
  __shared__ TT<C::N, C::NQ> tb;
  __shared__ double s_u[EPB][NEQ][ND];
  // traces interpolated along the first tangential direction to the face Gauss
  // abscissae: [face c][variable][qi][b], own and neighbour (its face frame)
  __shared__ double s_to[ES][2][NEQ][NQ][NB];
  __shared__ double s_tn[ES][2][NEQ][NQ][NB];
  __shared__ double s_work[EPB][WORK > LIFT ? WORK : LIFT];
  __shared__ double s_lift[LIFT_IN_WORK || !SOLVE ? 1 : EPB][LIFT_IN_WORK || !SOLVE ? 1 : LIFT];  // face fluxes lifted to face nodes
  __shared__ double s_nrm[ES][NF][DIM + 2];   // unit normal, flux sign * |n|, 0.5 / |unit normal|
  __shared__ double s_geo[EPB][GD];
  __shared__ int s_nb[ES][NF];                // by MFEM local face
  __shared__ int s_code[ES][NF];
  __shared__ int s_slot[ROLE == 1 ? EPB : 1][NF];      // receiver slot of each face's neighbour
  __shared__ double s_fb[ROLE == 2 ? EPB : 1][ROLE == 2 ? FB : 1];  // receivers: [lf][v][face node]

  const int tid = threadIdx.x;
  const int team = tid / T, lane = tid % T;
  const int64_t NE = A.NE;
  const int64_t NDOFS = NE * ND;
  const int64_t blk0 = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t blk = A.order ? (int64_t)A.order[blk0] : blk0;
  const int64_t e0 = blk * EPB;  // ROLE 0: the block's first element
  const int64_t li = blk0 * EPB + team;  // ROLES 1, 2: this team's index in the element list
  const int64_t e = ROLE == 0 ? e0 + team : (li < A.nlist ? (int64_t)A.elist[li] : (int64_t)0);
  const bool active = ROLE == 0 ? e < NE : li < A.nlist;

  // ---- phase 0: tables, element state, connectivity, normals and geometry
  // into LDS. Every load is issued before the first LDS store: one memory round
  // trip, plus one for the normal's geometry, which needs the connectivity.
  static_assert(NF <= T, "one lane per local face");
  CopyRegs<sizeof(TT<N, NQ>) / sizeof(double), BLOCK> tbr;
  tbr.load(gtb->packed, tid);
  StateRegs<EPB, NEQ, ND, BLOCK> str;
  if constexpr (ROLE == 0) str.load(A.x, NDOFS, e0, NE, tid);
  else str.load_list(A.x, NDOFS, A.elist, blk0 * EPB, A.nlist, tid);
  // limiter on load: factor and means of the element whose node this lane holds
  const bool lim_load = ROLE == 0 && A.lim_f != nullptr;
  double lf_own = 1.0, lm_own[NEQ];
  if (lim_load) {
    const int64_t el = e0 + tid / ND;
    const bool ok = tid < EPB * ND && el < NE;
    lf_own = ok ? A.lim_f[el] : 1.0;
#pragma unroll
    for (int q = 0; q < NEQ; q++) lm_own[q] = ok ? A.lim_mean[q * NE + el] : 0.0;
  }
  // receivers: the block's lifted face fluxes, contiguous in A.fbuf
  constexpr int FBP = ROLE == 2 ? (EPB * FB + BLOCK - 1) / BLOCK : 1;
  double fbr[FBP];
  if constexpr (ROLE == 2) {
    const int64_t f0 = blk0 * EPB * (int64_t)FB;
    const int64_t fend = (A.nlist < blk0 * EPB + EPB ? A.nlist : blk0 * EPB + EPB) * (int64_t)FB;
#pragma unroll
    for (int p = 0; p < FBP; p++) {
      const int64_t i = f0 + tid + p * BLOCK;
      fbr[p] = (tid + p * BLOCK < EPB * FB && i < fend) ? A.fbuf[i] : 0.0;
    }
  }
  constexpr int GDP = (GD + T - 1) / T;
  double gv[GDP];
#pragma unroll
  for (int p = 0; p < GDP; p++) gv[p] = (active && lane + p * T < GD) ? A.geo[(lane + p * T) * NE + e] : 0.0;
  // connectivity of local face i = lane and Elem1's unit normal on it (the
  // neighbour's geometry when the neighbour is Elem1)
  const bool face_lane = SOLVE && active && lane < NF;
  const int nb0 = face_lane ? A.nbr[lane * NE + e] : -1;
  const int cd0 = face_lane ? A.fcode[lane * NE + e] : 0;
  if constexpr (ROLE == 1) {
    if (face_lane) s_slot[team][lane] = nb0 >= 0 ? A.fslot[nb0] : -1;
  }
  if (face_lane) {
    const int i = lane;
    const int nb = nb0;
    const int code = nb >= 0 ? cd0 : 0;
    s_nb[team][i] = nb;
    s_code[team][i] = code;
    int r, c;
    rc_of<DIM>(i, r, c);
    int rn, cn;
    rc_of<DIM>(code & 7, rn, cn);
    const bool from_nb = nb >= 0 && !((code >> 6) & 1);
    const double sign = from_nb ? 1.0 : -1.0;
    const double sg = from_nb ? (cn ? 1.0 : -1.0) : (c ? 1.0 : -1.0);
    const int64_t ge = from_nb ? (int64_t)nb : e;
    const int gr = from_nb ? rn : r;
    double an[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) an[d] = sg * A.geo[(gr * DIM + d) * NE + ge];
    // |n| as the plain Euclidean norm (face normals are O(h^2): the scaling of
    // MFEM's Norml2 against overflow changes nothing but the last bit)
    double nm2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) nm2 += an[d] * an[d];
    const double nmag = sqrt_ws(nm2);
    const double inv = recip(nmag);
    const int f = 2 * r + c;
    double nu2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const double nu = an[d] * inv;
      s_nrm[team][f][d] = nu;
      nu2 += nu * nu;
    }
    s_nrm[team][f][DIM] = sign * nmag;
    s_nrm[team][f][DIM + 1] = 0.5 * recip(sqrt_ws(nu2));
  }
  tbr.store(reinterpret_cast<double*>(&tb), tid);
  if (lim_load) {
#pragma unroll
    for (int q = 0; q < NEQ; q++) str.v[q] = lm_own[q] + lf_own * (str.v[q] - lm_own[q]);
  }
  str.store(s_u, tid);
#pragma unroll
  for (int p = 0; p < GDP; p++)
    if (active && lane + p * T < GD) s_geo[team][lane + p * T] = gv[p];
  if constexpr (ROLE == 2) {
#pragma unroll
    for (int p = 0; p < FBP; p++)
      if (tid + p * BLOCK < EPB * FB) (&s_fb[0][0])[tid + p * BLOCK] = fbr[p];
  }
  __syncthreads();

  const int j = lane;
  const bool node_lane = active && lane < ND;
  int jj[3] = {0, 0, 0};
  if (DIM == 2) {
    jj[0] = j % N;
    jj[1] = j / N;
  } else {
    jj[0] = j % N;
    jj[1] = (j / N) % N;
    jj[2] = j / (N * N);
  }
  double z[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; v++) z[v] = 0.0;
  double lmax = 0.0;
  double* work = s_work[team];
  double* lift = LIFT_IN_WORK ? work : s_lift[LIFT_IN_WORK ? 0 : team];  // [c][v][face node]

  // ---- volume term: sum_r D^T (w * adj(J) row r . F(u)) along axis r
  if (!A.face_only) {
    double G[DIM][NEQ];
    if (node_lane) {
      double st[NEQ];
#pragma unroll
      for (int v = 0; v < NEQ; v++) st[v] = s_u[team][v][j];
      const SidePrim<DIM> q = side_prim<DIM>(st, g);
      if (!A.no_lambda) lmax = sqrt_ws(q.m2) * q.inv + side_sound<DIM>(st, q, g);
      double aj[CV ? DIM * DIM : 1];  // curved: adj(J) at this node
      if constexpr (CV) {
        const double* cg = A.cgeo + e * CL::size + CL::adj + j * DIM * DIM;
#pragma unroll
        for (int k = 0; k < DIM * DIM; k++) aj[k] = cg[k];
      }
      double wj = tb.w[jj[0]] * tb.w[jj[1]];
      if (DIM == 3) wj *= tb.w[jj[2]];
      // contravariant fluxes G^r = w_j adj(J)(r,:) . F(u): with m.a_r = sum_d a_rd m_d,
      // G^r = (m.a_r, (m.a_r) u + p a_r, (m.a_r) H)
#pragma unroll
      for (int r = 0; r < DIM; r++) {
        double ar[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) ar[d] = wj * (CV ? aj[r * DIM + d] : s_geo[team][r * DIM + d]);
        double ma = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) ma += st[1 + d] * ar[d];
        const double mai = ma * q.inv;
        G[r][0] = ma;
#pragma unroll
        for (int d = 0; d < DIM; d++) G[r][1 + d] = mai * st[1 + d] + q.p * ar[d];
        G[r][NEQ - 1] = mai * (st[NEQ - 1] + q.p);
      }
    }
#pragma unroll
    for (int r = 0; r < DIM; r++) {
      if (node_lane) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) work[v * ND + j] = G[r][v];
      }
      __syncthreads();
      if (node_lane) {
#pragma unroll
        for (int a = 0; a < N; a++) {
          const double dd = tb.D[a][jj[r]];
          const int q = vnode<DIM, N>(r, a, r == 0 ? jj[1] : jj[0], r == 2 ? jj[1] : jj[2]);
#pragma unroll
          for (int v = 0; v < NEQ; v++) z[v] += dd * work[v * ND + q];
        }
      }
      __syncthreads();
    }
  }

  if constexpr (SOLVE) {
    // ---- face terms, one axis at a time. Phases per axis r: points(r) |
    // lift(r) + traces(r+1) | node accumulation(r) + points(r+1) ..., so each
    // axis costs two barriers.
    bool bad = false;
    // Table values that depend on the lane but not on the axis, read from LDS
    // once (the phases below are one pass of the team when ONE_PASS; otherwise
    // they read the tables per item).
    constexpr bool ONE_PASS = 2 * NEQ * N <= T && 2 * NFP <= T;
    double ec[N];         // traces: E[c][k] of the lane's face side
    double bqj[N];        // points: B[qj][b] of the lane's point
    double bqa[NQ];       // lift: B[qi][a] of the lane's row
    double wq2 = 0.0;     // points: the lane's face weight
    if (ONE_PASS) {
      const int ct = lane / (NEQ * NB);
      const int at = (lane - (lane / (NEQ * N)) * (NEQ * N)) % N;
      const int qt = lane % NFP, qit = qt % NQ, qjt = DIM == 3 ? qt / NQ : 0;
  #pragma unroll
      for (int k = 0; k < N; k++) ec[k] = tb.E[ct < 2 ? ct : 0][k];
  #pragma unroll
      for (int b = 0; b < N; b++) bqj[b] = tb.B[qjt][b];
  #pragma unroll
      for (int q = 0; q < NQ; q++) bqa[q] = tb.B[q][at];
      wq2 = DIM == 3 ? tb.wq[qit] * tb.wq[qjt] : tb.wq[qit];
    }
    // own normal-direction traces on faces (r,0), (r,1) and the neighbour's
    // traces in its own face frame (read from global memory), then interpolated
    // along the first tangential direction a to the Gauss abscissae qi. One lane
    // per (face, variable, second tangential index b).
    auto traces = [&](const int r) {
      for (int it = lane; it < 2 * NEQ * NB; it += T) {
        if (!active) break;
        const int c = it / (NEQ * NB), rem = it - c * (NEQ * NB);
        const int v = rem / NB, b = rem - v * NB;
        const int lf = lf_of<DIM>(r, c);
        const int nb = s_nb[team][lf];
        double vals[N][N];
        int cn = 0;
        double nf_f = 1.0, nf_m = 0.0;  // limiter on load: the neighbour's factor and mean
        if (nb >= 0) {
          int rn, sk, sa, sb;
          if constexpr (AL) {
            rn = r;
            cn = 1 - c;
          } else {
            rc_of<DIM>(s_code[team][lf] & 7, rn, cn);
          }
          vstrides<DIM, N>(rn, sk, sa, sb);
          const double* xn = A.x + v * NDOFS + (int64_t)nb * ND + b * sb;
  #pragma unroll
          for (int a = 0; a < N; a++)
  #pragma unroll
            for (int k = 0; k < N; k++) vals[a][k] = xn[a * sa + k * sk];
          if (lim_load) {
            nf_f = A.lim_f[nb];
            nf_m = A.lim_mean[v * NE + nb];
          }
        }
        double tr[N];
  #pragma unroll
        for (int a = 0; a < N; a++) {
          double t = 0.0;
  #pragma unroll
          for (int k = 0; k < N; k++) t += s_u[team][v][vnode<DIM, N>(r, k, a, b)] * (ONE_PASS ? ec[k] : tb.E[c][k]);
          tr[a] = t;
        }
  #pragma unroll
        for (int qi = 0; qi < NQ; qi++) {
          double tt = 0.0;
  #pragma unroll
          for (int a = 0; a < N; a++) tt += tr[a] * kt.B[qi][a];
          s_to[team][c][v][qi][b] = tt;
        }
        if (nb >= 0) {
  #pragma unroll
          for (int a = 0; a < N; a++) {
            double t = 0.0;
  #pragma unroll
            for (int k = 0; k < N; k++) t += vals[a][k] * tb.E[cn][k];
            // the trace of the limited state (the trace of a constant is the constant)
            tr[a] = lim_load ? nf_m + nf_f * (t - nf_m) : t;
          }
  #pragma unroll
          for (int qi = 0; qi < NQ; qi++) {
            double tt = 0.0;
  #pragma unroll
            for (int a = 0; a < N; a++) tt += tr[a] * kt.B[qi][a];
            s_tn[team][c][v][qi][b] = tt;
          }
        }
      }
    };
    traces(0);
    __syncthreads();
  #pragma unroll
    for (int r = 0; r < DIM; r++) {
      // face points: traces, Riemann solve, weighted flux
      for (int it = lane; it < 2 * NFP; it += T) {
        if (!active) break;
        const int c = it / NFP, q = it - c * NFP;
        const int qi = q % NQ, qj = DIM == 3 ? q / NQ : 0;
        const int f = 2 * r + c;
        const int lf = lf_of<DIM>(r, c);
        const int nb = s_nb[team][lf];
        double um[NEQ], up[NEQ];
  #pragma unroll
        for (int v = 0; v < NEQ; v++) {
          double t = 0.0;
          if (DIM == 3) {
  #pragma unroll
            for (int b = 0; b < N; b++) t += s_to[team][c][v][qi][b] * (ONE_PASS ? bqj[b] : tb.B[qj][b]);
          } else {
            t = s_to[team][c][v][qi][0];
          }
          um[v] = t;
        }
        double nrm[DIM];
        double snm, inm;  // -|n| on Elem1's side, +|n| on Elem2's; 0.5 / |unit normal|
        if constexpr (CV) {
          // Elem1's normal at this point: the own record, or the neighbour's
          // at the point in its face frame
          const int code = nb >= 0 ? s_code[team][lf] : 0;
          const bool nb_first = nb >= 0 && !((code >> 6) & 1);
          const int64_t eo = nb_first ? (int64_t)nb : e;
          const int lfo = nb_first ? (code & 7) : lf;
          const int qo = nb_first ? map_point<DIM, NQ>(code, qi, qj) : q;
          const double* fn = A.cgeo + eo * CL::size + CL::fn + (lfo * NFP + qo) * 5;
  #pragma unroll
          for (int d = 0; d < DIM; d++) nrm[d] = fn[d];
          snm = nb_first ? fn[DIM] : -fn[DIM];
          inm = fn[DIM + 1];
        } else {
  #pragma unroll
          for (int d = 0; d < DIM; d++) nrm[d] = s_nrm[team][f][d];
          snm = s_nrm[team][f][DIM];
          inm = s_nrm[team][f][DIM + 1];
        }
        double fl[NEQ];
        double factor = -1.0;
        bool ok = true;
        double lam = 0.0;
        if (!BC || nb >= 0) {
          const int q2 = AL ? q : map_point<DIM, NQ>(s_code[team][lf], qi, qj);
          const int pi = q2 % NQ, pj = DIM == 3 ? q2 / NQ : 0;
  #pragma unroll
          for (int v = 0; v < NEQ; v++) {
            double t = 0.0;
            if (DIM == 3) {
  #pragma unroll
              for (int b = 0; b < N; b++) t += s_tn[team][c][v][pi][b] * ((AL && ONE_PASS) ? bqj[b] : tb.B[pj][b]);
            } else {
              t = s_tn[team][c][v][pi][0];
            }
            up[v] = t;
          }
          factor = snm < 0.0 ? -1.0 : 1.0;
          const bool elem1 = factor < 0.0;
          double sL[NEQ], sR[NEQ];
  #pragma unroll
          for (int v = 0; v < NEQ; v++) {
            sL[v] = elem1 ? um[v] : up[v];
            sR[v] = elem1 ? up[v] : um[v];
          }
          lam = riemann_t<DIM, RS>(g, sL, sR, nrm, fl, ok, inm);
        } else if constexpr (BC) {
          const int slot = -1 - nb;
          if (slot < g.num_bc) {
            bc_state<DIM>(g, slot, um, nrm, up);
            if (g.bc_type[slot] == RKDG_BC_CHAR_OUTLET) lam = char_outlet_flux<DIM>(g, up, nrm, fl);
            else lam = riemann_t<DIM, RS>(g, um, up, nrm, fl, ok, inm);
          } else {
  #pragma unroll
            for (int v = 0; v < NEQ; v++) fl[v] = 0.0;
          }
        }
        if (!ok || lam < 0.0) {
          if (!bad) {
            const bool nf = nonfinite_nodes(A.x, NDOFS, e, ND, NEQ) || nonfinite_nodes(A.x, NDOFS, nb, ND, NEQ);
            flag_error(A.err, factor < 0.0 ? e : nb, factor < 0.0 ? (nb >= 0 ? nb : -1) : e, lf,
                       nf ? kErrNonFinite : kErrUnphysical);
          }
          bad = true;
        } else {
          lmax = lam > lmax ? lam : lmax;
        }
        const double wgt = (ONE_PASS ? wq2 : (DIM == 3 ? tb.wq[qi] * tb.wq[qj] : tb.wq[qi])) * snm;  // flux sign * w_q * |n|
  #pragma unroll
        for (int v = 0; v < NEQ; v++) work[(c * NEQ + v) * RW + q] = fl[v] * wgt;
      }
      __syncthreads();
      // lift to face nodes, sum-factorised: one lane per (face, variable, a)
      // forms W(qj) = sum_qi F(qi, qj) B(qi, a) once and the N nodes (a, b) from it
      for (int it = lane; it < 2 * NEQ * N; it += T) {
        if (!active) break;
        const int c = it / (NEQ * N), rem = it - c * (NEQ * N);
        const int v = rem / N, a = rem - v * N;
        const double* w = work + (c * NEQ + v) * RW;
        if (DIM == 3) {
          double W[NQ];
  #pragma unroll
          for (int qj = 0; qj < NQ; qj++) {
            double tt = 0.0;
  #pragma unroll
            for (int qi = 0; qi < NQ; qi++) tt += w[qi + NQ * qj] * (ONE_PASS ? bqa[qi] : tb.B[qi][a]);
            W[qj] = tt;
          }
  #pragma unroll
          for (int b = 0; b < N; b++) {
            double L = 0.0;
  #pragma unroll
            for (int qj = 0; qj < NQ; qj++) L += W[qj] * kt.B[qj][b];
            lift[(c * NEQ + v) * NFN + a + N * b] = L;
            if constexpr (ROLE == 1) {
              // the receiver's share: the same flux with the opposite sign,
              // at the same face node in the receiver's face frame
              if (!A.face_only) {
                const int lf = lf_of<DIM>(r, c);
                const int code = s_code[team][lf];
                A.fbuf[((int64_t)s_slot[team][lf] * NF + (code & 7)) * (NEQ * NFN) + v * NFN +
                       map_point<DIM, N>(code, a, b)] = -L;
              }
            }
          }
        } else {
          double L = 0.0;
  #pragma unroll
          for (int qi = 0; qi < NQ; qi++) L += w[qi] * tb.B[qi][a];
          lift[(c * NEQ + v) * NFN + a] = L;
          if constexpr (ROLE == 1) {
            if (!A.face_only) {
              const int lf = lf_of<DIM>(r, c);
              const int code = s_code[team][lf];
              A.fbuf[((int64_t)s_slot[team][lf] * NF + (code & 7)) * (NEQ * NFN) + v * NFN +
                     map_point<DIM, N>(code, a, 0)] = -L;
            }
          }
        }
      }
      if (r + 1 < DIM) traces(r + 1);
      __syncthreads();
      if (node_lane) {
        const int k = jj[r];
        const int a = r == 0 ? jj[1] : jj[0];
        const int b = DIM == 3 ? (r == 2 ? jj[1] : jj[2]) : 0;
        const int fn = DIM == 3 ? a + N * b : a;
        const double e0v = tb.E[0][k], e1v = tb.E[1][k];
  #pragma unroll
        for (int v = 0; v < NEQ; v++) z[v] += lift[v * NFN + fn] * e0v + lift[(NEQ + v) * NFN + fn] * e1v;
      }
    }
  } else {
    // receivers: the lifted face fluxes of the solver elements, axis by axis
    // as a solver element adds its own (the same formula and order)
    if (node_lane) {
#pragma unroll
      for (int r = 0; r < DIM; r++) {
        const int k = jj[r];
        const int a = r == 0 ? jj[1] : jj[0];
        const int b = DIM == 3 ? (r == 2 ? jj[1] : jj[2]) : 0;
        const int fn = DIM == 3 ? a + N * b : a;
        const double e0v = tb.E[0][k], e1v = tb.E[1][k];
        const double* l0 = &s_fb[team][lf_of<DIM>(r, 0) * (NEQ * NFN) + fn];
        const double* l1 = &s_fb[team][lf_of<DIM>(r, 1) * (NEQ * NFN) + fn];
#pragma unroll
        for (int v = 0; v < NEQ; v++) z[v] += l0[v * NFN] * e0v + l1[v * NFN] * e1v;
      }
    }
  }

  // ---- inverse mass (affine: diagonal det(J) * w_j) and stage epilogue
  if (!A.face_only) {
    double outv[NEQ];
    double detJ = 1.0;
    if (node_lane && A.ncslot) {
      // the element's share of its non-conforming faces (nc_face_kernel)
      const int ns = A.ncslot[e];
      if (ns >= 0) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) z[v] += A.ncbuf[((int64_t)ns * NEQ + v) * ND + j];
      }
    }
    if constexpr (CV) {
      // curved element: z <- M^-1 z with the element's dense inverse mass
      __syncthreads();  // the face phases' reads of work are done
      if (node_lane) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) work[v * ND + j] = z[v];
      }
      __syncthreads();
      if (node_lane) {
        // column j of the record's M^-1 (stored transposed): the lanes' loads
        // of one k are consecutive; a partial unroll keeps the row out of VGPRs
        const double* mi = A.cgeo + e * CL::size + CL::minv + j;
        double y[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) y[v] = 0.0;
#pragma unroll 4
        for (int k = 0; k < ND; k++) {
          const double m = mi[k * ND];
#pragma unroll
          for (int v = 0; v < NEQ; v++) y[v] += m * work[v * ND + k];
        }
#pragma unroll
        for (int v = 0; v < NEQ; v++) z[v] = y[v];
      }
    }
    if (node_lane) {
      detJ = s_geo[team][GD - 1];
      double iw = tb.invw[jj[0]] * tb.invw[jj[1]];
      if (DIM == 3) iw *= tb.invw[jj[2]];
      const double im = CV ? 1.0 : iw * recip(detJ);
      // every base is loaded before the first store (an output may be a base)
      // addresses: the workgroup's block of each variable (uniform, scalar
      // registers) plus a small 32-bit lane offset; with an element list the
      // lane offset is the element's own node index
      const int64_t eb = ROLE == 0 ? e0 * ND : 0;
      const unsigned lo = ROLE == 0 ? (unsigned)(team * ND + j) : (unsigned)e * (unsigned)ND + (unsigned)j;
      double u0[NEQ];  // the stage input at this node (the base of outputs without one)
      if (lim_load) {  // the unlimited stage value
#pragma unroll
        for (int v = 0; v < NEQ; v++) u0[v] = sbase(A.x + (v * NDOFS + eb))[lo];
      } else {
#pragma unroll
        for (int v = 0; v < NEQ; v++) u0[v] = s_u[team][v][j];
      }
      double base[4][NEQ];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const double* bp = m < A.nout ? A.outs[m].base : nullptr;
        if (bp) {
#pragma unroll
          for (int v = 0; v < NEQ; v++) base[m][v] = sbase(bp + (v * NDOFS + eb))[lo];
        } else {
#pragma unroll
          for (int v = 0; v < NEQ; v++) base[m][v] = u0[v];
        }
      }
#pragma unroll
      for (int v = 0; v < NEQ; v++) outv[v] = 0.0;
#pragma unroll
      for (int m = 0; m < 4; m++) {
        if (m < A.nout) {
          double* op = A.outs[m].out;
          const double cf = A.outs[m].coef;
          const double wb = A.outs[m].wb, wi = A.outs[m].wi;
#pragma unroll
          for (int v = 0; v < NEQ; v++) {
            const double o = wi == 0.0 ? base[m][v] + cf * (z[v] * im) : wb * base[m][v] + wi * (u0[v] + cf * (z[v] * im));
            sbase(op + (v * NDOFS + eb))[lo] = o;
            if (m == 0) outv[v] = o;  // means of outs[0] (the host puts the selected output first)
          }
        }
      }
      if (A.k) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) sbase(A.k + (v * NDOFS + eb))[lo] = z[v] * im;
      }
    }
    if (A.means) {
      // element means of the selected output: sum_k (w_k detJ) u_k / sum_k w_k detJ;
      // det J is constant on an affine element and cancels
      __syncthreads();
      if (node_lane) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) work[v * ND + j] = outv[v];
      }
      __syncthreads();
      if (active && lane < NEQ) {
        double s = 0.0;
        if constexpr (CV) {
          // curved element: the integrals of the basis as weights (Averager)
          const double* mw = A.cgeo + e * CL::size + CL::mw;
          double ws = 0.0;
          for (int k = 0; k < ND; k++) ws += mw[k];
          for (int k = 0; k < ND; k++) s += mw[k] * work[lane * ND + k];
          A.means[lane * NE + e] = s / ws;
        } else {
#pragma unroll
          for (int k = 0; k < ND; k++) s += ka.wk[k] * work[lane * ND + k];
          A.means[lane * NE + e] = s / ka.wsum;
        }
      }
    }
  }

  // ---- max wave speed: wave reduction, one atomic per wave
  if (!A.no_lambda) {
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(lmax, off, 64);
      lmax = o > lmax ? o : lmax;
    }
    if (tid == 0) atomic_max_pos(A.lam, lmax);
  }
}

// ---------------------------------------------------------------------------
// element means (Averager::computeMeanValues): one team per element, lanes on
// nodes (coalesced loads), sum_k w_k u_k / sum_k w_k by a shuffle reduction
// over the team (det J is constant on an affine element and cancels)
template <int DIM, int P>
__global__ __launch_bounds__(64) void tensor_means_kernel(const MeansArgs A,
                                                          const TensorTables* __restrict__ gtb) {
  using C = TC<DIM, P>;
  constexpr int N = C::N, ND = C::ND, NEQ = C::NEQ, T = C::T, EPB = C::EPB;
  const int tid = threadIdx.x;
  const int team = tid / T, lane = tid % T;
  const int64_t NE = A.NE;
  const int64_t NDOFS = NE * ND;
  const int64_t e = (int64_t)blockIdx.x * EPB + team;
  const bool on = e < NE && lane < ND;
  double wk = 0.0, s[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; v++) s[v] = on ? A.x[v * NDOFS + e * ND + lane] : 0.0;
  if (on) {
    if (DIM == 3 && A.cgeo) {
      // curved hexahedra: the integrals of the basis (Averager)
      wk = A.cgeo[e * CurvedLayout<P>::size + CurvedLayout<P>::mw + lane];
    } else {
      wk = gtb->w[lane % N] * gtb->w[(lane / N) % N];
      if (DIM == 3) wk *= gtb->w[lane / (N * N)];
    }
  }
#pragma unroll
  for (int v = 0; v < NEQ; v++) s[v] *= wk;
#pragma unroll
  for (int off = T / 2; off > 0; off >>= 1) {
    wk += __shfl_xor(wk, off, 64);
#pragma unroll
    for (int v = 0; v < NEQ; v++) s[v] += __shfl_xor(s[v], off, 64);
  }
  if (e < NE && lane == 0) {
#pragma unroll
    for (int v = 0; v < NEQ; v++) A.means[v * NE + e] = s[v] / wk;
  }
}

// ---------------------------------------------------------------------------
// Limiter::update for one element per team: stencil bounds of the means over
// the face neighbours, the indicator at two Gauss points per face direction of
// every interior face, the limiter function, then LimiterMultiplier (with the
// last-hope reset at the reference vertices) or LimiterFinDiff. Only the
// indicator of the variable that decides is formed: the multiplier uses the
// density's (minValues = yMin[rho], indicator_bj.cpp:193), FinDiff the last
// variable's (limiter_findiff.cpp:7-8).
template <int DIM, int P>
__global__ __launch_bounds__(64) void tensor_limit_kernel(const LimitArgs A, const SchemeParams g,
                                                          const TensorTables* __restrict__ gtb) {
  using C = TC<DIM, P>;
  constexpr int N = C::N, ND = C::ND, NEQ = C::NEQ, NFN = C::NFN, NF = C::NF, NIP = C::NIP;
  constexpr int T = C::T, EPB = C::EPB, BLOCK = C::BLOCK;
  constexpr int NPT = NF * NIP;  // indicator points per element
  constexpr int NVERT = C::NVERT;

  __shared__ TT<C::N, C::NQ> tb;
  __shared__ double s_u[EPB][NEQ][ND];
  __shared__ double s_tr[EPB][NF][NFN];   // traces of the deciding variable
  constexpr int NB = DIM == 3 ? N : 1;
  // last hope: every variable on the x faces interpolated to y = 0, 1: [cx][cy][v][z node]
  __shared__ double s_ty[EPB][2][2][NEQ][NB];
  __shared__ double s_mean[EPB][NEQ];  // the element's means
  __shared__ double s_nbm[EPB][NF];    // deciding variable's means of the face neighbours (lowest)
  __shared__ double s_nbh[EPB][NF];    // ... highest (several neighbours across a non-conforming face)

  const int tid = threadIdx.x;
  const int team = tid / T, lane = tid % T;
  const int64_t NE = A.NE;
  const int64_t NDOFS = NE * ND;
  const int64_t e0 = xcd_remap(blockIdx.x, gridDim.x) * EPB;
  const int64_t e = e0 + team;
  const bool active = e < NE;
  if (g.limiter == RKDG_LIM_NONE) {
    if (A.troubled && active && lane == 0) A.troubled[e] = 0;
    if (A.factor && active && lane == 0) A.factor[e] = 1.0;
    return;
  }
  const int vsel = g.limiter == RKDG_LIM_FINDIFF ? NEQ - 1 : 0;

  // every load of the setup is issued before the first LDS store
  CopyRegs<sizeof(TT<N, C::NQ>) / sizeof(double), BLOCK> tbr;
  tbr.load(gtb->packed, tid);
  StateRegs<EPB, NEQ, ND, BLOCK> str;
  str.load(A.x, NDOFS, e0, NE, tid);
  // the element's means (lane v < NEQ) and the deciding variable's means of
  // the face neighbours (lane lf < NF): one load instruction each, combined
  // through LDS after the barrier (a wave-wide load per value would cost the
  // memory pipeline 64 lanes for two distinct addresses)
  static_assert(NEQ <= T && NF <= T, "one lane per variable and per face");
  const bool ind_bj = g.indicator == RKDG_IND_BJ;
  const double mown = (active && lane < NEQ) ? A.means[lane * NE + e] : 0.0;
  const int nbl = (active && ind_bj && lane < NF) ? A.nbr[lane * NE + e] : -1;
  double nml = nbl >= 0 ? A.means[vsel * NE + nbl] : 0.0;
  double nmh = nml;  // a non-conforming face: the bounds over its neighbours
  bool has_nb = nbl >= 0;
  if (A.ncnb && nbl <= kNcCode0) {
    const int k = kNcCode0 - nbl;
    nml = 1e300;
    nmh = -1e300;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int n = A.ncnb[4 * k + q];
      if (n >= 0) {
        const double a = A.means[vsel * NE + n];
        nml = a < nml ? a : nml;
        nmh = a > nmh ? a : nmh;
      }
    }
    has_nb = true;
  }
  const unsigned long long nbb = __ballot(has_nb);
  const int nbmask = (int)((nbb >> (team * T)) & ((1ull << NF) - 1ull));  // bit lf: face lf has a neighbour
  // Venkatakrishnan eps^2 = (K h)^3 with h = |element|^(1/dim): h^3 = det J
  // (3D) or det J^(3/2) (2D) for affine elements
  double eps2 = 0.0;
  if (active && g.phi == RKDG_PHI_VENKAT && g.venkat_k > 0.0) {
    const double dj = A.geo[DIM * DIM * NE + e];
    const double k3 = g.venkat_k * g.venkat_k * g.venkat_k;
    eps2 = k3 * (DIM == 3 ? dj : dj * sqrt(dj));
  }
  tbr.store(reinterpret_cast<double*>(&tb), tid);
  str.store(s_u, tid);
  if (lane < NEQ) s_mean[team][lane] = mown;
  if (lane < NF) {
    s_nbm[team][lane] = nml;
    s_nbh[team][lane] = nmh;
  }
  __syncthreads();
  double mean[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; v++) mean[v] = s_mean[team][v];
  // stencil bounds (cell + face neighbours) of the deciding variable's means
  double mI = 1e9, MI = -1e9;
  if (ind_bj && active) {
    const double mv = mean[vsel];
    mI = mv;
    MI = mv;
#pragma unroll
    for (int lf = 0; lf < NF; lf++) {
      if ((nbmask >> lf) & 1) {
        const double a = s_nbm[team][lf], b = s_nbh[team][lf];
        mI = a < mI ? a : mI;
        MI = b > MI ? b : MI;
      }
    }
  }

  // normal-direction traces of the deciding variable on every face
  if (ind_bj) {
    for (int it = lane; it < NF * NFN; it += T) {
      if (!active) break;
      const int f = it / NFN, fn = it - f * NFN;
      const int r = f / 2, c = f % 2;
      const int a = fn % N, b = fn / N;
      int sk, sa, sb;
      vstrides<DIM, N>(r, sk, sa, sb);
      const double* uv = &s_u[team][vsel][a * sa + b * sb];
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < N; k++) t += uv[k * sk] * tb.E[c][k];
      s_tr[team][f][fn] = t;
    }
  }
  if constexpr (P == 1) {
    // linearize: the multiplier and the last hope work on the element's linear
    // reconstruction (the indicator above reads the state as it is)
    if (g.linearize && g.limiter == RKDG_LIM_MULTI) {
      double lin[NEQ];
      const bool nl = active && lane < ND;
      if (nl) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) lin[v] = linearize_q1<DIM>(&s_u[team][v][0], mean[v], lane);
      }
      __syncthreads();
      if (nl) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) s_u[team][v][lane] = lin[v];
      }
      __syncthreads();
    }
  }
  if (g.limiter == RKDG_LIM_MULTI && g.last_hope) {
    // sum-factorised vertex values, first two directions: one lane per
    // (x side cx, variable, z node b) forms the x-face traces of its row and
    // interpolates them to y = 0 and y = 1
    for (int it = lane; it < 2 * NEQ * NB; it += T) {
      if (!active) break;
      const int cx = it / (NEQ * NB), rem = it - cx * (NEQ * NB);
      const int v = rem / NB, b = rem - v * NB;
      double tr[N];
#pragma unroll
      for (int a = 0; a < N; a++) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < N; k++) t += s_u[team][v][vnode<DIM, N>(0, k, a, b)] * tb.E[cx][k];
        tr[a] = t;
      }
#pragma unroll
      for (int cy = 0; cy < 2; cy++) {
        double tt = 0.0;
#pragma unroll
        for (int a = 0; a < N; a++) tt += tr[a] * tb.E[cy][a];
        s_ty[team][cx][cy][v][b] = tt;
      }
    }
  }
  __syncthreads();

  // indicator values at the points of interior faces
  double ylane = 1e9;  // this lane's minimum over its indicator points
  if (ind_bj) {
    for (int it = lane; it < NPT; it += T) {
      if (!active) break;
      const int f = it / NIP, ip = it - f * NIP;
      const int r = f / 2, c = f % 2;
      const int lf = lf_of<DIM>(r, c);
      double yc = 1e9;
      if ((nbmask >> lf) & 1) {
        const int pi = ip % 2, pj = ip / 2;
        double u = 0.0;
        if (DIM == 3) {
#pragma unroll
          for (int b = 0; b < N; b++) {
            double tt = 0.0;
#pragma unroll
            for (int a = 0; a < N; a++) tt += s_tr[team][f][a + N * b] * tb.BI[pi][a];
            u += tt * tb.BI[pj][b];
          }
        } else {
#pragma unroll
          for (int a = 0; a < N; a++) u += s_tr[team][f][a] * tb.BI[pi][a];
        }
        const double diff = u - mean[vsel];
        const double y = indicator_y(diff, mean[vsel], mI, MI);
        if (g.phi == RKDG_PHI_BJ) {
          yc = y < 1.0 ? y : 1.0;
        } else if (g.phi == RKDG_PHI_VENKAT) {
          if (g.venkat_k > 0.0) {
            const double d2 = diff * diff;
            yc = (d2 * (y * y + 2.0 * y) + eps2) / (d2 * (y * y + y + 2.0) + eps2);
          } else {
            yc = (y * y + 2.0 * y) / (y * y + y + 2.0);
          }
        } else {
          const double yStar = 1.5;
          const double yRel = y / yStar;
          yc = y < 1.0 ? y + (3.0 - 2.0 * yStar) * yRel * yRel + (yStar - 2.0) * yRel * yRel * yRel : 1.0;
        }
      }
      ylane = yc < ylane ? yc : ylane;
    }
  }

  // minimum over the team's lanes (min is exact, so the order does not matter)
#pragma unroll
  for (int off = T / 2; off > 0; off >>= 1) {
    const double o = __shfl_xor(ylane, off, 64);
    ylane = o < ylane ? o : ylane;
  }
  double yr = ylane;
  if (g.indicator == RKDG_IND_NOWHERE) yr = 1.0;
  else if (g.indicator == RKDG_IND_EVERYWHERE) yr = 0.0;

  // limit: each node lane owns one node
  const int j = lane;
  const bool node_lane = active && lane < ND;
  double u[NEQ];
  if (node_lane) {
#pragma unroll
    for (int v = 0; v < NEQ; v++) u[v] = s_u[team][v][j];
  }
  int trouble = 0;
  double feff = 1.0;  // the factor of the limited state mean + feff (u - mean)
  if (g.limiter == RKDG_LIM_MULTI) {
    if (yr < 1.0) trouble = 1;
    // u <- mean + y (u - mean) applied nd times (limiter_multiplier.cpp:24-38)
    // is u <- mean + y^nd (u - mean); y^nd by repeated squaring
    double f = yr;
    if (g.pow_nd) {
      double acc = 1.0, b = yr;
      for (int e2 = ND; e2 > 0; e2 >>= 1) {
        if (e2 & 1) acc *= b;
        b *= b;
      }
      f = acc;
    }
    feff = f;
    if (node_lane) {
#pragma unroll
      for (int v = 0; v < NEQ; v++) u[v] = mean[v] + f * (u[v] - mean[v]);
    }
    if (g.last_hope) {
      bool bad_v = false;
      // the limited state at the reference vertices: mean + f (u(vertex) - mean),
      // u(vertex) from the two-direction partial sums (last direction here)
      for (int it = lane; it < NVERT; it += T) {
        if (!active) break;
        const int cx = it & 1, cy = (it >> 1) & 1, cz = (it >> 2) & 1;
        double vs[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
          double s = 0.0;
#pragma unroll
          for (int b = 0; b < NB; b++) {
            const double tt = s_ty[team][cx][cy][v][b];
            s += DIM == 3 ? tt * tb.E[cz][b] : tt;
          }
          vs[v] = mean[v] + f * (s - mean[v]);
        }
        if (!physical<DIM>(vs, g)) bad_v = true;
      }
      const unsigned long long bm = __ballot(bad_v);
      const unsigned long long tm = T == 64 ? ~0ull : (((1ull << T) - 1ull) << (team * T));
      if (active && (bm & tm)) {
        trouble |= 2;
        feff = 0.0;
        if (node_lane) {
#pragma unroll
          for (int v = 0; v < NEQ; v++) u[v] = mean[v];
        }
      }
    }
  } else if (g.limiter == RKDG_LIM_FINDIFF) {
    if (yr < 0.999) {
      trouble = 1;
      feff = 0.0;
      if (node_lane) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) u[v] = mean[v];
      }
    }
  }
  if (A.factor) {
    if (active && lane == 0) A.factor[e] = feff;
  } else if (node_lane) {
#pragma unroll
    for (int v = 0; v < NEQ; v++) A.x[v * NDOFS + e * ND + j] = u[v];
  }
  if (A.troubled && active && lane == 0) A.troubled[e] = (int8_t)trouble;
}

// ---------------------------------------------------------------------------
__global__ void fill_kernel(double* p, double v, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void copy_kernel(double* dst, const double* src, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

hipError_t launch_fill(double* p, double v, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  fill_kernel<<<(n + 255) / 256, 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}
hipError_t launch_copy(double* dst, const double* src, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  copy_kernel<<<(n + 255) / 256, 256, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

#define RKDG_TENSOR_CASES(X) X(2, 1) X(2, 2) X(2, 3) X(2, 4) X(2, 5) X(3, 1) X(3, 2) X(3, 3)

bool tensor_supported(int dim, int p) {
#define X(D, PP) \
  if (dim == D && p == PP) return true;
  RKDG_TENSOR_CASES(X)
#undef X
  return false;
}

int tensor_elems_per_block(int dim, int p) {
#define X(D, PP) \
  if (dim == D && p == PP) return TC<D, PP>::EPB;
  RKDG_TENSOR_CASES(X)
#undef X
  return 1;
}

template <int D, int PP, int RS>
hipError_t launch_rhs_rs(const RhsArgs& a, const SchemeParams& g, const TensorTables* t,
                         const TensorTables& th, bool bc, hipStream_t s) {
  using C = TC<D, PP>;
  const int64_t nb = (a.NE + C::EPB - 1) / C::EPB;
  TArg<D, C::N, C::NQ> ka;
  std::memcpy(&ka.t, th.packed, sizeof(ka.t));
  ka.wsum = 0.0;
  for (int k = 0; k < C::ND; k++) {
    const int i0 = k % C::N, i1 = (k / C::N) % C::N, i2 = k / (C::N * C::N);
    double w = th.w[i0] * th.w[i1];
    if (D == 3) w *= th.w[i2];
    ka.wk[k] = w;
    ka.wsum += w;
  }
  if (a.role != 0) {
    // two-coloured meshes have no boundary faces (the host checks)
    if (bc) return hipErrorInvalidValue;
    const int64_t nl = (a.nlist + C::EPB - 1) / C::EPB;
    if (nl == 0) return hipSuccess;
    if (a.role == 1) tensor_rhs_kernel<D, PP, RS, false, 1, false><<<(unsigned)nl, C::BLOCK, 0, s>>>(a, g, t, ka);
    else tensor_rhs_kernel<D, PP, RS, false, 2, false><<<(unsigned)nl, C::BLOCK, 0, s>>>(a, g, t, ka);
    return hipGetLastError();
  }
  if constexpr (D == 3) {
    if (a.cgeo) {
      if (bc) tensor_rhs_kernel<D, PP, RS, true, 0, false, true><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t, ka);
      else tensor_rhs_kernel<D, PP, RS, false, 0, false, true><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t, ka);
      return hipGetLastError();
    }
  }
  if (a.aligned) {
    if (bc) tensor_rhs_kernel<D, PP, RS, true, 0, true><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t, ka);
    else tensor_rhs_kernel<D, PP, RS, false, 0, true><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t, ka);
    return hipGetLastError();
  }
  if (bc) tensor_rhs_kernel<D, PP, RS, true, 0, false><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t, ka);
  else tensor_rhs_kernel<D, PP, RS, false, 0, false><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t, ka);
  return hipGetLastError();
}

template <int D, int PP>
hipError_t launch_rhs_dp(const RhsArgs& a, const SchemeParams& g, const TensorTables* t,
                         const TensorTables& th, bool bc, hipStream_t s) {
  switch (g.riemann) {
    case RKDG_RS_RUSANOV: return launch_rhs_rs<D, PP, RKDG_RS_RUSANOV>(a, g, t, th, bc, s);
    case RKDG_RS_LLF: return launch_rhs_rs<D, PP, RKDG_RS_LLF>(a, g, t, th, bc, s);
    case RKDG_RS_HLL: return launch_rhs_rs<D, PP, RKDG_RS_HLL>(a, g, t, th, bc, s);
    default: return launch_rhs_rs<D, PP, RKDG_RS_HLLC>(a, g, t, th, bc, s);
  }
}

hipError_t launch_tensor_rhs(int dim, int p, const RhsArgs& a, const SchemeParams& g,
                             const TensorTables* t, const TensorTables& th, bool has_bdr, hipStream_t s) {
#define X(D, PP) \
  if (dim == D && p == PP) return launch_rhs_dp<D, PP>(a, g, t, th, has_bdr, s);
  RKDG_TENSOR_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_tensor_limit(int dim, int p, const LimitArgs& a, const SchemeParams& g,
                               const TensorTables* t, hipStream_t s) {
#define X(D, PP)                                                           \
  if (dim == D && p == PP) {                                               \
    using C = TC<D, PP>;                                                   \
    const int64_t nb = (a.NE + C::EPB - 1) / C::EPB;                       \
    tensor_limit_kernel<D, PP><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t); \
    return hipGetLastError();                                              \
  }
  RKDG_TENSOR_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_tensor_means(int dim, int p, const MeansArgs& a, const TensorTables* t,
                               hipStream_t s) {
#define X(D, PP)                                                                      \
  if (dim == D && p == PP) {                                                          \
    using C = TC<D, PP>;                                                              \
    tensor_means_kernel<D, PP><<<(unsigned)((a.NE + C::EPB - 1) / C::EPB), C::BLOCK, 0, s>>>(a, t); \
    return hipGetLastError();                                                         \
  }
  RKDG_TENSOR_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

}  // namespace rkdg
