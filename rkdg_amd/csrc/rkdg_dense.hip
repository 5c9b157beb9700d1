// rkdg_dense.hip — gfx950 kernels for small 2D elements whose reference
// operators are kept as dense matrices: affine triangles P_p (the
// L2_TriangleElement nodal basis) and non-affine (bilinear) quadrilaterals Q_p.
// Same path as rkdg_tensor.hip: FE_Evolution::Mult
// (src/fe_evolution/fe_evolution.cpp:102-216) fused with the RK-stage epilogue,
// element means (src/averager/averager.cpp:84-109) and the limiter
// (src/limiter/limiter.cpp:22-94, src/indicator/indicator_bj.cpp:15-200,
// src/limiter/limiter_multiplier.cpp:12-99, src/limiter/limiter_findiff.cpp:3-33).
//
// Volume term: z_j += sum_r sum_k K_r(j,k) G^r_k with G^r_k = adj(J)(r,:) . F(u_k)
// (nodal flux, GetFlux fe_evolution.cpp:219-245) and K_r(j,k) the exact integral
// of phi_k d_r phi_j (triangles; DomainIntegrator with intorder 2p is exact) or
// the collocated w_k d_r phi_j(x_k) with adj(J) taken at node k (Q_p on
// Gauss-Legendre nodes). Faces: per-face trace tables Phi_f[q][k]; the
// neighbour's trace is evaluated at the mirrored point of its own face. Mass:
// M^-1 = Mhat^-1 / det J (affine triangles) or a per-element dense inverse
// (bilinear quads).
#include "rkdg_dense.hpp"

namespace rkdg {

// one work block (EPB elements, one wavefront) of the right-hand side
template <int GEOM, int P, int RS, bool BC>
__device__ __forceinline__ void dense_rhs_block(const RhsArgs& A, const SchemeParams& g, int64_t vb,
                                                DenseTab<GEOM, P>& tb) {
  using C = DC<GEOM, P>;
  constexpr int ND = C::ND, NEQ = C::NEQ, NF = C::NF, NQ = C::NQ, NPT = C::NPT;
  constexpr int T = C::T, EPB = C::EPB, BLOCK = C::BLOCK, GS = C::GSIZE;
  constexpr bool TRI = C::TRI;

  __shared__ double s_u[EPB][NEQ][ND];
  __shared__ double s_g[EPB][2][NEQ][ND];     // contravariant nodal fluxes, then z
  __shared__ double s_f[EPB][NPT][NEQ];       // weighted face fluxes
  __shared__ double s_geo[EPB][GS];
  __shared__ int s_nb[EPB][NF], s_code[EPB][NF];

  const int tid = threadIdx.x;
  const int team = tid / T, lane = tid % T;
  const int64_t NE = A.NE;
  const int64_t NDOFS = NE * ND;
  const int64_t e0 = vb * EPB;
  const int64_t e = e0 + team;
  const int ins = (A.inst && e < NE) ? A.inst[e] : 0;
  const bool active = e < NE && (!A.active || A.active[ins]);

  // every load of the block's setup is issued before the first LDS store
  static_assert(NF <= T, "one lane per face");
  StateRegs<EPB, NEQ, ND, BLOCK> str;
  str.load(A.x, NDOFS, e0, NE, tid);
  constexpr int GSP = (GS + T - 1) / T;
  double gv[GSP];
#pragma unroll
  for (int p = 0; p < GSP; p++) gv[p] = (active && lane + p * T < GS) ? A.geo[(int64_t)(lane + p * T) * NE + e] : 0.0;
  const bool face_lane = active && lane < NF;
  const int nb0 = face_lane ? A.nbr[lane * NE + e] : -1;
  const int cd0 = face_lane ? A.fcode[lane * NE + e] : 0;
  str.store(s_u, tid);
  if (face_lane) {
    s_nb[team][lane] = nb0;
    s_code[team][lane] = nb0 >= 0 ? cd0 : 0;
  }
#pragma unroll
  for (int p = 0; p < GSP; p++)
    if (active && lane + p * T < GS) s_geo[team][lane + p * T] = gv[p];
  __syncthreads();

  const int j = lane;
  const bool node_lane = active && lane < ND;
  double lmax = 0.0;
  bool bad = false;

  // ---- volume: nodal flux, contravariant components
  if (!A.face_only && node_lane) {
    double st[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) st[v] = s_u[team][v][j];
    const SidePrim<2> q = side_prim<2>(st, g);
    lmax = sqrt_ws(q.m2) * q.inv + side_sound<2>(st, q, g);
    const double* adj = TRI ? &s_geo[team][C::G_ADJ] : &s_geo[team][C::G_ADJ + 4 * j];
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const double a0 = adj[r * 2 + 0], a1 = adj[r * 2 + 1];
      const double ma = st[1] * a0 + st[2] * a1;
      const double mai = ma * q.inv;
      s_g[team][r][0][j] = ma;
      s_g[team][r][1][j] = mai * st[1] + q.p * a0;
      s_g[team][r][2][j] = mai * st[2] + q.p * a1;
      s_g[team][r][3][j] = mai * (st[3] + q.p);
    }
  }
  // ---- face points: traces, Riemann solve, weighted flux
  for (int it = lane; it < NPT; it += T) {
    if (!active) break;
    const int f = it / NQ, q = it - f * NQ;
    const int nb = s_nb[team][f];
    double um[NEQ], up[NEQ], fl[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < ND; k++) t += s_u[team][v][k] * tb.Phi[f][q][k];
      um[v] = t;
    }
    double an[2];
    double factor = -1.0;
    bool ok = true;
    double lam = 0.0;
    const int code = s_code[team][f];
    const bool elem1 = nb < 0 || ((code >> 6) & 1);
    if (elem1) {
      an[0] = s_geo[team][C::G_NRM + 2 * f];
      an[1] = s_geo[team][C::G_NRM + 2 * f + 1];
    } else {
      const int fn = code & 7;
      an[0] = A.geo[(int64_t)(C::G_NRM + 2 * fn) * NE + nb];
      an[1] = A.geo[(int64_t)(C::G_NRM + 2 * fn + 1) * NE + nb];
    }
    const double normag = sqrt_ws(an[0] * an[0] + an[1] * an[1]);
    const double inv = 1.0 / norml2_scaled<2>(an);
    const double nrm[2] = {an[0] * inv, an[1] * inv};
    if (!BC || nb >= 0) {
      const int fn = code & 7;
      const int q2 = ((code >> 3) & 2) ? NQ - 1 - q : q;
      const double* xn = A.x + (int64_t)nb * ND;
#pragma unroll
      for (int v = 0; v < NEQ; v++) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < ND; k++) t += xn[v * NDOFS + k] * tb.Phi[fn][q2][k];
        up[v] = t;
      }
      double sL[NEQ], sR[NEQ];
#pragma unroll
      for (int v = 0; v < NEQ; v++) {
        sL[v] = elem1 ? um[v] : up[v];
        sR[v] = elem1 ? up[v] : um[v];
      }
      if (!elem1) factor = 1.0;
      lam = riemann_t<2, RS>(g, sL, sR, nrm, fl, ok);
    } else if constexpr (BC) {
      const int slot = -1 - nb;
      if (slot < g.num_bc) {
        const double* Pover = A.bcp ? A.bcp + ((int64_t)ins * RKDG_MAX_BC + slot) * 8 : nullptr;
        bc_state<2>(g, slot, um, nrm, up, Pover);
        if (g.bc_type[slot] == RKDG_BC_CHAR_OUTLET) lam = char_outlet_flux<2>(g, up, nrm, fl);
        else lam = riemann_t<2, RS>(g, um, up, nrm, fl, ok);
      } else {
#pragma unroll
        for (int v = 0; v < NEQ; v++) fl[v] = 0.0;
      }
    }
    if (!ok || lam < 0.0) {
      if (!bad) {
        const int64_t other = nb >= 0 ? (int64_t)nb : -1;
        const bool mine = factor < 0.0;
        const bool nf = nonfinite_nodes(A.x, NDOFS, e, ND, NEQ) || nonfinite_nodes(A.x, NDOFS, other, ND, NEQ);
        flag_error(A.err, mine ? e : other, mine ? other : e, f, nf ? kErrNonFinite : kErrUnphysical);
      }
      bad = true;
    } else {
      lmax = lam > lmax ? lam : lmax;
    }
    const double wgt = factor * tb.wq[q] * normag;
#pragma unroll
    for (int v = 0; v < NEQ; v++) s_f[team][it][v] = fl[v] * wgt;
  }
  __syncthreads();

  // ---- node lanes: volume contraction + lift
  double z[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; v++) z[v] = 0.0;
  if (node_lane) {
    if (!A.face_only) {
#pragma unroll
      for (int r = 0; r < 2; r++)
#pragma unroll
        for (int k = 0; k < ND; k++) {
          const double kk = tb.K[r][j][k];
#pragma unroll
          for (int v = 0; v < NEQ; v++) z[v] += kk * s_g[team][r][v][k];
        }
    }
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const double ph = tb.Phi[f][q][j];
#pragma unroll
        for (int v = 0; v < NEQ; v++) z[v] += s_f[team][f * NQ + q][v] * ph;
      }
  }
  if (!A.face_only) {
    __syncthreads();
    if (node_lane) {
#pragma unroll
      for (int v = 0; v < NEQ; v++) s_g[team][0][v][j] = z[v];
    }
    __syncthreads();
    // ---- inverse mass and stage epilogue
    double outv[NEQ];
    double mk = 0.0;
    if (node_lane) {
      double y[NEQ];
#pragma unroll
      for (int v = 0; v < NEQ; v++) y[v] = 0.0;
      if (TRI) {
        const double idet = recip(s_geo[team][C::G_DET]);
#pragma unroll
        for (int k = 0; k < ND; k++) {
          const double mi = tb.Minv[j][k];
#pragma unroll
          for (int v = 0; v < NEQ; v++) y[v] += mi * s_g[team][0][v][k];
        }
#pragma unroll
        for (int v = 0; v < NEQ; v++) y[v] *= idet;
        mk = tb.mhat[j] * s_geo[team][C::G_DET];
      } else {
#pragma unroll
        for (int k = 0; k < ND; k++) {
          const double mi = s_geo[team][C::G_MINV + j * ND + k];
#pragma unroll
          for (int v = 0; v < NEQ; v++) y[v] += mi * s_g[team][0][v][k];
        }
        mk = s_geo[team][C::G_MW + j];
      }
      const double dtm = A.dt_inst ? A.dt_inst[ins] : 1.0;
      // every base is loaded before the first store (an output may be its own base)
      const int64_t i0 = e * ND + j;
      double u0[NEQ];
#pragma unroll
      for (int v = 0; v < NEQ; v++) u0[v] = s_u[team][v][j];
      double base[4][NEQ];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const double* bp = m < A.nout ? A.outs[m].base : nullptr;
        if (bp) {
          bp += i0;
#pragma unroll
          for (int v = 0; v < NEQ; v++) base[m][v] = bp[v * NDOFS];
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
          double* op = A.outs[m].out + i0;
          const double cf = A.outs[m].coef * dtm;
          const double wb = A.outs[m].wb, wi = A.outs[m].wi;
#pragma unroll
          for (int v = 0; v < NEQ; v++) {
            const double o = wi == 0.0 ? base[m][v] + cf * y[v] : wb * base[m][v] + wi * (u0[v] + cf * y[v]);
            op[v * NDOFS] = o;
            if (m == A.means_of) outv[v] = o;
          }
        }
      }
      if (A.k) {
        double* kp = A.k + i0;
#pragma unroll
        for (int v = 0; v < NEQ; v++) kp[v * NDOFS] = y[v];
      }
    }
    if (A.means) {
      __syncthreads();
      if (node_lane) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) s_g[team][1][v][j] = outv[v];
        s_g[team][0][0][j] = mk;
      }
      __syncthreads();
      if (active && lane < NEQ) {
        double s = 0.0, ms = 0.0;
        for (int k = 0; k < ND; k++) {
          const double w = s_g[team][0][0][k];
          s += w * s_g[team][1][lane][k];
          ms += w;
        }
        A.means[lane * NE + e] = s / ms;
      }
    }
  }
  // ---- max wave speed: per instance slot
  for (int off = T / 2; off > 0; off >>= 1) {
    const double o = __shfl_xor(lmax, off, 64);
    lmax = o > lmax ? o : lmax;
  }
  const int64_t elast = e0 + EPB - 1 < NE ? e0 + EPB - 1 : NE - 1;
  const bool uniform = !A.inst || (e0 < NE && A.inst[e0] == A.inst[elast]);
  if (uniform) {
    for (int off = 32; off >= T; off >>= 1) {
      const double o = __shfl_xor(lmax, off, 64);
      lmax = o > lmax ? o : lmax;
    }
    const int i0 = A.inst && e0 < NE ? A.inst[e0] : 0;
    if (tid == 0) atomic_max_pos(A.lam + (A.inst ? i0 : 0), lmax);
  } else if (lane == 0 && active) {
    atomic_max_pos(A.lam + ins, lmax);
  }
  __syncthreads();  // LDS is reused by the next work block
}

// persistent: the grid strides over the work blocks
template <int GEOM, int P, int RS, bool BC>
__global__ __launch_bounds__(64, 3) void dense_rhs_kernel(const RhsArgs A, const SchemeParams g,
                                                       const double* __restrict__ gtab, int64_t nvb) {
  __shared__ DenseTab<GEOM, P> tb;
  stage_dense<GEOM, P>(tb, gtab);
  __syncthreads();
  for (int64_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
    // keep table reads inside the iteration (hoisted, they would hold registers across the loop)
    asm volatile("" ::: "memory");
    dense_rhs_block<GEOM, P, RS, BC>(A, g, vb, tb);
  }
}

template <int GEOM, int P>
__global__ __launch_bounds__(256) void dense_means_kernel(const MeansArgs A, const double* __restrict__ gtab) {
  using C = DC<GEOM, P>;
  constexpr int ND = C::ND, NEQ = C::NEQ;
  const DenseTab<GEOM, P>* tb = reinterpret_cast<const DenseTab<GEOM, P>*>(gtab);
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= A.NE) return;
  const int64_t NDOFS = A.NE * ND;
  double s[NEQ], ms = 0.0;
#pragma unroll
  for (int v = 0; v < NEQ; v++) s[v] = 0.0;
  for (int k = 0; k < ND; k++) {
    const double mk = C::TRI ? tb->mhat[k] * A.geo[(int64_t)C::G_DET * A.NE + e]
                             : A.geo[(int64_t)(C::G_MW + k) * A.NE + e];
    ms += mk;
#pragma unroll
    for (int v = 0; v < NEQ; v++) s[v] += mk * A.x[v * NDOFS + e * ND + k];
  }
#pragma unroll
  for (int v = 0; v < NEQ; v++) A.means[v * A.NE + e] = s[v] / ms;
}

// Only the deciding variable's indicator is formed (density for the
// multiplier, indicator_bj.cpp:193; the last variable for FinDiff,
// limiter_findiff.cpp:7-8).
template <int GEOM, int P>
__global__ __launch_bounds__(64, 4) void dense_limit_kernel(const LimitArgs A, const SchemeParams g,
                                                         const double* __restrict__ gtab) {
  using C = DC<GEOM, P>;
  constexpr int ND = C::ND, NEQ = C::NEQ, NF = C::NF, NV = C::NV;
  constexpr int T = C::T, EPB = C::EPB, BLOCK = C::BLOCK;
  constexpr int NPT = NF * 2;

  __shared__ DenseTab<GEOM, P> tb;
  __shared__ double s_u[EPB][NEQ][ND];
  __shared__ double s_mean[EPB][NEQ];  // the element's means
  __shared__ double s_nbm[EPB][NF];    // deciding variable's means of the face neighbours
  __shared__ double s_area[EPB][ND];   // node mass weights (triangles: det J in [0])

  const int tid = threadIdx.x;
  const int team = tid / T, lane = tid % T;
  const int64_t NE = A.NE;
  const int64_t NDOFS = NE * ND;
  const int64_t e0 = xcd_remap_d(blockIdx.x, gridDim.x) * EPB;
  const int64_t e = e0 + team;
  const bool active = e < NE && (!A.active || A.active[A.inst[e]]);
  const int vsel = g.limiter == RKDG_LIM_FINDIFF ? NEQ - 1 : 0;

  // every load of the setup is issued before the first LDS store
  CopyRegs<sizeof(DenseTab<GEOM, P>) / sizeof(double), BLOCK> tbr;
  tbr.load(gtab, tid);
  StateRegs<EPB, NEQ, ND, BLOCK> str;
  str.load(A.x, NDOFS, e0, NE, tid);
  // the element's means (lane v < NEQ), the deciding variable's means of the
  // face neighbours (lane f < NF) and the node mass weights (lane k < ND): one
  // load instruction each, combined through LDS after the barrier
  static_assert(NEQ <= T && NF <= T && ND <= T, "one lane per variable, face and node");
  const bool ind_bj = g.indicator == RKDG_IND_BJ;
  const bool venkat = g.phi == RKDG_PHI_VENKAT && g.venkat_k > 0.0;
  const double mown = (active && lane < NEQ) ? A.means[lane * NE + e] : 0.0;
  const int nbl = (active && ind_bj && lane < NF) ? A.nbr[lane * NE + e] : -1;
  const double nml = nbl >= 0 ? A.means[vsel * NE + nbl] : 0.0;
  const unsigned long long nbb = __ballot(nbl >= 0);
  const int nbmask = (int)((nbb >> (team * T)) & ((1ull << NF) - 1ull));  // bit f: face f has a neighbour
  double arl = 0.0;
  if (active && venkat) {
    if (C::TRI) arl = lane == 0 ? A.geo[(int64_t)C::G_DET * NE + e] : 0.0;
    else arl = lane < ND ? A.geo[(int64_t)(C::G_MW + lane) * NE + e] : 0.0;
  }
  tbr.store(reinterpret_cast<double*>(&tb), tid);
  str.store(s_u, tid);
  if (lane < NEQ) s_mean[team][lane] = mown;
  if (lane < NF) s_nbm[team][lane] = nml;
  if (lane < ND) s_area[team][lane] = arl;
  __syncthreads();
  double mean[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; v++) mean[v] = s_mean[team][v];
  double mI = 1e9, MI = -1e9;
  if (active && ind_bj) {
    const double mv = mean[vsel];
    mI = mv;
    MI = mv;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      if ((nbmask >> f) & 1) {
        const double a = s_nbm[team][f];
        mI = a < mI ? a : mI;
        MI = a > MI ? a : MI;
      }
    }
  }
  double eps2 = 0.0;
  if (active && venkat) {
    // (K h)^3 with h = |element|^(1/2)
    double area = 0.0;
    if (C::TRI) area = 0.5 * s_area[team][0];
    else
#pragma unroll
      for (int k = 0; k < ND; k++) area += s_area[team][k];
    const double h = sqrt(area);
    const double kh = g.venkat_k * h;
    eps2 = kh * kh * kh;
  }
  double ylane = 1e9;  // this lane's minimum over its indicator points
  if (ind_bj) {
    for (int it = lane; it < NPT; it += T) {
      if (!active) break;
      const int f = it / 2, ip = it % 2;
      double yc = 1e9;
      if ((nbmask >> f) & 1) {
        double u = 0.0;
#pragma unroll
        for (int k = 0; k < ND; k++) u += s_u[team][vsel][k] * tb.PhiI[f][ip][k];
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
  if constexpr (P == 1 && !C::TRI) {
    // linearize: the multiplier and the last hope work on the element's linear
    // reconstruction (triangles are skipped, limiter_multiplier.cpp:29)
    if (g.linearize && g.limiter == RKDG_LIM_MULTI) {
      double lin[NEQ];
      const bool nl = active && lane < ND;
      if (nl) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) lin[v] = linearize_q1<2>(&s_u[team][v][0], mean[v], lane);
      }
      __syncthreads();
      if (nl) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) s_u[team][v][lane] = lin[v];
      }
      __syncthreads();
    }
  }
  const int j = lane;
  const bool node_lane = active && lane < ND;
  double u[NEQ];
  if (node_lane) {
#pragma unroll
    for (int v = 0; v < NEQ; v++) u[v] = s_u[team][v][j];
  }
  int trouble = 0;
  if (g.limiter == RKDG_LIM_MULTI) {
    if (yr < 1.0) trouble = 1;
    double f = yr;
    if (g.pow_nd) {
      double acc = 1.0, b = yr;
      for (int e2 = ND; e2 > 0; e2 >>= 1) {
        if (e2 & 1) acc *= b;
        b *= b;
      }
      f = acc;
    }
    if (node_lane) {
#pragma unroll
      for (int v = 0; v < NEQ; v++) u[v] = mean[v] + f * (u[v] - mean[v]);
    }
    if (g.last_hope) {
      bool bad_v = false;
      for (int it = lane; it < NV; it += T) {
        if (!active) break;
        double vs[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < ND; k++) s += s_u[team][v][k] * tb.PhiV[it][k];
          vs[v] = mean[v] + f * (s - mean[v]);
        }
        if (!physical<2>(vs, g)) bad_v = true;
      }
      // any unphysical vertex of the team's element (one ballot, no barrier)
      const unsigned long long bm = __ballot(bad_v);
      const unsigned long long tm = T == 64 ? ~0ull : (((1ull << T) - 1ull) << (team * T));
      if (active && (bm & tm)) {
        trouble |= 2;
        if (node_lane) {
#pragma unroll
          for (int v = 0; v < NEQ; v++) u[v] = mean[v];
        }
      }
    }
  } else if (g.limiter == RKDG_LIM_FINDIFF) {
    if (yr < 0.999) {
      trouble = 1;
      if (node_lane) {
#pragma unroll
        for (int v = 0; v < NEQ; v++) u[v] = mean[v];
      }
    }
  }
  if (node_lane && (g.limiter != RKDG_LIM_NONE || A.out)) {
    double* dst = A.out ? A.out : A.x;
#pragma unroll
    for (int v = 0; v < NEQ; v++) dst[v * NDOFS + e * ND + j] = u[v];
  }
  if (A.troubled && active && lane == 0) A.troubled[e] = (int8_t)trouble;
}

// ---------------------------------------------------------------------------
bool dense_supported(int geom, int p) {
#define X(G, PP) \
  if (geom == G && p == PP) return true;
  RKDG_DENSE_CASES(X)
#undef X
  return false;
}

int dense_table_size(int geom, int p) {
#define X(G, PP) \
  if (geom == G && p == PP) return (int)(sizeof(DenseTab<G, PP>) / sizeof(double));
  RKDG_DENSE_CASES(X)
#undef X
  return 0;
}

int dense_geo_size(int geom, int p) {
#define X(G, PP) \
  if (geom == G && p == PP) return DC<G, PP>::GSIZE;
  RKDG_DENSE_CASES(X)
#undef X
  return 0;
}

template <int G, int PP, int RS>
hipError_t dense_rhs_rs(const RhsArgs& a, const SchemeParams& g, const double* t, bool bc, hipStream_t s) {
  using C = DC<G, PP>;
  const int64_t nvb = (a.NE + C::EPB - 1) / C::EPB;
  const int64_t grid = a.grid_cap > 0 && a.grid_cap < nvb ? a.grid_cap : nvb;
  if (bc) dense_rhs_kernel<G, PP, RS, true><<<(unsigned)grid, C::BLOCK, 0, s>>>(a, g, t, nvb);
  else dense_rhs_kernel<G, PP, RS, false><<<(unsigned)grid, C::BLOCK, 0, s>>>(a, g, t, nvb);
  return hipGetLastError();
}


template <int G, int PP>
hipError_t dense_rhs_gp(const RhsArgs& a, const SchemeParams& g, const double* t, bool bc, hipStream_t s) {
  switch (g.riemann) {
    case RKDG_RS_RUSANOV: return dense_rhs_rs<G, PP, RKDG_RS_RUSANOV>(a, g, t, bc, s);
    case RKDG_RS_LLF: return dense_rhs_rs<G, PP, RKDG_RS_LLF>(a, g, t, bc, s);
    case RKDG_RS_HLL: return dense_rhs_rs<G, PP, RKDG_RS_HLL>(a, g, t, bc, s);
    default: return dense_rhs_rs<G, PP, RKDG_RS_HLLC>(a, g, t, bc, s);
  }
}

hipError_t launch_dense_rhs(int geom, int p, const RhsArgs& a, const SchemeParams& g, const double* t,
                            bool has_bdr, hipStream_t s) {
#define X(G, PP) \
  if (geom == G && p == PP) return dense_rhs_gp<G, PP>(a, g, t, has_bdr, s);
  RKDG_DENSE_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_dense_limit(int geom, int p, const LimitArgs& a, const SchemeParams& g, const double* t,
                              hipStream_t s) {
#define X(G, PP)                                                                \
  if (geom == G && p == PP) {                                                   \
    using C = DC<G, PP>;                                                        \
    const int64_t nb = (a.NE + C::EPB - 1) / C::EPB;                            \
    dense_limit_kernel<G, PP><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t);       \
    return hipGetLastError();                                                   \
  }
  RKDG_DENSE_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_dense_means(int geom, int p, const MeansArgs& a, const double* t, hipStream_t s) {
#define X(G, PP)                                                                         \
  if (geom == G && p == PP) {                                                            \
    dense_means_kernel<G, PP><<<(unsigned)((a.NE + 255) / 256), 256, 0, s>>>(a, t);    \
    return hipGetLastError();                                                            \
  }
  RKDG_DENSE_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

}  // namespace rkdg
