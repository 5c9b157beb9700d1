// rkdg_adjoint.hip — backward (adjoint) kernels of the 2D dense RKDG path:
// the transpose of the right-hand side FE_Evolution::Mult
// (src/fe_evolution/fe_evolution.cpp:102-216; volume flux GetFlux :219-245,
// face loop face_integrator.cpp:11-276, boundary_integrator.cpp:12-194), of the
// limiter Limiter::update (limiter.cpp:22-94 with indicator_bj.cpp:15-200 and
// limiter_multiplier.cpp:12-99), and of the objective (pressure integrated over
// the faces of one boundary patch).
//
// The right-hand-side transpose is gather-only: every element recomputes the
// face fluxes it shares with its neighbours (their states and adjoints are read,
// never written), so no atomics touch the state adjoint. For a face point with
// flux F(uL, uR, n) entering z_e1 with weight -w|n| and z_e2 with +w|n|, the flux
// adjoint is w|n| (zbar_e2 . phi2 - zbar_e1 . phi1), with zbar = M^-1 ybar; each
// side applies the Riemann solver's vector-Jacobian product and keeps its own
// part. The limiter transpose is two passes: pass A differentiates each
// element's own update and records the adjoint of the stencil bound it used
// (which belongs to the stencil element that attained it); pass B gathers those
// records over the stencil and spreads the mean adjoints to the nodes.
#include "rkdg_adjoint.hpp"
#include "rkdg_dense.hpp"

namespace rkdg {

// ---------------------------------------------------------------------------
// right-hand side transpose: xbar = addend + (dR/dx)^T kbar
template <int GEOM, int P, int RS, bool BC>
__global__ __launch_bounds__(64) void dense_rhs_adj_kernel(const RhsAdjArgs A, const SchemeParams g,
                                                           const double* __restrict__ gtab) {
  using C = DC<GEOM, P>;
  constexpr int ND = C::ND, NEQ = C::NEQ, NF = C::NF, NQ = C::NQ, NPT = C::NPT;
  constexpr int T = C::T, EPB = C::EPB, BLOCK = C::BLOCK, GS = C::GSIZE;
  constexpr bool TRI = C::TRI;

  __shared__ DenseTab<GEOM, P> tb;
  __shared__ double s_u[EPB][NEQ][ND];
  __shared__ double s_kb[EPB][NEQ][ND];
  __shared__ double s_zb[EPB][NEQ][ND];
  __shared__ double s_ft[EPB][NPT][NEQ];
  __shared__ double s_geo[EPB][GS];
  __shared__ int s_nb[EPB][NF], s_code[EPB][NF];

  const int tid = threadIdx.x;
  const int team = tid / T, lane = tid % T;
  const int64_t NE = A.NE;
  const int64_t NDOFS = NE * ND;
  const int64_t e0 = xcd_remap_d(blockIdx.x, gridDim.x) * EPB;
  const int64_t e = e0 + team;
  const bool active = e < NE;

  // every load of the states is issued before the first LDS store
  CopyRegs<sizeof(DenseTab<GEOM, P>) / sizeof(double), BLOCK> tbr;
  tbr.load(gtab, tid);
  StateRegs<EPB, NEQ, ND, BLOCK> sx, skb;
  sx.load(A.x, NDOFS, e0, NE, tid, 1.0);
  skb.load(A.kbar, NDOFS, e0, NE, tid, 0.0);
  tbr.store(reinterpret_cast<double*>(&tb), tid);
  sx.store(s_u, tid);
  skb.store(s_kb, tid);
  for (int i = lane; active && i < NF; i += T) {
    const int nb = A.nbr[i * NE + e];
    s_nb[team][i] = nb;
    s_code[team][i] = nb >= 0 ? A.fcode[i * NE + e] : 0;
  }
  for (int i = lane; active && i < GS; i += T) s_geo[team][i] = A.geo[(int64_t)i * NE + e];
  __syncthreads();

  const int j = lane;
  const bool node_lane = active && lane < ND;
  // zbar = M^-1 kbar (the inverse mass matrix is symmetric)
  if (node_lane) {
    double zb[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) zb[v] = 0.0;
    if (TRI) {
      const double idet = 1.0 / s_geo[team][C::G_DET];
#pragma unroll
      for (int k = 0; k < ND; k++) {
        const double mi = tb.Minv[j][k] * idet;
#pragma unroll
        for (int v = 0; v < NEQ; v++) zb[v] += mi * s_kb[team][v][k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < ND; k++) {
        const double mi = s_geo[team][C::G_MINV + j * ND + k];
#pragma unroll
        for (int v = 0; v < NEQ; v++) zb[v] += mi * s_kb[team][v][k];
      }
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) s_zb[team][v][j] = zb[v];
  }
  __syncthreads();

  // ---- volume: ubar_j += sum_r (dG^r/du)^T (sum_i K_r[i][j] zbar_i)
  double ub[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; v++) ub[v] = 0.0;
  if (node_lane) {
    double st[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) st[v] = s_u[team][v][j];
    const double* adjm = TRI ? &s_geo[team][C::G_ADJ] : &s_geo[team][C::G_ADJ + 4 * j];
#pragma unroll
    for (int r = 0; r < 2; r++) {
      double Gb[NEQ];
#pragma unroll
      for (int v = 0; v < NEQ; v++) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < ND; i++) t += tb.K[r][i][j] * s_zb[team][v][i];
        Gb[v] = t;
      }
      const double a[2] = {adjm[r * 2 + 0], adjm[r * 2 + 1]};
      adj::volume_flux_vjp(st, a, Gb, g, ub);
    }
  }

  // ---- face points: flux adjoint from both sides, Riemann VJP, own part
  for (int it = lane; it < NPT; it += T) {
    if (!active) break;
    const int f = it / NQ, q = it - f * NQ;
    const int nb = s_nb[team][f];
    const int code = s_code[team][f];
    double um[NEQ], zo[NEQ], utr[NEQ];
#pragma unroll
    for (int v = 0; v < NEQ; v++) {
      double t = 0.0, z = 0.0;
#pragma unroll
      for (int k = 0; k < ND; k++) {
        t += s_u[team][v][k] * tb.Phi[f][q][k];
        z += s_zb[team][v][k] * tb.Phi[f][q][k];
      }
      um[v] = t;
      zo[v] = z;
      utr[v] = 0.0;
    }
    const bool elem1 = nb < 0 || ((code >> 6) & 1);
    double an[2];
    if (elem1) {
      an[0] = s_geo[team][C::G_NRM + 2 * f];
      an[1] = s_geo[team][C::G_NRM + 2 * f + 1];
    } else {
      const int fn = code & 7;
      an[0] = A.geo[(int64_t)(C::G_NRM + 2 * fn) * NE + nb];
      an[1] = A.geo[(int64_t)(C::G_NRM + 2 * fn + 1) * NE + nb];
    }
    const double normag = sqrt(an[0] * an[0] + an[1] * an[1]);
    const double inv = 1.0 / norml2_scaled<2>(an);
    const double nrm[2] = {an[0] * inv, an[1] * inv};
    const double wn = tb.wq[q] * normag;
    if (!BC || nb >= 0) {
      const int fn = code & 7;
      const int q2 = ((code >> 3) & 2) ? NQ - 1 - q : q;
      double up[NEQ], zn[NEQ];
      // neighbour trace and zbar_nb . phi_nb with zbar_nb = Minv_nb kbar_nb
      double tcoef[ND];
      if (TRI) {
        const double idet = 1.0 / A.geo[(int64_t)C::G_DET * NE + nb];
#pragma unroll
        for (int k = 0; k < ND; k++) {
          double t = 0.0;
#pragma unroll
          for (int i = 0; i < ND; i++) t += tb.Phi[fn][q2][i] * tb.Minv[i][k];
          tcoef[k] = t * idet;
        }
      } else {
#pragma unroll
        for (int k = 0; k < ND; k++) {
          double t = 0.0;
#pragma unroll
          for (int i = 0; i < ND; i++) t += tb.Phi[fn][q2][i] * A.geo[(int64_t)(C::G_MINV + i * ND + k) * NE + nb];
          tcoef[k] = t;
        }
      }
      const double* xn = A.x + (int64_t)nb * ND;
      const double* kn = A.kbar + (int64_t)nb * ND;
#pragma unroll
      for (int v = 0; v < NEQ; v++) {
        double t = 0.0, z = 0.0;
#pragma unroll
        for (int k = 0; k < ND; k++) {
          t += xn[v * NDOFS + k] * tb.Phi[fn][q2][k];
          z += kn[v * NDOFS + k] * tcoef[k];
        }
        up[v] = t;
        zn[v] = z;
      }
      // own weight: -w|n| for Elem1, +w|n| for Elem2; the neighbour has the other sign
      const double so = elem1 ? -1.0 : 1.0;
      double Fb[NEQ], sLb[NEQ], sRb[NEQ], sL[NEQ], sR[NEQ];
#pragma unroll
      for (int v = 0; v < NEQ; v++) {
        Fb[v] = wn * so * (zo[v] - zn[v]);
        sL[v] = elem1 ? um[v] : up[v];
        sR[v] = elem1 ? up[v] : um[v];
        sLb[v] = 0.0;
        sRb[v] = 0.0;
      }
      adj::riemann_vjp<RS>(g, sL, sR, nrm, Fb, sLb, sRb);
#pragma unroll
      for (int v = 0; v < NEQ; v++) utr[v] = elem1 ? sLb[v] : sRb[v];
    } else if constexpr (BC) {
      const int slot = -1 - nb;
      if (slot < g.num_bc) {
        double up[NEQ], Fb[NEQ], sLb[NEQ], sRb[NEQ];
        bc_state<2>(g, slot, um, nrm, up);
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
          Fb[v] = -wn * zo[v];
          sLb[v] = 0.0;
          sRb[v] = 0.0;
        }
        adj::riemann_vjp<RS>(g, um, up, nrm, Fb, sLb, sRb);
        double Pb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        adj::bc_state_vjp(g, slot, um, nrm, sRb, sLb, Pb);
#pragma unroll
        for (int v = 0; v < NEQ; v++) utr[v] = sLb[v];
        if (A.pbar) {
#pragma unroll
          for (int k = 0; k < 8; k++)
            if (Pb[k] != 0.0) atomicAdd(A.pbar + slot * 8 + k, Pb[k]);
        }
      }
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) s_ft[team][it][v] = utr[v];
  }
  __syncthreads();
  if (node_lane) {
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const double ph = tb.Phi[f][q][j];
#pragma unroll
        for (int v = 0; v < NEQ; v++) ub[v] += ph * s_ft[team][f * NQ + q][v];
      }
#pragma unroll
    for (int v = 0; v < NEQ; v++) {
      const int64_t idx = v * NDOFS + e * ND + j;
      A.xbar[idx] = (A.addend ? A.addend[idx] : 0.0) + ub[v];
    }
  }
}

// ---------------------------------------------------------------------------
// limiter transpose, pass A (per element): forward decisions recomputed from
// the pre-limit state; vbar gets the element-local part, meanbar the local mean
// adjoints, and (send_to, send_val) the adjoint of the density stencil bound
// that produced the multiplier, owed to the stencil element that attained it.
template <int GEOM, int P>
__global__ __launch_bounds__(64) void dense_limit_adj_a_kernel(const LimAdjArgs A, const SchemeParams g,
                                                               const double* __restrict__ gtab) {
  using C = DC<GEOM, P>;
  constexpr int ND = C::ND, NEQ = C::NEQ, NF = C::NF, NV = C::NV;
  constexpr int T = C::T, EPB = C::EPB, BLOCK = C::BLOCK;
  constexpr int NPT = NF * 2;

  __shared__ DenseTab<GEOM, P> tb;
  __shared__ double s_u[EPB][NEQ][ND];
  __shared__ double s_ub[EPB][NEQ][ND];

  const int tid = threadIdx.x;
  const int team = tid / T, lane = tid % T;
  const int64_t NE = A.NE;
  const int64_t NDOFS = NE * ND;
  const int64_t e0 = xcd_remap_d(blockIdx.x, gridDim.x) * EPB;
  const int64_t e = e0 + team;
  const bool active = e < NE;

  // every load of the states is issued before the first LDS store
  CopyRegs<sizeof(DenseTab<GEOM, P>) / sizeof(double), BLOCK> tbr;
  tbr.load(gtab, tid);
  StateRegs<EPB, NEQ, ND, BLOCK> sx, sub;
  sx.load(A.x, NDOFS, e0, NE, tid, 1.0);
  sub.load(A.ubar, NDOFS, e0, NE, tid, 0.0);
  // the serial lane's element data, also loaded before the barrier and
  // branch-free (the neighbours' density means in one round trip after the
  // indices; a boundary face reads the element's own, which is ignored)
  const bool serial = active && lane == 0;
  double mean[NEQ];
  int nbs[NF];
  double nmd[NF], nml[NF];
  double mw[C::TRI ? 1 : ND];
#pragma unroll
  for (int v = 0; v < NEQ; v++) mean[v] = serial ? A.means[v * NE + e] : 0.0;
#pragma unroll
  for (int fc = 0; fc < NF; fc++) nbs[fc] = serial ? A.nbr[fc * NE + e] : -1;
#pragma unroll
  for (int fc = 0; fc < NF; fc++) {
    const int64_t ne = nbs[fc] >= 0 ? (int64_t)nbs[fc] : e;
    nmd[fc] = serial ? A.means[ne] : 0.0;
    nml[fc] = (serial && g.limiter == RKDG_LIM_FINDIFF) ? A.means[(NEQ - 1) * NE + ne] : 0.0;
  }
  if (C::TRI) mw[0] = serial ? A.geo[(int64_t)C::G_DET * NE + e] : 0.0;
  else
#pragma unroll
    for (int k = 0; k < (C::TRI ? 1 : ND); k++) mw[k] = serial ? A.geo[(int64_t)(C::G_MW + k) * NE + e] : 0.0;
  tbr.store(reinterpret_cast<double*>(&tb), tid);
  sx.store(s_u, tid);
  sub.store(s_ub, tid);
  // stencil bounds of the density mean with the element that attains them
  // (serial lane), shared through LDS with the team's lanes
  __shared__ double s_st[EPB][4];  // mI, MI, eps2, mean of the density
  __shared__ int s_nbok[EPB];      // bit fc: face fc has a neighbour
  const bool ind_bj = g.indicator == RKDG_IND_BJ;
  double mI = 1e9, MI = -1e9;
  int64_t amin = -1, amax = -1;
  double eps2 = 0.0;
  if (serial && ind_bj) {
    {
      const double a = mean[0];
      if (a < mI) { mI = a; amin = e; }
      if (a > MI) { MI = a; amax = e; }
    }
    int ok = 0;
#pragma unroll
    for (int fc = 0; fc < NF; fc++) {
      const int nb = nbs[fc];
      if (nb < 0) continue;
      ok |= 1 << fc;
      const double a = nmd[fc];
      if (a < mI) { mI = a; amin = nb; }
      if (a > MI) { MI = a; amax = nb; }
    }
    if (g.phi == RKDG_PHI_VENKAT && g.venkat_k > 0.0) {
      double area = 0.0;
      if (C::TRI) area = 0.5 * mw[0];
      else
        for (int k = 0; k < ND; k++) area += mw[k];
      const double kh = g.venkat_k * sqrt(area);
      eps2 = kh * kh * kh;
    }
    s_st[team][0] = mI;
    s_st[team][1] = MI;
    s_st[team][2] = eps2;
    s_st[team][3] = mean[0];
    s_nbok[team] = ok;
  }
  __syncthreads();
  // the indicator points of interior faces, one lane each, and the team's
  // argmin with the serial loop's tie-break (the first minimum in point order)
  double ycl = 1e9, dl = 0.0, yl = 1.0;
  int itl = NPT;
  if (active && ind_bj) {
    const double mIs = s_st[team][0], MIs = s_st[team][1], e2s = s_st[team][2], m0 = s_st[team][3];
    const int ok = s_nbok[team];
    for (int it = lane; it < NPT; it += T) {
      const int fc = it / 2, ip = it % 2;
      if (!((ok >> fc) & 1)) continue;
      double uu = 0.0;
#pragma unroll
      for (int k = 0; k < ND; k++) uu += s_u[team][0][k] * tb.PhiI[fc][ip][k];
      const double diff = uu - m0;
      const double y = indicator_y(diff, m0, mIs, MIs);
      double dy, dd;
      const double yc = adj::limiter_phi(g, y, diff, e2s, &dy, &dd);
      if (yc < ycl) {
        ycl = yc;
        dl = diff;
        yl = y;
        itl = it;
      }
    }
  }
#pragma unroll
  for (int off = T / 2; off > 0; off >>= 1) {
    const double oy = __shfl_xor(ycl, off, 64), od = __shfl_xor(dl, off, 64), oyy = __shfl_xor(yl, off, 64);
    const int oi = __shfl_xor(itl, off, 64);
    if (oy < ycl || (oy == ycl && oi < itl)) {
      ycl = oy;
      dl = od;
      yl = oyy;
      itl = oi;
    }
  }
  // one lane per element does the rest of the (small) serial recomputation
  if (serial) {
    // objective on the limited state of this element (adds to ubar)
    double u[NEQ][ND];
    double alpha = 1.0, f = 1.0;
    bool reset = false;
    double yMin0 = 1e9;
    int pstar = -1;
    double dstar = 0.0, ystar = 1.0;
    int branch = 0;  // +1: diff > 1e-3 (MI), -1: diff < -1e-3 (mI), 0: y = 1
    if (ind_bj) {
      if (ycl < yMin0) {
        yMin0 = ycl;
        pstar = itl;
        dstar = dl;
        ystar = yl;
        branch = dl > 1e-3 ? 1 : (dl < -1e-3 ? -1 : 0);
      }
    } else {
      yMin0 = g.indicator == RKDG_IND_NOWHERE ? 1.0 : 0.0;
    }
    alpha = yMin0;
    double fprime = 1.0;  // d f / d alpha
    if (g.limiter == RKDG_LIM_MULTI) {
      f = alpha;
      if (g.pow_nd) {
        double acc = 1.0, b = alpha;
        for (int e2 = ND; e2 > 0; e2 >>= 1) {
          if (e2 & 1) acc *= b;
          b *= b;
        }
        f = acc;
        double accp = 1.0;
        for (int k = 0; k < ND - 1; k++) accp *= alpha;
        fprime = ND * accp;
      }
      if (g.last_hope) {
        for (int vt = 0; vt < NV && !reset; vt++) {
          double vs[NEQ];
#pragma unroll
          for (int v = 0; v < NEQ; v++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < ND; k++) s += s_u[team][v][k] * tb.PhiV[vt][k];
            vs[v] = mean[v] + f * (s - mean[v]);
          }
          if (!physical<2>(vs, g)) reset = true;
        }
      }
    } else if (g.limiter == RKDG_LIM_FINDIFF) {
      // the forward's total-reset test uses the last variable's indicator
      double yl = 1e9;
      if (ind_bj) {
        for (int it = 0; it < NPT; it++) {
          const int fc = it / 2, ip = it % 2;
          if (nbs[fc] < 0) continue;
          double mIl = 1e9, MIl = -1e9;
          {
            const double a = mean[NEQ - 1];
            mIl = a < mIl ? a : mIl;
            MIl = a > MIl ? a : MIl;
          }
          for (int f2 = 0; f2 < NF; f2++) {
            if (nbs[f2] < 0) continue;
            const double a = nml[f2];
            mIl = a < mIl ? a : mIl;
            MIl = a > MIl ? a : MIl;
          }
          double uu = 0.0;
#pragma unroll
          for (int k = 0; k < ND; k++) uu += s_u[team][NEQ - 1][k] * tb.PhiI[fc][ip][k];
          const double diff = uu - mean[NEQ - 1];
          const double y = indicator_y(diff, mean[NEQ - 1], mIl, MIl);
          double dy, dd;
          const double yc = adj::limiter_phi(g, y, diff, eps2, &dy, &dd);
          yl = yc < yl ? yc : yl;
        }
      } else {
        yl = g.indicator == RKDG_IND_NOWHERE ? 1.0 : 0.0;
      }
      reset = yl < 0.999;
    }
    // limited state (needed by the fused objective)
#pragma unroll
    for (int v = 0; v < NEQ; v++)
#pragma unroll
      for (int k = 0; k < ND; k++) {
        const double x0 = s_u[team][v][k];
        u[v][k] = g.limiter == RKDG_LIM_NONE ? x0 : (reset ? mean[v] : (g.limiter == RKDG_LIM_MULTI ? mean[v] + f * (x0 - mean[v]) : x0));
      }
    // fused objective: ubar += coef * d/du of sum_q w|n| p(trace) on faces of obj_slot
    if (A.obj_slot >= 0) {
      for (int fc = 0; fc < NF; fc++) {
        if (nbs[fc] != -1 - A.obj_slot) continue;
        const double an0 = A.geo[(int64_t)(C::G_NRM + 2 * fc) * NE + e];
        const double an1 = A.geo[(int64_t)(C::G_NRM + 2 * fc + 1) * NE + e];
        const double normag = sqrt(an0 * an0 + an1 * an1);
        for (int q = 0; q < C::NQ; q++) {
          double tr[NEQ], dp[NEQ];
#pragma unroll
          for (int v = 0; v < NEQ; v++) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < ND; k++) t += u[v][k] * tb.Phi[fc][q][k];
            tr[v] = t;
          }
          adj::pressure_grad(tr, g, dp);
          const double w = A.obj_coef * tb.wq[q] * normag;
#pragma unroll
          for (int v = 0; v < NEQ; v++)
#pragma unroll
            for (int k = 0; k < ND; k++) s_ub[team][v][k] += w * dp[v] * tb.Phi[fc][q][k];
        }
      }
    }
    // transpose of u = mean + f (x - mean) (or u = mean after a reset, or u = x)
    double mb[NEQ];
    double Abar = 0.0;
#pragma unroll
    for (int v = 0; v < NEQ; v++) {
      double sub = 0.0;
#pragma unroll
      for (int k = 0; k < ND; k++) sub += s_ub[team][v][k];
      if (g.limiter == RKDG_LIM_NONE || (!reset && g.limiter == RKDG_LIM_FINDIFF)) {
        mb[v] = 0.0;
      } else if (reset) {
        mb[v] = sub;
#pragma unroll
        for (int k = 0; k < ND; k++) s_ub[team][v][k] = 0.0;
      } else {
        mb[v] = (1.0 - f) * sub;
#pragma unroll
        for (int k = 0; k < ND; k++) {
          Abar += s_ub[team][v][k] * (s_u[team][v][k] - mean[v]);
          s_ub[team][v][k] *= f;
        }
      }
    }
    int64_t send_to = -1;
    double send_val = 0.0;
    if (g.limiter == RKDG_LIM_MULTI && !reset && ind_bj && pstar >= 0) {
      const double alphabar = Abar * fprime;
      double dy, dd;
      adj::limiter_phi(g, ystar, dstar, eps2, &dy, &dd);
      const double ybar = branch != 0 ? alphabar * dy : 0.0;
      double diffbar = alphabar * dd;
      if (branch != 0) {
        diffbar += ybar * (-ystar / dstar);
        const double Bbar = ybar / dstar;
        mb[0] += -Bbar;  // y = (B - mean)/diff
        send_to = branch > 0 ? amax : amin;
        send_val = Bbar;
      }
      mb[0] += -diffbar;  // diff = u(p*) - mean
      const int fc = pstar / 2, ip = pstar % 2;
#pragma unroll
      for (int k = 0; k < ND; k++) s_ub[team][0][k] += diffbar * tb.PhiI[fc][ip][k];
    }
#pragma unroll
    for (int v = 0; v < NEQ; v++) A.meanbar[v * NE + e] = mb[v];
    A.send_to[e] = send_to;
    A.send_val[e] = send_val;
  }
  __syncthreads();
  if (active && lane < ND) {
#pragma unroll
    for (int v = 0; v < NEQ; v++) A.vbar[v * NDOFS + e * ND + lane] = s_ub[team][v][lane];
  }
}

// pass B: gather the bound adjoints owed to this element, spread mean adjoints
template <int GEOM, int P>
__global__ __launch_bounds__(256) void dense_limit_adj_b_kernel(const LimAdjArgs A, const double* __restrict__ gtab) {
  using C = DC<GEOM, P>;
  constexpr int ND = C::ND, NEQ = C::NEQ, NF = C::NF;
  const DenseTab<GEOM, P>* tb = reinterpret_cast<const DenseTab<GEOM, P>*>(gtab);
  // one thread per (element, node): the element's mean adjoint is gathered by
  // each of its node threads (cache hits after the first), so the node updates
  // below are coalesced across the wave
  const int64_t gid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t e = gid / ND;
  const int k = (int)(gid - e * ND);
  if (e >= A.NE) return;
  const int64_t NE = A.NE, NDOFS = NE * ND;
  double mb[NEQ];
#pragma unroll
  for (int v = 0; v < NEQ; v++) mb[v] = A.meanbar[v * NE + e];
  if (A.send_to[e] == e) mb[0] += A.send_val[e];
  // branch-free: the indices, then every neighbour's record, in flight together
  int nbs[NF];
#pragma unroll
  for (int fc = 0; fc < NF; fc++) nbs[fc] = A.nbr[fc * NE + e];
  int64_t st[NF];
  double sv[NF];
#pragma unroll
  for (int fc = 0; fc < NF; fc++) {
    const int64_t ne = nbs[fc] >= 0 ? (int64_t)nbs[fc] : e;
    st[fc] = A.send_to[ne];
    sv[fc] = A.send_val[ne];
  }
#pragma unroll
  for (int fc = 0; fc < NF; fc++)
    if (nbs[fc] >= 0 && st[fc] == e) mb[0] += sv[fc];
  double mw[ND], ms = 0.0;
#pragma unroll
  for (int q = 0; q < ND; q++) {
    mw[q] = C::TRI ? tb->mhat[q] * A.geo[(int64_t)C::G_DET * NE + e] : A.geo[(int64_t)(C::G_MW + q) * NE + e];
    ms += mw[q];
  }
  double mwk = mw[0];
#pragma unroll
  for (int q = 1; q < ND; q++) mwk = k == q ? mw[q] : mwk;
#pragma unroll
  for (int v = 0; v < NEQ; v++) {
    const double c = mb[v] / ms;
    const int64_t i = v * NDOFS + e * ND + k;
    const double x = A.vbar[i] + c * mwk;
    A.vbar[i] = x;
    // the stage-combination adjoints fed by this vbar (the same update as
    // axpy_multi_kernel, fused here to save a launch per stage)
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (t < A.ax.n) {
        double* y = A.ax.y[t];
        y[i] = (A.ax.set[t] ? 0.0 : y[i]) + A.ax.c[t] * x;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// objective: J += coef * sum over faces of boundary slot obj_slot of
// sum_q w_q |n| p(trace); optionally xbar += coef * dJ/dx. One workgroup
// strides over the (few) elements that touch the slot and reduces in a fixed
// order, so J is the same double on every run (finite differences of J need that).
template <int GEOM, int P>
__global__ __launch_bounds__(256) void dense_objective_kernel(const ObjArgs A, const SchemeParams g,
                                                              const double* __restrict__ gtab) {
  using C = DC<GEOM, P>;
  constexpr int ND = C::ND, NEQ = C::NEQ, NF = C::NF, NQ = C::NQ;
  __shared__ double s_part[4];
  const DenseTab<GEOM, P>* tb = reinterpret_cast<const DenseTab<GEOM, P>*>(gtab);
  double jl = 0.0;
  const int64_t NE = A.NE, NDOFS = NE * ND;
  for (int t = threadIdx.x; t < A.nelems; t += blockDim.x) {
    const int64_t e = A.elems[t];
    for (int fc = 0; fc < NF; fc++) {
      if (A.nbr[fc * NE + e] != -1 - A.slot) continue;
      const double an0 = A.geo[(int64_t)(C::G_NRM + 2 * fc) * NE + e];
      const double an1 = A.geo[(int64_t)(C::G_NRM + 2 * fc + 1) * NE + e];
      const double normag = sqrt(an0 * an0 + an1 * an1);
      for (int q = 0; q < NQ; q++) {
        double tr[NEQ], dp[NEQ];
#pragma unroll
        for (int v = 0; v < NEQ; v++) {
          double t = 0.0;
#pragma unroll
          for (int k = 0; k < ND; k++) t += A.x[v * NDOFS + e * ND + k] * tb->Phi[fc][q][k];
          tr[v] = t;
        }
        const double p = adj::pressure_grad(tr, g, dp);
        const double w = A.coef * tb->wq[q] * normag;
        jl += w * p;
        if (A.xbar) {
#pragma unroll
          for (int v = 0; v < NEQ; v++)
#pragma unroll
            for (int k = 0; k < ND; k++) A.xbar[v * NDOFS + e * ND + k] += w * dp[v] * tb->Phi[fc][q][k];
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) jl += __shfl_xor(jl, off, 64);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = jl;
  __syncthreads();
  if (threadIdx.x == 0 && A.J) A.J[0] += ((s_part[0] + s_part[1]) + s_part[2]) + s_part[3];
}

// y_t = (set_t ? 0 : y_t) + c_t v for up to four targets
__global__ __launch_bounds__(256) void axpy_multi_kernel(const double* __restrict__ v, int64_t n, AxpyTargets t) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double x = v[i];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (k < t.n) {
        double* y = t.y[k];
        y[i] = (t.set[k] ? 0.0 : y[i]) + t.c[k] * x;
      }
    }
  }
}

// ---------------------------------------------------------------------------
template <int G, int PP, int RS>
hipError_t rhs_adj_rs(const RhsAdjArgs& a, const SchemeParams& g, const double* t, bool bc, hipStream_t s) {
  using C = DC<G, PP>;
  const int64_t nb = (a.NE + C::EPB - 1) / C::EPB;
  if (bc) dense_rhs_adj_kernel<G, PP, RS, true><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t);
  else dense_rhs_adj_kernel<G, PP, RS, false><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t);
  return hipGetLastError();
}

template <int G, int PP>
hipError_t rhs_adj_gp(const RhsAdjArgs& a, const SchemeParams& g, const double* t, bool bc, hipStream_t s) {
  switch (g.riemann) {
    case RKDG_RS_RUSANOV: return rhs_adj_rs<G, PP, RKDG_RS_RUSANOV>(a, g, t, bc, s);
    case RKDG_RS_LLF: return rhs_adj_rs<G, PP, RKDG_RS_LLF>(a, g, t, bc, s);
    case RKDG_RS_HLL: return rhs_adj_rs<G, PP, RKDG_RS_HLL>(a, g, t, bc, s);
    default: return rhs_adj_rs<G, PP, RKDG_RS_HLLC>(a, g, t, bc, s);
  }
}

hipError_t launch_dense_rhs_adj(int geom, int p, const RhsAdjArgs& a, const SchemeParams& g, const double* t,
                                bool has_bdr, hipStream_t s) {
#define X(G, PP) \
  if (geom == G && p == PP) return rhs_adj_gp<G, PP>(a, g, t, has_bdr, s);
  RKDG_DENSE_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_dense_limit_adj(int geom, int p, const LimAdjArgs& a, const SchemeParams& g, const double* t,
                                  hipStream_t s) {
#define X(G, PP)                                                                              \
  if (geom == G && p == PP) {                                                                 \
    using C = DC<G, PP>;                                                                      \
    const int64_t nb = (a.NE + C::EPB - 1) / C::EPB;                                          \
    dense_limit_adj_a_kernel<G, PP><<<(unsigned)nb, C::BLOCK, 0, s>>>(a, g, t);               \
    hipError_t err = hipGetLastError();                                                       \
    if (err != hipSuccess) return err;                                                        \
    dense_limit_adj_b_kernel<G, PP><<<(unsigned)((a.NE * C::ND + 255) / 256), 256, 0, s>>>(a, t); \
    return hipGetLastError();                                                                 \
  }
  RKDG_DENSE_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_dense_objective(int geom, int p, const ObjArgs& a, const SchemeParams& g, const double* t,
                                  hipStream_t s) {
#define X(G, PP)                                                                                 \
  if (geom == G && p == PP) {                                                                    \
    dense_objective_kernel<G, PP><<<1, 256, 0, s>>>(a, g, t);                                    \
    return hipGetLastError();                                                                    \
  }
  RKDG_DENSE_CASES(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_axpy_multi(const double* v, int64_t n, const AxpyTargets& t, hipStream_t s) {
  int64_t grid = (n + 255) / 256;
  if (grid > 8192) grid = 8192;
  axpy_multi_kernel<<<(unsigned)grid, 256, 0, s>>>(v, n, t);
  return hipGetLastError();
}

}  // namespace rkdg
