// rkdg_adjoint.hpp — hand-written reverse-mode derivatives (vector-Jacobian
// products) of the 2D physics in rkdg_physics.hpp, for the discrete adjoint of
// the RKDG step. Each *_vjp function takes the forward inputs and the adjoint of
// the forward outputs and ADDS the adjoint of the inputs. Branches follow the
// forward evaluation (the derivative of the active branch; at min/max and
// threshold switches the branch the forward took is differentiated).
//
// Forward counterparts: pressure/sound speed physics.cpp:101-140, flux
// physics.cpp:143-231, Toro/Einfeldt wave speeds physics.cpp:269-350, Riemann
// solvers rs_rusanov.cpp:6-40, rs_llf.cpp:6-37, rs_hll.cpp:6-50,
// rs_hllc.cpp:71-161, ghost states boundary_integrator_slip.cpp:7-17,
// boundary_integrator_open.cpp:9-12, boundary_integrator_supersonic_inlet.cpp:8-19,
// limiter functions indicator_bj.cpp:49-79.
#pragma once

#include "rkdg_physics.hpp"

namespace rkdg {
namespace adj {

constexpr int NE2 = 4;  // 2D conserved variables (rho, m_x, m_y, E)

// per-side adjoint accumulators: state, and the intermediate quantities whose
// adjoints are converted to state adjoints at the end (finish_side)
struct SideBar {
  double s[NE2];
  double p;     // pressure
  double c;     // sound speed
  double inv;   // 1/rho
};

RKDG_HD void zero(SideBar& b) {
  for (int i = 0; i < NE2; i++) b.s[i] = 0.0;
  b.p = b.c = b.inv = 0.0;
}

// pressure p = (g-1)(E - |m|^2/(2 rho)) [/(1 - b rho)], sound c = sqrt(g p / rho [/(1 - b rho)]);
// converts the p, c, inv adjoints into state adjoints.
RKDG_HD void finish_side(const double* s, const SchemeParams& g, SideBar& b) {
  const double rho = s[0];
  const double inv = 1.0 / rho;
  const double m2 = s[1] * s[1] + s[2] * s[2];
  const double cg = g.gamma - 1.0;
  const double pid = cg * (s[3] - 0.5 * m2 * inv);
  const double f = g.covolume != 0.0 ? 1.0 / (1.0 - rho * g.covolume) : 1.0;
  const double p = pid * f;
  if (b.c != 0.0) {
    const double a = g.gamma * p * inv * f;
    const double c = sqrt(a);
    const double abar = b.c / (2.0 * c);
    b.p += abar * g.gamma * inv * f;
    b.inv += abar * g.gamma * p * f;
    if (g.covolume != 0.0) b.s[0] += abar * g.gamma * p * inv * g.covolume * f * f;
  }
  if (b.p != 0.0) {
    const double pidbar = b.p * f;
    if (g.covolume != 0.0) b.s[0] += b.p * pid * g.covolume * f * f;
    b.s[3] += cg * pidbar;
    b.s[1] += -cg * pidbar * s[1] * inv;
    b.s[2] += -cg * pidbar * s[2] * inv;
    b.inv += -0.5 * cg * pidbar * m2;
  }
  b.s[0] += -b.inv * inv * inv;
  b.p = b.c = b.inv = 0.0;
}

// u = s[1]/s[0] (a velocity formed by division)
RKDG_HD void velocity_vjp(const double* s, double ubar, SideBar& b) {
  b.s[1] += ubar / s[0];
  b.s[0] += -ubar * s[1] / (s[0] * s[0]);
}

// contravariant volume flux G = a . F(u) at a node (the nodal flux of the
// volume term): G0 = m.a, G_{1+d} = (m.a) u_d + p a_d, G3 = (m.a) H.
RKDG_HD void volume_flux_vjp(const double* s, const double* a, const double* Gbar, const SchemeParams& g,
                             double* sbar) {
  SideBar b;
  zero(b);
  const double inv = 1.0 / s[0];
  const double m2 = s[1] * s[1] + s[2] * s[2];
  const double cg = g.gamma - 1.0;
  const double f = g.covolume != 0.0 ? 1.0 / (1.0 - s[0] * g.covolume) : 1.0;
  const double p = cg * (s[3] - 0.5 * m2 * inv) * f;
  const double ma = s[1] * a[0] + s[2] * a[1];
  const double H = (s[3] + p) * inv;
  double mabar = Gbar[0];
  for (int d = 0; d < 2; d++) {
    const double ud = s[1 + d] * inv;
    mabar += Gbar[1 + d] * ud;
    const double udbar = Gbar[1 + d] * ma;
    b.s[1 + d] += udbar * inv;
    b.inv += udbar * s[1 + d];
    b.p += Gbar[1 + d] * a[d];
  }
  mabar += Gbar[3] * H;
  const double Hbar = Gbar[3] * ma;
  b.s[3] += Hbar * inv;
  b.p += Hbar * inv;
  b.inv += Hbar * (s[3] + p);
  b.s[1] += mabar * a[0];
  b.s[2] += mabar * a[1];
  finish_side(s, g, b);
  for (int i = 0; i < NE2; i++) sbar[i] += b.s[i];
}

// x-flux of a rotated state f = (m, u m + p, u m_y, u (E + p)), u = m * (1/rho)
RKDG_HD void xflux_vjp(const double* s, double p, const double* fbar, SideBar& b) {
  const double inv = 1.0 / s[0];
  const double u = s[1] * inv;
  b.s[1] += fbar[0];
  const double ubar = fbar[1] * s[1] + fbar[2] * s[2] + fbar[3] * (s[3] + p);
  b.s[1] += fbar[1] * u;
  b.s[2] += fbar[2] * u;
  b.s[3] += fbar[3] * u;
  b.p += fbar[1] + fbar[3] * u;
  b.s[1] += ubar * inv;
  b.inv += ubar * s[1];
}

RKDG_HD double side_pressure(const double* s, const SchemeParams& g) {
  const double inv = 1.0 / s[0];
  const double m2 = s[1] * s[1] + s[2] * s[2];
  double p = (g.gamma - 1.0) * (s[3] - 0.5 * m2 * inv);
  if (g.covolume != 0.0) p /= (1.0 - s[0] * g.covolume);
  return p;
}
RKDG_HD double side_sound_speed(const double* s, double p, const SchemeParams& g) {
  double a = g.gamma * p / s[0];
  if (g.covolume != 0.0) a /= (1.0 - s[0] * g.covolume);
  return sqrt(a);
}

// Toro two-rarefaction (or Einfeldt) speeds of rotated states: adjoint of (SL, SR)
RKDG_HD void wave_speeds_vjp(const double* s1, const double* s2, double p1, double p2, const SchemeParams& g,
                             double SLbar, double SRbar, SideBar& b1, SideBar& b2) {
  const double cL = side_sound_speed(s1, p1, g), cR = side_sound_speed(s2, p2, g);
  const double uL = s1[1] / s1[0], uR = s2[1] / s2[0];
  double uLbar = 0.0, uRbar = 0.0;
  if (g.wave_speeds == RKDG_WS_EINFELDT) {
    const double sL = sqrt(s1[0]), sR = sqrt(s2[0]);
    const double sum = sL + sR;
    const double eta2 = 0.5 * sL * sR / (sum * sum);
    const double B = s1[1] / sL + s2[1] / sR;
    const double du = uR - uL;
    const double A = sL * (cL * cL) + sR * (cR * cR);
    const double Q = A / sum + eta2 * (du * du);
    const double cav = sqrt(Q);
    const double uavbar = SLbar + SRbar;
    const double cavbar = -SLbar + SRbar;
    const double Qbar = cavbar / (2.0 * cav);
    const double Abar = Qbar / sum;
    double sumbar = -Qbar * A / (sum * sum);
    const double eta2bar = Qbar * du * du;
    const double dubar = Qbar * eta2 * 2.0 * du;
    double sLbar = Abar * cL * cL, sRbar = Abar * cR * cR;
    b1.c += Abar * sL * 2.0 * cL;
    b2.c += Abar * sR * 2.0 * cR;
    uRbar += dubar;
    uLbar -= dubar;
    sLbar += eta2bar * 0.5 * sR / (sum * sum);
    sRbar += eta2bar * 0.5 * sL / (sum * sum);
    sumbar += -2.0 * eta2 / sum * eta2bar;
    const double Bbar = uavbar / sum;
    sumbar += -uavbar * B / (sum * sum);
    b1.s[1] += Bbar / sL;
    sLbar += -Bbar * s1[1] / (sL * sL);
    b2.s[1] += Bbar / sR;
    sRbar += -Bbar * s2[1] / (sR * sR);
    sLbar += sumbar;
    sRbar += sumbar;
    b1.s[0] += sLbar / (2.0 * sL);
    b2.s[0] += sRbar / (2.0 * sR);
  } else {
    const double gm = g.gamma;
    const double z = 0.5 * (gm - 1.0) / gm;
    const double pLz = pow(p1, z), pRz = pow(p2, z);
    const double num = cL + cR - 0.5 * (gm - 1.0) * (uR - uL);
    const double den = cL / pLz + cR / pRz;
    const double r = num / den;
    const double pStar = pow(r, 1.0 / z);
    const double kap = 0.5 * (gm + 1.0) / gm;
    const double qL = pStar > p1 ? sqrt(1.0 + kap * (pStar / p1 - 1.0)) : 1.0;
    const double qR = pStar > p2 ? sqrt(1.0 + kap * (pStar / p2 - 1.0)) : 1.0;
    uLbar += SLbar;
    const double qLbar = -SLbar * cL;
    b1.c += -SLbar * qL;
    uRbar += SRbar;
    const double qRbar = SRbar * cR;
    b2.c += SRbar * qR;
    double pStarbar = 0.0;
    if (pStar > p1) {
      const double t = qLbar * kap / (2.0 * qL);
      pStarbar += t / p1;
      b1.p += -t * pStar / (p1 * p1);
    }
    if (pStar > p2) {
      const double t = qRbar * kap / (2.0 * qR);
      pStarbar += t / p2;
      b2.p += -t * pStar / (p2 * p2);
    }
    const double rbar = pStarbar * pStar / (z * r);
    const double numbar = rbar / den;
    const double denbar = -rbar * num / (den * den);
    b1.c += numbar;
    b2.c += numbar;
    uRbar += -0.5 * (gm - 1.0) * numbar;
    uLbar += 0.5 * (gm - 1.0) * numbar;
    b1.c += denbar / pLz;
    b2.c += denbar / pRz;
    b1.p += denbar * cL * (-z) / (pLz * p1);
    b2.p += denbar * cR * (-z) / (pRz * p2);
  }
  velocity_vjp(s1, uLbar, b1);
  velocity_vjp(s2, uRbar, b2);
}

// 2D rotation by the unit normal (rs_basic.cpp:11-38): the momentum of a state
RKDG_HD void rot2(double* v, const double* n) {
  const double a = n[0] * v[1] + n[1] * v[2];
  const double c = -n[1] * v[1] + n[0] * v[2];
  v[1] = a;
  v[2] = c;
}
RKDG_HD void irot2(double* v, const double* n) {
  const double a = n[0] * v[1] - n[1] * v[2];
  const double c = n[1] * v[1] + n[0] * v[2];
  v[1] = a;
  v[2] = c;
}

// adjoint of flux = riemann(uL, uR, n) (n held fixed): adds to uLbar, uRbar
template <int RS>
RKDG_HD void riemann_vjp(const SchemeParams& g, const double* uL, const double* uR, const double* n,
                         const double* Fbar, double* uLbar, double* uRbar) {
  SideBar b1, b2;
  zero(b1);
  zero(b2);
  if constexpr (RS == RKDG_RS_RUSANOV) {
    const double* s1 = uL;
    const double* s2 = uR;
    const double p1 = side_pressure(s1, g), p2 = side_pressure(s2, g);
    const double c1 = side_sound_speed(s1, p1, g), c2 = side_sound_speed(s2, p2, g);
    const double inv1 = 1.0 / s1[0], inv2 = 1.0 / s2[0];
    const double q1 = s1[1] * s1[1] + s1[2] * s1[2], q2 = s2[1] * s2[1] + s2[2] * s2[2];
    const double sq1 = sqrt(q1), sq2 = sqrt(q2);
    const double m1 = sq1 * inv1 + c1, m2 = sq2 * inv2 + c2;
    const double mE = m1 > m2 ? m1 : m2;
    const double nm = n[0] * n[0] + n[1] * n[1];
    const double inm = 0.5 / sqrt(nm);
    double hmbar = 0.0;
    double f1bar[NE2], f2bar[NE2];
    for (int i = 0; i < NE2; i++) {
      f1bar[i] = Fbar[i] * inm;
      f2bar[i] = Fbar[i] * inm;
      hmbar += -Fbar[i] * (s2[i] - s1[i]);
      b2.s[i] += -0.5 * mE * Fbar[i];
      b1.s[i] += 0.5 * mE * Fbar[i];
    }
    const double mEbar = 0.5 * hmbar;
    // max(|u| + c) of the side that won
    SideBar& bw = m1 > m2 ? b1 : b2;
    const double* sw = m1 > m2 ? s1 : s2;
    const double sqw = m1 > m2 ? sq1 : sq2;
    const double invw = m1 > m2 ? inv1 : inv2;
    if (sqw > 0.0) {
      bw.s[1] += mEbar * invw * sw[1] / sqw;
      bw.s[2] += mEbar * invw * sw[2] / sqw;
    }
    bw.inv += mEbar * sqw;
    bw.c += mEbar;
    // F(u).n per side
    for (int k = 0; k < 2; k++) {
      const double* s = k ? s2 : s1;
      const double* fb = k ? f2bar : f1bar;
      SideBar& b = k ? b2 : b1;
      const double p = k ? p2 : p1;
      const double inv = k ? inv2 : inv1;
      const double vn = s[1] * n[0] + s[2] * n[1];
      const double a = vn * inv;
      double vnbar = fb[0];
      const double abar = fb[1] * s[1] + fb[2] * s[2] + fb[3] * (s[3] + p);
      b.s[1] += fb[1] * a;
      b.s[2] += fb[2] * a;
      b.p += fb[1] * n[0] + fb[2] * n[1] + fb[3] * a;
      b.s[3] += fb[3] * a;
      vnbar += abar * inv;
      b.inv += abar * vn;
      b.s[1] += vnbar * n[0];
      b.s[2] += vnbar * n[1];
    }
    finish_side(s1, g, b1);
    finish_side(s2, g, b2);
    for (int i = 0; i < NE2; i++) {
      uLbar[i] += b1.s[i];
      uRbar[i] += b2.s[i];
    }
    return;
  } else {
    double s1[NE2], s2[NE2], Fr[NE2];
    for (int i = 0; i < NE2; i++) {
      s1[i] = uL[i];
      s2[i] = uR[i];
      Fr[i] = Fbar[i];
    }
    rot2(s1, n);
    rot2(s2, n);
    rot2(Fr, n);  // adjoint of the inverse rotation of the flux
    const double p1 = side_pressure(s1, g), p2 = side_pressure(s2, g);
    double f1[NE2], f2[NE2];
    {
      const double u1 = s1[1] / s1[0], u2 = s2[1] / s2[0];
      f1[0] = s1[1];
      f2[0] = s2[1];
      f1[1] = u1 * s1[1] + p1;
      f2[1] = u2 * s2[1] + p2;
      f1[2] = u1 * s1[2];
      f2[2] = u2 * s2[2];
      f1[3] = u1 * (s1[3] + p1);
      f2[3] = u2 * (s2[3] + p2);
    }
    double SL, SR;
    wave_speeds<2>(s1, s2, p1, p2, g, SL, SR);
    double f1bar[NE2] = {0, 0, 0, 0}, f2bar[NE2] = {0, 0, 0, 0};
    double SLbar = 0.0, SRbar = 0.0;
    if constexpr (RS == RKDG_RS_LLF) {
      const double mE = fmax(fabs(SL), fabs(SR));
      double mEbar = 0.0;
      for (int i = 0; i < NE2; i++) {
        f1bar[i] = 0.5 * Fr[i];
        f2bar[i] = 0.5 * Fr[i];
        mEbar += -0.5 * Fr[i] * (s2[i] - s1[i]);
        b2.s[i] += -0.5 * mE * Fr[i];
        b1.s[i] += 0.5 * mE * Fr[i];
      }
      if (fabs(SL) >= fabs(SR)) SLbar += mEbar * (SL >= 0.0 ? 1.0 : -1.0);
      else SRbar += mEbar * (SR >= 0.0 ? 1.0 : -1.0);
    } else if constexpr (RS == RKDG_RS_HLL) {
      if (SL >= 0) {
        for (int i = 0; i < NE2; i++) f1bar[i] = Fr[i];
      } else if (SR <= 0) {
        for (int i = 0; i < NE2; i++) f2bar[i] = Fr[i];
      } else {
        const double inv = 1.0 / (SR - SL);
        double invbar = 0.0;
        for (int i = 0; i < NE2; i++) {
          const double d = s2[i] - s1[i];
          const double N = SR * f1[i] - SL * f2[i] + SL * SR * d;
          const double Nbar = Fr[i] * inv;
          invbar += Fr[i] * N;
          SRbar += Nbar * (f1[i] + SL * d);
          SLbar += Nbar * (-f2[i] + SR * d);
          f1bar[i] = SR * Nbar;
          f2bar[i] = -SL * Nbar;
          b2.s[i] += SL * SR * Nbar;
          b1.s[i] -= SL * SR * Nbar;
        }
        SRbar += -invbar * inv * inv;
        SLbar += invbar * inv * inv;
      }
    } else {  // HLLC
      if (SL >= 0) {
        for (int i = 0; i < NE2; i++) f1bar[i] = Fr[i];
      } else if (SR <= 0) {
        for (int i = 0; i < NE2; i++) f2bar[i] = Fr[i];
      } else {
        const double gL = SL - s1[1] / s1[0];
        const double gR = SR - s2[1] / s2[0];
        const double num = p2 - p1 + s1[1] * gL - s2[1] * gR;
        const double den = s1[0] * gL - s2[0] * gR;
        const double sStar = num / den;
        const bool left = sStar > 0.0;
        const double* st = left ? s1 : s2;
        const double* fK = left ? f1 : f2;
        double* fKbar = left ? f1bar : f2bar;
        SideBar& bK = left ? b1 : b2;
        const double pK = left ? p1 : p2;
        const double SK = left ? SL : SR;
        const double gK = left ? gL : gR;
        const double Wv = sStar - st[1] / st[0];
        const double inner = pK + st[0] * gK * Wv;
        const double pterm = SK * inner;
        const double inv = 1.0 / (SK - sStar);
        double invbar = 0.0, sStarbar = 0.0, SKbar = 0.0, ptermbar = 0.0;
        for (int i = 0; i < NE2; i++) {
          const double D = i == 1 ? 1.0 : (i == 3 ? sStar : 0.0);
          const double Q = sStar * (st[i] * SK - fK[i]) + D * pterm;
          const double Qbar = Fr[i] * inv;
          invbar += Fr[i] * Q;
          sStarbar += Qbar * (st[i] * SK - fK[i]);
          bK.s[i] += Qbar * sStar * SK;
          SKbar += Qbar * sStar * st[i];
          fKbar[i] += -Qbar * sStar;
          ptermbar += D * Qbar;
          if (i == 3) sStarbar += Qbar * pterm;
        }
        SKbar += -invbar * inv * inv;
        sStarbar += invbar * inv * inv;
        SKbar += ptermbar * inner;
        const double innerbar = ptermbar * SK;
        bK.p += innerbar;
        bK.s[0] += innerbar * gK * Wv;
        double gKbar = innerbar * st[0] * Wv;
        const double Wvbar = innerbar * st[0] * gK;
        sStarbar += Wvbar;
        bK.s[1] += -Wvbar / st[0];
        bK.s[0] += Wvbar * st[1] / (st[0] * st[0]);
        const double numbar = sStarbar / den;
        const double denbar = -sStarbar * num / (den * den);
        b2.p += numbar;
        b1.p -= numbar;
        b1.s[1] += numbar * gL;
        double gLbar = numbar * s1[1];
        b2.s[1] -= numbar * gR;
        double gRbar = -numbar * s2[1];
        b1.s[0] += denbar * gL;
        gLbar += denbar * s1[0];
        b2.s[0] -= denbar * gR;
        gRbar -= denbar * s2[0];
        if (left) {
          gLbar += gKbar;
          SLbar += SKbar;
        } else {
          gRbar += gKbar;
          SRbar += SKbar;
        }
        gKbar = 0.0;
        SLbar += gLbar;
        b1.s[1] -= gLbar / s1[0];
        b1.s[0] += gLbar * s1[1] / (s1[0] * s1[0]);
        SRbar += gRbar;
        b2.s[1] -= gRbar / s2[0];
        b2.s[0] += gRbar * s2[1] / (s2[0] * s2[0]);
      }
    }
    if (SLbar != 0.0 || SRbar != 0.0) wave_speeds_vjp(s1, s2, p1, p2, g, SLbar, SRbar, b1, b2);
    xflux_vjp(s1, p1, f1bar, b1);
    xflux_vjp(s2, p2, f2bar, b2);
    finish_side(s1, g, b1);
    finish_side(s2, g, b2);
    irot2(b1.s, n);  // adjoint of the rotation of the inputs
    irot2(b2.s, n);
    for (int i = 0; i < NE2; i++) {
      uLbar[i] += b1.s[i];
      uRbar[i] += b2.s[i];
    }
  }
}

RKDG_HD void riemann_vjp_rt(const SchemeParams& g, const double* uL, const double* uR, const double* n,
                            const double* Fbar, double* uLbar, double* uRbar) {
  switch (g.riemann) {
    case RKDG_RS_RUSANOV: riemann_vjp<RKDG_RS_RUSANOV>(g, uL, uR, n, Fbar, uLbar, uRbar); break;
    case RKDG_RS_LLF: riemann_vjp<RKDG_RS_LLF>(g, uL, uR, n, Fbar, uLbar, uRbar); break;
    case RKDG_RS_HLL: riemann_vjp<RKDG_RS_HLL>(g, uL, uR, n, Fbar, uLbar, uRbar); break;
    default: riemann_vjp<RKDG_RS_HLLC>(g, uL, uR, n, Fbar, uLbar, uRbar); break;
  }
}

// boundary types whose ghost state has an adjoint here
RKDG_HD bool bc_has_vjp(int type) {
  return type == RKDG_BC_SLIP || type == RKDG_BC_OUTLET || type == RKDG_BC_SUPERSONIC_INLET;
}

// adjoint of s2 = bc_state(s1): adds to s1bar and to the parameter adjoint Pbar[8]
RKDG_HD void bc_state_vjp(const SchemeParams& g, int bi, const double* s1, const double* n, const double* s2bar,
                          double* s1bar, double* Pbar, const double* Pover = nullptr) {
  const double* P = Pover ? Pover : g.bc_param[bi];
  switch (g.bc_type[bi]) {
    case RKDG_BC_SLIP: {
      double t[NE2] = {s2bar[0], s2bar[1], s2bar[2], s2bar[3]};
      rot2(t, n);
      t[1] = -t[1];
      irot2(t, n);
      for (int i = 0; i < NE2; i++) s1bar[i] += t[i];
      break;
    }
    case RKDG_BC_OUTLET:
      for (int i = 0; i < NE2; i++) s1bar[i] += s2bar[i];
      break;
    case RKDG_BC_SUPERSONIC_INLET: {
      const double rho = P[0], u = P[1], v = P[2];
      if (Pbar) {
        Pbar[0] += s2bar[0] + s2bar[1] * u + s2bar[2] * v + s2bar[3] * 0.5 * (u * u + v * v);
        Pbar[1] += s2bar[1] * rho + s2bar[3] * rho * u;
        Pbar[2] += s2bar[2] * rho + s2bar[3] * rho * v;
        Pbar[4] += s2bar[3] / (g.gamma - 1.0);
      }
      break;
    }
    default:
      break;
  }
}

// pressure of a state and its gradient (the objective integrand)
RKDG_HD double pressure_grad(const double* s, const SchemeParams& g, double* dp) {
  SideBar b;
  zero(b);
  b.p = 1.0;
  finish_side(s, g, b);
  for (int i = 0; i < NE2; i++) dp[i] = b.s[i];
  return side_pressure(s, g);
}

// limiter function phi(y) of the BJ-family indicator (indicator_bj.cpp:49-79)
// and its partial derivatives in y and in diff (the Venkatakrishnan form with
// eps^2 = (K h)^3 depends on diff^2)
RKDG_HD double limiter_phi(const SchemeParams& g, double y, double diff, double eps2, double* dy, double* dd) {
  *dy = 0.0;
  *dd = 0.0;
  if (g.phi == RKDG_PHI_BJ) {
    if (y < 1.0) {
      *dy = 1.0;
      return y;
    }
    return 1.0;
  }
  if (g.phi == RKDG_PHI_VENKAT) {
    if (g.venkat_k > 0.0) {
      const double d2 = diff * diff;
      const double Nv = d2 * (y * y + 2.0 * y) + eps2;
      const double Dv = d2 * (y * y + y + 2.0) + eps2;
      *dy = d2 * ((2.0 * y + 2.0) * Dv - (2.0 * y + 1.0) * Nv) / (Dv * Dv);
      *dd = 2.0 * diff * ((y * y + 2.0 * y) * Dv - (y * y + y + 2.0) * Nv) / (Dv * Dv);
      return Nv / Dv;
    }
    const double Nv = y * y + 2.0 * y, Dv = y * y + y + 2.0;
    *dy = ((2.0 * y + 2.0) * Dv - (2.0 * y + 1.0) * Nv) / (Dv * Dv);
    return Nv / Dv;
  }
  const double yStar = 1.5;
  if (y < 1.0) {
    const double yRel = y / yStar;
    *dy = 1.0 + 2.0 * (3.0 - 2.0 * yStar) * yRel / yStar + 3.0 * (yStar - 2.0) * yRel * yRel / yStar;
    return y + (3.0 - 2.0 * yStar) * yRel * yRel + (yStar - 2.0) * yRel * yRel * yRel;
  }
  return 1.0;
}

}  // namespace adj
}  // namespace rkdg
