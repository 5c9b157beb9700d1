// rkdg_physics.hpp — device (and host) physics for the RKDG kernels:
// EoS, physical fluxes, wave-speed estimates, Riemann solvers and ghost states.
// Formulas follow src/physics/physics.cpp, src/riemannsolvers/rs_*.cpp and
// src/bc/boundary_integrator_*.cpp of the reference; each function cites its line.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rkdg_capi.h"

#define RKDG_HD __host__ __device__ __forceinline__

namespace rkdg {

struct SchemeParams {
  int riemann, wave_speeds, hllc_abs;
  double gamma, gas_constant, covolume;
  int num_bc;
  int bc_attr[RKDG_MAX_BC];
  int bc_type[RKDG_MAX_BC];
  double bc_param[RKDG_MAX_BC][8];
  int indicator, phi;
  double venkat_k;
  int limiter, last_hope, pow_nd;
  int linearize;  // multiplier: linear reconstruction first (p = 1 tensor elements)
};

// 1/x on the device: v_rcp_f64 refined by two Newton steps (correct to about
// one ulp). Per side, 1/rho is formed once and reused where the reference
// divides by rho in each formula; the differences are rounding-level.
RKDG_HD double recip(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  r = fma(fma(-x, r, 1.0), r, r);
  return r;
#else
  return 1.0 / x;
#endif
}

// Square root on the device (wave speeds, sound speeds, boundary states):
// v_rsq_f64 refined by one Goldschmidt step (g = x r, h = r/2, e = 1/2 - h g)
// and one Newton correction of g, about half the instructions of the correctly
// rounded sqrt and within an ulp of it; negative input gives NaN as sqrt does.
// The seed is rsq(x + 2^-1022): for x >= 2^-969 the sum is x itself, and at
// x = 0 the seed stays finite so every step below yields exactly 0 (no
// select on x == 0). Positive x below 2^-969 gets a less accurate result. The
// host build uses the library sqrt.
RKDG_HD double sqrt_ws(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  const double r = __builtin_amdgcn_rsq(x + 0x1p-1022);
  double gq = x * r, h = 0.5 * r;
  const double e = fma(-h, gq, 0.5);
  gq = fma(gq, e, gq);
  h = fma(h, e, h);
  const double d = fma(-gq, gq, x);
  gq = fma(d, h, gq);
  return gq;
#else
  return sqrt(x);
#endif
}

// x^y for x > 0 on the device (the two-rarefaction pressure estimate, the
// isentropic boundary states):
// log x from frexp and the series log m = 2 atanh(s), s = (m-1)/(m+1), with m
// in [1/sqrt2, sqrt2) (|s| < 0.172, terms to s^21), ln 2 split in a high part
// with trailing zero bits and a low part; exp t by t = k ln2 + r, |r| <= ln2/2,
// and the Taylor polynomial of degree 13. Relative error within ~1e-14 of the
// library pow (the error of y log x grows with |y log x|); a third of its
// instructions. The host build uses the library pow.
RKDG_HD double pow_ws(double x, double y) {
#ifdef __HIP_DEVICE_COMPILE__
  int ex;
  double m = frexp(x, &ex);  // x = m 2^ex, m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    ex -= 1;
  }
  const double s = (m - 1.0) * recip(m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, s2, 1.0 / 19.0);
  p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  const double lm = 2.0 * s + 2.0 * s * s2 * p;  // log m
  const double LN2_HI = 6.93147180369123816490e-01, LN2_LO = 1.90821492927058770002e-10;
  const double lx = fma((double)ex, LN2_HI, fma((double)ex, LN2_LO, lm));
  const double t = y * lx;
  const double k = rint(t * 1.44269504088896340736);  // t / ln 2
  const double r = fma(-k, LN2_LO, fma(-k, LN2_HI, t));
  double q = 1.0 / 6227020800.0;  // 1/13!
  q = fma(q, r, 1.0 / 479001600.0);
  q = fma(q, r, 1.0 / 39916800.0);
  q = fma(q, r, 1.0 / 3628800.0);
  q = fma(q, r, 1.0 / 362880.0);
  q = fma(q, r, 1.0 / 40320.0);
  q = fma(q, r, 1.0 / 5040.0);
  q = fma(q, r, 1.0 / 720.0);
  q = fma(q, r, 1.0 / 120.0);
  q = fma(q, r, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  q = fma(q, r, 1.0);
  q = fma(q, r, 1.0);
  return ldexp(q, (int)k);
#else
  return pow(x, y);
#endif
}

// physics.cpp:101-112 (covolume EoS; b == 0 skips the exact division by 1)
template <int DIM>
RKDG_HD double pressure(const double* s, const SchemeParams& g) {
  double v2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; d++) v2 += s[1 + d] * s[1 + d];
  v2 /= s[0];
  double p = (g.gamma - 1.0) * (s[1 + DIM] - 0.5 * v2);
  if (g.covolume != 0.0) p /= (1.0 - s[0] * g.covolume);
  return p;
}

// physics.cpp:134-140
template <int DIM>
RKDG_HD double sound_speed_p(const double* s, double p, const SchemeParams& g) {
  double a = g.gamma * p / s[0];
  if (g.covolume != 0.0) a /= (1.0 - s[0] * g.covolume);
  return sqrt_ws(a);
}

// physics.cpp:53-97 (pressure check without covolume, :84)
template <int DIM>
RKDG_HD bool physical_say(const double* s, const SchemeParams& g) {
  double den = s[0], E = s[1 + DIM];
  if (!(den >= 0.0)) return false;
  if (!(E > 0.0)) return false;
  double v2 = 0.0;
#pragma unroll
  for (int i = 0; i < DIM; i++) v2 += s[1 + i] * s[1 + i];
  v2 /= den;
  double p = (g.gamma - 1.0) * (E - 0.5 * v2);
  return p > 0.0;
}

// physics.cpp:6-50
template <int DIM>
RKDG_HD bool physical(const double* s, const SchemeParams& g) {
  double den = s[0], E = s[1 + DIM];
  if (!(den >= 0.0)) return false;
  if (!(E > 0.0)) return false;
  double p = pressure<DIM>(s, g);
  return p > 0.0;
}

// physics.cpp:252-266: |u| + c
template <int DIM>
RKDG_HD double max_char_speed(const double* s, const SchemeParams& g) {
  double v2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; d++) v2 += s[1 + d] * s[1 + d];
  v2 /= s[0];
  double p = (g.gamma - 1.0) * (s[1 + DIM] - 0.5 * v2);
  if (g.covolume != 0.0) p /= (1.0 - s[0] * g.covolume);
  double c = sound_speed_p<DIM>(s, p, g);
  return sqrt_ws(v2 / s[0]) + c;
}

// physics.cpp:143-168; f[k*DIM + d]
template <int DIM>
RKDG_HD void flux_tensor(const double* s, const SchemeParams& g, double* f) {
  const double den = s[0];
  const double p = pressure<DIM>(s, g);
  const double inv = 1.0 / den;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    f[0 * DIM + d] = s[1 + d];
#pragma unroll
    for (int i = 0; i < DIM; i++) f[(1 + i) * DIM + d] = s[1 + i] * s[1 + d] * inv;
    f[(1 + d) * DIM + d] += p;
  }
  const double H = (s[1 + DIM] + p) * inv;
#pragma unroll
  for (int d = 0; d < DIM; d++) f[(1 + DIM) * DIM + d] = s[1 + d] * H;
}

// physics.cpp:171-195
template <int DIM>
RKDG_HD void flux_dot_n(const double* s, const double* n, double p, const SchemeParams& g, double* fn) {
  const double inv = 1.0 / s[0];
  double vn = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; d++) vn += s[1 + d] * n[d];
  fn[0] = vn;
#pragma unroll
  for (int d = 0; d < DIM; d++) fn[1 + d] = vn * s[1 + d] * inv + p * n[d];
  fn[1 + DIM] = vn * ((s[1 + DIM] + p) * inv);
}

// physics.cpp:198-231 (x-flux of a rotated state)
template <int DIM>
RKDG_HD void flux_f(const double* s, double p, double* f) {
  const double inv = 1.0 / s[0];
  f[0] = s[1];
  f[1] = s[1] * s[1] * inv + p;
#pragma unroll
  for (int d = 1; d < DIM; d++) f[d + 1] = s[1] * s[d + 1] * inv;
  f[1 + DIM] = s[1] * ((s[1 + DIM] + p) * inv);
}

// rs_basic.cpp:11-38
template <int DIM>
RKDG_HD void rotate(double* s, const double* n) {
  double* m = s + 1;
  if (DIM == 2) {
    double mx = n[0] * m[0] + n[1] * m[1];
    double my = -n[1] * m[0] + n[0] * m[1];
    m[0] = mx;
    m[1] = my;
  } else {
    const double sp = n[2];
    const double cp = sqrt_ws(1 - sp * sp);
    const double st = fabs(cp) > 1e-6 ? n[1] / cp : 1.0;
    const double ct = fabs(cp) > 1e-6 ? n[0] / cp : 0.0;
    double mx = cp * ct * m[0] + cp * st * m[1] + sp * m[2];
    double my = -st * m[0] + ct * m[1];
    double mz = -sp * ct * m[0] - sp * st * m[1] + cp * m[2];
    m[0] = mx;
    m[1] = my;
    m[2] = mz;
  }
}

// rs_basic.cpp:40-67
template <int DIM>
RKDG_HD void inverse_rotate(double* s, const double* n) {
  double* m = s + 1;
  if (DIM == 2) {
    double mx = n[0] * m[0] - n[1] * m[1];
    double my = n[1] * m[0] + n[0] * m[1];
    m[0] = mx;
    m[1] = my;
  } else {
    const double sp = n[2];
    const double cp = sqrt_ws(1 - sp * sp);
    const double st = fabs(cp) > 1e-6 ? n[1] / cp : 1.0;
    const double ct = fabs(cp) > 1e-6 ? n[0] / cp : 0.0;
    double mx = cp * ct * m[0] - st * m[1] - sp * ct * m[2];
    double my = cp * st * m[0] + ct * m[1] - sp * st * m[2];
    double mz = sp * m[0] + cp * m[2];
    m[0] = mx;
    m[1] = my;
    m[2] = mz;
  }
}

// Toro two-rarefaction speeds (physics.cpp:309-350) or Einfeldt (:269-296) on
// rotated states; lo/hi = lambdaF[0], lambdaF[dim+1].
template <int DIM>
RKDG_HD void wave_speeds(const double* s1, const double* s2, double p1, double p2,
                         const SchemeParams& g, double& lo, double& hi) {
  const double cL = sound_speed_p<DIM>(s1, p1, g), cR = sound_speed_p<DIM>(s2, p2, g);
  const double uL = s1[1] / s1[0], uR = s2[1] / s2[0];
  if (g.wave_speeds == RKDG_WS_EINFELDT) {
    const double sL = sqrt_ws(s1[0]), sR = sqrt_ws(s2[0]);
    const double sum = sL + sR;
    const double eta2 = 0.5 * sL * sR / (sum * sum);
    const double uav = (s1[1] / sL + s2[1] / sR) / sum;
    const double cav = sqrt_ws((sL * (cL * cL) + sR * (cR * cR)) / sum + eta2 * ((uR - uL) * (uR - uL)));
    lo = uav - cav;
    hi = uav + cav;
    return;
  }
  const double gm = g.gamma;
  const double z = 0.5 * (gm - 1.0) / gm;
  const double pStar = pow_ws((cL + cR - 0.5 * (gm - 1.0) * (uR - uL)) / (cL / pow_ws(p1, z) + cR / pow_ws(p2, z)),
                              1.0 / z);
  const double qL = pStar > p1 ? sqrt_ws(1.0 + 0.5 * (gm + 1.0) * (pStar / p1 - 1.0) / gm) : 1.0;
  const double qR = pStar > p2 ? sqrt_ws(1.0 + 0.5 * (gm + 1.0) * (pStar / p2 - 1.0) / gm) : 1.0;
  lo = uL - qL * cL;
  hi = uR + qR * cR;
}

template <int DIM>
struct SidePrim {
  double inv, m2, p, pid;  // 1/rho, |rho u|^2, pressure, covolume-free pressure
};
template <int DIM>
RKDG_HD SidePrim<DIM> side_prim(const double* s, const SchemeParams& g) {
  SidePrim<DIM> q;
  q.inv = recip(s[0]);
  double m2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; d++) m2 += s[1 + d] * s[1 + d];
  q.m2 = m2;
  q.pid = (g.gamma - 1.0) * (s[1 + DIM] - 0.5 * (m2 * q.inv));
  q.p = g.covolume != 0.0 ? q.pid / (1.0 - s[0] * g.covolume) : q.pid;
  return q;
}
template <int DIM>
RKDG_HD double side_sound(const double* s, const SidePrim<DIM>& q, const SchemeParams& g) {
  double a = g.gamma * q.p * q.inv;
  if (g.covolume != 0.0) a /= (1.0 - s[0] * g.covolume);
  return sqrt_ws(a);
}
// StateIsPhysicalSay (physics.cpp:53-97) from the side primitives
template <int DIM>
RKDG_HD bool side_ok(const double* s, const SidePrim<DIM>& q) {
  return (s[0] >= 0.0) && (s[1 + DIM] > 0.0) && (q.pid > 0.0);
}

// Numerical flux through the unit normal n (rs_rusanov.cpp:6-40, rs_llf.cpp:6-37,
// rs_hll.cpp:6-50, rs_hllc.cpp:71-161 with getFStar :39-67). The states are
// taken by value (the reference rotates its inputs in place; nothing observes
// it). Returns the solver's wave speed; ok=false flags an unphysical input.
// inm: 0.5 / |n| when the caller has it per face (Rusanov only; 0 = form it here)
template <int DIM, int RS>
RKDG_HD double riemann_t(const SchemeParams& g, const double* uL, const double* uR, const double* n,
                         double* flux, bool& ok, double inm_face = 0.0) {
  constexpr int NE = DIM + 2;
  double s1[NE], s2[NE], f1[NE], f2[NE];
#pragma unroll
  for (int i = 0; i < NE; i++) {
    s1[i] = uL[i];
    s2[i] = uR[i];
  }
  if constexpr (RS == RKDG_RS_RUSANOV) {
    const SidePrim<DIM> q1 = side_prim<DIM>(s1, g), q2 = side_prim<DIM>(s2, g);
    ok = side_ok<DIM>(s1, q1) && side_ok<DIM>(s2, q2);
    const double m1 = sqrt_ws(q1.m2) * q1.inv + side_sound<DIM>(s1, q1, g);
    const double m2 = sqrt_ws(q2.m2) * q2.inv + side_sound<DIM>(s2, q2, g);
    const double mE = m1 > m2 ? m1 : m2;
    double vn1 = 0.0, vn2 = 0.0, nm = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      vn1 += s1[1 + d] * n[d];
      vn2 += s2[1 + d] * n[d];
      nm += n[d] * n[d];
    }
    const double a1 = vn1 * q1.inv, a2 = vn2 * q2.inv;
    f1[0] = vn1;
    f2[0] = vn2;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      f1[1 + d] = a1 * s1[1 + d] + q1.p * n[d];
      f2[1 + d] = a2 * s2[1 + d] + q2.p * n[d];
    }
    f1[1 + DIM] = a1 * (s1[1 + DIM] + q1.p);
    f2[1 + DIM] = a2 * (s2[1 + DIM] + q2.p);
    // the unit normal's own norm (rs_rusanov.cpp:27-32); 1 up to rounding
    const double inm = inm_face != 0.0 ? inm_face : 0.5 * recip(sqrt_ws(nm));
    const double hm = 0.5 * mE;
#pragma unroll
    for (int i = 0; i < NE; i++) flux[i] = (f1[i] + f2[i]) * inm - hm * (s2[i] - s1[i]);
    return mE;
  } else {
  rotate<DIM>(s1, n);
  rotate<DIM>(s2, n);
  const SidePrim<DIM> q1 = side_prim<DIM>(s1, g), q2 = side_prim<DIM>(s2, g);
  ok = side_ok<DIM>(s1, q1) && side_ok<DIM>(s2, q2);
  const double p1 = q1.p, p2 = q2.p;
  {
    const double u1 = s1[1] * q1.inv, u2 = s2[1] * q2.inv;
    f1[0] = s1[1];
    f2[0] = s2[1];
    f1[1] = u1 * s1[1] + p1;
    f2[1] = u2 * s2[1] + p2;
#pragma unroll
    for (int d = 1; d < DIM; d++) {
      f1[d + 1] = u1 * s1[d + 1];
      f2[d + 1] = u2 * s2[d + 1];
    }
    f1[1 + DIM] = u1 * (s1[1 + DIM] + p1);
    f2[1 + DIM] = u2 * (s2[1 + DIM] + p2);
  }
  double SL, SR;
  wave_speeds<DIM>(s1, s2, p1, p2, g, SL, SR);
  double mE;
  if constexpr (RS == RKDG_RS_LLF) {
    mE = fmax(fabs(SL), fabs(SR));
#pragma unroll
    for (int i = 0; i < NE; i++) flux[i] = 0.5 * (f1[i] + f2[i]) - 0.5 * mE * (s2[i] - s1[i]);
  } else if constexpr (RS == RKDG_RS_HLL) {
    mE = fmax(fabs(SL), fabs(SR));
    if (SL >= 0) {
#pragma unroll
      for (int i = 0; i < NE; i++) flux[i] = f1[i];
    } else if (SR <= 0) {
#pragma unroll
      for (int i = 0; i < NE; i++) flux[i] = f2[i];
    } else {
      const double inv = 1.0 / (SR - SL);
#pragma unroll
      for (int i = 0; i < NE; i++) flux[i] = (SR * f1[i] - SL * f2[i] + SL * SR * (s2[i] - s1[i])) * inv;
    }
  } else {  // HLLC
    mE = g.hllc_abs ? fmax(fabs(SL), fabs(SR)) : fmax(SL, SR);
    if (SL >= 0) {
#pragma unroll
      for (int i = 0; i < NE; i++) flux[i] = f1[i];
    } else if (SR <= 0) {
#pragma unroll
      for (int i = 0; i < NE; i++) flux[i] = f2[i];
    } else {
      const double cL = SL - s1[1] / s1[0];
      const double cR = SR - s2[1] / s2[0];
      const double sStar = (p2 - p1 + s1[1] * cL - s2[1] * cR) / (s1[0] * cL - s2[0] * cR);
      const bool left = sStar > 0.0;
      const double* st = left ? s1 : s2;
      const double* fK = left ? f1 : f2;
      const double pK = left ? p1 : p2;
      const double SK = left ? SL : SR;
      const double cK = left ? cL : cR;
      const double pterm = SK * (pK + st[0] * cK * (sStar - st[1] / st[0]));
      const double inv = 1.0 / (SK - sStar);
#pragma unroll
      for (int i = 0; i < NE; i++) {
        const double D = i == 1 ? 1.0 : (i == DIM + 1 ? sStar : 0.0);
        flux[i] = (sStar * (st[i] * SK - fK[i]) + D * pterm) * inv;
      }
    }
  }
  inverse_rotate<DIM>(flux, n);
  return mE;
  }
}

// runtime-dispatched form (host unit tests and cold paths)
template <int DIM>
RKDG_HD double riemann(const SchemeParams& g, const double* uL, const double* uR, const double* n,
                       double* flux, bool& ok) {
  switch (g.riemann) {
    case RKDG_RS_RUSANOV: return riemann_t<DIM, RKDG_RS_RUSANOV>(g, uL, uR, n, flux, ok);
    case RKDG_RS_LLF: return riemann_t<DIM, RKDG_RS_LLF>(g, uL, uR, n, flux, ok);
    case RKDG_RS_HLL: return riemann_t<DIM, RKDG_RS_HLL>(g, uL, uR, n, flux, ok);
    default: return riemann_t<DIM, RKDG_RS_HLLC>(g, uL, uR, n, flux, ok);
  }
}

// Ghost state for boundary slot bi (src/bc/boundary_integrator_*.cpp
// computeRightState); n is the unit outward normal.
template <int DIM>
RKDG_HD void bc_state(const SchemeParams& g, int bi, const double* s1, const double* n, double* s2,
                      const double* Pover = nullptr) {
  constexpr int NE = DIM + 2;
  const double* P = Pover ? Pover : g.bc_param[bi];
  const double gm = g.gamma, R = g.gas_constant;
#pragma unroll
  for (int i = 0; i < NE; i++) s2[i] = s1[i];
  switch (g.bc_type[bi]) {
    case RKDG_BC_SLIP:  // boundary_integrator_slip.cpp:7-17
      rotate<DIM>(s2, n);
      s2[1] *= -1;
      inverse_rotate<DIM>(s2, n);
      break;
    case RKDG_BC_OUTLET:  // boundary_integrator_open.cpp:9-12
      break;
    case RKDG_BC_SUPERSONIC_INLET: {  // ..._supersonic_inlet.cpp:8-19; state: case_manager.cpp:736-750
      const double rho = P[0], u = P[1], v = P[2], w = P[3], p = P[4];
      s2[0] = rho;
      s2[1] = rho * u;
      s2[2] = rho * v;
      if (DIM == 3) s2[3] = rho * w;
      s2[NE - 1] = p / (gm - 1.) + 0.5 * rho * (u * u + v * v);
      break;
    }
    case RKDG_BC_SUBSONIC_INLET_TOTAL_PRESSURE:    // ..._subsonic_inlet_total_pressure.cpp:8-38
    case RKDG_BC_SUBSONIC_INLET_FIXED_PRESSURE: {  // ..._subsonic_inlet_fixed_pressure.cpp:8-38
      const double pTot = P[0], TTot = P[1];
      double rhoU2 = s2[1] * s2[1] + s2[2] * s2[2];
      if (DIM == 3) rhoU2 += s2[3] * s2[3];
      rhoU2 /= s2[0];
      const double MTTotSq = rhoU2 / s2[0] / gm / R / TTot;
      const double MSq = MTTotSq / (1.0 - 0.5 * (gm - 1.0) * MTTotSq);
      const double p = g.bc_type[bi] == RKDG_BC_SUBSONIC_INLET_TOTAL_PRESSURE
                           ? pTot * pow_ws(1.0 + 0.5 * (gm - 1.0) * MSq, -gm / (gm - 1.0))
                           : pTot;
      const double T = TTot / (1.0 + 0.5 * (gm - 1.0) * MSq);
      const double rho = p / R / T;
      s2[NE - 1] = p / (gm - 1.0) + 0.5 * rhoU2 / s2[0] * rho;
      s2[1] = s2[1] / s2[0] * rho;
      s2[2] = s2[2] / s2[0] * rho;
      if (DIM == 3) s2[3] = s2[3] / s2[0] * rho;
      s2[0] = rho;
      break;
    }
    case RKDG_BC_FIXED_PRESSURE_OUTLET: {  // ..._open_fixed_pressure.cpp:9-32
      const double pFix = P[0];
      double rhoU2 = s2[1] * s2[1] + s2[2] * s2[2];
      if (DIM == 3) rhoU2 += s2[3] * s2[3];
      const double U2 = rhoU2 / s2[0] / s2[0];
      const double epsIn = (s2[NE - 1] - 0.5 * rhoU2) / s2[0];
      const double rho = pFix / (gm - 1.0) / epsIn;
      s2[NE - 1] = pFix / (gm - 1.0) + 0.5 * rho * U2;
      s2[1] = s1[1] / s1[0] * rho;
      s2[2] = s1[2] / s1[0] * rho;
      if (DIM == 3) s2[3] = s2[3] / s1[0] * rho;
      s2[0] = rho;
      break;
    }
    case RKDG_BC_TOTAL_PRESSURE_OUTLET: {  // ..._open_total_pressure.cpp:9-42 (normal index fixed)
      const double pTotal = P[0], TFix = P[1];
      double r2U2 = s2[1] * s2[1] + s2[2] * s2[2];
      if (DIM == 3) r2U2 += s2[3] * s2[3];
      double rhoN = s2[1] * n[0] + s2[2] * n[1];
      if (DIM == 3) rhoN += s2[3] * n[2];
      const double U2 = r2U2 / s2[0] / s2[0];
      const double epsIn = rhoN > 0 ? (s2[NE - 1] - 0.5 * r2U2) / s2[0] : R / (gm - 1.0) * TFix;
      double v2 = r2U2 / s2[0];
      double pin = pressure<DIM>(s2, g);
      const double M = sqrt_ws(v2 / s2[0]) / sound_speed_p<DIM>(s2, pin, g);
      double pr = M > 1 ? pTotal * pow_ws(1.0 + 0.5 * (gm - 1.0) * M * M, -gm / (gm - 1.0)) : pTotal;
      const double rho = M > 1 ? pr / (gm - 1.0) / epsIn : pr / (0.5 * U2 + (gm - 1.0) * epsIn);
      pr = M > 1 ? pr : pr - 0.5 * U2 * rho;
      s2[NE - 1] = pr / (gm - 1.0) + 0.5 * rho * U2;
      s2[1] *= rho / s2[0];
      s2[2] *= rho / s2[0];
      if (DIM == 3) s2[3] *= rho / s2[0];
      s2[0] = rho;
      break;
    }
    case RKDG_BC_CHAR_OUTLET: {  // boundary_integrator_char_outlet.cpp:14-51 with its ctor :6-12
      // the fixed state as case_manager.cpp:781-799 builds it (energy without w^2)
      double fs[NE];
      fs[0] = P[0];
      fs[1] = P[0] * P[1];
      fs[2] = P[0] * P[2];
      if (DIM == 3) fs[3] = P[0] * P[3];
      fs[NE - 1] = P[4] / (gm - 1.) + 0.5 * P[0] * (P[1] * P[1] + P[2] * P[2]);
      const double cOut = sound_speed_p<DIM>(fs, pressure<DIM>(fs, g), g);
      const double UOut = fs[1] / fs[0];
      const double cIn = sound_speed_p<DIM>(s1, pressure<DIM>(s1, g), g);
      // state2 starts as state1 (the reference rotates it uninitialised)
      rotate<DIM>(s2, n);
      const double UIn = s2[1] / s2[0];
      const double UB = 0.5 * (UIn + UOut) - (cOut - cIn) / (gm - 1.0);
      const double cB = -0.25 * (gm - 1.0) * (UOut - UIn) + 0.5 * (cOut + cIn);
      const double sB = s2[1] < 0 ? cOut * cOut / pow_ws(fs[0], gm - 1.0) / gm   // inflow
                                  : cIn * cIn / pow_ws(s2[0], gm - 1.0) / gm;    // outflow
      const double rhoB = pow_ws(cB * cB / sB / gm, 1.0 / (gm - 1.0));
      const double pB = rhoB * cB * cB / gm;
      s2[0] = rhoB;
      // normal momentum rhoB UB (the reference stores the velocity UB in the
      // momentum slot); the tangential momenta stay, so the tangential
      // velocities in the energy are those momenta over rhoB, as there
      s2[1] = rhoB * UB;
      const double vt = s2[2] / s2[0], wt = DIM == 3 ? s2[3] / s2[0] : 0.0;
      s2[NE - 1] = pB / (gm - 1.0) * (1.0 - rhoB * g.covolume) + 0.5 * rhoB * (UB * UB + vt * vt + wt * wt);
      inverse_rotate<DIM>(s2, n);
      break;
    }
    default:
      break;
  }
}

// boundary_integrator_char_outlet.cpp:153-189: the characteristic outlet's flux
// is the physical flux of its ghost state through the unit normal (no Riemann
// solver); returns |u|+c of the ghost
template <int DIM>
RKDG_HD double char_outlet_flux(const SchemeParams& g, const double* s2, const double* n, double* fl) {
  flux_dot_n<DIM>(s2, n, pressure<DIM>(s2, g), g, fl);
  return max_char_speed<DIM>(s2, g);
}

// Euclidean norm by the scaled (hypot-style) recurrence of MFEM Vector::Norml2,
// which the reference uses to normalise face normals (face_integrator.cpp:165).
// A vector with one non-zero entry a gets exactly |a|.
template <int DIM>
RKDG_HD double norml2_scaled(const double* v) {
  double scale = 0.0, sum = 0.0;
#pragma unroll
  for (int i = 0; i < DIM; i++) {
    if (v[i] != 0.0) {
      const double a = fabs(v[i]);
      if (scale <= a) {
        const double r = scale / a;
        sum = 1.0 + sum * (r * r);
        scale = a;
      } else {
        const double r = a / scale;
        sum += r * r;
      }
    }
  }
  return scale * sqrt_ws(sum);
}

// indicator_bj.cpp:49-79: ratio y from the stencil bounds, then the limiter
// function (Michalak active in the reference; BJ and Venkatakrishnan selectable).
RKDG_HD double indicator_y(double diff, double mean, double mI, double MI) {
  if (diff > 1e-3) return (MI - mean) / diff;
  if (diff < -1e-3) return (mI - mean) / diff;
  return 1.0;
}

}  // namespace rkdg
