// rkdg_selftest.cpp — host entry points of include/rkdg_selftest.h: the 2D
// physics and its hand-written derivatives evaluated on the host, so unit tests
// can compare the vector-Jacobian products with finite differences without a GPU.
#include "../../include/rkdg_selftest.h"

#include "rkdg_adjoint.hpp"

using namespace rkdg;

namespace {

SchemeParams params(const rkdg_scheme_desc* s) {
  SchemeParams g{};
  g.riemann = s->riemann;
  g.wave_speeds = s->wave_speeds;
  g.hllc_abs = s->hllc_abs;
  g.gamma = s->gamma;
  g.gas_constant = s->gas_constant;
  g.covolume = s->covolume;
  g.num_bc = s->num_bc;
  for (int i = 0; i < RKDG_MAX_BC; i++) {
    g.bc_attr[i] = s->bc_attr[i];
    g.bc_type[i] = s->bc_type[i];
    for (int k = 0; k < 8; k++) g.bc_param[i][k] = s->bc_param[i][k];
  }
  g.indicator = s->indicator;
  g.phi = s->phi;
  g.venkat_k = s->venkat_k;
  g.limiter = s->limiter;
  g.last_hope = s->last_hope;
  g.pow_nd = s->pow_nd;
  return g;
}

}  // namespace

extern "C" {

double rkdg_selftest_riemann(const rkdg_scheme_desc* s, const double* uL, const double* uR, const double* n,
                             double* flux) {
  const SchemeParams g = params(s);
  bool ok = true;
  return riemann<2>(g, uL, uR, n, flux, ok);
}

void rkdg_selftest_riemann_vjp(const rkdg_scheme_desc* s, const double* uL, const double* uR, const double* n,
                               const double* Fbar, double* uLbar, double* uRbar) {
  const SchemeParams g = params(s);
  adj::riemann_vjp_rt(g, uL, uR, n, Fbar, uLbar, uRbar);
}

void rkdg_selftest_volume_flux(const rkdg_scheme_desc* s, const double* u, const double* a, double* G) {
  const SchemeParams g = params(s);
  const SidePrim<2> q = side_prim<2>(u, g);
  const double ma = u[1] * a[0] + u[2] * a[1];
  const double mai = ma * q.inv;
  G[0] = ma;
  G[1] = mai * u[1] + q.p * a[0];
  G[2] = mai * u[2] + q.p * a[1];
  G[3] = mai * (u[3] + q.p);
}

void rkdg_selftest_volume_flux_vjp(const rkdg_scheme_desc* s, const double* u, const double* a,
                                   const double* Gbar, double* ubar) {
  const SchemeParams g = params(s);
  adj::volume_flux_vjp(u, a, Gbar, g, ubar);
}

void rkdg_selftest_bc_state(const rkdg_scheme_desc* s, int bi, const double* u, const double* n, double* ghost) {
  const SchemeParams g = params(s);
  bc_state<2>(g, bi, u, n, ghost);
}

void rkdg_selftest_bc_state_vjp(const rkdg_scheme_desc* s, int bi, const double* u, const double* n,
                                const double* gbar, double* ubar, double* Pbar) {
  const SchemeParams g = params(s);
  adj::bc_state_vjp(g, bi, u, n, gbar, ubar, Pbar);
}

double rkdg_selftest_limiter_phi(const rkdg_scheme_desc* s, double y, double diff, double eps2, double* d) {
  const SchemeParams g = params(s);
  return adj::limiter_phi(g, y, diff, eps2, &d[1], &d[2]);
}

double rkdg_selftest_pressure(const rkdg_scheme_desc* s, const double* u, double* dp) {
  const SchemeParams g = params(s);
  return adj::pressure_grad(u, g, dp);
}

}  // extern "C"
