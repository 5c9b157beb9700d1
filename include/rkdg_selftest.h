/*
 * rkdg_selftest.h — host-only entry points of librkdg.so for unit tests of the
 * hand-written derivatives (no GPU needed). No counterpart in the reference.
 * Every function evaluates the same host/device code the kernels use, on the
 * host, for one 2D point: the forward map and its vector-Jacobian product.
 */
#ifndef RKDG_SELFTEST_H
#define RKDG_SELFTEST_H

#include "rkdg_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* flux[4] = F(uL, uR, n) of the scheme's Riemann solver; returns the wave speed */
double rkdg_selftest_riemann(const rkdg_scheme_desc* s, const double* uL, const double* uR, const double* n,
                             double* flux);
/* uLbar, uRbar (+=) = (dF/duL)^T Fbar, (dF/duR)^T Fbar */
void rkdg_selftest_riemann_vjp(const rkdg_scheme_desc* s, const double* uL, const double* uR, const double* n,
                               const double* Fbar, double* uLbar, double* uRbar);
/* G[4] = a . F(u) (contravariant nodal flux) and its VJP */
void rkdg_selftest_volume_flux(const rkdg_scheme_desc* s, const double* u, const double* a, double* G);
void rkdg_selftest_volume_flux_vjp(const rkdg_scheme_desc* s, const double* u, const double* a,
                                   const double* Gbar, double* ubar);
/* ghost state of boundary slot bi and its VJP (Pbar[8]: parameter adjoint) */
void rkdg_selftest_bc_state(const rkdg_scheme_desc* s, int bi, const double* u, const double* n, double* ghost);
void rkdg_selftest_bc_state_vjp(const rkdg_scheme_desc* s, int bi, const double* u, const double* n,
                                const double* gbar, double* ubar, double* Pbar);
/* limiter function phi(y; diff, eps2) with its partial derivatives out[1], out[2] */
double rkdg_selftest_limiter_phi(const rkdg_scheme_desc* s, double y, double diff, double eps2, double* d);
/* pressure and its gradient */
double rkdg_selftest_pressure(const rkdg_scheme_desc* s, const double* u, double* dp);

#ifdef __cplusplus
}
#endif
#endif
