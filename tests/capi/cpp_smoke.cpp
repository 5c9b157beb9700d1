// A C++ caller of librkdg.so through the class mirror include/rkdg.hpp:
// RiemannSolver::Eval and BoundaryIntegrator::computeRightState on the host,
// then (unless --plugins-only) FE_Evolution::Mult, Limiter::update and two
// ExplicitRKLimitedSolver::Step calls on device buffers, and the
// UnphysicalState exception. Results go to OUT_DIR for the test that checks
// them against the oracle.
//
//   cpp_smoke OUT_DIR [--plugins-only]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "rkdg.hpp"
#include "smoke_case.h"

static void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw rkdg::Error(RKDG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int main(int argc, char** argv) try {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUT_DIR [--plugins-only]\n", argv[0]);
    return 2;
  }
  const char* out = argv[1];
  const bool plugins_only = argc > 2 && std::strcmp(argv[2], "--plugins-only") == 0;
  static smoke_case c;
  smoke_setup(&c, 2);

  // plug-ins: fluxes of seeded state pairs through a skew normal, and the
  // ghost states of the four boundary slots
  rkdg::RiemannSolver rs(c.scheme, 2);
  const double nor[2] = {0.3, -0.7};
  const double pairs[3][2][4] = {{{1.0, 0.2, -0.1, 2.6}, {0.6, -0.1, 0.3, 1.9}},
                                 {{2.0, 1.0, 0.5, 6.0}, {2.0, 1.0, 0.5, 6.0}},
                                 {{0.125, 0.0, 0.0, 0.25}, {1.0, 0.0, 0.0, 2.5}}};
  std::vector<double> plug;
  for (auto& pr : pairs) {
    double f[4];
    const double lam = rs.Eval(pr[0], pr[1], nor, f);
    plug.insert(plug.end(), f, f + 4);
    plug.push_back(lam);
  }
  const double neg[4] = {-1.0, 0.0, 0.0, 1.0};
  double fneg[4];
  plug.push_back(rs.Eval(neg, pairs[0][0], nor, fneg));  // -1: unphysical
  for (int slot = 0; slot < c.scheme.num_bc; slot++) {
    rkdg::BoundaryIntegrator bi(c.scheme, 2, slot);
    double s2[4];
    bi.computeRightState(pairs[0][0], s2, nor);
    plug.insert(plug.end(), s2, s2 + 4);
  }
  if (smoke_write(out, "plugins.f64", plug.data(), sizeof(double) * plug.size())) return 1;
  std::printf("[cpp_smoke] plug-ins: %zu values\n", plug.size());
  if (plugins_only) return 0;

  rkdg::FE_Evolution f(c.mesh, c.scheme);
  rkdg::Limiter lim(f);
  const int64_t size = f.Width(), n = f.info().num_dofs;
  std::vector<double> xyz = f.NodeCoords(), x(size), y(size), xs(size);
  smoke_state(xyz.data(), n, x.data());
  double *dx = nullptr, *dy = nullptr;
  hip_check(hipMalloc(&dx, sizeof(double) * size), "hipMalloc");
  hip_check(hipMalloc(&dy, sizeof(double) * size), "hipMalloc");
  hip_check(hipMemcpy(dx, x.data(), sizeof(double) * size, hipMemcpyHostToDevice), "hipMemcpy");
  f.Mult(dx, dy);
  const double lam = f.max_char_speed;
  hip_check(hipMemcpy(y.data(), dy, sizeof(double) * size, hipMemcpyDeviceToHost), "hipMemcpy");
  lim.update(dx);
  const double a[3] = {1.0, 0.25, 0.25}, b[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}, cc[2] = {1.0, 0.5};
  rkdg::ExplicitRKLimitedSolver ode(3, a, b, cc, lim);
  ode.Init(f);
  double t = 0.0, dt = 0.1 * f.hmin() / lam / (2 * f.order() + 1);
  for (int s = 0; s < 2; s++) {
    ode.Step(dx, t, dt);
    t += dt;
  }
  hip_check(hipMemcpy(xs.data(), dx, sizeof(double) * size, hipMemcpyDeviceToHost), "hipMemcpy");
  // the unphysical state: negative density at a node of element 13
  std::vector<double> bad = x;
  bad[13 * f.info().nodes_per_element] = -1.0;
  hip_check(hipMemcpy(dx, bad.data(), sizeof(double) * size, hipMemcpyHostToDevice), "hipMemcpy");
  int64_t e1 = -2, e2 = -2;
  try {
    f.Mult(dx, dy);
  } catch (const rkdg::UnphysicalState& u) {
    e1 = u.elem1, e2 = u.elem2;
    std::printf("[cpp_smoke] caught: %s\n", u.what());
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  const double scal[5] = {lam, f.max_char_speed, dt, (double)e1, (double)e2};
  int w = smoke_write(out, "cpp_y.f64", y.data(), sizeof(double) * size);
  w |= smoke_write(out, "cpp_xs.f64", xs.data(), sizeof(double) * size);
  w |= smoke_write(out, "cpp_x.f64", x.data(), sizeof(double) * size);
  w |= smoke_write(out, "cpp_scalars.f64", scal, sizeof(scal));
  if (w) return 1;
  std::printf("[cpp_smoke] lambda %.17g dt %.17g unphysical pair %lld %lld: OK\n", lam, dt, (long long)e1,
              (long long)e2);
  return e1 >= 0 ? 0 : 3;
} catch (const std::exception& e) {
  std::fprintf(stderr, "[cpp_smoke] error: %s\n", e.what());
  return 1;
}
