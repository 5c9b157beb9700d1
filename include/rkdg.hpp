/* rkdg.hpp -- C++ mirror of the reference's operator classes over the C-ABI
 * of include/rkdg_capi.h (header only; link librkdg.so).
 *
 * The class names, constructor arguments and call semantics follow the
 * reference so that a caller ports line by line; vectors are plain device
 * pointers in the byNODES layout x[v*NE*nd + e*nd + i] (MFEM's Vector and
 * DenseMatrix are not part of this build):
 *
 *   rkdg::FE_Evolution             src/fe_evolution/fe_evolution.hpp:27-82
 *       Mult(x, y); max_char_speed (the reference's global, fe_evolution.hpp:10);
 *       UpdateMesh after re-meshing (UpdateInverseMassMatrix/UpdateAfluxPointer)
 *   rkdg::Limiter                  src/limiter/limiter.hpp:24-116 (update)
 *   rkdg::ExplicitRKLimitedSolver  src/ode_solver/rk_explicit_limited.hpp:24-60
 *       Init(f); Step(x, t, dt) -- Step does not advance t (the driver does)
 *   rkdg::RiemannSolver            src/riemannsolvers/rs_basic.hpp:18-51
 *       Eval(state1, state2, nor, flux) -> max wave speed, -1 when unphysical
 *   rkdg::BoundaryIntegrator       src/bc/boundary_integrator.hpp:63
 *       computeRightState(state1, state2, nor)
 *
 * Errors: a failing C-ABI call throws rkdg::Error (what() is rkdg_last_error());
 * an unphysical face state found by a kernel throws rkdg::UnphysicalState with
 * the element pair, where the reference prints the pair and exits
 * (face_integrator.cpp:214-218).
 */
#ifndef RKDG_HPP
#define RKDG_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rkdg_capi.h"

namespace rkdg {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

class UnphysicalState : public Error {
 public:
  UnphysicalState(int code, int64_t e1, int64_t e2)
      : Error(code, "unphysical state at the faces between elements " + std::to_string(e1) + " and " +
                        std::to_string(e2)),
        elem1(e1), elem2(e2) {}
  int64_t elem1, elem2;
};

inline void check(int rc) {
  if (rc != RKDG_OK) {
    const char* m = rkdg_last_error();
    throw Error(rc, m ? m : "rkdg error");
  }
}

/* Numerical flux plug-in: the scheme's Riemann solver and wave-speed estimate. */
class RiemannSolver {
 public:
  RiemannSolver(const rkdg_scheme_desc& scheme, int dim) : s_(scheme), dim_(dim) {}
  /* flux per unit normal through nor (normalized here, as the reference's
   * face integrator does before Eval); returns the wave speed, or -1 for an
   * unphysical input as the reference does */
  double Eval(const double* state1, const double* state2, const double* nor, double* flux) const {
    double lam = -1.0;
    const int rc = rkdg_riemann_eval(&s_, dim_, state1, state2, nor, flux, &lam);
    if (rc == RKDG_E_UNPHYSICAL) return -1.0;
    check(rc);
    return lam;
  }

 private:
  rkdg_scheme_desc s_;
  int dim_;
};

/* Ghost state of one boundary patch (the scheme's boundary slot). */
class BoundaryIntegrator {
 public:
  BoundaryIntegrator(const rkdg_scheme_desc& scheme, int dim, int slot) : s_(scheme), dim_(dim), slot_(slot) {}
  void computeRightState(const double* state1, double* state2, const double* nor) const {
    check(rkdg_bc_right_state(&s_, dim_, slot_, state1, nor, state2));
  }

 private:
  rkdg_scheme_desc s_;
  int dim_, slot_;
};

class FE_Evolution {
 public:
  FE_Evolution(const rkdg_mesh_desc& mesh, const rkdg_scheme_desc& scheme, int flags = 0)
      : scheme_(scheme), order_(mesh.order) {
    check(rkdg_create_ex(&mesh, &scheme, flags, &ctx_));
    check(rkdg_get_info(ctx_, &info_));
  }
  ~FE_Evolution() {
    if (ctx_) rkdg_destroy(ctx_);
  }
  FE_Evolution(const FE_Evolution&) = delete;
  FE_Evolution& operator=(const FE_Evolution&) = delete;

  /* y = M^-1 [faces + volume](x), DEVICE x and y; updates max_char_speed */
  void Mult(const double* x, double* y) {
    double lam = 0.0;
    throw_on_state(rkdg_rhs(ctx_, x, y, &lam));
    max_char_speed = lam;
  }
  /* A.Mult(sol, z) of the first dt (main.cpp:473-484): face wave speeds only */
  double FaceMaxCharSpeed(const double* x) {
    double lam = 0.0;
    throw_on_state(rkdg_face_lambda(ctx_, x, &lam));
    return lam;
  }
  /* after a re-mesh (main.cpp:310-311, 610-611): rebuild on the new mesh */
  void UpdateMesh(const rkdg_mesh_desc& mesh) {
    rkdg_ctx* next = nullptr;
    check(rkdg_create(&mesh, &scheme_, &next));
    rkdg_destroy(ctx_);
    ctx_ = next;
    order_ = mesh.order;
    check(rkdg_get_info(ctx_, &info_));
  }
  double hmin() const {
    double h = 0.0;
    check(rkdg_hmin(ctx_, &h));
    return h;
  }
  int64_t Width() const { return (int64_t)info_.num_equations * info_.num_dofs; }
  std::vector<double> NodeCoords() const {
    std::vector<double> xyz((size_t)info_.num_dofs * info_.dim);
    check(rkdg_node_coords(ctx_, xyz.data()));
    return xyz;
  }
  const rkdg_info& info() const { return info_; }
  rkdg_ctx* ctx() const { return ctx_; }
  int order() const { return order_; }

  void throw_on_state(int rc) const {
    if (rc == RKDG_E_UNPHYSICAL || rc == RKDG_E_NAN) {
      int64_t pair[2] = {-1, -1};
      rkdg_error_elements(ctx_, pair);
      throw UnphysicalState(rc, pair[0], pair[1]);
    }
    check(rc);
  }

  double max_char_speed = 0.0;

 private:
  rkdg_ctx* ctx_ = nullptr;
  rkdg_scheme_desc scheme_;
  int order_;
  rkdg_info info_{};
};

class Limiter {
 public:
  explicit Limiter(FE_Evolution& f) : f_(f) {}
  /* in place on a DEVICE state; troubled: DEVICE int8[NE] or nullptr */
  void update(double* x, int8_t* troubled = nullptr) { f_.throw_on_state(rkdg_limit(f_.ctx(), x, troubled)); }

 private:
  FE_Evolution& f_;
};

class ExplicitRKLimitedSolver {
 public:
  /* Butcher tableau consumed with the packed a[l++] rule: a holds the strictly
   * lower triangle row by row (s(s-1)/2 entries), b the s weights, c the nodes.
   * The context applies its scheme's limiter after each stage and at the end of
   * the step (rk_explicit_limited.cpp:63,81); the limiter argument is kept for
   * the reference's constructor signature. */
  ExplicitRKLimitedSolver(int s, const double* a, const double* b, const double* c, Limiter& limiter)
      : s_(s), a_(a, a + (s * (s - 1)) / 2), b_(b, b + s), c_(c, c + (s > 1 ? s - 1 : 0)), lim_(limiter) {
    a_.push_back(0.0);
    c_.push_back(0.0);
  }
  void Init(FE_Evolution& f) { f_ = &f; }
  Limiter& limiter() { return lim_; }
  /* advances the DEVICE state x by dt; t and dt are left as they are */
  void Step(double* x, double& t, double& dt) {
    if (!f_) throw Error(RKDG_E_ARG, "ExplicitRKLimitedSolver::Step before Init");
    double lam = 0.0;
    f_->throw_on_state(rkdg_step(f_->ctx(), x, t, dt, s_, a_.data(), b_.data(), c_.data(), &lam));
    f_->max_char_speed = lam;
  }

 private:
  int s_;
  std::vector<double> a_, b_, c_;
  Limiter& lim_;
  FE_Evolution* f_ = nullptr;
};

}  // namespace rkdg

#endif
