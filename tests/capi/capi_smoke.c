/* A C caller of librkdg.so through include/rkdg_capi.h alone, the way a
 * ctypes/cgo/JNI binding calls it: create -> info -> node coordinates ->
 * right-hand side -> limiter -> two RK steps -> error check -> destroy, on
 * host arrays (the *_host entry points), then a deliberately unphysical state
 * whose error comes back as RKDG_E_UNPHYSICAL with the element pair. The
 * arrays are written to OUT_DIR for the test that checks them against the
 * oracle.
 *
 *   capi_smoke OUT_DIR [ORDER]
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "smoke_case.h"

#define CHECK(call)                                                                 \
  do {                                                                              \
    int rc_ = (call);                                                               \
    if (rc_ != RKDG_OK) {                                                           \
      fprintf(stderr, "%s failed: %d %s\n", #call, rc_, rkdg_last_error());         \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s OUT_DIR [ORDER]\n", argv[0]);
    return 2;
  }
  const char* out = argv[1];
  const int order = argc > 2 ? atoi(argv[2]) : 2;
  static smoke_case c;
  smoke_setup(&c, order);
  rkdg_ctx* ctx = NULL;
  CHECK(rkdg_create(&c.mesh, &c.scheme, &ctx));
  rkdg_info info;
  CHECK(rkdg_get_info(ctx, &info));
  printf("[capi_smoke] NE %lld nd %d neq %d interior faces %lld boundary faces %lld\n",
         (long long)info.num_elements, info.nodes_per_element, info.num_equations,
         (long long)info.num_interior_faces, (long long)info.num_boundary_faces);
  const int64_t n = info.num_dofs, size = info.state_size;
  double* xyz = malloc(sizeof(double) * n * 2);
  double* x = malloc(sizeof(double) * size);
  double* y = malloc(sizeof(double) * size);
  int8_t* tr = malloc((size_t)info.num_elements);
  CHECK(rkdg_node_coords(ctx, xyz));
  smoke_state(xyz, n, x);
  double lam = 0.0, lam_s = 0.0, hmin = 0.0;
  CHECK(rkdg_rhs_host(ctx, x, y, &lam));
  CHECK(rkdg_hmin(ctx, &hmin));
  /* limit a copy, then two SSP-RK3 steps from the limited state */
  double* xl = malloc(sizeof(double) * size);
  for (int64_t k = 0; k < size; k++) xl[k] = x[k];
  CHECK(rkdg_limit_host(ctx, xl, tr));
  double* xs = malloc(sizeof(double) * size);
  for (int64_t k = 0; k < size; k++) xs[k] = xl[k];
  const double a[3] = {1.0, 0.25, 0.25}, b[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}, cc[3] = {1.0, 0.5, 0.0};
  const double dt = 0.1 * hmin / lam / (2 * order + 1);
  for (int s = 0; s < 2; s++) CHECK(rkdg_step_host(ctx, xs, 0.0, dt, 3, a, b, cc, &lam_s));
  int64_t pair[2] = {-1, -1};
  CHECK(rkdg_check_error(ctx, pair));
  /* an unphysical state: negative density at one node of element 13 */
  double* bad = malloc(sizeof(double) * size);
  for (int64_t k = 0; k < size; k++) bad[k] = x[k];
  bad[13 * info.nodes_per_element] = -1.0;
  const int rc_bad = rkdg_rhs_host(ctx, bad, y + 0 * size, NULL);
  printf("[capi_smoke] unphysical state: rc %d (%s)\n", rc_bad, rc_bad ? rkdg_last_error() : "");
  /* the right-hand side again for the output (y was overwritten above) */
  CHECK(rkdg_rhs_host(ctx, x, y, &lam));
  const double scal[4] = {lam, lam_s, dt, (double)rc_bad};
  int w = 0;
  w |= smoke_write(out, "coords.f64", c.coords, sizeof(c.coords));
  w |= smoke_write(out, "verts.i64", c.verts, sizeof(c.verts));
  w |= smoke_write(out, "bverts.i64", c.bverts, sizeof(c.bverts));
  w |= smoke_write(out, "battr.i32", c.battr, sizeof(c.battr));
  w |= smoke_write(out, "x.f64", x, sizeof(double) * size);
  w |= smoke_write(out, "y.f64", y, sizeof(double) * size);
  w |= smoke_write(out, "xl.f64", xl, sizeof(double) * size);
  w |= smoke_write(out, "tr.i8", tr, (size_t)info.num_elements);
  w |= smoke_write(out, "xs.f64", xs, sizeof(double) * size);
  w |= smoke_write(out, "scalars.f64", scal, sizeof(scal));
  CHECK(rkdg_destroy(ctx));
  free(xyz), free(x), free(y), free(tr), free(xl), free(xs), free(bad);
  if (w) {
    fprintf(stderr, "could not write to %s\n", out);
    return 1;
  }
  printf("[capi_smoke] lambda %.17g after steps %.17g dt %.17g: OK\n", lam, lam_s, dt);
  return rc_bad == RKDG_E_UNPHYSICAL ? 0 : 3;
}
