// rkdg_host.cpp — host side of the RKDG path: flattens a mesh into the SoA
// device arrays the kernels read (connectivity, face orientation codes,
// geometric factors), builds the 1D basis/quadrature tables, owns device
// buffers and implements the C-ABI of include/rkdg_capi.h.
//
// The mesh processing restates what the reference obtains from MFEM
// (face-neighbour tables, Elem1/Elem2 and Loc1/Loc2 face maps, element
// Jacobians), see src/eq_terms/face_integrator.cpp:85-165 and
// src/limiter/limiter.cpp:98-243.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rkdg_capi.h"
#include "rkdg_kernels.hpp"

using namespace rkdg;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return fail(RKDG_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Gauss-Legendre nodes/weights on [0,1] (Newton on the Legendre recurrence).
void gl_rule(int n, double* x, double* w) {
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; k++) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = z; p0 = 1.0; }
      double dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; k++) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) { p1 = z; p0 = 1.0; }
    double dp = n * (z * p1 - p0) / (z * z - 1.0);
    double wi = 1.0 / ((1.0 - z * z) * dp * dp);
    x[i] = 0.5 * (1.0 - z);
    x[n - 1 - i] = 0.5 * (1.0 + z);
    w[i] = w[n - 1 - i] = wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
}

// Lagrange basis psi_b on nodes xs, value and derivative at y
double lag(const double* xs, int n, int b, double y) {
  double v = 1.0;
  for (int j = 0; j < n; j++)
    if (j != b) v *= (y - xs[j]) / (xs[b] - xs[j]);
  return v;
}
double dlag(const double* xs, int n, int b, double y) {
  double s = 0.0;
  for (int m = 0; m < n; m++) {
    if (m == b) continue;
    double t = 1.0 / (xs[b] - xs[m]);
    for (int j = 0; j < n; j++)
      if (j != b && j != m) t *= (y - xs[j]) / (xs[b] - xs[j]);
    s += t;
  }
  return s;
}

// MFEM reference vertices: quad (0,0),(1,0),(1,1),(0,1); hex bottom then top.
int quad_vertex(int ix, int iy) { return iy ? (ix ? 2 : 3) : (ix ? 1 : 0); }
int hex_vertex(int ix, int iy, int iz) { return iz * 4 + quad_vertex(ix, iy); }

void rc_of_lf(int dim, int lf, int& r, int& c) {
  if (dim == 3) {
    const int rr[6] = {2, 1, 0, 1, 0, 2}, cc[6] = {0, 0, 1, 1, 0, 1};
    r = rr[lf];
    c = cc[lf];
  } else {
    const int rr[4] = {1, 0, 1, 0}, cc[4] = {0, 1, 1, 0};
    r = rr[lf];
    c = cc[lf];
  }
}

// reference vertex of local face lf at face-frame corner (sa, sb); the face
// frame's axes are the two tangential reference axes in increasing order.
int face_corner_vertex(int dim, int lf, int sa, int sb) {
  int r, c;
  rc_of_lf(dim, lf, r, c);
  int idx[3] = {0, 0, 0};
  idx[r] = c;
  if (dim == 2) {
    idx[1 - r] = sa;
    return quad_vertex(idx[0], idx[1]);
  }
  int t1 = r == 0 ? 1 : 0, t2 = r == 2 ? 1 : 2;
  idx[t1] = sa;
  idx[t2] = sb;
  return hex_vertex(idx[0], idx[1], idx[2]);
}

struct FaceKey {
  std::array<int64_t, 4> k;
  int64_t e;
  int lf;
  bool operator<(const FaceKey& o) const {
    if (k != o.k) return k < o.k;
    if (e != o.e) return e < o.e;
    return lf < o.lf;
  }
};

}  // namespace

struct rkdg_ctx {
  int geom = 0, p = 1, dim = 2, nv = 4, nd = 4, neq = 4, nf = 4, nvf = 2;
  int64_t NE = 0, ndofs = 0;
  int affine = 1;
  SchemeParams sp{};
  TensorTables tb{};
  hipStream_t stream = nullptr;
  double* d_geo = nullptr;
  int* d_nbr = nullptr;
  int* d_fcode = nullptr;
  double* d_lam = nullptr;    // per-stage wave speeds [8]
  long long* d_err = nullptr;
  double* d_means = nullptr;
  TensorTables* d_tb = nullptr;
  double* d_stage[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<double> coords;  // [NE][nv][dim]
  std::vector<double> geo_h;   // [g][NE]
  int64_t n_int = 0, n_bdr = 0;
  double hmin = 0.0;
  // instrumentation: event pairs around kernel launches
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_rhs, ev_lim;  // (start, stop) pool indices
};

namespace {

void free_ctx(rkdg_ctx* c) {
  if (!c) return;
  (void)hipFree(c->d_geo);
  (void)hipFree(c->d_nbr);
  (void)hipFree(c->d_fcode);
  (void)hipFree(c->d_lam);
  (void)hipFree(c->d_err);
  (void)hipFree(c->d_means);
  (void)hipFree(c->d_tb);
  for (auto& p : c->d_stage) (void)hipFree(p);
  for (auto& e : c->ev_pool) (void)hipEventDestroy(e);
  delete c;
}

double min_singular(int d, const double* J) {
  double A[9] = {0};
  for (int i = 0; i < d; i++)
    for (int j = 0; j < d; j++) {
      double s = 0;
      for (int k = 0; k < d; k++) s += J[k * d + i] * J[k * d + j];
      A[i * d + j] = s;
    }
  double lmin;
  if (d == 2) {
    double tr = A[0] + A[3], dt = A[0] * A[3] - A[1] * A[2];
    lmin = tr / 2 - std::sqrt(std::max(0.0, tr * tr / 4 - dt));
  } else {
    double p1 = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    double q = (A[0] + A[4] + A[8]) / 3;
    if (p1 == 0) {
      lmin = std::min(A[0], std::min(A[4], A[8]));
    } else {
      double p2 = (A[0] - q) * (A[0] - q) + (A[4] - q) * (A[4] - q) + (A[8] - q) * (A[8] - q) + 2 * p1;
      double p = std::sqrt(p2 / 6);
      double B[9];
      for (int i = 0; i < 9; i++) B[i] = (A[i] - (i % 4 == 0 ? q : 0)) / p;
      double r = (B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                  B[2] * (B[3] * B[7] - B[4] * B[6])) / 2;
      double phi = r <= -1 ? M_PI / 3 : (r >= 1 ? 0 : std::acos(r) / 3);
      lmin = q + 2 * p * std::cos(phi + 2 * M_PI / 3);
    }
  }
  return std::sqrt(std::max(0.0, lmin));
}

int build_ctx(rkdg_ctx* c, const rkdg_mesh_desc* m, const rkdg_scheme_desc* s) {
  if (m->geom == RKDG_GEOM_TRI) return fail(RKDG_E_UNSUPPORTED, "triangle elements are not built into this library");
  if (m->geom != RKDG_GEOM_QUAD && m->geom != RKDG_GEOM_HEX) return fail(RKDG_E_ARG, "unknown geometry");
  c->geom = m->geom;
  c->p = m->order;
  c->dim = m->geom == RKDG_GEOM_HEX ? 3 : 2;
  if (!tensor_supported(c->dim, c->p)) return fail(RKDG_E_UNSUPPORTED, "polynomial order not instantiated");
  c->nv = c->dim == 3 ? 8 : 4;
  c->nvf = c->dim == 3 ? 4 : 2;
  c->nf = 2 * c->dim;
  const int N = c->p + 1;
  c->nd = c->dim == 3 ? N * N * N : N * N;
  c->neq = c->dim + 2;
  c->NE = m->num_elements;
  c->ndofs = c->NE * c->nd;
  if (c->NE <= 0) return fail(RKDG_E_ARG, "empty mesh");
  if (c->NE > (int64_t)2147483647 / 8) return fail(RKDG_E_ARG, "too many elements for 32-bit neighbour ids");
  if (s->num_bc < 0 || s->num_bc > RKDG_MAX_BC) return fail(RKDG_E_ARG, "num_bc out of range");
  const int dim = c->dim, nv = c->nv, NE_ = (int)c->NE;
  const int64_t NE = c->NE;
  c->coords.assign(m->elem_coords, m->elem_coords + NE * nv * dim);

  // scheme
  SchemeParams& g = c->sp;
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

  // 1D tables
  TensorTables& t = c->tb;
  std::memset(&t, 0, sizeof(t));
  gl_rule(N, t.x, t.w);
  for (int a = 0; a < N; a++) t.invw[a] = 1.0 / t.w[a];
  for (int a = 0; a < N; a++)
    for (int b = 0; b < N; b++) t.D[a][b] = dlag(t.x, N, b, t.x[a]);
  for (int b = 0; b < N; b++) {
    t.E[0][b] = lag(t.x, N, b, 0.0);
    t.E[1][b] = lag(t.x, N, b, 1.0);
  }
  const int NQ = dim == 3 ? c->p + 2 : c->p + 1;
  double sq[kMaxQ], wq[kMaxQ];
  gl_rule(NQ, sq, wq);
  for (int q = 0; q < NQ; q++) {
    t.wq[q] = wq[q];
    for (int b = 0; b < N; b++) t.B[q][b] = lag(t.x, N, b, sq[q]);
  }
  double si[2], wi[2];
  gl_rule(2, si, wi);
  for (int q = 0; q < 2; q++)
    for (int b = 0; b < N; b++) t.BI[q][b] = lag(t.x, N, b, si[q]);

  // geometry (affine tensor elements): adj(J) row-major, det(J)
  const int G = dim * dim + 1;
  c->geo_h.assign((size_t)G * NE, 0.0);
  double hmin = 1e300;
  for (int64_t e = 0; e < NE; e++) {
    const double* X = &c->coords[e * nv * dim];
    double J[9];
    int cols[3];
    if (dim == 2) { cols[0] = 1; cols[1] = 3; }
    else { cols[0] = 1; cols[1] = 3; cols[2] = 4; }
    for (int i = 0; i < dim; i++)
      for (int r = 0; r < dim; r++) J[i * dim + r] = X[cols[r] * dim + i] - X[0 * dim + i];
    // affinity: every vertex equals X0 + J * (reference vertex)
    double scale = 0.0;
    for (int i = 0; i < dim * dim; i++) scale = std::max(scale, std::fabs(J[i]));
    for (int v = 0; v < nv; v++) {
      int ix = (v == 1 || v == 2 || v == 5 || v == 6) ? 1 : 0;
      int iy = (v == 2 || v == 3 || v == 6 || v == 7) ? 1 : 0;
      int iz = v >= 4 ? 1 : 0;
      int ref[3] = {ix, iy, iz};
      for (int i = 0; i < dim; i++) {
        double xv = X[0 * dim + i];
        for (int r = 0; r < dim; r++) xv += J[i * dim + r] * ref[r];
        if (std::fabs(xv - X[v * dim + i]) > 1e-10 * scale) c->affine = 0;
      }
    }
    double A[9], det;
    if (dim == 2) {
      A[0] = J[3]; A[1] = -J[1]; A[2] = -J[2]; A[3] = J[0];
      det = J[0] * J[3] - J[1] * J[2];
    } else {
      A[0] = J[4] * J[8] - J[5] * J[7];
      A[1] = J[2] * J[7] - J[1] * J[8];
      A[2] = J[1] * J[5] - J[2] * J[4];
      A[3] = J[5] * J[6] - J[3] * J[8];
      A[4] = J[0] * J[8] - J[2] * J[6];
      A[5] = J[2] * J[3] - J[0] * J[5];
      A[6] = J[3] * J[7] - J[4] * J[6];
      A[7] = J[1] * J[6] - J[0] * J[7];
      A[8] = J[0] * J[4] - J[1] * J[3];
      det = J[0] * A[0] + J[1] * A[3] + J[2] * A[6];
    }
    if (!(det > 0.0)) return fail(RKDG_E_ARG, "element " + std::to_string(e) + " has non-positive Jacobian");
    for (int i = 0; i < dim * dim; i++) c->geo_h[i * NE + e] = A[i];
    c->geo_h[(G - 1) * NE + e] = det;
    hmin = std::min(hmin, min_singular(dim, J));
  }
  if (!c->affine) return fail(RKDG_E_UNSUPPORTED, "non-affine tensor elements are not built into this library");
  c->hmin = hmin;

  // faces: sort element faces by their sorted topological vertex sets
  const int nf = c->nf, nvf = c->nvf;
  std::vector<FaceKey> keys((size_t)NE * nf);
  for (int64_t e = 0; e < NE; e++)
    for (int lf = 0; lf < nf; lf++) {
      FaceKey& k = keys[e * nf + lf];
      k.k = {-1, -1, -1, -1};
      for (int sa = 0; sa < 2; sa++)
        for (int sb = 0; sb < (dim == 3 ? 2 : 1); sb++) {
          int v = face_corner_vertex(dim, lf, sa, sb);
          k.k[sa + 2 * sb] = m->elem_verts[e * nv + v];
        }
      std::sort(k.k.begin(), k.k.begin() + nvf);
      k.e = e;
      k.lf = lf;
    }
  std::sort(keys.begin(), keys.end());
  std::vector<int> nbr((size_t)nf * NE, -1 - RKDG_MAX_BC), fcode((size_t)nf * NE, 0);
  auto corner_topo = [&](int64_t e, int lf, int sa, int sb) {
    return m->elem_verts[e * nv + face_corner_vertex(dim, lf, sa, sb)];
  };
  auto orient = [&](int64_t ea, int lfa, int64_t eb, int lfb) -> int {
    for (int o = 0; o < 8; o++) {
      if (dim == 2 && (o & 5)) continue;
      bool okk = true;
      for (int sa = 0; sa < 2 && okk; sa++)
        for (int sb = 0; sb < (dim == 3 ? 2 : 1) && okk; sb++) {
          int s2 = sa, t2 = sb;
          if (o & 1) std::swap(s2, t2);
          if (o & 2) s2 = 1 - s2;
          if (o & 4) t2 = 1 - t2;
          if (corner_topo(eb, lfb, s2, t2) != corner_topo(ea, lfa, sa, sb)) okk = false;
        }
      if (okk) return o;
    }
    return -1;
  };
  std::vector<std::pair<std::array<int64_t, 4>, std::pair<int64_t, int>>> open;
  size_t i = 0;
  while (i < keys.size()) {
    size_t j = i + 1;
    while (j < keys.size() && keys[j].k == keys[i].k) j++;
    if (j - i > 2) return fail(RKDG_E_ARG, "face shared by more than two elements");
    if (j - i == 2) {
      const FaceKey& a = keys[i];
      const FaceKey& b = keys[i + 1];
      if (a.e == b.e) return fail(RKDG_E_ARG, "element adjacent to itself (periodic direction too short)");
      int oab = orient(a.e, a.lf, b.e, b.lf), oba = orient(b.e, b.lf, a.e, a.lf);
      if (oab < 0 || oba < 0) return fail(RKDG_E_ARG, "could not match face orientation");
      nbr[a.lf * NE + a.e] = (int)b.e;
      nbr[b.lf * NE + b.e] = (int)a.e;
      fcode[a.lf * NE + a.e] = b.lf | (oab << 3) | (1 << 6);  // a has the lower id: Elem1
      fcode[b.lf * NE + b.e] = a.lf | (oba << 3);
      c->n_int++;
    } else {
      open.push_back({keys[i].k, {keys[i].e, keys[i].lf}});
    }
    i = j;
  }
  // boundary elements -> open faces
  std::vector<std::pair<std::array<int64_t, 4>, int>> bk(m->num_bdr);
  for (int64_t b = 0; b < m->num_bdr; b++) {
    std::array<int64_t, 4> k = {-1, -1, -1, -1};
    for (int q = 0; q < nvf; q++) k[q] = m->bdr_verts[b * nvf + q];
    std::sort(k.begin(), k.begin() + nvf);
    bk[b] = {k, m->bdr_attr[b]};
  }
  std::sort(bk.begin(), bk.end());
  for (auto& of : open) {
    auto it = std::lower_bound(bk.begin(), bk.end(), std::make_pair(of.first, -2147483647));
    int slot = RKDG_MAX_BC;  // no boundary integrator: no flux (NonlinearForm skips it)
    if (it != bk.end() && it->first == of.first) {
      for (int q = 0; q < g.num_bc; q++)
        if (g.bc_attr[q] == it->second) slot = q;
    }
    nbr[of.second.second * NE + of.second.first] = -1 - slot;
    c->n_bdr++;
  }
  (void)NE_;

  HIP_TRY(hipMalloc(&c->d_geo, sizeof(double) * G * NE));
  HIP_TRY(hipMemcpy(c->d_geo, c->geo_h.data(), sizeof(double) * G * NE, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_nbr, sizeof(int) * nf * NE));
  HIP_TRY(hipMemcpy(c->d_nbr, nbr.data(), sizeof(int) * nf * NE, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_fcode, sizeof(int) * nf * NE));
  HIP_TRY(hipMemcpy(c->d_fcode, fcode.data(), sizeof(int) * nf * NE, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_lam, sizeof(double) * 8));
  HIP_TRY(hipMemset(c->d_lam, 0, sizeof(double) * 8));
  HIP_TRY(hipMalloc(&c->d_err, sizeof(long long) * 4));
  HIP_TRY(hipMemset(c->d_err, 0, sizeof(long long) * 4));
  HIP_TRY(hipMalloc(&c->d_means, sizeof(double) * c->neq * NE));
  HIP_TRY(hipMalloc(&c->d_tb, sizeof(TensorTables)));
  HIP_TRY(hipMemcpy(c->d_tb, &c->tb, sizeof(TensorTables), hipMemcpyHostToDevice));
  return RKDG_OK;
}

constexpr size_t kMaxProfiled = 8192;

// returns (start, stop) event indices or (-1, -1) when profiling is off/full
std::pair<int, int> prof_begin(rkdg_ctx* c, std::vector<std::pair<int, int>>& list) {
  if (!c->prof || list.size() >= kMaxProfiled) return {-1, -1};
  size_t need = (c->ev_rhs.size() + c->ev_lim.size() + 1) * 2;
  while (c->ev_pool.size() < need) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return {-1, -1};
    c->ev_pool.push_back(e);
  }
  std::pair<int, int> pr{(int)need - 2, (int)need - 1};
  (void)hipEventRecord(c->ev_pool[pr.first], c->stream);
  list.push_back(pr);
  return pr;
}
void prof_end(rkdg_ctx* c, std::pair<int, int> pr) {
  if (pr.second >= 0) (void)hipEventRecord(c->ev_pool[pr.second], c->stream);
}

hipError_t rhs_launch(rkdg_ctx* c, const RhsArgs& a) {
  auto pr = prof_begin(c, c->ev_rhs);
  hipError_t e = launch_tensor_rhs(c->dim, c->p, a, c->sp, c->d_tb, c->n_bdr > 0, c->stream);
  prof_end(c, pr);
  return e;
}

RhsArgs base_args(rkdg_ctx* c, const double* x) {
  RhsArgs a{};
  a.x = x;
  a.k = nullptr;
  a.nout = 0;
  a.means = nullptr;
  a.means_of = -1;
  a.geo = c->d_geo;
  a.nbr = c->d_nbr;
  a.fcode = c->d_fcode;
  a.NE = c->NE;
  a.lam = c->d_lam;
  a.err = c->d_err;
  a.face_only = 0;
  return a;
}

int check_device_error(rkdg_ctx* c, int64_t* info) {
  long long h[4];
  HIP_TRY(hipMemcpyAsync(h, c->d_err, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (h[0]) {
    if (info) {
      info[0] = h[1];
      info[1] = h[2];
    }
    HIP_TRY(hipMemsetAsync(c->d_err, 0, sizeof(long long) * 4, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return fail(RKDG_E_UNPHYSICAL, "unphysical state at a face between elements " + std::to_string(h[1]) +
                                       " and " + std::to_string(h[2]));
  }
  return RKDG_OK;
}

int read_lambda(rkdg_ctx* c, int slot, double* lam) {
  HIP_TRY(hipMemcpyAsync(lam, c->d_lam + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return check_device_error(c, nullptr);
}

int limit_impl(rkdg_ctx* c, double* x, int8_t* troubled, bool have_means) {
  if (!have_means) {
    MeansArgs ma{x, c->d_means, c->d_geo, c->NE};
    HIP_TRY(launch_tensor_means(c->dim, c->p, ma, c->d_tb, c->stream));
  }
  LimitArgs la{x, c->d_means, c->d_geo, c->d_nbr, c->d_fcode, c->NE, troubled};
  auto pr = prof_begin(c, c->ev_lim);
  hipError_t le = launch_tensor_limit(c->dim, c->p, la, c->sp, c->d_tb, c->stream);
  prof_end(c, pr);
  HIP_TRY(le);
  return RKDG_OK;
}

int ensure_stage_buffers(rkdg_ctx* c, int n) {
  const int64_t sz = (int64_t)c->neq * c->ndofs;
  for (int i = 0; i < n; i++)
    if (!c->d_stage[i]) HIP_TRY(hipMalloc(&c->d_stage[i], sizeof(double) * sz));
  return RKDG_OK;
}

}  // namespace

// ============================================================================
extern "C" {

const char* rkdg_last_error(void) { return g_last_error.c_str(); }

int rkdg_create(const rkdg_mesh_desc* mesh, const rkdg_scheme_desc* scheme, rkdg_ctx** out) {
  if (!mesh || !scheme || !out) return fail(RKDG_E_ARG, "null argument");
  *out = nullptr;
  rkdg_ctx* c = new rkdg_ctx();
  int rc = build_ctx(c, mesh, scheme);
  if (rc != RKDG_OK) {
    std::string keep = g_last_error;
    free_ctx(c);
    g_last_error = keep;
    return rc;
  }
  *out = c;
  return RKDG_OK;
}

int rkdg_destroy(rkdg_ctx* ctx) {
  if (ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    free_ctx(ctx);
  }
  return RKDG_OK;
}

int rkdg_get_info(rkdg_ctx* c, rkdg_info* info) {
  if (!c || !info) return fail(RKDG_E_ARG, "null argument");
  info->dim = c->dim;
  info->num_equations = c->neq;
  info->nodes_per_element = c->nd;
  info->faces_per_element = c->nf;
  info->num_elements = c->NE;
  info->num_dofs = c->ndofs;
  info->state_size = (int64_t)c->neq * c->ndofs;
  info->num_interior_faces = c->n_int;
  info->num_boundary_faces = c->n_bdr;
  info->affine = c->affine;
  return RKDG_OK;
}

int rkdg_set_stream(rkdg_ctx* c, void* stream) {
  if (!c) return fail(RKDG_E_ARG, "null context");
  c->stream = (hipStream_t)stream;
  return RKDG_OK;
}

int rkdg_rhs(rkdg_ctx* c, const double* x, double* y, double* lambda_max) {
  if (!c || !x || !y) return fail(RKDG_E_ARG, "null argument");
  HIP_TRY(launch_fill(c->d_lam, 0.0, 1, c->stream));
  RhsArgs a = base_args(c, x);
  a.k = y;
  HIP_TRY(rhs_launch(c, a));
  if (lambda_max) return read_lambda(c, 0, lambda_max);
  return RKDG_OK;
}

int rkdg_face_lambda(rkdg_ctx* c, const double* x, double* lambda_max) {
  if (!c || !x || !lambda_max) return fail(RKDG_E_ARG, "null argument");
  HIP_TRY(launch_fill(c->d_lam, 0.0, 1, c->stream));
  RhsArgs a = base_args(c, x);
  a.face_only = 1;
  HIP_TRY(rhs_launch(c, a));
  return read_lambda(c, 0, lambda_max);
}

int rkdg_limit(rkdg_ctx* c, double* x, int8_t* troubled) {
  if (!c || !x) return fail(RKDG_E_ARG, "null argument");
  return limit_impl(c, x, troubled, false);
}

int rkdg_means(rkdg_ctx* c, const double* x, double* means) {
  if (!c || !x || !means) return fail(RKDG_E_ARG, "null argument");
  MeansArgs ma{x, means, c->d_geo, c->NE};
  HIP_TRY(launch_tensor_means(c->dim, c->p, ma, c->d_tb, c->stream));
  return RKDG_OK;
}

// ExplicitRKLimitedSolver::Step (rk_explicit_limited.cpp:30-82). Stage i reads
// its (limited) input and produces, in the epilogue of the RHS kernel, every
// pending combination that needs k_i: W_j += (a_ji dt) k_i for later stages j
// and acc += (b_i dt) k_i, in the reference's association order. W_1 (stage 0)
// and the final x are the only vectors written whole; the means of the next
// stage input come out of the same epilogue for the limiter.
int rkdg_step(rkdg_ctx* c, double* x, double t, double dt, int s, const double* a, const double* b,
              const double* cc, double* lambda_max) {
  (void)t;
  (void)cc;
  if (!c || !x || !a || !b) return fail(RKDG_E_ARG, "null argument");
  if (s < 1 || s > 4) return fail(RKDG_E_ARG, "1 <= s <= 4 stages supported");
  // buffers: d_stage[0..s-2] = W_1..W_{s-1}, d_stage[3] = accumulator
  int rc = ensure_stage_buffers(c, 4);
  if (rc) return rc;
  double* W[4] = {nullptr, c->d_stage[0], c->d_stage[1], c->d_stage[2]};
  double* acc = c->d_stage[3];
  HIP_TRY(launch_fill(c->d_lam, 0.0, 8, c->stream));
  auto A_ = [&](int i, int j) { return a[i * (i - 1) / 2 + j]; };
  for (int i = 0; i < s; i++) {
    const double* in = i == 0 ? x : W[i];
    RhsArgs ra = base_args(c, in);
    ra.lam = c->d_lam + i;
    int n = 0;
    for (int j = i + 1; j < s; j++) {
      ra.outs[n].out = W[j];
      ra.outs[n].base = i == 0 ? nullptr : W[j];
      ra.outs[n].coef = A_(j, i) * dt;
      n++;
    }
    const bool last = i == s - 1;
    ra.outs[n].out = last ? (s == 1 ? acc : x) : acc;
    ra.outs[n].base = i == 0 ? nullptr : acc;
    ra.outs[n].coef = b[i] * dt;
    n++;
    ra.nout = n;
    ra.means = c->d_means;
    ra.means_of = last ? n - 1 : 0;  // next stage input W_{i+1} is outs[0]; final is the last
    HIP_TRY(rhs_launch(c, ra));
    if (last && s == 1) HIP_TRY(launch_copy(x, acc, (int64_t)c->neq * c->ndofs, c->stream));
    double* lim = last ? x : W[i + 1];
    rc = limit_impl(c, lim, nullptr, true);
    if (rc) return rc;
  }
  if (lambda_max) return read_lambda(c, s - 1, lambda_max);
  return RKDG_OK;
}

int rkdg_hmin(rkdg_ctx* c, double* hmin) {
  if (!c || !hmin) return fail(RKDG_E_ARG, "null argument");
  *hmin = c->hmin;
  return RKDG_OK;
}

int rkdg_node_coords(rkdg_ctx* c, double* xyz) {
  if (!c || !xyz) return fail(RKDG_E_ARG, "null argument");
  const int N = c->p + 1, dim = c->dim;
  const int64_t NE = c->NE;
  for (int64_t e = 0; e < NE; e++) {
    const double* X = &c->coords[e * c->nv * dim];
    for (int n = 0; n < c->nd; n++) {
      int ix = n % N, iy = (n / N) % N, iz = n / (N * N);
      double xi[3] = {c->tb.x[ix], c->tb.x[iy], dim == 3 ? c->tb.x[iz] : 0.0};
      for (int d = 0; d < dim; d++) {
        double v = 0.0;
        for (int q = 0; q < c->nv; q++) {
          int rx = (q == 1 || q == 2 || q == 5 || q == 6) ? 1 : 0;
          int ry = (q == 2 || q == 3 || q == 6 || q == 7) ? 1 : 0;
          int rz = q >= 4 ? 1 : 0;
          double Nq = (rx ? xi[0] : 1 - xi[0]) * (ry ? xi[1] : 1 - xi[1]);
          if (dim == 3) Nq *= rz ? xi[2] : 1 - xi[2];
          v += Nq * X[q * dim + d];
        }
        xyz[(e * c->nd + n) * dim + d] = v;
      }
    }
  }
  return RKDG_OK;
}

int rkdg_profile_enable(rkdg_ctx* c, int enable) {
  if (!c) return fail(RKDG_E_ARG, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->prof = enable != 0;
  c->ev_rhs.clear();
  c->ev_lim.clear();
  return RKDG_OK;
}

int rkdg_profile_read(rkdg_ctx* c, double* rhs_ms, int64_t* rhs_n, double* lim_ms, int64_t* lim_n) {
  if (!c) return fail(RKDG_E_ARG, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  double tr = 0.0, tl = 0.0;
  for (auto& pr : c->ev_rhs) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev_pool[pr.first], c->ev_pool[pr.second]));
    tr += ms;
  }
  for (auto& pr : c->ev_lim) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev_pool[pr.first], c->ev_pool[pr.second]));
    tl += ms;
  }
  if (rhs_ms) *rhs_ms = tr;
  if (rhs_n) *rhs_n = (int64_t)c->ev_rhs.size();
  if (lim_ms) *lim_ms = tl;
  if (lim_n) *lim_n = (int64_t)c->ev_lim.size();
  return RKDG_OK;
}

int rkdg_check_error(rkdg_ctx* c, int64_t* info) {
  if (!c) return fail(RKDG_E_ARG, "null context");
  return check_device_error(c, info);
}

}  // extern "C"
