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
#include <limits>
#include <map>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rkdg_capi.h"
#include "rkdg_adjoint.hpp"
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
  int64_t err_pair[2] = {-1, -1};  // element pair of the last reported device error
  double* d_means = nullptr;
  // tensor contexts: the second means buffer and the per-element limiter
  // factors of the steps that apply the limiter on load
  double* d_means2 = nullptr;
  double* d_fac = nullptr;
  long long* h_pin = nullptr;    // pinned host words: device error word [4], then the wave speeds [count]
  int* d_order = nullptr;        // tensor contexts: work blocks in Morton order of their centroids
  // two-coloured tensor meshes (bipartite face graph, no boundary faces): the
  // solver and receiver element lists (each in Morton order of the element
  // centroids), each element's receiver slot and the receivers' face fluxes
  int rb = 0;
  int64_t n_solve = 0, n_recv = 0;
  int* d_elist_solve = nullptr;
  int* d_elist_recv = nullptr;
  int* d_fslot = nullptr;
  double* d_fbuf = nullptr;
  TensorTables* d_tb = nullptr;
  int dense = 0;                 // 1: dense small-element kernels (triangles, bilinear quads)
  double* d_dtab = nullptr;      // dense reference tables
  double* d_shu_coef = nullptr;  // Shu indicator: extrapolation weights [(nf+1)*nd][NE]
  int8_t* d_shu_flag = nullptr;  // Shu indicator: troubled elements of the last pass
  std::vector<double> ref_nodes; // reference node coordinates [nd][dim] (dense path)
  double* d_stage[4] = {nullptr, nullptr, nullptr, nullptr};
  double* scratch = nullptr;     // this device's static scratch array (rkdg_step_noalloc)
  // diagnostics: element volumes, the means and per-block partial sums of the
  // totals, the basis table of the last interpolation
  double* d_vol = nullptr;
  double* d_dmeans = nullptr;
  double* d_partial = nullptr;
  double* d_phi = nullptr;
  int phi_cap = 0;
  std::vector<double> coords;  // [NE][nv][dim]
  std::vector<double> geo_h;   // [g][NE]
  int64_t n_int = 0, n_bdr = 0;
  double hmin = 0.0;
  std::vector<double> h_elem;  // GetElementSize(e, 1) per element
  // instrumentation: event pairs around kernel launches
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_rhs, ev_lim;  // (start, stop) pool indices
  // batched ensembles: count disjoint instances in one union mesh
  int count = 1;
  int force_dense = 0;
  int two_colour = 0;            // the solver/receiver form where the mesh allows it
  // non-conforming faces (tensor contexts): the descriptor while the context
  // is built, the element faces coded kNcCode0 - k (slave and master sides)
  // with their neighbours, the slave-face records, the elements with such
  // faces (their items: record * 2 + side) and their flux buffer
  // curved (trilinear) hexahedra: the CurvedLayout records [e][size] and the volumes
  int curved = 0;
  double* d_cgeo = nullptr;
  std::vector<double> vol_h;
  const rkdg_nc_desc* nc_in = nullptr;
  int64_t n_nc = 0, n_ncf = 0, n_nce = 0;
  int* d_ncnb = nullptr;
  int* d_ncslot = nullptr;
  double* d_ncbuf = nullptr;
  int* d_ncrec = nullptr;
  double* d_ncgeo = nullptr;
  int* d_nce = nullptr;
  int* d_ncptr = nullptr;
  int* d_ncitem = nullptr;
  int aligned = 0;               // tensor mesh: every interior face joins (r, c) to (r, 1-c), identity orientation
  int grid_cap = 0;              // persistent grid of the dense right-hand-side kernel
  std::vector<int64_t> inst_off; // [count+1] element offsets
  int* d_inst = nullptr;         // [NE] element -> instance
  double* d_bcp = nullptr;       // [count][RKDG_MAX_BC][8]
  int* d_active = nullptr;       // [count]
  double *d_t = nullptr, *d_dt = nullptr, *d_dtr = nullptr, *d_tf = nullptr, *d_hmin = nullptr;
  long long* d_steps = nullptr;
  int* d_nact = nullptr;
  double cfl = 0.0;
  long long max_steps = 0;
  hipStream_t bstream = nullptr;    // stream of batch steps (graph capture needs a created stream)
  hipGraphExec_t gexec = nullptr;   // captured batch step
  std::vector<double> gkey;         // what the graph was captured for: x, tableau
  int use_graph = 1;
  // adjoint work space (dense 2D path)
  double* d_ameans = nullptr;     // means of a pre-limit state [neq*NE]
  double* d_meanbar = nullptr;    // [neq*NE]
  int64_t* d_send_to = nullptr;   // [NE]
  double* d_send_val = nullptr;   // [NE]
  double* d_pbar = nullptr;       // [RKDG_MAX_BC*8]
  double* d_J = nullptr;          // [1]
  std::vector<double*> d_adj;     // adjoint vectors: u0bar, kbar[4], tmp[2]
  std::vector<double*> d_rec;     // recorded steps of a segment
  std::vector<double*> d_cp;      // checkpoints
  std::vector<int> nbr_h;         // host copy of the neighbour table [lf][NE]
  int obj_slot_cached = -1;       // boundary slot of d_obj_elems
  int* d_obj_elems = nullptr;     // elements with a face on that slot
  int n_obj_elems = 0;
};

namespace {

// hipFree of a null pointer is a no-op but still initialises the runtime: a
// context torn down after a failed argument check touches no device
template <typename Tp>
void dfree(Tp* p) {
  if (p) (void)hipFree((void*)p);
}

void free_ctx(rkdg_ctx* c) {
  if (!c) return;
  dfree(c->d_geo);
  dfree(c->d_nbr);
  dfree(c->d_fcode);
  dfree(c->d_lam);
  dfree(c->d_err);
  if (c->h_pin) (void)hipHostFree(c->h_pin);
  dfree(c->d_means);
  dfree(c->d_means2);
  dfree(c->d_fac);
  dfree(c->d_cgeo);
  dfree(c->d_ncnb);
  dfree(c->d_ncslot);
  dfree(c->d_ncbuf);
  dfree(c->d_ncrec);
  dfree(c->d_ncgeo);
  dfree(c->d_nce);
  dfree(c->d_ncptr);
  dfree(c->d_ncitem);
  dfree(c->d_order);
  dfree(c->d_elist_solve);
  dfree(c->d_elist_recv);
  dfree(c->d_fslot);
  dfree(c->d_fbuf);
  dfree(c->d_tb);
  dfree(c->d_dtab);
  dfree(c->d_shu_coef);
  if (c->d_shu_flag) (void)hipFree(c->d_shu_flag);
  for (auto& p : c->d_stage) dfree(p);
  for (auto& e : c->ev_pool) (void)hipEventDestroy(e);
  dfree(c->d_inst);
  dfree(c->d_bcp);
  dfree(c->d_active);
  dfree(c->d_t);
  dfree(c->d_dt);
  dfree(c->d_dtr);
  dfree(c->d_tf);
  dfree(c->d_hmin);
  dfree(c->d_steps);
  dfree(c->d_nact);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->bstream) (void)hipStreamDestroy(c->bstream);
  dfree(c->d_ameans);
  dfree(c->d_meanbar);
  dfree(c->d_send_to);
  dfree(c->d_send_val);
  dfree(c->d_pbar);
  dfree(c->d_J);
  for (auto p : c->d_adj) dfree(p);
  for (auto p : c->d_rec) dfree(p);
  for (auto p : c->d_cp) dfree(p);
  dfree(c->d_obj_elems);
  dfree(c->d_vol);
  dfree(c->d_dmeans);
  dfree(c->d_partial);
  dfree(c->d_phi);
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

// ---------------------------------------------------------------------------
// An element is taken as affine (parallelogram, parallelepiped) when its last
// vertices sit where the affine map through the first ones puts them, to
// within rounding of the coordinates: 1e-13 of the element's size or 16 ulps
// of its largest coordinate. A larger defect is geometry the reference maps
// exactly (bilinear/trilinear Jacobians per point); treating it as affine
// would move the right-hand side by about defect / size.
static double affine_tol(double size, const double* X, int n) {
  double big = 0.0;
  for (int k = 0; k < n; k++) big = std::max(big, std::fabs(X[k]));
  return std::max(1e-13 * size, 16.0 * big * std::numeric_limits<double>::epsilon());
}

// Dense small-element path: affine triangles P_p and bilinear quadrilaterals Q_p.

bool invert_small(int n, const double* A, double* Ai) {
  std::vector<double> M(A, A + n * n), I(n * n, 0.0);
  for (int i = 0; i < n; i++) I[i * n + i] = 1.0;
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(M[r * n + c]) > std::fabs(M[piv * n + c])) piv = r;
    if (M[piv * n + c] == 0.0) return false;
    for (int k = 0; k < n; k++) {
      std::swap(M[c * n + k], M[piv * n + k]);
      std::swap(I[c * n + k], I[piv * n + k]);
    }
    const double inv = 1.0 / M[c * n + c];
    for (int k = 0; k < n; k++) {
      M[c * n + k] *= inv;
      I[c * n + k] *= inv;
    }
    for (int r = 0; r < n; r++) {
      if (r == c || M[r * n + c] == 0.0) continue;
      const double f = M[r * n + c];
      for (int k = 0; k < n; k++) {
        M[r * n + k] -= f * M[c * n + k];
        I[r * n + k] -= f * I[c * n + k];
      }
    }
  }
  std::memcpy(Ai, I.data(), sizeof(double) * n * n);
  return true;
}

// nodal basis of a reference element: values and gradients at a point
struct RefBasis {
  int tri = 0, p = 1, nd = 4;
  std::vector<double> x1;                 // 1D GL nodes (quads)
  std::vector<std::array<int, 2>> mono;   // triangles: monomial exponents
  std::vector<double> coef;               // triangles: coef[m*nd + k]
  std::vector<double> nodes;              // [nd][2]
  void init(int is_tri, int order) {
    tri = is_tri;
    p = order;
    const int N = p + 1;
    x1.assign(N, 0.0);
    std::vector<double> w(N);
    gl_rule(N, x1.data(), w.data());
    if (!tri) {
      nd = N * N;
      nodes.resize(2 * nd);
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          nodes[2 * (i + N * j)] = x1[i];
          nodes[2 * (i + N * j) + 1] = x1[j];
        }
      return;
    }
    // L2_TriangleElement nodes from Gauss-Legendre open points (j outer, i inner)
    nd = N * (N + 1) / 2;
    nodes.clear();
    for (int j = 0; j <= p; j++)
      for (int i = 0; i + j <= p; i++) {
        const double wsum = x1[i] + x1[j] + x1[p - i - j];
        nodes.push_back(x1[i] / wsum);
        nodes.push_back(x1[j] / wsum);
      }
    for (int a = 0; a <= p; a++)
      for (int b = 0; a + b <= p; b++) mono.push_back({a, b});
    std::vector<double> V(nd * nd);
    for (int k = 0; k < nd; k++)
      for (int m2 = 0; m2 < nd; m2++)
        V[k * nd + m2] = std::pow(nodes[2 * k], mono[m2][0]) * std::pow(nodes[2 * k + 1], mono[m2][1]);
    coef.assign(nd * nd, 0.0);
    invert_small(nd, V.data(), coef.data());
  }
  void eval(double x, double y, double* phi, double* dx, double* dy) const {
    if (tri) {
      for (int k = 0; k < nd; k++) {
        double v = 0, gx = 0, gy = 0;
        for (int m2 = 0; m2 < nd; m2++) {
          const int a = mono[m2][0], b = mono[m2][1];
          const double c2 = coef[m2 * nd + k];
          v += c2 * std::pow(x, a) * std::pow(y, b);
          if (a > 0) gx += c2 * a * std::pow(x, a - 1) * std::pow(y, b);
          if (b > 0) gy += c2 * b * std::pow(x, a) * std::pow(y, b - 1);
        }
        phi[k] = v;
        if (dx) dx[k] = gx;
        if (dy) dy[k] = gy;
      }
      return;
    }
    const int N = p + 1;
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const int k = i + N * j;
        const double lx = lag(x1.data(), N, i, x), ly = lag(x1.data(), N, j, y);
        phi[k] = lx * ly;
        if (dx) dx[k] = dlag(x1.data(), N, i, x) * ly;
        if (dy) dy[k] = lx * dlag(x1.data(), N, j, y);
      }
  }
};

// exact rules on the reference triangle (collapsed Gauss-Legendre) and square
void tri_rule(int n, std::vector<double>& pts, std::vector<double>& w) {
  std::vector<double> x(n), wx(n);
  gl_rule(n, x.data(), wx.data());
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      pts.push_back(x[i]);
      pts.push_back((1.0 - x[i]) * x[j]);
      w.push_back(wx[i] * wx[j] * (1.0 - x[i]));
    }
}
void quad_rule(int n, std::vector<double>& pts, std::vector<double>& w) {
  std::vector<double> x(n), wx(n);
  gl_rule(n, x.data(), wx.data());
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      pts.push_back(x[i]);
      pts.push_back(x[j]);
      w.push_back(wx[i] * wx[j]);
    }
}

const int kTriEdge[3][2] = {{0, 1}, {1, 2}, {2, 0}};
const int kQuadEdge[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
const double kTriV[3][2] = {{0, 0}, {1, 0}, {0, 1}};
const double kQuadV[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};

// bilinear quad map: Jacobian at (x, y) from vertex coordinates X[4][2]; J row-major.
// Written on the edge vectors, dX/dx = (1-y)(X1-X0) + y(X2-X3) and
// dX/dy = (1-x)(X3-X0) + x(X2-X1): the differences of neighbouring vertices are
// exact, so J keeps full relative precision however far the element sits from
// the origin. The vertex-weighted sum sum_v X_v dN_v (MFEM's PointMat * dshape)
// loses the ratio |X| / size in relative precision (1e-14 on a mesh of
// 5e-4 elements at x = 0.05), which the cancelling volume and face terms of
// the right-hand side amplify.
void bilinear_jac(const double* X, double x, double y, double* J) {
  for (int i = 0; i < 2; i++) {
    J[i * 2 + 0] = (1 - y) * (X[2 + i] - X[i]) + y * (X[4 + i] - X[6 + i]);
    J[i * 2 + 1] = (1 - x) * (X[6 + i] - X[i]) + x * (X[4 + i] - X[2 + i]);
  }
}

int build_dense(rkdg_ctx* c, const rkdg_mesh_desc* m) {
  const bool tri = m->geom == RKDG_GEOM_TRI;
  if (!dense_supported(m->geom, c->p)) return fail(RKDG_E_UNSUPPORTED, "polynomial order not instantiated for this element");
  c->dense = 1;
  c->dim = 2;
  c->neq = 4;
  c->nv = tri ? 3 : 4;
  c->nvf = 2;
  c->nf = tri ? 3 : 4;
  RefBasis rb;
  rb.init(tri, c->p);
  const int nd = rb.nd, nf = c->nf, p = c->p;
  c->nd = nd;
  c->ndofs = c->NE * nd;
  c->ref_nodes = rb.nodes;
  const int64_t NE = c->NE;
  const int NQ = p + 1;

  // reference tables, laid out exactly as DenseTab<GEOM, P>
  const int tsize = dense_table_size(m->geom, p);
  std::vector<double> tab(tsize, 0.0);
  size_t o = 0;
  double* K = &tab[o];
  o += 2 * nd * nd;
  double* Phi = &tab[o];
  o += nf * NQ * nd;
  double* WQ = &tab[o];
  o += NQ;
  double* Minv = &tab[o];
  o += nd * nd;
  double* mhat = &tab[o];
  o += nd;
  double* PhiI = &tab[o];
  o += nf * 2 * nd;
  double* PhiV = &tab[o];
  o += c->nv * nd;
  if ((int)o != tsize) return fail(RKDG_E_ARG, "dense table layout mismatch");
  std::vector<double> phi(nd), dx(nd), dy(nd);
  std::vector<double> gw1(p + 1), gx1(p + 1);
  gl_rule(p + 1, gx1.data(), gw1.data());
  if (tri) {
    std::vector<double> pts, w;
    tri_rule(p + 2, pts, w);
    std::vector<double> M(nd * nd, 0.0);
    for (size_t q = 0; q < w.size(); q++) {
      rb.eval(pts[2 * q], pts[2 * q + 1], phi.data(), dx.data(), dy.data());
      for (int j2 = 0; j2 < nd; j2++)
        for (int k = 0; k < nd; k++) {
          K[(0 * nd + j2) * nd + k] += w[q] * phi[k] * dx[j2];
          K[(1 * nd + j2) * nd + k] += w[q] * phi[k] * dy[j2];
          M[j2 * nd + k] += w[q] * phi[j2] * phi[k];
        }
      for (int k = 0; k < nd; k++) mhat[k] += w[q] * phi[k];
    }
    invert_small(nd, M.data(), Minv);
  } else {
    // collocated weak derivative on Gauss-Legendre nodes: K_r(j,k) = w_k d_r phi_j(x_k)
    for (int k = 0; k < nd; k++) {
      const int i = k % (p + 1), jq = k / (p + 1);
      const double wk = gw1[i] * gw1[jq];
      rb.eval(rb.nodes[2 * k], rb.nodes[2 * k + 1], phi.data(), dx.data(), dy.data());
      for (int j2 = 0; j2 < nd; j2++) {
        K[(0 * nd + j2) * nd + k] = wk * dx[j2];
        K[(1 * nd + j2) * nd + k] = wk * dy[j2];
      }
    }
  }
  std::vector<double> sq(NQ), wq(NQ);
  gl_rule(NQ, sq.data(), wq.data());
  for (int q = 0; q < NQ; q++) WQ[q] = wq[q];
  double s2[2], w2[2];
  gl_rule(2, s2, w2);
  const double(*RV)[2] = tri ? kTriV : kQuadV;
  const int(*ED)[2] = tri ? kTriEdge : kQuadEdge;
  for (int f = 0; f < nf; f++) {
    const double* a = RV[ED[f][0]];
    const double* b = RV[ED[f][1]];
    for (int q = 0; q < NQ; q++) {
      rb.eval(a[0] + sq[q] * (b[0] - a[0]), a[1] + sq[q] * (b[1] - a[1]), phi.data(), nullptr, nullptr);
      for (int k = 0; k < nd; k++) Phi[(f * NQ + q) * nd + k] = phi[k];
    }
    for (int q = 0; q < 2; q++) {
      rb.eval(a[0] + s2[q] * (b[0] - a[0]), a[1] + s2[q] * (b[1] - a[1]), phi.data(), nullptr, nullptr);
      for (int k = 0; k < nd; k++) PhiI[(f * 2 + q) * nd + k] = phi[k];
    }
  }
  for (int v = 0; v < c->nv; v++) {
    rb.eval(RV[v][0], RV[v][1], phi.data(), nullptr, nullptr);
    for (int k = 0; k < nd; k++) PhiV[v * nd + k] = phi[k];
  }

  // per-element geometry
  const int GS = dense_geo_size(m->geom, p);
  const int G_NRM = tri ? 5 : 4 * nd + nd * nd + nd;
  c->geo_h.assign((size_t)GS * NE, 0.0);
  std::vector<double> Mq, wqq;
  std::vector<double> mpts;
  if (!tri) quad_rule(p + 2, mpts, Mq);
  c->h_elem.assign(NE, 0.0);
  double hmin = 1e300;
  for (int64_t e = 0; e < NE; e++) {
    const double* X = &c->coords[e * c->nv * 2];
    auto G = [&](int gi) -> double& { return c->geo_h[(size_t)gi * NE + e]; };
    if (tri) {
      const double J[4] = {X[2] - X[0], X[4] - X[0], X[3] - X[1], X[5] - X[1]};
      const double det = J[0] * J[3] - J[1] * J[2];
      if (!(det > 0.0)) return fail(RKDG_E_ARG, "element " + std::to_string(e) + " has non-positive Jacobian");
      G(0) = J[3];
      G(1) = -J[1];
      G(2) = -J[2];
      G(3) = J[0];
      G(4) = det;
      // GetElementSize(i,1) on the equilateral (perfect) triangle
      const double s3 = std::sqrt(3.0);
      const double Pm[4] = {1.0, -1.0 / s3, 0.0, 2.0 / s3};
      double Q[4] = {J[0] * Pm[0] + J[1] * Pm[2], J[0] * Pm[1] + J[1] * Pm[3],
                     J[2] * Pm[0] + J[3] * Pm[2], J[2] * Pm[1] + J[3] * Pm[3]};
      c->h_elem[e] = min_singular(2, Q);
      hmin = std::min(hmin, c->h_elem[e]);
    } else {
      std::vector<double> Mm(nd * nd, 0.0), mw(nd, 0.0);
      for (int k = 0; k < nd; k++) {
        double J[4];
        bilinear_jac(X, rb.nodes[2 * k], rb.nodes[2 * k + 1], J);
        G(4 * k + 0) = J[3];
        G(4 * k + 1) = -J[1];
        G(4 * k + 2) = -J[2];
        G(4 * k + 3) = J[0];
        if (!(J[0] * J[3] - J[1] * J[2] > 0.0))
          return fail(RKDG_E_ARG, "element " + std::to_string(e) + " has non-positive Jacobian");
      }
      for (size_t q = 0; q < Mq.size(); q++) {
        double J[4];
        bilinear_jac(X, mpts[2 * q], mpts[2 * q + 1], J);
        const double wdet = Mq[q] * (J[0] * J[3] - J[1] * J[2]);
        rb.eval(mpts[2 * q], mpts[2 * q + 1], phi.data(), nullptr, nullptr);
        for (int i2 = 0; i2 < nd; i2++) {
          mw[i2] += wdet * phi[i2];
          for (int k = 0; k < nd; k++) Mm[i2 * nd + k] += wdet * phi[i2] * phi[k];
        }
      }
      std::vector<double> Mi(nd * nd);
      invert_small(nd, Mm.data(), Mi.data());
      for (int i2 = 0; i2 < nd * nd; i2++) G(4 * nd + i2) = Mi[i2];
      for (int i2 = 0; i2 < nd; i2++) G(4 * nd + nd * nd + i2) = mw[i2];
      double Jc[4];
      bilinear_jac(X, 0.5, 0.5, Jc);
      c->h_elem[e] = min_singular(2, Jc);
      hmin = std::min(hmin, c->h_elem[e]);
    }
    for (int f = 0; f < nf; f++) {
      const double* a = &X[ED[f][0] * 2];
      const double* b = &X[ED[f][1] * 2];
      const double t0 = b[0] - a[0], t1 = b[1] - a[1];
      G(G_NRM + 2 * f) = t1;
      G(G_NRM + 2 * f + 1) = -t0;
    }
  }
  c->hmin = hmin;
  c->affine = tri ? 1 : 0;
  // 1D Gauss-Legendre nodes for node coordinates of quads
  gl_rule(p + 1, c->tb.x, c->tb.w);

  // edges: pair by sorted vertex ids; flip when the neighbour's edge runs the other way
  std::vector<FaceKey> keys((size_t)NE * nf);
  for (int64_t e = 0; e < NE; e++)
    for (int lf = 0; lf < nf; lf++) {
      FaceKey& k = keys[e * nf + lf];
      k.k = {m->elem_verts[e * c->nv + ED[lf][0]], m->elem_verts[e * c->nv + ED[lf][1]], -1, -1};
      if (k.k[0] > k.k[1]) std::swap(k.k[0], k.k[1]);
      k.e = e;
      k.lf = lf;
    }
  std::sort(keys.begin(), keys.end());
  std::vector<int> nbr((size_t)nf * NE, -1 - RKDG_MAX_BC), fcode((size_t)nf * NE, 0);
  std::vector<std::pair<std::array<int64_t, 4>, std::pair<int64_t, int>>> open;
  for (size_t i = 0; i < keys.size();) {
    size_t j2 = i + 1;
    while (j2 < keys.size() && keys[j2].k == keys[i].k) j2++;
    if (j2 - i > 2) return fail(RKDG_E_ARG, "edge shared by more than two elements");
    if (j2 - i == 2) {
      const FaceKey& a = keys[i];
      const FaceKey& b = keys[i + 1];
      if (a.e == b.e) return fail(RKDG_E_ARG, "element adjacent to itself (periodic direction too short)");
      const bool same = m->elem_verts[a.e * c->nv + ED[a.lf][0]] == m->elem_verts[b.e * c->nv + ED[b.lf][0]];
      const int o2 = same ? 0 : 2;
      nbr[a.lf * NE + a.e] = (int)b.e;
      nbr[b.lf * NE + b.e] = (int)a.e;
      fcode[a.lf * NE + a.e] = b.lf | (o2 << 3) | (1 << 6);
      fcode[b.lf * NE + b.e] = a.lf | (o2 << 3);
      c->n_int++;
    } else {
      open.push_back({keys[i].k, {keys[i].e, keys[i].lf}});
    }
    i = j2;
  }
  std::vector<std::pair<std::array<int64_t, 4>, int>> bk(m->num_bdr);
  for (int64_t b = 0; b < m->num_bdr; b++) {
    std::array<int64_t, 4> k = {m->bdr_verts[b * 2], m->bdr_verts[b * 2 + 1], -1, -1};
    if (k[0] > k[1]) std::swap(k[0], k[1]);
    bk[b] = {k, m->bdr_attr[b]};
  }
  std::sort(bk.begin(), bk.end());
  for (auto& of : open) {
    auto it = std::lower_bound(bk.begin(), bk.end(), std::make_pair(of.first, -2147483647));
    int slot = RKDG_MAX_BC;
    if (it != bk.end() && it->first == of.first)
      for (int q = 0; q < c->sp.num_bc; q++)
        if (c->sp.bc_attr[q] == it->second) slot = q;
    nbr[of.second.second * NE + of.second.first] = -1 - slot;
    c->n_bdr++;
  }
  HIP_TRY(hipMalloc(&c->d_geo, sizeof(double) * GS * NE));
  HIP_TRY(hipMemcpy(c->d_geo, c->geo_h.data(), sizeof(double) * GS * NE, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_nbr, sizeof(int) * nf * NE));
  HIP_TRY(hipMemcpy(c->d_nbr, nbr.data(), sizeof(int) * nf * NE, hipMemcpyHostToDevice));
  c->nbr_h = nbr;
  HIP_TRY(hipMalloc(&c->d_fcode, sizeof(int) * nf * NE));
  HIP_TRY(hipMemcpy(c->d_fcode, fcode.data(), sizeof(int) * nf * NE, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_dtab, sizeof(double) * tsize));
  HIP_TRY(hipMemcpy(c->d_dtab, tab.data(), sizeof(double) * tsize, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_lam, sizeof(double) * 8 * c->count));
  HIP_TRY(hipMemset(c->d_lam, 0, sizeof(double) * 8 * c->count));
  HIP_TRY(hipMalloc(&c->d_err, sizeof(long long) * 4));
  HIP_TRY(hipMemset(c->d_err, 0, sizeof(long long) * 4));
  HIP_TRY(hipMalloc(&c->d_means, sizeof(double) * c->neq * NE));
  return RKDG_OK;
}

// Two-colouring of the face graph of a tensor mesh without boundary faces
// (breadth-first, every face joining the two colours). On success the lists of
// the tensor kernel's two-colour form are built and uploaded; a mesh with an
// odd cycle (e.g. a periodic direction of odd length) keeps the one-kernel form.
int build_two_colour(rkdg_ctx* c, const std::vector<int>& nbr) {
  const int64_t NE = c->NE;
  const int nf = c->nf;
  std::vector<int8_t> col((size_t)NE, -1);
  std::vector<int64_t> queue;
  queue.reserve((size_t)NE);
  for (int64_t s0 = 0; s0 < NE; s0++) {
    if (col[s0] >= 0) continue;
    col[s0] = 0;
    queue.clear();
    queue.push_back(s0);
    for (size_t q = 0; q < queue.size(); q++) {
      const int64_t e = queue[q];
      for (int lf = 0; lf < nf; lf++) {
        const int n = nbr[(size_t)lf * NE + e];
        if (n < 0) return RKDG_OK;  // a boundary face: not this form
        if (col[n] < 0) {
          col[n] = (int8_t)(1 - col[e]);
          queue.push_back(n);
        } else if (col[n] == col[e]) {
          return RKDG_OK;  // odd cycle
        }
      }
    }
  }
  // Morton order of the element centroids (as the work blocks of the one-kernel form)
  const int dim = c->dim, nv = c->nv;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  std::vector<double> cen((size_t)NE * dim, 0.0);
  for (int64_t e = 0; e < NE; e++)
    for (int i = 0; i < dim; i++) {
      double m = 0.0;
      for (int v = 0; v < nv; v++) m += c->coords[(e * nv + v) * dim + i];
      cen[e * dim + i] = m / nv;
      lo[i] = std::min(lo[i], cen[e * dim + i]);
      hi[i] = std::max(hi[i], cen[e * dim + i]);
    }
  const int bits = dim == 3 ? 10 : 16;
  const double cells = (double)((1u << bits) - 1);
  std::vector<std::pair<uint64_t, int>> key[2];
  for (int64_t e = 0; e < NE; e++) {
    uint64_t code = 0;
    for (int i = 0; i < dim; i++) {
      const double span = hi[i] > lo[i] ? hi[i] - lo[i] : 1.0;
      const uint64_t q = (uint64_t)((cen[e * dim + i] - lo[i]) / span * cells + 0.5);
      for (int k = 0; k < bits; k++) code |= ((q >> k) & 1ull) << (k * dim + i);
    }
    key[col[e]].push_back({code, (int)e});
  }
  for (auto& k : key) std::stable_sort(k.begin(), k.end());
  std::vector<int> solve(key[0].size()), recv(key[1].size()), slot((size_t)NE, -1);
  for (size_t i = 0; i < key[0].size(); i++) solve[i] = key[0][i].second;
  for (size_t i = 0; i < key[1].size(); i++) {
    recv[i] = key[1][i].second;
    slot[recv[i]] = (int)i;
  }
  c->n_solve = (int64_t)solve.size();
  c->n_recv = (int64_t)recv.size();
  const int N = c->p + 1;
  const int64_t fb = (int64_t)nf * c->neq * (dim == 3 ? N * N : N);
  HIP_TRY(hipMalloc(&c->d_elist_solve, sizeof(int) * std::max<size_t>(1, solve.size())));
  HIP_TRY(hipMemcpy(c->d_elist_solve, solve.data(), sizeof(int) * solve.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_elist_recv, sizeof(int) * std::max<size_t>(1, recv.size())));
  HIP_TRY(hipMemcpy(c->d_elist_recv, recv.data(), sizeof(int) * recv.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_fslot, sizeof(int) * NE));
  HIP_TRY(hipMemcpy(c->d_fslot, slot.data(), sizeof(int) * NE, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_fbuf, sizeof(double) * std::max<int64_t>(1, c->n_recv * fb)));
  c->rb = 1;
  return RKDG_OK;
}

// rkdg_scheme_desc -> SchemeParams, with the checks of the scheme's plug-in
// choices (the reference's case manager stops on an unknown name,
// case_manager.cpp:523-612, 614-820)
static int scheme_params(const rkdg_scheme_desc* s, SchemeParams& g) {
  if (s->num_bc < 0 || s->num_bc > RKDG_MAX_BC) return fail(RKDG_E_ARG, "num_bc out of range");
  if (s->riemann < RKDG_RS_RUSANOV || s->riemann > RKDG_RS_HLLC) return fail(RKDG_E_ARG, "unknown Riemann solver");
  if (s->wave_speeds != RKDG_WS_TORO && s->wave_speeds != RKDG_WS_EINFELDT)
    return fail(RKDG_E_ARG, "unknown wave-speed estimate");
  if (s->indicator < RKDG_IND_NOWHERE || s->indicator > RKDG_IND_SHU) return fail(RKDG_E_ARG, "unknown indicator");
  if (s->phi < RKDG_PHI_MICHALAK || s->phi > RKDG_PHI_VENKAT) return fail(RKDG_E_ARG, "unknown limiter function");
  if (s->limiter < RKDG_LIM_NONE || s->limiter > RKDG_LIM_MULTI) return fail(RKDG_E_ARG, "unknown limiter");
  if (!(s->gamma > 1.0)) return fail(RKDG_E_ARG, "gamma must exceed 1");
  if (!(s->covolume >= 0.0)) return fail(RKDG_E_ARG, "covolume must be non-negative");
  for (int i = 0; i < s->num_bc; i++)
    if (s->bc_type[i] < RKDG_BC_SLIP || s->bc_type[i] > RKDG_BC_CHAR_OUTLET)
      return fail(RKDG_E_ARG, "unknown boundary condition type in slot " + std::to_string(i));
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
  g.linearize = s->linearize ? 1 : 0;
  return RKDG_OK;
}

// Jacobian of the trilinear map of hexahedron X (MFEM vertex order) at xi:
// dx/dxi_r as the bilinear blend, in the other two coordinates, of the four
// edges along axis r (the same map as PointMat * dshape, without its rounding
// on elements small against their coordinates)
void trilinear_jac(const double* X, const double* xi, double* J) {
  for (int r = 0; r < 3; r++) {
    const int s1 = r == 0 ? 1 : 0, s2 = r == 2 ? 1 : 2;
    for (int i = 0; i < 3; i++) J[i * 3 + r] = 0.0;
    for (int b2 = 0; b2 < 2; b2++)
      for (int b1 = 0; b1 < 2; b1++) {
        const double w = (b1 ? xi[s1] : 1.0 - xi[s1]) * (b2 ? xi[s2] : 1.0 - xi[s2]);
        int i0[3] = {0, 0, 0};
        i0[s1] = b1;
        i0[s2] = b2;
        int i1[3] = {i0[0], i0[1], i0[2]};
        i1[r] = 1;
        const int v0 = hex_vertex(i0[0], i0[1], i0[2]), v1 = hex_vertex(i1[0], i1[1], i1[2]);
        for (int i = 0; i < 3; i++) J[i * 3 + r] += w * (X[v1 * 3 + i] - X[v0 * 3 + i]);
      }
  }
}
void adj3(const double* J, double* A) {
  A[0] = J[4] * J[8] - J[5] * J[7];
  A[1] = J[2] * J[7] - J[1] * J[8];
  A[2] = J[1] * J[5] - J[2] * J[4];
  A[3] = J[5] * J[6] - J[3] * J[8];
  A[4] = J[0] * J[8] - J[2] * J[6];
  A[5] = J[2] * J[3] - J[0] * J[5];
  A[6] = J[3] * J[7] - J[4] * J[6];
  A[7] = J[1] * J[6] - J[0] * J[7];
  A[8] = J[0] * J[4] - J[1] * J[3];
}

// CurvedLayout records of curved (trilinear) hexahedra of order p, as the
// reference builds them on any element: the collocated volume term takes
// adj(J) at the nodes (DomainIntegrator's rule of order 2p has the p + 1
// Gauss-Legendre points per axis, domain_integrator.cpp:22-24), the mass
// matrix MassIntegrator's rule of order 2p + OrderW (OrderW = 2,
// fe_evolution.cpp:42-69), the means Averager's order p + OrderW
// (averager.cpp:84-109), the face normals CalcOrtho of the face Jacobian at
// each point of the face rule (face_integrator.cpp:144-152). Returns the
// records [e][size]; vol receives the volumes.
template <int P>
bool curved_records(const rkdg_ctx* c, std::vector<double>& rec, std::vector<double>& vol) {
  using CL = CurvedLayout<P>;
  constexpr int N = CL::N, ND = CL::ND, NQ = CL::NQ, NFP = CL::NFP;
  const int64_t NE = c->NE;
  rec.assign((size_t)NE * CL::size, 0.0);
  vol.assign((size_t)NE, 0.0);
  double xn[kMaxN], wn[kMaxN];
  gl_rule(N, xn, wn);
  auto rule = [](int order, std::vector<double>& x, std::vector<double>& w) {
    const int n = (order | 1) / 2 + 1;
    x.resize(n);
    w.resize(n);
    gl_rule(n, x.data(), w.data());
  };
  std::vector<double> xm, wm, xa, wa;
  rule(2 * P + 2, xm, wm);
  rule(P + 2, xa, wa);
  double sq[kMaxQ], wq[kMaxQ];
  gl_rule(NQ, sq, wq);
  auto shape = [&](const double* xi, double* phi) {
    double L[3][kMaxN];
    for (int d = 0; d < 3; d++)
      for (int a = 0; a < N; a++) L[d][a] = lag(xn, N, a, xi[d]);
    for (int k = 0; k < ND; k++) phi[k] = L[0][k % N] * L[1][(k / N) % N] * L[2][k / (N * N)];
  };
  std::vector<double> M((size_t)ND * ND), Mi((size_t)ND * ND), phi(ND);
  for (int64_t e = 0; e < NE; e++) {
    const double* X = &c->coords[(size_t)e * 24];
    double* R = &rec[(size_t)e * CL::size];
    double J[9], A[9];
    for (int k = 0; k < ND; k++) {
      const double xi[3] = {xn[k % N], xn[(k / N) % N], xn[k / (N * N)]};
      trilinear_jac(X, xi, J);
      adj3(J, A);
      for (int q = 0; q < 9; q++) R[CL::adj + k * 9 + q] = A[q];
    }
    std::fill(M.begin(), M.end(), 0.0);
    double v = 0.0;
    const int nm = (int)xm.size();
    for (int q3 = 0; q3 < nm; q3++)
      for (int q2 = 0; q2 < nm; q2++)
        for (int q1 = 0; q1 < nm; q1++) {
          const double xi[3] = {xm[q1], xm[q2], xm[q3]};
          trilinear_jac(X, xi, J);
          const double dj = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                            J[2] * (J[3] * J[7] - J[4] * J[6]);
          if (!(dj > 0.0)) return false;
          const double w = dj * wm[q1] * wm[q2] * wm[q3];
          v += w;
          shape(xi, phi.data());
          for (int i = 0; i < ND; i++)
            for (int j = 0; j < ND; j++) M[(size_t)i * ND + j] += w * phi[i] * phi[j];
        }
    if (!invert_small(ND, M.data(), Mi.data())) return false;
    // transposed: the kernel's lane j reads entry (j, k) at k * ND + j
    for (int i = 0; i < ND; i++)
      for (int k = 0; k < ND; k++) R[CL::minv + k * ND + i] = Mi[(size_t)i * ND + k];
    vol[e] = v;
    R[CL::vol] = v;
    const int na = (int)xa.size();
    for (int q3 = 0; q3 < na; q3++)
      for (int q2 = 0; q2 < na; q2++)
        for (int q1 = 0; q1 < na; q1++) {
          const double xi[3] = {xa[q1], xa[q2], xa[q3]};
          trilinear_jac(X, xi, J);
          const double dj = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                            J[2] * (J[3] * J[7] - J[4] * J[6]);
          const double w = dj * wa[q1] * wa[q2] * wa[q3];
          shape(xi, phi.data());
          for (int k = 0; k < ND; k++) R[CL::mw + k] += w * phi[k];
        }
    for (int lf = 0; lf < 6; lf++) {
      int r, cc;
      rc_of_lf(3, lf, r, cc);
      const int t1 = r == 0 ? 1 : 0, t2 = r == 2 ? 1 : 2;
      for (int qj = 0; qj < NQ; qj++)
        for (int qi = 0; qi < NQ; qi++) {
          double xi[3];
          xi[r] = cc;
          xi[t1] = sq[qi];
          xi[t2] = sq[qj];
          trilinear_jac(X, xi, J);
          adj3(J, A);
          double an[3], n2 = 0.0;
          for (int d = 0; d < 3; d++) {
            an[d] = (cc ? 1.0 : -1.0) * A[r * 3 + d];
            n2 += an[d] * an[d];
          }
          const double nm2 = std::sqrt(n2);
          double* F = R + CL::fn + ((size_t)lf * NFP + qi + NQ * qj) * 5;
          // the reference's unit normal: nor * (1 / Norml2(nor))
          // (face_integrator.cpp:165), as the affine kernels form it; an
          // axis-aligned normal may come out as 1 - ulp, which the 3D rotation
          // of the Riemann solvers (rs_basic.cpp:21-28) turns into a 1e-8-level
          // term: taking the same steps keeps that term the reference's
          const double inv = 1.0 / norml2_scaled<3>(an);
          double u2 = 0.0;
          for (int d = 0; d < 3; d++) {
            F[d] = an[d] * inv;
            u2 += F[d] * F[d];
          }
          F[3] = nm2;
          F[4] = 0.5 / std::sqrt(u2);
        }
    }
  }
  return true;
}

// MFEM local face vertices of hexahedra (quadrilateral edges: kQuadEdge)
const int kHexFace[6][4] = {{3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
int mfem_face_vertex(int dim, int lf, int q) { return dim == 2 ? kQuadEdge[lf][q] : kHexFace[lf][q]; }
void ref_vertex(int dim, int v, double* xi) {
  const int b = v & 3;
  xi[0] = (b == 1 || b == 2) ? 1.0 : 0.0;
  xi[1] = (b == 2 || b == 3) ? 1.0 : 0.0;
  if (dim == 3) xi[2] = v >= 4 ? 1.0 : 0.0;
}
// bilinear (linear) weights of the corners of a face at parameter t
void face_shape(int dim, const double* t, double* Nw) {
  if (dim == 2) {
    Nw[0] = 1.0 - t[0];
    Nw[1] = t[0];
  } else {
    Nw[0] = (1.0 - t[0]) * (1.0 - t[1]);
    Nw[1] = t[0] * (1.0 - t[1]);
    Nw[2] = t[0] * t[1];
    Nw[3] = (1.0 - t[0]) * t[1];
  }
}

// Non-conforming faces of a tensor mesh (rkdg_nc_desc): codes for the slave
// and master element faces (kNcCode0 - k, with the neighbours across each),
// and, per slave face, the points of its Gauss-Legendre rule (the rule of the
// conforming faces) in the slave's and the master's reference coordinates
// (through the point matrix: slave face parameter -> master face parameter ->
// master reference coordinates), the weights times |n| and the slave's unit
// outward normal, as the reference's face loop takes them on a slave face
// (Loc1, Loc2, CalcOrtho of the slave face, face_integrator.cpp:44-150).
int mark_nc_faces(rkdg_ctx* c, const rkdg_mesh_desc* m,
                  const std::vector<std::pair<std::array<int64_t, 4>, std::pair<int64_t, int>>>& open,
                  std::map<std::pair<int64_t, int>, int>& ncface) {
  const rkdg_nc_desc* d = c->nc_in;
  const int dim = c->dim, nf = c->nf, nvf = c->nvf, fd = dim - 1;
  const int64_t NE = c->NE;
  if (!d->slave || !d->slave_face || !d->master || !d->master_face || !d->param)
    return fail(RKDG_E_ARG, "null array in the non-conforming face descriptor");
  std::map<std::pair<int64_t, int>, int> is_open;
  for (auto& of : open) is_open[{of.second.first, of.second.second}] = 1;
  std::vector<std::array<int, 4>> nbs;
  auto code_of = [&](int64_t e, int lf) -> int {
    auto it = ncface.find({e, lf});
    if (it != ncface.end()) return it->second;
    const int k = (int)nbs.size();
    ncface[{e, lf}] = k;
    nbs.push_back({-1, -1, -1, -1});
    return k;
  };
  auto add_nb = [&](int k, int64_t n) -> bool {
    for (int q = 0; q < 4; q++) {
      if (nbs[k][q] == (int)n) return true;
      if (nbs[k][q] < 0) {
        nbs[k][q] = (int)n;
        return true;
      }
    }
    return false;
  };
  const int NQ = dim == 3 ? c->p + 2 : c->p + 1;
  const int NP = dim == 3 ? NQ * NQ : NQ;
  const int PR = 3 * dim + 1;  // per point: slave xi, master xi, w |n|, unit normal
  const int RG = NP * PR;
  double sq[kMaxQ], wq[kMaxQ];
  gl_rule(NQ, sq, wq);
  std::vector<double> geo((size_t)d->count * RG);
  std::vector<int> rec((size_t)d->count * 3);
  for (int64_t k = 0; k < d->count; k++) {
    const int64_t es = d->slave[k], em = d->master[k];
    const int ls = d->slave_face[k], lm = d->master_face[k];
    const std::string at = "non-conforming face record " + std::to_string(k);
    if (es < 0 || es >= NE || em < 0 || em >= NE || es == em || ls < 0 || ls >= nf || lm < 0 || lm >= nf)
      return fail(RKDG_E_ARG, at + " out of range");
    if (!is_open.count({es, ls})) return fail(RKDG_E_ARG, at + ": the slave face has a conforming neighbour");
    if (!is_open.count({em, lm})) return fail(RKDG_E_ARG, at + ": the master face has a conforming neighbour");
    const int ks = code_of(es, ls), km = code_of(em, lm);
    if (nbs[ks][0] >= 0 && nbs[ks][0] != (int)em) return fail(RKDG_E_ARG, at + ": a slave face with two masters");
    add_nb(ks, em);
    if (!add_nb(km, es)) return fail(RKDG_E_ARG, at + ": more than four slaves on a master face");
    double par[4][2];
    for (int q = 0; q < nvf; q++)
      for (int j = 0; j < fd; j++) {
        par[q][j] = d->param[(k * nvf + q) * fd + j];
        if (!(par[q][j] >= 0.0 && par[q][j] <= 1.0)) return fail(RKDG_E_ARG, at + ": face parameter outside [0, 1]");
      }
    rec[3 * k] = (int)es;
    rec[3 * k + 1] = (int)em;
    rec[3 * k + 2] = ls;
    int rs, cs;
    rc_of_lf(dim, ls, rs, cs);
    const int t1 = rs == 0 ? 1 : 0, t2 = rs == 2 ? 1 : 2;
    double R0[3], R1[3], R3[3], Rm[4][3];
    ref_vertex(dim, mfem_face_vertex(dim, ls, 0), R0);
    ref_vertex(dim, mfem_face_vertex(dim, ls, 1), R1);
    if (dim == 3) ref_vertex(dim, mfem_face_vertex(dim, ls, 3), R3);
    for (int q = 0; q < nvf; q++) ref_vertex(dim, mfem_face_vertex(dim, lm, q), Rm[q]);
    // the slave face's area normal: sign * row rs of adj(J), constant on an
    // affine element, per point on a curved one
    double an[3], nn = 0.0;
    for (int dd = 0; dd < dim; dd++) {
      an[dd] = (cs ? 1.0 : -1.0) * c->geo_h[(size_t)(rs * dim + dd) * NE + es];
      nn += an[dd] * an[dd];
    }
    nn = std::sqrt(nn);
    double* G = &geo[(size_t)k * RG];
    for (int qj = 0; qj < (dim == 3 ? NQ : 1); qj++)
      for (int qi = 0; qi < NQ; qi++) {
        const int q = qi + NQ * qj;
        double xs[3] = {0, 0, 0};
        xs[rs] = cs;
        xs[t1] = sq[qi];
        if (dim == 3) xs[t2] = sq[qj];
        double sg[2] = {0, 0}, Nw[4], pi[2] = {0, 0}, xm[3] = {0, 0, 0};
        for (int dd = 0; dd < dim; dd++) {
          sg[0] += (xs[dd] - R0[dd]) * (R1[dd] - R0[dd]);
          if (dim == 3) sg[1] += (xs[dd] - R0[dd]) * (R3[dd] - R0[dd]);
        }
        face_shape(dim, sg, Nw);
        for (int cc = 0; cc < nvf; cc++)
          for (int j = 0; j < fd; j++) pi[j] += Nw[cc] * par[cc][j];
        face_shape(dim, pi, Nw);
        for (int cc = 0; cc < nvf; cc++)
          for (int dd = 0; dd < dim; dd++) xm[dd] += Nw[cc] * Rm[cc][dd];
        double* gp = G + q * PR;
        double ap[3] = {an[0], an[1], an[2]}, np_ = nn;
        if (c->curved) {
          double J[9], A[9];
          trilinear_jac(&c->coords[(size_t)es * 24], xs, J);
          adj3(J, A);
          np_ = 0.0;
          for (int dd = 0; dd < 3; dd++) {
            ap[dd] = (cs ? 1.0 : -1.0) * A[rs * 3 + dd];
            np_ += ap[dd] * ap[dd];
          }
          np_ = std::sqrt(np_);
        }
        for (int dd = 0; dd < dim; dd++) {
          gp[dd] = xs[dd];
          gp[dim + dd] = xm[dd];
        }
        gp[2 * dim] = (dim == 3 ? wq[qi] * wq[qj] : wq[qi]) * np_;
        // the reference's unit normal nor * (1 / Norml2(nor)), as the element kernels form it
        const double inv = 1.0 / (dim == 3 ? norml2_scaled<3>(ap) : norml2_scaled<2>(ap));
        for (int dd = 0; dd < dim; dd++) gp[2 * dim + 1 + dd] = ap[dd] * inv;
      }
  }
  // elements with non-conforming faces, their items in record order
  std::map<int64_t, std::vector<int>> items;
  for (int64_t k = 0; k < d->count; k++) {
    items[d->slave[k]].push_back((int)(2 * k));
    items[d->master[k]].push_back((int)(2 * k + 1));
  }
  std::vector<int> elem, ptr(1, 0), item, slot((size_t)NE, -1);
  for (auto& kv : items) {
    slot[kv.first] = (int)elem.size();
    elem.push_back((int)kv.first);
    for (int it : kv.second) item.push_back(it);
    ptr.push_back((int)item.size());
  }
  std::vector<int> nbv;
  for (auto& a : nbs)
    for (int q = 0; q < 4; q++) nbv.push_back(a[q]);
  c->n_nc = d->count;
  c->n_ncf = (int64_t)nbs.size();
  c->n_nce = (int64_t)elem.size();
  auto up = [&](auto** dst, const auto& v) -> hipError_t {
    using T = typename std::decay<decltype(v[0])>::type;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), sizeof(T) * std::max<size_t>(1, v.size()));
    if (e == hipSuccess && !v.empty()) e = hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
    return e;
  };
  HIP_TRY(up(&c->d_ncnb, nbv));
  HIP_TRY(up(&c->d_ncrec, rec));
  HIP_TRY(up(&c->d_ncgeo, geo));
  HIP_TRY(up(&c->d_nce, elem));
  HIP_TRY(up(&c->d_ncptr, ptr));
  HIP_TRY(up(&c->d_ncitem, item));
  HIP_TRY(up(&c->d_ncslot, slot));
  HIP_TRY(hipMalloc(&c->d_ncbuf, sizeof(double) * std::max<int64_t>(1, c->n_nce * c->neq * c->nd)));
  return RKDG_OK;
}

int build_ctx(rkdg_ctx* c, const rkdg_mesh_desc* m, const rkdg_scheme_desc* s) {
  if (m->geom != RKDG_GEOM_QUAD && m->geom != RKDG_GEOM_HEX && m->geom != RKDG_GEOM_TRI)
    return fail(RKDG_E_ARG, "unknown geometry");
  c->geom = m->geom;
  c->p = m->order;
  c->dim = m->geom == RKDG_GEOM_HEX ? 3 : 2;
  if (m->geom != RKDG_GEOM_TRI && !tensor_supported(c->dim, c->p))
    return fail(RKDG_E_UNSUPPORTED, "polynomial order not instantiated");
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
  {
    const int rc = scheme_params(s, c->sp);
    if (rc) return rc;
  }
  SchemeParams& g = c->sp;
  const int dim = c->dim, nv = c->nv, NE_ = (int)c->NE;
  const int64_t NE = c->NE;
  if (m->geom == RKDG_GEOM_TRI) c->nv = 3;
  c->coords.assign(m->elem_coords, m->elem_coords + NE * c->nv * dim);

  // triangles and non-parallelogram quadrilaterals take the dense path (any
  // quadrilateral mesh when asked for)
  bool dense = m->geom == RKDG_GEOM_TRI || (m->geom == RKDG_GEOM_QUAD && c->force_dense);
  if (c->force_dense && m->geom == RKDG_GEOM_HEX)
    return fail(RKDG_E_UNSUPPORTED, "hexahedra have no dense kernels");
  if (m->geom == RKDG_GEOM_QUAD) {
    for (int64_t e = 0; e < NE && !dense; e++) {
      const double* X = &c->coords[e * 8];
      const double sc = std::fabs(X[2] - X[0]) + std::fabs(X[3] - X[1]) + std::fabs(X[6] - X[0]) + std::fabs(X[7] - X[1]);
      // the tensor kernels take the element as a parallelogram; the reference
      // maps it bilinearly, and the right-hand side moves by about 40x the
      // relative geometry defect (the cancelling volume and face terms), so
      // only defects of 1e-14 of the element size stay on the tensor path
      for (int i = 0; i < 2; i++)
        if (std::fabs(X[4 + i] - (X[2 + i] + X[6 + i] - X[i])) > 1e-14 * sc) dense = true;
    }
  }
  if (dense && c->nc_in && c->nc_in->count > 0)
    return fail(RKDG_E_UNSUPPORTED, "non-conforming faces are built for affine quadrilaterals and hexahedra");
  if (dense) return build_dense(c, m);

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
  {
    int k = 0;
    for (int a = 0; a < N; a++) t.packed[k++] = t.w[a];
    for (int a = 0; a < N; a++) t.packed[k++] = t.invw[a];
    for (int a = 0; a < N; a++)
      for (int b = 0; b < N; b++) t.packed[k++] = t.D[a][b];
    for (int c2 = 0; c2 < 2; c2++)
      for (int b = 0; b < N; b++) t.packed[k++] = t.E[c2][b];
    for (int q = 0; q < NQ; q++)
      for (int b = 0; b < N; b++) t.packed[k++] = t.B[q][b];
    for (int q = 0; q < NQ; q++) t.packed[k++] = t.wq[q];
    for (int q = 0; q < 2; q++)
      for (int b = 0; b < N; b++) t.packed[k++] = t.BI[q][b];
  }

  // geometry (affine tensor elements): adj(J) row-major, det(J)
  const int G = dim * dim + 1;
  c->geo_h.assign((size_t)G * NE, 0.0);
  c->h_elem.assign(NE, 0.0);
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
        if (std::fabs(xv - X[v * dim + i]) > affine_tol(scale, X, nv * dim)) c->affine = 0;
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
    c->h_elem[e] = min_singular(dim, J);
    hmin = std::min(hmin, c->h_elem[e]);
  }
  if (!c->affine) {
    // curved (trilinear) hexahedra: the per-element records of the CV kernels;
    // geo keeps adj(J) at the centre and the volume in the det(J) slot (the
    // limiter's Venkatakrishnan eps^2 = (K h)^3 with h^3 the volume), h the
    // smallest singular value of J at the centre (GetElementSize)
    if (dim != 3) return fail(RKDG_E_UNSUPPORTED, "non-affine quadrilaterals take the dense path");
    std::vector<double> rec;
    bool ok = false;
    if (c->p == 1) ok = curved_records<1>(c, rec, c->vol_h);
    else if (c->p == 2) ok = curved_records<2>(c, rec, c->vol_h);
    else if (c->p == 3) ok = curved_records<3>(c, rec, c->vol_h);
    if (!ok) return fail(RKDG_E_ARG, "a curved hexahedron with a non-positive Jacobian");
    hmin = 1e300;
    for (int64_t e = 0; e < NE; e++) {
      const double xi[3] = {0.5, 0.5, 0.5};
      double J[9], A[9];
      trilinear_jac(&c->coords[(size_t)e * 24], xi, J);
      adj3(J, A);
      for (int i = 0; i < 9; i++) c->geo_h[(size_t)i * NE + e] = A[i];
      c->geo_h[(size_t)9 * NE + e] = c->vol_h[e];
      c->h_elem[e] = min_singular(3, J);
      hmin = std::min(hmin, c->h_elem[e]);
    }
    HIP_TRY(hipMalloc(&c->d_cgeo, sizeof(double) * rec.size()));
    HIP_TRY(hipMemcpy(c->d_cgeo, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice));
    c->curved = 1;
  }
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
  std::map<std::pair<int64_t, int>, int> ncface;  // (element, local face) -> k of the code kNcCode0 - k
  if (c->nc_in && c->nc_in->count > 0) {
    const int rc = mark_nc_faces(c, m, open, ncface);
    if (rc) return rc;
  }
  for (auto& of : open) {
    const auto nf_it = ncface.find({of.second.first, of.second.second});
    if (nf_it != ncface.end()) {
      nbr[of.second.second * NE + of.second.first] = kNcCode0 - nf_it->second;
      continue;
    }
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
  // aligned (structured) meshes: every interior face is the neighbour's
  // opposite face of the same axis with the identity orientation
  c->aligned = c->curved ? 0 : 1;  // the curved kernels take any orientation
  for (int64_t e = 0; e < NE && c->aligned; e++)
    for (int lf = 0; lf < nf; lf++) {
      if (nbr[(size_t)lf * NE + e] < 0) continue;
      int r, cc;
      rc_of_lf(dim, lf, r, cc);
      int rn, cn;
      const int code = fcode[(size_t)lf * NE + e];
      rc_of_lf(dim, code & 7, rn, cn);
      if (rn != r || cn != 1 - cc || ((code >> 3) & 7) != 0) {
        c->aligned = 0;
        break;
      }
    }

  HIP_TRY(hipMalloc(&c->d_geo, sizeof(double) * G * NE));
  HIP_TRY(hipMemcpy(c->d_geo, c->geo_h.data(), sizeof(double) * G * NE, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_nbr, sizeof(int) * nf * NE));
  HIP_TRY(hipMemcpy(c->d_nbr, nbr.data(), sizeof(int) * nf * NE, hipMemcpyHostToDevice));
  c->nbr_h = nbr;
  HIP_TRY(hipMalloc(&c->d_fcode, sizeof(int) * nf * NE));
  HIP_TRY(hipMemcpy(c->d_fcode, fcode.data(), sizeof(int) * nf * NE, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_lam, sizeof(double) * 8 * c->count));
  HIP_TRY(hipMemset(c->d_lam, 0, sizeof(double) * 8 * c->count));
  HIP_TRY(hipMalloc(&c->d_err, sizeof(long long) * 4));
  HIP_TRY(hipMemset(c->d_err, 0, sizeof(long long) * 4));
  HIP_TRY(hipMalloc(&c->d_means, sizeof(double) * c->neq * NE));
  HIP_TRY(hipMalloc(&c->d_means2, sizeof(double) * c->neq * NE));
  HIP_TRY(hipMalloc(&c->d_fac, sizeof(double) * NE));
  {
    // Traversal order of the RHS work blocks (EPB consecutive elements each):
    // Morton order of the block centroids on a 2^10 (3D) / 2^16 (2D) grid of
    // the bounding box, so that face neighbours are processed close in time.
    const int epb = tensor_elems_per_block(dim, c->p);
    const int64_t nb = (NE + epb - 1) / epb;
    const int nv = c->nv;
    std::vector<double> cen((size_t)nb * dim, 0.0);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t b = 0; b < nb; b++) {
      int cnt = 0;
      for (int64_t e = b * epb; e < std::min<int64_t>(NE, (b + 1) * epb); e++, cnt++)
        for (int v = 0; v < nv; v++)
          for (int i = 0; i < dim; i++) cen[b * dim + i] += c->coords[(e * nv + v) * dim + i];
      for (int i = 0; i < dim; i++) {
        cen[b * dim + i] /= (double)(cnt * nv);
        lo[i] = std::min(lo[i], cen[b * dim + i]);
        hi[i] = std::max(hi[i], cen[b * dim + i]);
      }
    }
    const int bits = dim == 3 ? 10 : 16;
    const double cells = (double)((1u << bits) - 1);
    std::vector<std::pair<uint64_t, int>> key((size_t)nb);
    for (int64_t b = 0; b < nb; b++) {
      uint64_t code = 0;
      for (int i = 0; i < dim; i++) {
        const double span = hi[i] > lo[i] ? hi[i] - lo[i] : 1.0;
        const uint64_t q = (uint64_t)((cen[b * dim + i] - lo[i]) / span * cells + 0.5);
        for (int k = 0; k < bits; k++) code |= ((q >> k) & 1ull) << (k * dim + i);
      }
      key[b] = {code, (int)b};
    }
    std::stable_sort(key.begin(), key.end());
    std::vector<int> order((size_t)nb);
    for (int64_t b = 0; b < nb; b++) order[b] = key[b].second;
    HIP_TRY(hipMalloc(&c->d_order, sizeof(int) * nb));
    HIP_TRY(hipMemcpy(c->d_order, order.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMalloc(&c->d_tb, sizeof(TensorTables)));
  HIP_TRY(hipMemcpy(c->d_tb, &c->tb, sizeof(TensorTables), hipMemcpyHostToDevice));
  if (c->n_bdr == 0 && c->n_ncf == 0 && !c->curved && c->two_colour) {
    int rc = build_two_colour(c, nbr);
    if (rc) return rc;
  }
  return RKDG_OK;
}

constexpr size_t kMaxProfiled = 8192;

// Restores a context field on every exit path of an entry point that switches
// it temporarily (the batch stream, the profiling flag, the tuner's grid cap).
template <typename Tp>
struct Restore {
  Tp& ref;
  Tp saved;
  explicit Restore(Tp& r) : ref(r), saved(r) {}
  Restore(const Restore&) = delete;
  Restore& operator=(const Restore&) = delete;
  ~Restore() { ref = saved; }
};

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

hipError_t rhs_launch(rkdg_ctx* c, const RhsArgs& a0) {
  // the tensor kernel forms the means of outs[0]: move the selected output
  // there (the outputs are independent, every base is read before any store)
  RhsArgs a = a0;
  if (a.means && a.means_of > 0 && a.means_of < a.nout) {
    std::swap(a.outs[0], a.outs[a.means_of]);
    a.means_of = 0;
  }
  auto pr = prof_begin(c, c->ev_rhs);
  hipError_t e;
  if (c->dense) {
    e = launch_dense_rhs(c->geom, c->p, a, c->sp, c->d_dtab, c->n_bdr > 0, c->stream);
  } else if (c->rb) {
    // two-colour form: the solver elements (every face flux once, the
    // receiver's share to fbuf), then the receivers (their share read back)
    a.fslot = c->d_fslot;
    a.fbuf = c->d_fbuf;
    a.role = 1;
    a.elist = c->d_elist_solve;
    a.nlist = c->n_solve;
    e = launch_tensor_rhs(c->dim, c->p, a, c->sp, c->d_tb, c->tb, false, c->stream);
    if (e == hipSuccess && !a.face_only) {
      a.role = 2;
      a.elist = c->d_elist_recv;
      a.nlist = c->n_recv;
      e = launch_tensor_rhs(c->dim, c->p, a, c->sp, c->d_tb, c->tb, false, c->stream);
    }
  } else {
    e = hipSuccess;
    if (c->n_nce > 0) {
      NcArgs na{};
      na.x = a.x;
      na.NE = c->NE;
      na.n_elem = c->n_nce;
      na.elem = c->d_nce;
      na.ptr = c->d_ncptr;
      na.item = c->d_ncitem;
      na.rec = c->d_ncrec;
      na.geo = c->d_ncgeo;
      na.out = c->d_ncbuf;
      na.lam = a.lam;
      na.err = a.err;
      na.no_lambda = a.no_lambda && !a.face_only;
      e = launch_nc_faces(c->dim, c->p, na, c->sp, c->d_tb, c->stream);
      a.ncslot = c->d_ncslot;
      a.ncbuf = c->d_ncbuf;
    }
    if (e == hipSuccess)
      e = launch_tensor_rhs(c->dim, c->p, a, c->sp, c->d_tb, c->tb, c->n_bdr + c->n_ncf > 0, c->stream);
  }
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
  a.inst = c->d_inst;
  a.bcp = c->d_bcp;
  a.grid_cap = c->grid_cap;
  a.order = c->dense ? nullptr : c->d_order;
  a.aligned = c->dense ? 0 : c->aligned;
  a.cgeo = c->d_cgeo;
  return a;
}

// pinned host words for the per-step readbacks (a copy to pageable memory is
// staged by the runtime and costs tens of microseconds per step)
hipError_t ensure_pinned(rkdg_ctx* c) {
  if (c->h_pin) return hipSuccess;
  return hipHostMalloc(reinterpret_cast<void**>(&c->h_pin), sizeof(long long) * 4 + sizeof(double) * c->count,
                       hipHostMallocDefault);
}

int check_device_error(rkdg_ctx* c, int64_t* info) {
  HIP_TRY(ensure_pinned(c));
  long long* h = c->h_pin;
  HIP_TRY(hipMemcpyAsync(h, c->d_err, sizeof(long long) * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (h[0]) {
    const long long kind = h[0], e1 = h[1], e2 = h[2];
    c->err_pair[0] = e1;
    c->err_pair[1] = e2;
    if (info) {
      info[0] = e1;
      info[1] = e2;
    }
    HIP_TRY(launch_fill(reinterpret_cast<double*>(c->d_err), 0.0, 4, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (kind == (long long)kErrNonFinite)
      return fail(RKDG_E_NAN, "non-finite state at a face between elements " + std::to_string(e1) + " and " +
                                  std::to_string(e2));
    return fail(RKDG_E_UNPHYSICAL, "unphysical state at a face between elements " + std::to_string(e1) +
                                       " and " + std::to_string(e2));
  }
  return RKDG_OK;
}

// wave speed of stage slot; on a batch, the maximum over the instances
int read_lambda(rkdg_ctx* c, int slot, double* lam) {
  HIP_TRY(ensure_pinned(c));
  double* h = reinterpret_cast<double*>(c->h_pin + 4);
  HIP_TRY(hipMemcpyAsync(h, c->d_lam + (int64_t)slot * c->count, sizeof(double) * c->count,
                         hipMemcpyDeviceToHost, c->stream));
  const int rc = check_device_error(c, nullptr);
  *lam = *std::max_element(h, h + c->count);
  return rc;
}

int limit_impl(rkdg_ctx* c, double* x, int8_t* troubled, bool have_means, const int* active = nullptr) {
  if (!have_means) {
    MeansArgs ma{x, c->d_means, c->d_geo, c->NE, c->d_cgeo};
    HIP_TRY(c->dense ? launch_dense_means(c->geom, c->p, ma, c->d_dtab, c->stream)
                     : launch_tensor_means(c->dim, c->p, ma, c->d_tb, c->stream));
  }
  if (c->sp.indicator == RKDG_IND_SHU) {
    auto pr = prof_begin(c, c->ev_lim);
    hipError_t le = launch_shu_indicator(x, c->d_means, c->d_shu_coef, c->d_nbr, c->NE, c->nd, c->neq, c->nf,
                                         c->d_shu_flag, troubled, c->stream);
    if (le == hipSuccess) le = launch_shu_apply(x, c->d_means, c->d_shu_flag, c->NE, c->nd, c->neq, c->stream);
    prof_end(c, pr);
    HIP_TRY(le);
    return RKDG_OK;
  }
  LimitArgs la{x, c->d_means, c->d_geo, c->d_nbr, c->d_fcode, c->NE, troubled, c->d_inst, active, nullptr,
               nullptr, c->d_ncnb};
  auto pr = prof_begin(c, c->ev_lim);
  hipError_t le = c->dense ? launch_dense_limit(c->geom, c->p, la, c->sp, c->d_dtab, c->stream)
                           : launch_tensor_limit(c->dim, c->p, la, c->sp, c->d_tb, c->stream);
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

// The limiter can run as a factor pass followed by limiting on load in the
// next stage kernel when it maps each element to mean + f (u - mean) (the
// multiplier without linearize, FinDiff) and the one-kernel tensor RHS runs.
bool limit_on_load(const rkdg_ctx* c) {
  return !c->dense && !c->rb && c->n_nc == 0 && c->d_fac && c->d_means2 && c->sp.indicator != RKDG_IND_SHU &&
         !c->sp.linearize && c->sp.limiter != RKDG_LIM_NONE;
}

// factor pass of the limiter over x (left as it is): every element's factor
// into c->d_fac, from the means of x in `means`
int limit_factor(rkdg_ctx* c, const double* x, const double* means) {
  LimitArgs la{const_cast<double*>(x), means, c->d_geo, c->d_nbr, c->d_fcode, c->NE, nullptr, nullptr, nullptr,
               nullptr, c->d_fac, c->d_ncnb};
  auto pr = prof_begin(c, c->ev_lim);
  const hipError_t le = launch_tensor_limit(c->dim, c->p, la, c->sp, c->d_tb, c->stream);
  prof_end(c, pr);
  HIP_TRY(le);
  return RKDG_OK;
}

// the three-stage SSP-RK3 tableau exactly: a = (1; 1/4, 1/4), b = (1/6, 1/6, 2/3)
bool is_ssprk3(int s, const double* a, const double* b) {
  return s == 3 && a[0] == 1.0 && a[1] == 0.25 && a[2] == 0.25 && b[0] == 1.0 / 6.0 && b[1] == 1.0 / 6.0 &&
         b[2] == 2.0 / 3.0;
}

// ExplicitRKLimitedSolver::Step (rk_explicit_limited.cpp:30-82). Stage i reads
// its (limited) input and produces, in the epilogue of the RHS kernel, every
// pending combination that needs k_i: W_j += (a_ji dt) k_i for later stages j
// and acc += (b_i dt) k_i, in the reference's association order. W_1 (stage 0)
// and the final x are the only vectors written whole; the means of the next
// stage input come out of the same epilogue for the limiter. buf[0..s-2] hold
// W_1..W_{s-1}, buf[3] the accumulator; every one is written before it is read
// (stage 0 has no bases), so nothing needs clearing. The last stage's wave
// speed lands in d_lam[s-1] when want_lambda.
//
// Where limit_on_load holds, the limiter of a stage input W_i is a factor pass
// (it writes one factor per element, not W_i) and the stage kernel limits W_i
// as it loads it; the final x is limited in place as before. The stage inputs
// then stay unlimited in memory, which for SSP-RK3 allows the stage sums of
// the Shu-Osher form, an exact rearrangement of the tableau (W_1 = x + dt k_0
// gives x + dt (k_0 + k_1)/4 = 3/4 x + 1/4 (W_1 + dt k_1) = W_2 and
// x + dt (k_0/6 + k_1/6 + 2 k_2/3) = 1/3 x + 2/3 (W_2 + dt k_2)): each stage
// reads its input and x and writes one vector, 8 vectors per step against 12.
// The means alternate between two buffers (a stage reads the means of its
// input while it writes those of its output); the last stage's are in d_means.
int step_impl(rkdg_ctx* c, double* x, double dt, int s, const double* a, const double* b, double* const* buf,
              bool want_lambda) {
  double* W[4] = {nullptr, buf[0], buf[1], buf[2]};
  double* acc = buf[3];
  const bool onload = limit_on_load(c);
  const bool so = onload && is_ssprk3(s, a, b);
  HIP_TRY(launch_fill(c->d_lam, 0.0, 8 * c->count, c->stream));
  auto A_ = [&](int i, int j) { return a[i * (i - 1) / 2 + j]; };
  auto mbuf = [&](int i) { return (!onload || (s - 1 - i) % 2 == 0) ? c->d_means : c->d_means2; };
  for (int i = 0; i < s; i++) {
    const double* in = i == 0 ? x : W[i];
    RhsArgs ra = base_args(c, in);
    ra.lam = c->d_lam + (int64_t)i * c->count;
    const bool last = i == s - 1;
    ra.no_lambda = (!last || !want_lambda) ? 1 : 0;  // only the last stage's wave speed is read
    if (onload && i > 0) {
      ra.lim_f = c->d_fac;
      ra.lim_mean = mbuf(i - 1);
    }
    int n = 0;
    if (so) {
      const double wb[3] = {0.0, 0.75, 1.0 / 3.0}, wi[3] = {0.0, 0.25, 2.0 / 3.0};
      ra.outs[0].out = last ? x : W[i + 1];
      ra.outs[0].base = i == 0 ? nullptr : x;
      ra.outs[0].coef = dt;
      ra.outs[0].wb = wb[i];
      ra.outs[0].wi = wi[i];
      n = 1;
    } else {
      for (int j = i + 1; j < s; j++) {
        ra.outs[n].out = W[j];
        ra.outs[n].base = i == 0 ? nullptr : W[j];
        ra.outs[n].coef = A_(j, i) * dt;
        n++;
      }
      ra.outs[n].out = last ? (s == 1 ? acc : x) : acc;
      ra.outs[n].base = i == 0 ? nullptr : acc;
      ra.outs[n].coef = b[i] * dt;
      n++;
    }
    ra.nout = n;
    ra.means = mbuf(i);
    ra.means_of = last ? n - 1 : 0;  // next stage input W_{i+1} is outs[0]; final is the last
    HIP_TRY(rhs_launch(c, ra));
    if (last && s == 1) HIP_TRY(launch_copy(x, acc, (int64_t)c->neq * c->ndofs, c->stream));
    const int rc = (last || !onload) ? limit_impl(c, last ? x : W[i + 1], nullptr, true)
                                     : limit_factor(c, W[i + 1], mbuf(i));
    if (rc) return rc;
  }
  return RKDG_OK;
}

BatchTime batch_time(rkdg_ctx* c, const double* lam) {
  BatchTime b{};
  b.count = c->count;
  b.t = c->d_t;
  b.dt = c->d_dt;
  b.dt_real = c->d_dtr;
  b.t_final = c->d_tf;
  b.hmin = c->d_hmin;
  b.lam = lam;
  b.active = c->d_active;
  b.steps = c->d_steps;
  b.n_active = c->d_nact;
  b.cfl = c->cfl;
  b.p = c->p;
  b.max_steps = c->max_steps;
  return b;
}

// one batch step on c->stream: per-instance dt, stages with the per-instance
// dt scaling every coefficient, limiter after each stage, dt/t/active update
int batch_step_launch(rkdg_ctx* c, double* x, int s, const double* a, const double* b) {
  double* W[4] = {nullptr, c->d_stage[0], c->d_stage[1], c->d_stage[2]};
  double* acc = c->d_stage[3];
  HIP_TRY(launch_fill(c->d_lam, 0.0, (int64_t)s * c->count, c->stream));
  BatchTime bt = batch_time(c, c->d_lam + (int64_t)(s - 1) * c->count);
  HIP_TRY(launch_batch_pre(bt, c->stream));
  auto A_ = [&](int i, int j) { return a[i * (i - 1) / 2 + j]; };
  for (int i = 0; i < s; i++) {
    const double* in = i == 0 ? x : W[i];
    RhsArgs ra = base_args(c, in);
    ra.lam = c->d_lam + (int64_t)i * c->count;
    ra.dt_inst = c->d_dtr;
    ra.active = c->d_active;
    int n = 0;
    for (int j = i + 1; j < s; j++) {
      ra.outs[n].out = W[j];
      ra.outs[n].base = i == 0 ? nullptr : W[j];
      ra.outs[n].coef = A_(j, i);
      n++;
    }
    const bool last = i == s - 1;
    ra.outs[n].out = last ? x : acc;
    ra.outs[n].base = i == 0 ? nullptr : acc;
    ra.outs[n].coef = b[i];
    n++;
    ra.nout = n;
    ra.means = c->d_means;
    ra.means_of = last ? n - 1 : 0;
    HIP_TRY(rhs_launch(c, ra));
    int rc = limit_impl(c, last ? x : W[i + 1], nullptr, true, c->d_active);
    if (rc) return rc;
  }
  HIP_TRY(launch_batch_post(bt, c->stream));
  return RKDG_OK;
}


// ---------------------------------------------------------------------------
// adjoint helpers (dense 2D path)

int adjoint_supported(rkdg_ctx* c) {
  if (!c->dense || c->dim != 2) return fail(RKDG_E_UNSUPPORTED, "the adjoint is built for the dense 2D kernels "
                                                                "(triangles, or quadrilaterals created with RKDG_CREATE_DENSE)");
  for (int q = 0; q < c->sp.num_bc; q++)
    if (!adj::bc_has_vjp(c->sp.bc_type[q]))
      return fail(RKDG_E_UNSUPPORTED, "boundary type " + std::to_string(c->sp.bc_type[q]) + " has no adjoint");
  return RKDG_OK;
}

int alloc_vec(double** p, int64_t n) {
  if (*p) return RKDG_OK;
  HIP_TRY(hipMalloc(p, sizeof(double) * n));
  return RKDG_OK;
}

int ensure_adjoint_space(rkdg_ctx* c) {
  const int64_t NE = c->NE, sz = (int64_t)c->neq * c->ndofs;
  if (!c->d_ameans) {
    HIP_TRY(hipMalloc(&c->d_ameans, sizeof(double) * c->neq * NE));
    HIP_TRY(hipMalloc(&c->d_meanbar, sizeof(double) * c->neq * NE));
    HIP_TRY(hipMalloc(&c->d_send_to, sizeof(int64_t) * NE));
    HIP_TRY(hipMalloc(&c->d_send_val, sizeof(double) * NE));
    HIP_TRY(hipMalloc(&c->d_pbar, sizeof(double) * RKDG_MAX_BC * 8));
    HIP_TRY(hipMalloc(&c->d_J, sizeof(double)));
  }
  while (c->d_adj.size() < 7) {
    double* p = nullptr;
    HIP_TRY(hipMalloc(&p, sizeof(double) * sz));
    c->d_adj.push_back(p);
  }
  return RKDG_OK;
}

int rhs_adj_launch(rkdg_ctx* c, const double* x, const double* kbar, double* xbar, const double* addend,
                   double* pbar) {
  RhsAdjArgs a{x, kbar, xbar, addend, pbar, c->d_geo, c->d_nbr, c->d_fcode, c->NE};
  HIP_TRY(launch_dense_rhs_adj(c->geom, c->p, a, c->sp, c->d_dtab, c->n_bdr > 0, c->stream));
  return RKDG_OK;
}

// vbar = (dL/dv)^T ubar (+ objective of the limited state when obj_slot >= 0);
// means: element means of v (computed here when null)
int limit_adj_launch(rkdg_ctx* c, const double* v, const double* means, const double* ubar, double* vbar,
                     int obj_slot, double obj_coef, const AxpyTargets* ax = nullptr) {
  if (!means) {
    MeansArgs ma{v, c->d_ameans, c->d_geo, c->NE};
    HIP_TRY(launch_dense_means(c->geom, c->p, ma, c->d_dtab, c->stream));
    means = c->d_ameans;
  }
  LimAdjArgs a{v, means, ubar, vbar, c->d_meanbar, c->d_send_to, c->d_send_val, c->d_geo, c->d_nbr, c->d_fcode,
               c->NE, obj_slot, obj_coef, AxpyTargets{}};
  if (ax) a.ax = *ax;
  HIP_TRY(launch_dense_limit_adj(c->geom, c->p, a, c->sp, c->d_dtab, c->stream));
  return RKDG_OK;
}

// one recorded step: the vectors of a step a reverse pass needs
struct StepRec {
  double* x;          // step input (limited)
  double* v[4];       // pre-limit stage states v_1..v_s (v_s: before the final limiter)
  double* w[4];       // limited stage inputs W_1..W_{s-1}
  double* mean[4];    // element means of v_1..v_s
};

int rec_vectors(int s) { return 1 + s + (s - 1); }

StepRec rec_at(rkdg_ctx* c, int s, size_t k) {
  // layout per recorded step: x, v_1..v_s, w_1..w_{s-1}, then means (s blocks of neq*NE)
  StepRec r{};
  const size_t per = rec_vectors(s) + 1;  // + one vector-sized block holding the means
  double** b = &c->d_rec[k * per];
  r.x = b[0];
  for (int i = 0; i < s; i++) r.v[i] = b[1 + i];
  for (int i = 0; i < s - 1; i++) r.w[i] = b[1 + s + i];
  for (int i = 0; i < s; i++) r.mean[i] = b[per - 1] + (size_t)i * c->neq * c->NE;
  return r;
}

int ensure_records(rkdg_ctx* c, int s, int64_t steps) {
  const size_t per = rec_vectors(s) + 1;
  const int64_t sz = (int64_t)c->neq * c->ndofs;
  if ((int64_t)s * c->neq * c->NE > sz) return fail(RKDG_E_UNSUPPORTED, "element means do not fit the record block");
  while (c->d_rec.size() < per * (size_t)steps) {
    double* p = nullptr;
    HIP_TRY(hipMalloc(&p, sizeof(double) * sz));
    c->d_rec.push_back(p);
  }
  return RKDG_OK;
}

// forward step from r.x recording the stage states; the limited output goes to xout
int step_record(rkdg_ctx* c, const StepRec& r, double* xout, double dt, int s, const double* a, const double* b) {
  auto A_ = [&](int i, int j) { return a[i * (i - 1) / 2 + j]; };
  HIP_TRY(launch_fill(c->d_lam, 0.0, 8 * c->count, c->stream));
  for (int i = 0; i < s; i++) {
    const double* in = i == 0 ? r.x : r.w[i - 1];
    RhsArgs ra = base_args(c, in);
    ra.lam = c->d_lam + (int64_t)i * c->count;
    int n = 0;
    for (int j = i + 1; j < s; j++) {
      ra.outs[n].out = r.v[j - 1];
      ra.outs[n].base = i == 0 ? nullptr : r.v[j - 1];
      ra.outs[n].coef = A_(j, i) * dt;
      n++;
    }
    const bool last = i == s - 1;
    ra.outs[n].out = r.v[s - 1];
    ra.outs[n].base = i == 0 ? nullptr : r.v[s - 1];
    ra.outs[n].coef = b[i] * dt;
    n++;
    ra.nout = n;
    ra.means = last ? r.mean[s - 1] : r.mean[i];
    ra.means_of = last ? n - 1 : 0;
    HIP_TRY(rhs_launch(c, ra));
    const double* vin = last ? r.v[s - 1] : r.v[i];
    double* dst = last ? xout : r.w[i];
    LimitArgs la{const_cast<double*>(vin), ra.means, c->d_geo, c->d_nbr, c->d_fcode, c->NE, nullptr, nullptr,
                 nullptr, dst};
    HIP_TRY(launch_dense_limit(c->geom, c->p, la, c->sp, c->d_dtab, c->stream));
  }
  return RKDG_OK;
}

// reverse of one recorded step: xbar = (dStep/dx)^T ybar, where ybar is the
// adjoint of the step output plus, when obj_slot >= 0, obj_coef times the
// objective gradient at the step output; pbar accumulates.
int step_adjoint(rkdg_ctx* c, const StepRec& r, double dt, int s, const double* a, const double* b,
                 const double* ybar, double* xbar, int obj_slot, double obj_coef) {
  auto A_ = [&](int i, int j) { return a[i * (i - 1) / 2 + j]; };
  const int64_t sz = (int64_t)c->neq * c->ndofs;
  double* u0bar = c->d_adj[0];
  double* kb[4] = {c->d_adj[1], c->d_adj[2], c->d_adj[3], c->d_adj[4]};
  double* vb = c->d_adj[5];
  double* Wb = c->d_adj[6];
  // the adjoint of each limiter feeds the stage combinations through the
  // targets its pass B updates (u0bar, kbar_i)
  {
    // u0bar = vb; kbar_i = dt b_i vb
    AxpyTargets t{};
    t.n = 0;
    t.y[t.n] = u0bar; t.c[t.n] = 1.0; t.set[t.n] = 1; t.n++;
    for (int i = 0; i < s && t.n < 4; i++) { t.y[t.n] = kb[i]; t.c[t.n] = dt * b[i]; t.set[t.n] = 1; t.n++; }
    int rc = limit_adj_launch(c, r.v[s - 1], r.mean[s - 1], ybar, vb, obj_slot, obj_coef, &t);
    if (rc) return rc;
    if (s == 4) {
      AxpyTargets t2{};
      t2.n = 1; t2.y[0] = kb[3]; t2.c[0] = dt * b[3]; t2.set[0] = 1;
      HIP_TRY(launch_axpy_multi(vb, sz, t2, c->stream));
    }
  }
  int rc = RKDG_OK;
  for (int j = s - 1; j >= 1; j--) {
    rc = rhs_adj_launch(c, r.w[j - 1], kb[j], Wb, nullptr, c->d_pbar);
    if (rc) return rc;
    AxpyTargets t{};
    t.y[t.n] = u0bar; t.c[t.n] = 1.0; t.set[t.n] = 0; t.n++;
    for (int i = 0; i < j; i++) { t.y[t.n] = kb[i]; t.c[t.n] = dt * A_(j, i); t.set[t.n] = 0; t.n++; }
    rc = limit_adj_launch(c, r.v[j - 1], r.mean[j - 1], Wb, vb, -1, 0.0, &t);
    if (rc) return rc;
  }
  return rhs_adj_launch(c, r.x, kb[0], xbar, u0bar, c->d_pbar);
}

int objective_launch(rkdg_ctx* c, const double* x, int slot, double coef, double* J, double* xbar) {
  if (c->obj_slot_cached != slot) {
    std::vector<int> el;
    for (int64_t e = 0; e < c->NE; e++)
      for (int f = 0; f < c->nf; f++)
        if (c->nbr_h[(size_t)f * c->NE + e] == -1 - slot) {
          el.push_back((int)e);
          break;
        }
    (void)hipFree(c->d_obj_elems);
    c->d_obj_elems = nullptr;
    if (!el.empty()) {
      HIP_TRY(hipMalloc(&c->d_obj_elems, sizeof(int) * el.size()));
      HIP_TRY(hipMemcpy(c->d_obj_elems, el.data(), sizeof(int) * el.size(), hipMemcpyHostToDevice));
    }
    c->n_obj_elems = (int)el.size();
    c->obj_slot_cached = slot;
  }
  if (c->n_obj_elems == 0) return RKDG_OK;
  ObjArgs o{x, J, xbar, coef, slot, c->d_geo, c->d_nbr, c->NE, c->d_obj_elems, c->n_obj_elems};
  HIP_TRY(launch_dense_objective(c->geom, c->p, o, c->sp, c->d_dtab, c->stream));
  return RKDG_OK;
}

}  // namespace

// ============================================================================
// ---------------------------------------------------------------- Shu indicator weights
// Extrapolated means of the stencil cells over element t
// (Averager::computeStencilExtrapAverages, averager.cpp:152-215, and
// assembleShiftedElementMatrix, :218-273): for cell k of t's stencil, the
// weight of k's node j is
//   wlast_k / |k| * sum_q w_q phi^k_j(T_k^-1(T_t(xi_q)))
// with (w_q, xi_q) the rule of order p + OrderW on t's reference element,
// wlast_k the Jacobian determinant of k at the last point of that rule (the
// reference reads Trans.Weight() without setting a point), |k| the measure of
// k. T_k^-1 is the affine inverse (triangles, affine hexahedra) or Newton's
// method on the bilinear or trilinear map (quadrilaterals, curved hexahedra),
// not clipped to the element: the neighbour's polynomial is extrapolated over t.
namespace {
struct ShuGeom {
  int geom, dim, nv, p, nd;
  bool curved = false;         // trilinear hexahedra: Newton's method for T^-1
  RefBasis rb;                 // quadrilaterals and triangles
  std::vector<double> x1;      // 1D Gauss-Legendre nodes
  void map(const double* X, const double* xi, double* x) const {
    if (geom == RKDG_GEOM_TRI) {
      for (int d = 0; d < 2; d++) x[d] = X[d] + (X[2 + d] - X[d]) * xi[0] + (X[4 + d] - X[d]) * xi[1];
      return;
    }
    for (int d = 0; d < dim; d++) x[d] = 0.0;
    for (int v = 0; v < nv; v++) {
      const int rx = (v == 1 || v == 2 || v == 5 || v == 6) ? 1 : 0;
      const int ry = (v == 2 || v == 3 || v == 6 || v == 7) ? 1 : 0;
      const int rz = v >= 4 ? 1 : 0;
      double N = (rx ? xi[0] : 1 - xi[0]) * (ry ? xi[1] : 1 - xi[1]);
      if (dim == 3) N *= rz ? xi[2] : 1 - xi[2];
      for (int d = 0; d < dim; d++) x[d] += N * X[v * dim + d];
    }
  }
  // J[i*dim + r] = dx_i / dxi_r
  void jac(const double* X, const double* xi, double* J) const {
    if (geom == RKDG_GEOM_TRI) {
      J[0] = X[2] - X[0], J[1] = X[4] - X[0], J[2] = X[3] - X[1], J[3] = X[5] - X[1];
      return;
    }
    if (geom == RKDG_GEOM_QUAD) {
      bilinear_jac(X, xi[0], xi[1], J);
      return;
    }
    for (int i = 0; i < 9; i++) J[i] = 0.0;
    for (int v = 0; v < 8; v++) {
      const int R[3] = {(v == 1 || v == 2 || v == 5 || v == 6) ? 1 : 0, (v == 2 || v == 3 || v == 6 || v == 7) ? 1 : 0,
                        v >= 4 ? 1 : 0};
      for (int r = 0; r < 3; r++) {
        double t = R[r] ? 1.0 : -1.0;
        for (int q = 0; q < 3; q++)
          if (q != r) t *= R[q] ? xi[q] : 1.0 - xi[q];
        for (int i = 0; i < 3; i++) J[i * 3 + r] += X[v * 3 + i] * t;
      }
    }
  }
  static double det(int d, const double* J) {
    if (d == 2) return J[0] * J[3] - J[1] * J[2];
    return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) + J[2] * (J[3] * J[7] - J[4] * J[6]);
  }
  static void solve(int d, const double* J, const double* b, double* x) {
    double A[9];
    if (d == 2) {
      A[0] = J[3], A[1] = -J[1], A[2] = -J[2], A[3] = J[0];
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
    }
    const double dj = det(d, J);
    for (int i = 0; i < d; i++) {
      double t = 0.0;
      for (int k = 0; k < d; k++) t += A[i * d + k] * b[k];
      x[i] = t / dj;
    }
  }
  void inverse(const double* X, const double* x, double* xi) const {
    if (geom == RKDG_GEOM_TRI || (geom == RKDG_GEOM_HEX && !curved)) {
      // affine: the vertex differences along the reference axes
      const int cols[3] = {1, geom == RKDG_GEOM_TRI ? 2 : 3, 4};
      double J[9], b[3];
      for (int i = 0; i < dim; i++) {
        for (int r = 0; r < dim; r++) J[i * dim + r] = X[cols[r] * dim + i] - X[i];
        b[i] = x[i] - X[i];
      }
      solve(dim, J, b, xi);
      return;
    }
    for (int d = 0; d < dim; d++) xi[d] = 0.5;
    for (int it = 0; it < 60; it++) {
      double xc[3], r[3], J[9], dxi[3];
      map(X, xi, xc);
      for (int i = 0; i < dim; i++) r[i] = x[i] - xc[i];
      jac(X, xi, J);
      solve(dim, J, r, dxi);
      double step = 0.0;
      for (int d = 0; d < dim; d++) {
        xi[d] += dxi[d];
        step = std::max(step, std::fabs(dxi[d]));
      }
      if (step < 1e-15) break;
    }
  }
  void shape(const double* xi, double* phi) const {
    if (geom != RKDG_GEOM_HEX) {
      rb.eval(xi[0], xi[1], phi, nullptr, nullptr);
      return;
    }
    const int N = p + 1;
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++)
          phi[i + N * (j + N * k)] = lag(x1.data(), N, i, xi[0]) * lag(x1.data(), N, j, xi[1]) * lag(x1.data(), N, k, xi[2]);
  }
};
}  // namespace

int build_shu(rkdg_ctx* c) {
  ShuGeom G;
  G.geom = c->geom;
  G.dim = c->dim;
  G.nv = c->geom == RKDG_GEOM_TRI ? 3 : (c->dim == 3 ? 8 : 4);
  G.p = c->p;
  G.nd = c->nd;
  G.curved = c->curved != 0;
  if (c->geom != RKDG_GEOM_HEX) G.rb.init(c->geom == RKDG_GEOM_TRI, c->p);
  G.x1.assign(c->p + 1, 0.0);
  {
    std::vector<double> w(c->p + 1);
    gl_rule(c->p + 1, G.x1.data(), w.data());
  }
  // the rule of order p + OrderW (OrderW: triangle 0, bilinear quad 1, trilinear hex 2)
  const int order = c->p + (c->geom == RKDG_GEOM_TRI ? 0 : (c->dim == 3 ? 2 : 1));
  std::vector<double> pts, wts;
  if (c->geom == RKDG_GEOM_TRI) {
    tri_rule((order + 3) / 2, pts, wts);
  } else {
    const int n = (order | 1) / 2 + 1;
    std::vector<double> x(n), w(n);
    gl_rule(n, x.data(), w.data());
    if (c->dim == 2) {
      quad_rule(n, pts, wts);
    } else {
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            pts.push_back(x[i]), pts.push_back(x[j]), pts.push_back(x[k]);
            wts.push_back(w[i] * w[j] * w[k]);
          }
    }
  }
  const int nq = (int)wts.size(), dim = c->dim, nd = c->nd, nf = c->nf;
  const int64_t NE = c->NE;
  const int stride = G.nv * dim;
  // wlast_k / |k| per element
  std::vector<double> scale(NE);
  for (int64_t k = 0; k < NE; k++) {
    const double* Xk = &c->coords[k * stride];
    double vol = 0.0, wl = 0.0;
    for (int q = 0; q < nq; q++) {
      double J[9];
      G.jac(Xk, &pts[q * dim], J);
      const double dj = ShuGeom::det(dim, J);
      vol += dj * wts[q];
      wl = dj;
    }
    scale[k] = wl / vol;
  }
  std::vector<double> coef((size_t)(nf + 1) * nd * NE, 0.0);
  std::vector<double> phi(nd);
  for (int64_t t = 0; t < NE; t++) {
    const double* Xt = &c->coords[t * stride];
    for (int sl = 0; sl <= nf; sl++) {
      const int64_t k = sl == 0 ? t : (int64_t)c->nbr_h[(size_t)(sl - 1) * NE + t];
      if (k < 0) continue;
      const double* Xk = &c->coords[k * stride];
      std::vector<double> acc(nd, 0.0);
      for (int q = 0; q < nq; q++) {
        double xi[3];
        if (k == t) {
          for (int d = 0; d < dim; d++) xi[d] = pts[q * dim + d];
        } else {
          double xp[3];
          G.map(Xt, &pts[q * dim], xp);
          G.inverse(Xk, xp, xi);
        }
        G.shape(xi, phi.data());
        for (int j = 0; j < nd; j++) acc[j] += scale[k] * wts[q] * phi[j];
      }
      for (int j = 0; j < nd; j++) coef[((size_t)sl * nd + j) * NE + t] = acc[j];
    }
  }
  HIP_TRY(hipMalloc(&c->d_shu_coef, sizeof(double) * coef.size()));
  HIP_TRY(hipMemcpy(c->d_shu_coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&c->d_shu_flag, (size_t)NE));
  return RKDG_OK;
}

extern "C" {

const char* rkdg_last_error(void) { return g_last_error.c_str(); }

int rkdg_create(const rkdg_mesh_desc* mesh, const rkdg_scheme_desc* scheme, rkdg_ctx** out) {
  return rkdg_create_ex(mesh, scheme, 0, out);
}

int rkdg_create_ex(const rkdg_mesh_desc* mesh, const rkdg_scheme_desc* scheme, int flags, rkdg_ctx** out) {
  return rkdg_create_nc(mesh, scheme, flags, nullptr, out);
}

int rkdg_create_nc(const rkdg_mesh_desc* mesh, const rkdg_scheme_desc* scheme, int flags, const rkdg_nc_desc* nc,
                   rkdg_ctx** out) {
  if (!mesh || !scheme || !out) return fail(RKDG_E_ARG, "null argument");
  *out = nullptr;
  if (nc && nc->count < 0) return fail(RKDG_E_ARG, "negative count of non-conforming faces");
  if (nc && nc->count > 0 && scheme->indicator == RKDG_IND_SHU)
    return fail(RKDG_E_UNSUPPORTED, "the Shu indicator is not built for non-conforming faces");
  rkdg_ctx* c = new rkdg_ctx();
  c->nc_in = nc;
  c->force_dense = (flags & RKDG_CREATE_DENSE) ? 1 : 0;
  c->two_colour = (flags & RKDG_CREATE_TWO_COLOUR) ? 1 : 0;
  int rc = RKDG_OK;
  if (scheme->indicator == RKDG_IND_SHU && scheme->limiter != RKDG_LIM_FINDIFF)
    // IndicatorShu writes no minValues, which the multiplier reads
    // (indicator_shu.cpp:68-73, limiter_multiplier.cpp:35)
    rc = fail(RKDG_E_ARG, "the Shu indicator sets no multiplier: pair it with the FinDiff limiter");
  else if (scheme->linearize && scheme->limiter == RKDG_LIM_MULTI && mesh->order != 1 && mesh->geom != RKDG_GEOM_TRI)
    rc = fail(RKDG_E_ARG, "linearize reads nd points of an order p+1 rule: p = 1 quadrilaterals and hexahedra only");
  if (rc == RKDG_OK) rc = build_ctx(c, mesh, scheme);
  c->nc_in = nullptr;
  if (rc == RKDG_OK && c->sp.indicator == RKDG_IND_SHU) rc = build_shu(c);
  if (rc == RKDG_OK) {
    const hipError_t se = noalloc_scratch(&c->scratch);
    if (se != hipSuccess) rc = fail(RKDG_E_HIP, std::string("scratch array: ") + hipGetErrorString(se));
  }
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
  info->two_colour = c->rb;
  return RKDG_OK;
}

int rkdg_set_stream(rkdg_ctx* c, void* stream) {
  if (!c) return fail(RKDG_E_ARG, "null context");
  c->stream = (hipStream_t)stream;
  return RKDG_OK;
}

int rkdg_rhs(rkdg_ctx* c, const double* x, double* y, double* lambda_max) {
  if (!c || !x || !y) return fail(RKDG_E_ARG, "null argument");
  HIP_TRY(launch_fill(c->d_lam, 0.0, c->count, c->stream));
  RhsArgs a = base_args(c, x);
  a.k = y;
  HIP_TRY(rhs_launch(c, a));
  if (lambda_max) return read_lambda(c, 0, lambda_max);
  return RKDG_OK;
}

int rkdg_face_lambda(rkdg_ctx* c, const double* x, double* lambda_max) {
  if (!c || !x || !lambda_max) return fail(RKDG_E_ARG, "null argument");
  HIP_TRY(launch_fill(c->d_lam, 0.0, c->count, c->stream));
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
  MeansArgs ma{x, means, c->d_geo, c->NE, c->d_cgeo};
  HIP_TRY(c->dense ? launch_dense_means(c->geom, c->p, ma, c->d_dtab, c->stream)
                   : launch_tensor_means(c->dim, c->p, ma, c->d_tb, c->stream));
  return RKDG_OK;
}

// ExplicitRKLimitedSolver::Step (rk_explicit_limited.cpp:30-82), see step_impl.
int rkdg_step(rkdg_ctx* c, double* x, double t, double dt, int s, const double* a, const double* b,
              const double* cc, double* lambda_max) {
  (void)t;
  (void)cc;
  if (!c || !x || !a || !b) return fail(RKDG_E_ARG, "null argument");
  if (s < 1 || s > 4) return fail(RKDG_E_ARG, "1 <= s <= 4 stages supported");
  // buffers: d_stage[0..s-2] = W_1..W_{s-1}, d_stage[3] = accumulator
  int rc = ensure_stage_buffers(c, 4);
  if (rc) return rc;
  rc = step_impl(c, x, dt, s, a, b, c->d_stage, lambda_max != nullptr);
  if (rc) return rc;
  if (lambda_max) return read_lambda(c, s - 1, lambda_max);
  return RKDG_OK;
}

// SSP-RK3 in the Shu-Osher form with the limiter applied to each stage value
// (no counterpart in the reference, whose ExplicitRKLimitedSolver consumes
// the Butcher tableau; the two forms agree when the limiter does not act):
//   u1 = x + dt R(x); u2 = 3/4 x + 1/4 (u1 + dt R(u1)); x = 1/3 x + 2/3 (u2 + dt R(u2)),
// each limited. Each stage kernel reads its input and x and writes one vector
// (the last one x in place: a stage reads x only at its own elements), 8
// vectors per step against 12 for the Butcher form with its stored stage sums.
int rkdg_step_ssprk3(rkdg_ctx* c, double* x, double t, double dt, double* lambda_max) {
  (void)t;
  if (!c || !x) return fail(RKDG_E_ARG, "null argument");
  int rc = ensure_stage_buffers(c, 2);
  if (rc) return rc;
  double* U[2] = {c->d_stage[0], c->d_stage[1]};
  HIP_TRY(launch_fill(c->d_lam, 0.0, 8 * c->count, c->stream));
  const double wb[3] = {0.0, 0.75, 1.0 / 3.0}, wi[3] = {0.0, 0.25, 2.0 / 3.0};
  for (int i = 0; i < 3; i++) {
    const double* in = i == 0 ? x : U[i - 1];
    double* out = i == 2 ? x : U[i];
    RhsArgs ra = base_args(c, in);
    ra.lam = c->d_lam + (int64_t)i * c->count;
    ra.no_lambda = (i < 2 || !lambda_max) ? 1 : 0;
    ra.nout = 1;
    ra.outs[0].out = out;
    ra.outs[0].base = i == 0 ? nullptr : x;
    ra.outs[0].coef = dt;
    ra.outs[0].wb = wb[i];
    ra.outs[0].wi = wi[i];
    ra.means = c->d_means;
    ra.means_of = 0;
    HIP_TRY(rhs_launch(c, ra));
    rc = limit_impl(c, out, nullptr, true);
    if (rc) return rc;
  }
  if (lambda_max) return read_lambda(c, 2, lambda_max);
  return RKDG_OK;
}

// rkdg_step without any allocation: the stage vectors live in this device's
// static scratch array (rkdg_noalloc.hip), located once at rkdg_create. The
// wave speed stays on the device (copied to lambda_dev, a device double, when
// given) and no host synchronisation or pinned staging happens, so a caller
// that may not allocate (a loop under HIP graph capture, a serving loop) can
// call it. Calls that overlap on one GPU share the array and race.
int rkdg_step_noalloc(rkdg_ctx* c, double* x, double t, double dt, int s, const double* a, const double* b,
                      const double* cc, double* lambda_dev) {
  (void)t;
  (void)cc;
  if (!c || !x || !a || !b) return fail(RKDG_E_ARG, "null argument");
  if (s < 1 || s > 4) return fail(RKDG_E_ARG, "1 <= s <= 4 stages supported");
  if (!c->scratch) return fail(RKDG_E_HIP, "no scratch array on this context's device");
  const int64_t sz = (int64_t)c->neq * c->ndofs;
  const int64_t nvec = 4;
  if (sz > (int64_t)(kScratchDoubles / nvec))
    return fail(RKDG_E_ARG, "state of " + std::to_string(sz) + " doubles exceeds the allocation-free limit of " +
                                std::to_string(kScratchDoubles / nvec) + " (4 stage vectors in " +
                                std::to_string(kScratchDoubles * 8 >> 20) + " MiB of static scratch)");
  double* W[4];
  for (int i = 0; i < 4; i++) W[i] = c->scratch + i * sz;
  int rc = step_impl(c, x, dt, s, a, b, W, lambda_dev != nullptr);
  if (rc) return rc;
  if (lambda_dev) HIP_TRY(launch_copy(lambda_dev, c->d_lam + (int64_t)(s - 1) * c->count, 1, c->stream));
  return RKDG_OK;
}

// ---------------------------------------------------------------------------
// Post-step diagnostics

namespace {
constexpr int kTotalBlocks = 512;

double element_volume(const rkdg_ctx* c, int64_t e) {
  if (c->curved) return c->vol_h[(size_t)e];
  const double* X = &c->coords[(size_t)e * c->nv * c->dim];
  if (c->geom == RKDG_GEOM_TRI)
    return 0.5 * std::fabs((X[2] - X[0]) * (X[5] - X[1]) - (X[3] - X[1]) * (X[4] - X[0]));
  if (c->geom == RKDG_GEOM_QUAD) {  // straight-sided quadrilateral: the polygon's area
    double a = 0.0;
    for (int v = 0; v < 4; v++) {
      const int w = (v + 1) % 4;
      a += X[2 * v] * X[2 * w + 1] - X[2 * w] * X[2 * v + 1];
    }
    return 0.5 * std::fabs(a);
  }
  double J[9];  // affine hexahedron: |det J|, columns X1 - X0, X3 - X0, X4 - X0
  const int cols[3] = {1, 3, 4};
  for (int i = 0; i < 3; i++)
    for (int r = 0; r < 3; r++) J[i * 3 + r] = X[cols[r] * 3 + i] - X[i];
  return std::fabs(J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                   J[2] * (J[3] * J[7] - J[4] * J[6]));
}
}  // namespace

// Conservation totals: integral over the mesh of every conserved variable,
// sum_e |e| mean_e (the element means are exact quadratures, so this is
// sum_e sum_k w_k u_k of any order; check_total_energy.cpp:3-63 forms the last
// variable's total with the order-(p+1) rule at the nodes, exact for Q1 only).
int rkdg_totals(rkdg_ctx* c, const double* x, double* totals) {
  if (!c || !x || !totals) return fail(RKDG_E_ARG, "null argument");
  const int64_t NE = c->NE;
  if (!c->d_vol) {
    std::vector<double> vol((size_t)NE);
    for (int64_t e = 0; e < NE; e++) vol[e] = element_volume(c, e);
    HIP_TRY(hipMalloc(&c->d_vol, sizeof(double) * NE));
    HIP_TRY(hipMemcpy(c->d_vol, vol.data(), sizeof(double) * NE, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_dmeans, sizeof(double) * c->neq * NE));
    HIP_TRY(hipMalloc(&c->d_partial, sizeof(double) * c->neq * kTotalBlocks));
  }
  MeansArgs ma{x, c->d_dmeans, c->d_geo, NE, c->d_cgeo};
  HIP_TRY(c->dense ? launch_dense_means(c->geom, c->p, ma, c->d_dtab, c->stream)
                   : launch_tensor_means(c->dim, c->p, ma, c->d_tb, c->stream));
  HIP_TRY(launch_totals(c->d_dmeans, c->d_vol, NE, c->neq, c->d_partial, kTotalBlocks, c->stream));
  std::vector<double> part((size_t)c->neq * kTotalBlocks);
  HIP_TRY(hipMemcpyAsync(part.data(), c->d_partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int v = 0; v < c->neq; v++) {
    double s = 0.0;
    for (int b = 0; b < kTotalBlocks; b++) s += part[(size_t)b * c->neq + v];
    totals[v] = s;
  }
  return RKDG_OK;
}

// Primitive fields at the nodes (main.cpp:388-451): prim[q*NE*nd + i] for
// q = rho, u_0..u_{dim-1}, p, T; acc (nullable) += weight * the same (the time
// integrals UMean, pMean, TMean, rhoMean of main.cpp:705-730 use weight dt).
int rkdg_primitives(rkdg_ctx* c, const double* x, double* prim, double* acc, double weight) {
  if (!c || !x || (!prim && !acc)) return fail(RKDG_E_ARG, "null argument");
  HIP_TRY(launch_primitives(x, c->ndofs, c->dim, c->sp.gamma, c->sp.gas_constant, c->sp.covolume, prim, acc, weight,
                            c->stream));
  return RKDG_OK;
}

// The state at npts reference points of every element (host ref_pts [npts][dim]
// in the reference element: [0,1]^dim, or the unit triangle); out (DEVICE)
// [v][NE][npts]. Used for output at the vertices (ParaView) and probes.
int rkdg_interpolate(rkdg_ctx* c, const double* x, int npts, const double* ref_pts, double* out) {
  if (!c || !x || !ref_pts || !out) return fail(RKDG_E_ARG, "null argument");
  if (npts < 1 || npts > 4096) return fail(RKDG_E_ARG, "1 <= npts <= 4096");
  const int nd = c->nd, dim = c->dim;
  std::vector<double> phi((size_t)npts * nd);
  if (c->dense) {
    RefBasis rb;
    rb.init(c->geom == RKDG_GEOM_TRI ? 1 : 0, c->p);
    for (int q = 0; q < npts; q++) rb.eval(ref_pts[q * 2], ref_pts[q * 2 + 1], &phi[(size_t)q * nd], nullptr, nullptr);
  } else {
    const int N = c->p + 1;
    for (int q = 0; q < npts; q++)
      for (int k = 0; k < nd; k++) {
        double v = 1.0;
        int kk = k;
        for (int d = 0; d < dim; d++) {
          v *= lag(c->tb.x, N, kk % N, ref_pts[q * dim + d]);
          kk /= N;
        }
        phi[(size_t)q * nd + k] = v;
      }
  }
  if (c->phi_cap < npts * nd) {
    dfree(c->d_phi);
    c->d_phi = nullptr;
    HIP_TRY(hipMalloc(&c->d_phi, sizeof(double) * npts * nd));
    c->phi_cap = npts * nd;
  }
  HIP_TRY(hipMemcpyAsync(c->d_phi, phi.data(), sizeof(double) * phi.size(), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(launch_interp(x, c->NE, nd, c->neq, c->d_phi, npts, out, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));  // the host table must outlive the copy
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
  if (c->geom == RKDG_GEOM_TRI) {
    for (int64_t e = 0; e < NE; e++) {
      const double* X = &c->coords[e * 6];
      for (int n = 0; n < c->nd; n++) {
        const double xi = c->ref_nodes[2 * n], eta = c->ref_nodes[2 * n + 1];
        for (int d = 0; d < 2; d++)
          xyz[(e * c->nd + n) * 2 + d] = X[d] + (X[2 + d] - X[d]) * xi + (X[4 + d] - X[d]) * eta;
      }
    }
    return RKDG_OK;
  }
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

int rkdg_error_elements(rkdg_ctx* c, int64_t* pair) {
  if (!c || !pair) return fail(RKDG_E_ARG, "null argument");
  pair[0] = c->err_pair[0];
  pair[1] = c->err_pair[1];
  return RKDG_OK;
}

// ---------------------------------------------------------------------------
// Discrete adjoint

int rkdg_set_bc_params(rkdg_ctx* c, int slot, const double* params) {
  if (!c || !params) return fail(RKDG_E_ARG, "null argument");
  if (slot < 0 || slot >= c->sp.num_bc) return fail(RKDG_E_ARG, "boundary slot out of range");
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 8; k++) c->sp.bc_param[slot][k] = params[k];
  return RKDG_OK;
}

int rkdg_rhs_adjoint(rkdg_ctx* c, const double* x, const double* kbar, double* xbar, const double* addend,
                     double* bc_param_bar) {
  if (c && (c->sp.indicator == RKDG_IND_SHU || c->sp.linearize))
    return fail(RKDG_E_UNSUPPORTED, "the discrete adjoint differentiates the BJ-family limiter without linearize");
  if (!c || !x || !kbar || !xbar) return fail(RKDG_E_ARG, "null argument");
  int rc = adjoint_supported(c);
  if (rc) return rc;
  return rhs_adj_launch(c, x, kbar, xbar, addend, bc_param_bar);
}

int rkdg_limit_adjoint(rkdg_ctx* c, const double* x, const double* ubar, double* xbar) {
  if (c && (c->sp.indicator == RKDG_IND_SHU || c->sp.linearize))
    return fail(RKDG_E_UNSUPPORTED, "the discrete adjoint differentiates the BJ-family limiter without linearize");
  if (!c || !x || !ubar || !xbar) return fail(RKDG_E_ARG, "null argument");
  int rc = adjoint_supported(c);
  if (rc) return rc;
  rc = ensure_adjoint_space(c);
  if (rc) return rc;
  return limit_adj_launch(c, x, nullptr, ubar, xbar, -1, 0.0);
}

int rkdg_objective(rkdg_ctx* c, const double* x, int bc_slot, double coef, double* J, double* xbar) {
  if (!c || !x) return fail(RKDG_E_ARG, "null argument");
  if (!c->dense || c->dim != 2) return fail(RKDG_E_UNSUPPORTED, "objective is built for the dense 2D kernels");
  if (bc_slot < 0 || bc_slot >= c->sp.num_bc) return fail(RKDG_E_ARG, "boundary slot out of range");
  return objective_launch(c, x, bc_slot, coef, J, xbar);
}

int rkdg_step_adjoint(rkdg_ctx* c, const double* x, double dt, int s, const double* a, const double* b,
                      const double* cc, const double* ybar, double* xbar, double* bc_param_bar) {
  if (c && (c->sp.indicator == RKDG_IND_SHU || c->sp.linearize))
    return fail(RKDG_E_UNSUPPORTED, "the discrete adjoint differentiates the BJ-family limiter without linearize");
  (void)cc;
  if (!c || !x || !a || !b || !ybar || !xbar) return fail(RKDG_E_ARG, "null argument");
  if (s < 2 || s > 4) return fail(RKDG_E_ARG, "2 <= s <= 4 stages supported");
  int rc = adjoint_supported(c);
  if (!rc) rc = ensure_adjoint_space(c);
  if (!rc) rc = ensure_records(c, s, 1);
  if (!rc) rc = ensure_stage_buffers(c, 4);
  if (rc) return rc;
  StepRec r = rec_at(c, s, 0);
  HIP_TRY(launch_copy(r.x, x, (int64_t)c->neq * c->ndofs, c->stream));
  rc = step_record(c, r, c->d_stage[3], dt, s, a, b);
  if (rc) return rc;
  HIP_TRY(launch_fill(c->d_pbar, 0.0, RKDG_MAX_BC * 8, c->stream));
  rc = step_adjoint(c, r, dt, s, a, b, ybar, xbar, -1, 0.0);
  if (rc) return rc;
  if (bc_param_bar) {
    double h[RKDG_MAX_BC * 8];
    HIP_TRY(hipMemcpyAsync(h, c->d_pbar, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int k = 0; k < RKDG_MAX_BC * 8; k++) bc_param_bar[k] = h[k];
  }
  return check_device_error(c, nullptr);
}

int rkdg_forward_solve(rkdg_ctx* c, double* x, int64_t nsteps, const double* dts, int s, const double* a,
                       const double* b, const double* cc, int obj_slot, double* J) {
  if (!c || !x || !dts || !a || !b) return fail(RKDG_E_ARG, "null argument");
  if (obj_slot >= c->sp.num_bc) return fail(RKDG_E_ARG, "boundary slot out of range");
  if (obj_slot >= 0 && (!c->dense || c->dim != 2)) return fail(RKDG_E_UNSUPPORTED, "objective needs the dense 2D kernels");
  int rc = ensure_adjoint_space(c);
  if (rc) return rc;
  HIP_TRY(launch_fill(c->d_J, 0.0, 1, c->stream));
  for (int64_t n = 0; n < nsteps; n++) {
    rc = rkdg_step(c, x, 0.0, dts[n], s, a, b, cc, nullptr);
    if (rc) return rc;
    if (obj_slot >= 0) {
      rc = objective_launch(c, x, obj_slot, dts[n], c->d_J, nullptr);
      if (rc) return rc;
    }
  }
  double h = 0.0;
  HIP_TRY(hipMemcpyAsync(&h, c->d_J, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  rc = check_device_error(c, nullptr);
  if (J) *J = h;
  return rc;
}

// J = sum_{n=1..N} dt_{n-1} P(x_n) with x_{n+1} = Step(x_n, dt_n) (dt schedule
// given, held fixed), P = pressure integrated over the faces of boundary slot
// obj_slot. Forward pass with a checkpoint every K steps, then per segment
// (last first) a recomputation recording the stage states and the reverse
// sweep through it: 2N forward steps and N adjoint steps, K recorded steps and
// N/K checkpoints resident (K = ceil(sqrt(N)) by default).
int rkdg_adjoint_solve(rkdg_ctx* c, const double* x0, int64_t N, const double* dts, int s, const double* a,
                       const double* b, const double* cc, int obj_slot, int64_t every, double* J, double* x0bar,
                       double* bc_param_bar, double* xN) {
  if (c && (c->sp.indicator == RKDG_IND_SHU || c->sp.linearize))
    return fail(RKDG_E_UNSUPPORTED, "the discrete adjoint differentiates the BJ-family limiter without linearize");
  if (!c || !x0 || !dts || !a || !b || !x0bar) return fail(RKDG_E_ARG, "null argument");
  if (s < 2 || s > 4) return fail(RKDG_E_ARG, "2 <= s <= 4 stages supported");
  if (N < 1) return fail(RKDG_E_ARG, "nsteps must be positive");
  if (obj_slot < 0 || obj_slot >= c->sp.num_bc) return fail(RKDG_E_ARG, "boundary slot out of range");
  int rc = adjoint_supported(c);
  if (!rc) rc = ensure_adjoint_space(c);
  if (!rc) rc = ensure_stage_buffers(c, 4);
  if (rc) return rc;
  const int64_t K = every > 0 ? every : (int64_t)std::ceil(std::sqrt((double)N));
  const int64_t ncp = (N + K - 1) / K;
  rc = ensure_records(c, s, K);
  if (rc) return rc;
  const int64_t sz = (int64_t)c->neq * c->ndofs;
  while ((int64_t)c->d_cp.size() < ncp + 1) {
    double* p = nullptr;
    HIP_TRY(hipMalloc(&p, sizeof(double) * sz));
    c->d_cp.push_back(p);
  }
  double* x = c->d_cp[ncp];  // running state of the forward pass
  HIP_TRY(launch_copy(x, x0, sz, c->stream));
  HIP_TRY(launch_fill(c->d_J, 0.0, 1, c->stream));
  HIP_TRY(launch_fill(c->d_pbar, 0.0, RKDG_MAX_BC * 8, c->stream));
  for (int64_t n = 0; n < N; n++) {
    if (n % K == 0) HIP_TRY(launch_copy(c->d_cp[n / K], x, sz, c->stream));
    rc = rkdg_step(c, x, 0.0, dts[n], s, a, b, cc, nullptr);
    if (rc) return rc;
    rc = objective_launch(c, x, obj_slot, dts[n], c->d_J, nullptr);
    if (rc) return rc;
  }
  if (xN) HIP_TRY(launch_copy(xN, x, sz, c->stream));
  // reverse: lam holds the adjoint of the current step output (zero at the end:
  // the final objective term enters through the fused objective of the last step)
  double* lam = x0bar;
  double* lam2 = c->d_cp[ncp];
  HIP_TRY(launch_fill(lam, 0.0, sz, c->stream));
  for (int64_t seg = ncp - 1; seg >= 0; seg--) {
    const int64_t start = seg * K, end = std::min(N, start + K);
    for (int64_t n = start; n < end; n++) {
      StepRec r = rec_at(c, s, (size_t)(n - start));
      if (n == start) HIP_TRY(launch_copy(r.x, c->d_cp[seg], sz, c->stream));
      double* xout = n + 1 < end ? rec_at(c, s, (size_t)(n + 1 - start)).x : c->d_stage[3];
      rc = step_record(c, r, xout, dts[n], s, a, b);
      if (rc) return rc;
    }
    for (int64_t n = end - 1; n >= start; n--) {
      StepRec r = rec_at(c, s, (size_t)(n - start));
      // the objective term of x_{n+1} (weight dt_n) joins the adjoint of the step output
      rc = step_adjoint(c, r, dts[n], s, a, b, lam, lam2, obj_slot, dts[n]);
      if (rc) return rc;
      std::swap(lam, lam2);
    }
  }
  if (lam != x0bar) HIP_TRY(launch_copy(x0bar, lam, sz, c->stream));
  double hJ = 0.0;
  double hp[RKDG_MAX_BC * 8];
  HIP_TRY(hipMemcpyAsync(&hJ, c->d_J, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(hp, c->d_pbar, sizeof(hp), hipMemcpyDeviceToHost, c->stream));
  rc = check_device_error(c, nullptr);
  if (J) *J = hJ;
  if (bc_param_bar)
    for (int k = 0; k < RKDG_MAX_BC * 8; k++) bc_param_bar[k] = hp[k];
  return rc;
}

// ---------------------------------------------------------------------------
// Batched ensembles

int rkdg_batch_create(const rkdg_mesh_desc* meshes, const rkdg_scheme_desc* schemes, int count,
                      rkdg_ctx** out) {
  if (!meshes || !schemes || !out || count < 1) return fail(RKDG_E_ARG, "null argument or count < 1");
  *out = nullptr;
  const rkdg_mesh_desc& m0 = meshes[0];
  if (m0.geom == RKDG_GEOM_HEX) return fail(RKDG_E_UNSUPPORTED, "batched ensembles of hexahedra are not built");
  const int nv = m0.geom == RKDG_GEOM_TRI ? 3 : 4;
  const rkdg_scheme_desc& s0 = schemes[0];
  if (s0.indicator == RKDG_IND_SHU || s0.linearize)
    return fail(RKDG_E_UNSUPPORTED, "batched ensembles run the BJ-family indicator without linearize");
  // every instance shares the discretisation; only boundary parameters differ
  for (int i = 1; i < count; i++) {
    const rkdg_scheme_desc& si = schemes[i];
    bool same = si.riemann == s0.riemann && si.wave_speeds == s0.wave_speeds && si.hllc_abs == s0.hllc_abs &&
                si.gamma == s0.gamma && si.gas_constant == s0.gas_constant && si.covolume == s0.covolume &&
                si.num_bc == s0.num_bc && si.indicator == s0.indicator && si.phi == s0.phi &&
                si.venkat_k == s0.venkat_k && si.limiter == s0.limiter && si.last_hope == s0.last_hope &&
                si.pow_nd == s0.pow_nd;
    for (int q = 0; q < RKDG_MAX_BC && same; q++)
      same = si.bc_attr[q] == s0.bc_attr[q] && si.bc_type[q] == s0.bc_type[q];
    if (!same) return fail(RKDG_E_ARG, "instance " + std::to_string(i) + ": scheme differs beyond boundary parameters");
    if (meshes[i].geom != m0.geom || meshes[i].order != m0.order)
      return fail(RKDG_E_ARG, "instance " + std::to_string(i) + ": element type or order differs");
  }
  // union mesh: instances concatenated, vertex ids shifted apart
  std::vector<double> coords;
  std::vector<int64_t> verts, bverts;
  std::vector<int32_t> battr;
  std::vector<int64_t> off(count + 1, 0);
  int64_t vbase = 0;
  for (int i = 0; i < count; i++) {
    const rkdg_mesh_desc& m = meshes[i];
    if (m.num_elements <= 0) return fail(RKDG_E_ARG, "instance " + std::to_string(i) + ": empty mesh");
    off[i + 1] = off[i] + m.num_elements;
    coords.insert(coords.end(), m.elem_coords, m.elem_coords + m.num_elements * nv * 2);
    int64_t vmax = -1;
    for (int64_t k = 0; k < m.num_elements * nv; k++) {
      verts.push_back(m.elem_verts[k] + vbase);
      vmax = std::max(vmax, m.elem_verts[k]);
    }
    for (int64_t k = 0; k < m.num_bdr * 2; k++) {
      bverts.push_back(m.bdr_verts[k] + vbase);
      vmax = std::max(vmax, m.bdr_verts[k]);
    }
    battr.insert(battr.end(), m.bdr_attr, m.bdr_attr + m.num_bdr);
    vbase += vmax + 1;
  }
  rkdg_mesh_desc u = m0;
  u.num_elements = off[count];
  u.elem_coords = coords.data();
  u.elem_verts = verts.data();
  u.num_bdr = (int64_t)battr.size();
  u.bdr_verts = bverts.data();
  u.bdr_attr = battr.data();

  rkdg_ctx* c = new rkdg_ctx();
  c->count = count;
  c->force_dense = 1;
  auto bail = [&](int rc) {
    std::string keep = g_last_error;
    free_ctx(c);
    g_last_error = keep;
    return rc;
  };
  int rc = build_ctx(c, &u, &s0);
  if (rc != RKDG_OK) return bail(rc);
  c->inst_off = off;
  std::vector<int> inst(c->NE);
  std::vector<double> bcp((size_t)count * RKDG_MAX_BC * 8), hm(count, 1e300);
  for (int i = 0; i < count; i++) {
    for (int64_t e = off[i]; e < off[i + 1]; e++) {
      inst[e] = i;
      hm[i] = std::min(hm[i], c->h_elem[e]);
    }
    for (int q = 0; q < RKDG_MAX_BC; q++)
      for (int k = 0; k < 8; k++) bcp[((size_t)i * RKDG_MAX_BC + q) * 8 + k] = schemes[i].bc_param[q][k];
  }
  auto up = [&](auto** d, const auto* h, size_t n) -> hipError_t {
    hipError_t e = hipMalloc((void**)d, sizeof(**d) * n);
    if (e != hipSuccess) return e;
    return h ? hipMemcpy(*d, h, sizeof(**d) * n, hipMemcpyHostToDevice) : hipMemset(*d, 0, sizeof(**d) * n);
  };
  const double* nod = nullptr;
  const int* noi = nullptr;
  const long long* noll = nullptr;
  hipError_t he = hipSuccess;
  if (he == hipSuccess) he = up(&c->d_inst, inst.data(), (size_t)c->NE);
  if (he == hipSuccess) he = up(&c->d_bcp, bcp.data(), bcp.size());
  if (he == hipSuccess) he = up(&c->d_hmin, hm.data(), (size_t)count);
  if (he == hipSuccess) he = up(&c->d_active, noi, (size_t)count);
  if (he == hipSuccess) he = up(&c->d_t, nod, (size_t)count);
  if (he == hipSuccess) he = up(&c->d_dt, nod, (size_t)count);
  if (he == hipSuccess) he = up(&c->d_dtr, nod, (size_t)count);
  if (he == hipSuccess) he = up(&c->d_tf, nod, (size_t)count);
  if (he == hipSuccess) he = up(&c->d_steps, noll, (size_t)count);
  if (he == hipSuccess) he = up(&c->d_nact, noi, (size_t)1);
  if (he == hipSuccess) he = hipStreamCreate(&c->bstream);
  if (he != hipSuccess) return bail(fail(RKDG_E_HIP, std::string("batch buffers: ") + hipGetErrorString(he)));
  *out = c;
  return RKDG_OK;
}

int rkdg_batch_offsets(rkdg_ctx* c, int64_t* elem_offsets) {
  if (!c || !elem_offsets) return fail(RKDG_E_ARG, "null argument");
  if (c->inst_off.empty()) {
    elem_offsets[0] = 0;
    elem_offsets[1] = c->NE;
    return RKDG_OK;
  }
  std::copy(c->inst_off.begin(), c->inst_off.end(), elem_offsets);
  return RKDG_OK;
}

// Start of the time loop for every instance (main.cpp:265, 467-484): limit the
// projected state, then the first dt from the face wave speeds.
int rkdg_batch_init(rkdg_ctx* c, double* x, const double* t0, const double* t_final, double cfl,
                    int64_t max_steps) {
  if (!c || !x || !t0 || !t_final) return fail(RKDG_E_ARG, "null argument");
  if (!c->d_inst) return fail(RKDG_E_ARG, "not a batch context");
  if (!(cfl > 0.0)) return fail(RKDG_E_ARG, "cfl must be positive");
  c->cfl = cfl;
  c->max_steps = max_steps > 0 ? max_steps : 0;
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gkey.clear();
  }
  Restore<hipStream_t> keep_stream(c->stream);
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->stream = c->bstream;
  HIP_TRY(hipMemcpyAsync(c->d_tf, t_final, sizeof(double) * c->count, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->d_t, t0, sizeof(double) * c->count, hipMemcpyHostToDevice, c->stream));
  int rc = limit_impl(c, x, nullptr, false);
  if (rc) return rc;
  HIP_TRY(launch_fill(c->d_lam, 0.0, c->count, c->stream));
  RhsArgs a = base_args(c, x);
  a.face_only = 1;
  HIP_TRY(rhs_launch(c, a));
  BatchTime b = batch_time(c, c->d_lam);
  HIP_TRY(launch_batch_init(b, c->stream));
  return check_device_error(c, nullptr);
}

// Advances every active instance by up to nsteps steps (one HIP graph launch
// per step when graphs are on); n_active (host, nullable) receives the number
// of instances still running.
int rkdg_batch_advance(rkdg_ctx* c, double* x, int s, const double* a, const double* b, const double* cc,
                       int64_t nsteps, int* n_active) {
  (void)cc;
  if (!c || !x || !a || !b) return fail(RKDG_E_ARG, "null argument");
  if (!c->d_inst) return fail(RKDG_E_ARG, "not a batch context");
  if (s < 2 || s > 4) return fail(RKDG_E_ARG, "2 <= s <= 4 stages supported for batches");
  int rc = ensure_stage_buffers(c, 4);
  if (rc) return rc;
  Restore<hipStream_t> keep_stream(c->stream);
  Restore<bool> keep_prof(c->prof);
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->stream = c->bstream;
  std::vector<double> key;
  key.push_back((double)(uintptr_t)x);
  key.push_back(s);
  for (int i = 0; i < s * (s - 1) / 2; i++) key.push_back(a[i]);
  for (int i = 0; i < s; i++) key.push_back(b[i]);
  if (c->use_graph && (!c->gexec || key != c->gkey)) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->prof = false;  // events are not recorded into the graph
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = batch_step_launch(c, x, s, a, b);
    hipError_t ce = hipStreamEndCapture(c->stream, &graph);
    c->prof = keep_prof.saved;
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIP_TRY(ce);
    hipError_t ie = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(ie);
    c->gkey = key;
  }
  for (int64_t n = 0; n < nsteps; n++) {
    if (c->use_graph) {
      HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
    } else {
      rc = batch_step_launch(c, x, s, a, b);
      if (rc) return rc;
    }
  }
  int h = 0;
  HIP_TRY(hipMemcpyAsync(&h, c->d_nact, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  rc = check_device_error(c, nullptr);
  if (n_active) *n_active = h;
  return rc;
}

// Host copies of the per-instance time state; any pointer may be NULL.
int rkdg_batch_status(rkdg_ctx* c, double* t, double* dt, int64_t* steps, int* active) {
  if (!c) return fail(RKDG_E_ARG, "null context");
  if (!c->d_inst) return fail(RKDG_E_ARG, "not a batch context");
  const size_t n = c->count;
  HIP_TRY(hipStreamSynchronize(c->bstream));
  if (t) HIP_TRY(hipMemcpy(t, c->d_t, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (dt) HIP_TRY(hipMemcpy(dt, c->d_dt, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (steps) {
    std::vector<long long> h(n);
    HIP_TRY(hipMemcpy(h.data(), c->d_steps, sizeof(long long) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) steps[i] = h[i];
  }
  if (active) HIP_TRY(hipMemcpy(active, c->d_active, sizeof(int) * n, hipMemcpyDeviceToHost));
  return RKDG_OK;
}

// use_graph: 1 = capture a step once and replay it; grid_cap: persistent grid of
// the right-hand-side kernel (0 = one workgroup per work block, -1 = keep)
int rkdg_batch_set_options(rkdg_ctx* c, int use_graph, int grid_cap) {
  if (!c) return fail(RKDG_E_ARG, "null context");
  c->use_graph = use_graph ? 1 : 0;
  if (grid_cap >= 0 && grid_cap != c->grid_cap) {
    c->grid_cap = grid_cap;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gkey.clear();
  }
  return RKDG_OK;
}

// Autotuner: times the right-hand side of state x (read only) for each
// candidate persistent grid size (0 and the given list) with HIP events and
// keeps the fastest. Results: chosen cap and the per-candidate milliseconds.
int rkdg_tune_grid(rkdg_ctx* c, const double* x, const int* candidates, int ncand, int reps, int* best,
                   double* ms_out) {
  if (!c || !x || (ncand > 0 && !candidates)) return fail(RKDG_E_ARG, "null argument");
  int rc = ensure_stage_buffers(c, 4);
  if (rc) return rc;
  reps = std::max(reps, 1);
  double best_ms = 1e300;
  int best_cap = 0;
  {
    // events freed and the grid cap put back on every exit path
    struct Events {
      hipEvent_t e[2] = {nullptr, nullptr};
      ~Events() {
        for (auto ev : e)
          if (ev) (void)hipEventDestroy(ev);
      }
    } ev;
    Restore<int> keep_cap(c->grid_cap);
    HIP_TRY(hipEventCreate(&ev.e[0]));
    HIP_TRY(hipEventCreate(&ev.e[1]));
    for (int k = -1; k < ncand; k++) {
      const int cap = k < 0 ? 0 : candidates[k];
      c->grid_cap = cap;
      RhsArgs ra = base_args(c, x);
      ra.k = c->d_stage[3];
      HIP_TRY(rhs_launch(c, ra));  // warm-up
      HIP_TRY(hipEventRecord(ev.e[0], c->stream));
      for (int r = 0; r < reps; r++) HIP_TRY(rhs_launch(c, ra));
      HIP_TRY(hipEventRecord(ev.e[1], c->stream));
      HIP_TRY(hipEventSynchronize(ev.e[1]));
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
      const double per = (double)ms / reps;
      if (ms_out) ms_out[k + 1] = per;
      if (per < best_ms) {
        best_ms = per;
        best_cap = cap;
      }
    }
  }
  rc = check_device_error(c, nullptr);
  if (rc) return rc;
  if (best) *best = best_cap;
  return rkdg_batch_set_options(c, c->use_graph, best_cap);
}

// ---------------------------------------------------------------- host-pointer variants
// rkdg_rhs / rkdg_limit / rkdg_step on HOST arrays: the state is copied to
// device buffers of the call, the operator runs on the context's stream and
// the result is copied back before returning (for tests and small callers;
// the device-pointer entry points are the fast path).
namespace {
struct DevBuf {
  double* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};
}  // namespace

int rkdg_rhs_host(rkdg_ctx* c, const double* x, double* y, double* lambda_max) {
  if (!c || !x || !y) return fail(RKDG_E_ARG, "null argument");
  const size_t n = (size_t)c->neq * c->ndofs;
  DevBuf dx, dy;
  HIP_TRY(hipMalloc(&dx.p, sizeof(double) * n));
  HIP_TRY(hipMalloc(&dy.p, sizeof(double) * n));
  HIP_TRY(hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  double lam = 0.0;
  int rc = rkdg_rhs(c, dx.p, dy.p, &lam);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  rc = check_device_error(c, nullptr);
  if (rc) return rc;
  if (lambda_max) *lambda_max = lam;
  return RKDG_OK;
}

int rkdg_limit_host(rkdg_ctx* c, double* x, int8_t* troubled) {
  if (!c || !x) return fail(RKDG_E_ARG, "null argument");
  const size_t n = (size_t)c->neq * c->ndofs;
  DevBuf dx;
  int8_t* dt = nullptr;
  HIP_TRY(hipMalloc(&dx.p, sizeof(double) * n));
  if (troubled) HIP_TRY(hipMalloc(&dt, (size_t)c->NE));
  struct FreeFlags {
    int8_t* q;
    ~FreeFlags() {
      if (q) (void)hipFree(q);
    }
  } keep{dt};
  HIP_TRY(hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (dt) HIP_TRY(hipMemsetAsync(dt, 0, (size_t)c->NE, c->stream));
  int rc = rkdg_limit(c, dx.p, dt);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  if (dt) HIP_TRY(hipMemcpyAsync(troubled, dt, (size_t)c->NE, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return check_device_error(c, nullptr);
}

int rkdg_step_host(rkdg_ctx* c, double* x, double t, double dt, int s, const double* a, const double* b,
                   const double* cc, double* lambda_max) {
  if (!c || !x) return fail(RKDG_E_ARG, "null argument");
  const size_t n = (size_t)c->neq * c->ndofs;
  DevBuf dx;
  HIP_TRY(hipMalloc(&dx.p, sizeof(double) * n));
  HIP_TRY(hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  double lam = 0.0;
  int rc = rkdg_step(c, dx.p, t, dt, s, a, b, cc, &lam);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  rc = check_device_error(c, nullptr);
  if (rc) return rc;
  if (lambda_max) *lambda_max = lam;
  return RKDG_OK;
}

// ---------------------------------------------------------------- physics plug-ins on the host
// The device code's Riemann solvers and boundary states, compiled for the
// host: RiemannSolver::Eval (rs_basic.hpp:49-50) and
// BoundaryIntegrator::computeRightState (boundary_integrator.hpp:63).
int rkdg_riemann_eval(const rkdg_scheme_desc* s, int dim, const double* state1, const double* state2,
                      const double* nor, double* flux, double* max_speed) {
  if (!s || !state1 || !state2 || !nor || !flux) return fail(RKDG_E_ARG, "null argument");
  if (dim != 2 && dim != 3) return fail(RKDG_E_ARG, "dim must be 2 or 3");
  SchemeParams g;
  std::memset(&g, 0, sizeof(g));
  int rc = scheme_params(s, g);
  if (rc) return rc;
  double n[3] = {nor[0], nor[1], dim == 3 ? nor[2] : 0.0};
  const double len = dim == 3 ? norml2_scaled<3>(n) : norml2_scaled<2>(n);
  if (!(len > 0.0)) return fail(RKDG_E_ARG, "zero normal");
  for (int i = 0; i < dim; i++) n[i] /= len;
  bool ok = true;
  const double lam = dim == 3 ? riemann<3>(g, state1, state2, n, flux, ok) : riemann<2>(g, state1, state2, n, flux, ok);
  if (max_speed) *max_speed = ok ? lam : -1.0;
  return ok && lam >= 0.0 ? RKDG_OK : fail(RKDG_E_UNPHYSICAL, "unphysical state in the Riemann solver");
}

int rkdg_bc_right_state(const rkdg_scheme_desc* s, int dim, int slot, const double* state1, const double* nor,
                        double* state2) {
  if (!s || !state1 || !nor || !state2) return fail(RKDG_E_ARG, "null argument");
  if (dim != 2 && dim != 3) return fail(RKDG_E_ARG, "dim must be 2 or 3");
  SchemeParams g;
  std::memset(&g, 0, sizeof(g));
  int rc = scheme_params(s, g);
  if (rc) return rc;
  if (slot < 0 || slot >= g.num_bc) return fail(RKDG_E_ARG, "boundary slot out of range");
  double n[3] = {nor[0], nor[1], dim == 3 ? nor[2] : 0.0};
  const double len = dim == 3 ? norml2_scaled<3>(n) : norml2_scaled<2>(n);
  if (!(len > 0.0)) return fail(RKDG_E_ARG, "zero normal");
  for (int i = 0; i < dim; i++) n[i] /= len;
  if (dim == 3) bc_state<3>(g, slot, state1, n, state2);
  else bc_state<2>(g, slot, state1, n, state2);
  return RKDG_OK;
}

}  // extern "C"
