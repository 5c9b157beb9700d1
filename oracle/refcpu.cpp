// refcpu.cpp — CPU restatement of the reference explicit RKDG path.
// TEST INFRASTRUCTURE ONLY (see refcpu.h): the parity oracle and the CPU baseline.
// Every function cites the reference file:line it follows. MFEM (external to the
// reference, absent here) behaviour is restated from its published algorithms:
// nodal Lagrange bases on Gauss-Legendre points, IntRules point counts
// (n = (order|1)/2+1 per axis), Loc1/Loc2 face maps, CalcOrtho, MassIntegrator,
// GetElementAverages, GetElementSize(.,1).

#include "refcpu.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

// The arithmetic type of the restatement. The default build is double, as the
// reference computes; REF_REAL=long double builds a wider copy of the same
// operator (librefcpu_ld.so), used to measure how far each double evaluation
// is from the operator evaluated with more precision. The C interface stays
// double in both.
#ifndef REF_REAL
#define REF_REAL double
#endif
using real = REF_REAL;

namespace {

// ----------------------------------------------------------------------------
// 1D Gauss-Legendre rule on [0,1] (MFEM QuadratureFunctions1D::GaussLegendre):
// Newton iteration on P_n, points ascending, weights sum to 1.
void gauss_legendre(int n, real* x, real* w) {
  for (int i = 0; i < (n + 1) / 2; i++) {
    real z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    real p1 = 0, p0 = 0, dp = 0;
    for (int it = 0; it < 100; it++) {
      p0 = 1.0;
      p1 = z;
      if (n == 1) p1 = z;
      for (int k = 2; k <= n; k++) {
        real p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = z; p0 = 1.0; }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      real dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    // re-evaluate derivative at the converged root for the weight
    p0 = 1.0;
    p1 = z;
    for (int k = 2; k <= n; k++) {
      real p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) { p1 = z; p0 = 1.0; }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    real wi = 1.0 / ((1.0 - z * z) * dp * dp);  // half of the [-1,1] weight
    x[i] = 0.5 * (1.0 - z);
    x[n - 1 - i] = 0.5 * (1.0 + z);
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
}

// Lagrange basis on given nodes (product form: exact Kronecker delta at nodes).
struct Lagrange1D {
  int n = 0;
  std::vector<real> x;
  void eval(real y, real* u) const {
    for (int i = 0; i < n; i++) {
      real v = 1.0;
      for (int j = 0; j < n; j++)
        if (j != i) v *= (y - x[j]) / (x[i] - x[j]);
      u[i] = v;
    }
  }
  void deriv(real y, real* d) const {
    for (int i = 0; i < n; i++) {
      real s = 0.0;
      for (int m = 0; m < n; m++) {
        if (m == i) continue;
        real t = 1.0 / (x[i] - x[m]);
        for (int j = 0; j < n; j++)
          if (j != i && j != m) t *= (y - x[j]) / (x[i] - x[j]);
        s += t;
      }
      d[i] = s;
    }
  }
};

bool invert_dense(int n, const real* A, real* Ainv) {
  // Gauss-Jordan with partial pivoting (stands in for MFEM DenseMatrixInverse LU)
  std::vector<real> M(A, A + n * n), I(n * n, 0.0);
  for (int i = 0; i < n; i++) I[i * n + i] = 1.0;
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (std::fabs(M[r * n + c]) > std::fabs(M[piv * n + c])) piv = r;
    if (M[piv * n + c] == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < n; k++) {
        std::swap(M[c * n + k], M[piv * n + k]);
        std::swap(I[c * n + k], I[piv * n + k]);
      }
    real inv = 1.0 / M[c * n + c];
    for (int k = 0; k < n; k++) {
      M[c * n + k] *= inv;
      I[c * n + k] *= inv;
    }
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      real f = M[r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; k++) {
        M[r * n + k] -= f * M[c * n + k];
        I[r * n + k] -= f * I[c * n + k];
      }
    }
  }
  std::memcpy(Ainv, I.data(), sizeof(real) * n * n);
  return true;
}

// ----------------------------------------------------------------------------
// Reference-element data (MFEM Geometry tables).
const real kQuadVerts[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
const real kTriVerts[3][2] = {{0, 0}, {1, 0}, {0, 1}};
const real kHexVerts[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
const int kQuadEdges[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
const int kTriEdges[3][2] = {{0, 1}, {1, 2}, {2, 0}};
const int kHexFaces[6][4] = {{3, 2, 1, 0}, {0, 1, 5, 4}, {1, 2, 6, 5},
                             {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};

struct RefFE {
  int geom = 0, dim = 2, p = 1, nd = 4, nv = 4, nf = 4, nvf = 2;
  Lagrange1D b1;
  // triangle P_p: monomial exponents and coefficient matrix (coef[m*nd+k])
  std::vector<std::array<int, 2>> mono;
  std::vector<real> coef;
  std::vector<real> nodes;  // nd*dim

  void init(int g, int order) {
    geom = g;
    p = order;
    if (g == REF_GEOM_QUAD) { dim = 2; nv = 4; nf = 4; nvf = 2; nd = (p + 1) * (p + 1); }
    else if (g == REF_GEOM_HEX) { dim = 3; nv = 8; nf = 6; nvf = 4; nd = (p + 1) * (p + 1) * (p + 1); }
    else { dim = 2; nv = 3; nf = 3; nvf = 2; nd = (p + 1) * (p + 2) / 2; }
    std::vector<real> gx(p + 1), gw(p + 1);
    gauss_legendre(p + 1, gx.data(), gw.data());
    b1.n = p + 1;
    b1.x = gx;
    nodes.assign(nd * dim, 0.0);
    if (g == REF_GEOM_QUAD) {
      for (int j = 0; j <= p; j++)
        for (int i = 0; i <= p; i++) {
          int o = i + (p + 1) * j;
          nodes[o * 2 + 0] = gx[i];
          nodes[o * 2 + 1] = gx[j];
        }
    } else if (g == REF_GEOM_HEX) {
      for (int k = 0; k <= p; k++)
        for (int j = 0; j <= p; j++)
          for (int i = 0; i <= p; i++) {
            int o = i + (p + 1) * (j + (p + 1) * k);
            nodes[o * 3 + 0] = gx[i];
            nodes[o * 3 + 1] = gx[j];
            nodes[o * 3 + 2] = gx[k];
          }
    } else {
      // L2_TriangleElement nodes from GL open points: (op[i]/w, op[j]/w),
      // w = op[i]+op[j]+op[p-i-j]; j outer, i inner.
      int o = 0;
      for (int j = 0; j <= p; j++)
        for (int i = 0; i + j <= p; i++) {
          real w = gx[i] + gx[j] + gx[p - i - j];
          nodes[o * 2 + 0] = gx[i] / w;
          nodes[o * 2 + 1] = gx[j] / w;
          o++;
        }
      for (int a = 0; a <= p; a++)
        for (int b = 0; a + b <= p; b++) mono.push_back({a, b});
      std::vector<real> V(nd * nd);
      for (int k = 0; k < nd; k++)
        for (int m = 0; m < nd; m++)
          V[k * nd + m] = std::pow(nodes[k * 2], mono[m][0]) * std::pow(nodes[k * 2 + 1], mono[m][1]);
      coef.assign(nd * nd, 0.0);
      invert_dense(nd, V.data(), coef.data());  // V c = I  -> c = V^-1, c[m*nd+k]
    }
  }

  void shape(const real* xi, real* phi) const {
    if (geom == REF_GEOM_TRI) {
      real mv[64];
      for (int m = 0; m < nd; m++) mv[m] = std::pow(xi[0], mono[m][0]) * std::pow(xi[1], mono[m][1]);
      for (int k = 0; k < nd; k++) {
        real s = 0;
        for (int m = 0; m < nd; m++) s += mv[m] * coef[m * nd + k];
        phi[k] = s;
      }
      return;
    }
    real ux[16], uy[16], uz[16];
    b1.eval(xi[0], ux);
    b1.eval(xi[1], uy);
    int n = p + 1;
    if (geom == REF_GEOM_QUAD) {
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) phi[i + n * j] = ux[i] * uy[j];
    } else {
      b1.eval(xi[2], uz);
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) phi[i + n * (j + n * k)] = ux[i] * uy[j] * uz[k];
    }
  }
  // dphi[k*dim + r] = d phi_k / d xi_r
  void dshape(const real* xi, real* dphi) const {
    if (geom == REF_GEOM_TRI) {
      for (int k = 0; k < nd; k++) {
        real dx = 0, dy = 0;
        for (int m = 0; m < nd; m++) {
          int a = mono[m][0], b = mono[m][1];
          real c = coef[m * nd + k];
          if (a > 0) dx += c * a * std::pow(xi[0], a - 1) * std::pow(xi[1], b);
          if (b > 0) dy += c * b * std::pow(xi[0], a) * std::pow(xi[1], b - 1);
        }
        dphi[k * 2 + 0] = dx;
        dphi[k * 2 + 1] = dy;
      }
      return;
    }
    int n = p + 1;
    real ux[16], uy[16], uz[16], dx[16], dy[16], dz[16];
    b1.eval(xi[0], ux);
    b1.eval(xi[1], uy);
    b1.deriv(xi[0], dx);
    b1.deriv(xi[1], dy);
    if (geom == REF_GEOM_QUAD) {
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          int o = i + n * j;
          dphi[o * 2 + 0] = dx[i] * uy[j];
          dphi[o * 2 + 1] = ux[i] * dy[j];
        }
    } else {
      b1.eval(xi[2], uz);
      b1.deriv(xi[2], dz);
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            int o = i + n * (j + n * k);
            dphi[o * 3 + 0] = dx[i] * uy[j] * uz[k];
            dphi[o * 3 + 1] = ux[i] * dy[j] * uz[k];
            dphi[o * 3 + 2] = ux[i] * uy[j] * dz[k];
          }
    }
  }
  const real* vert(int v) const {
    if (geom == REF_GEOM_QUAD) return kQuadVerts[v];
    if (geom == REF_GEOM_TRI) return kTriVerts[v];
    return kHexVerts[v];
  }
  int face_vert(int f, int m) const {
    if (geom == REF_GEOM_QUAD) return kQuadEdges[f][m];
    if (geom == REF_GEOM_TRI) return kTriEdges[f][m];
    return kHexFaces[f][m];
  }
  // MFEM ElementTransformation::OrderW for linear-geometry elements
  int order_w() const {
    if (geom == REF_GEOM_TRI) return 0;   // (1-1)*2
    if (geom == REF_GEOM_QUAD) return 1;  // 1*2-1
    return 2;                             // 1*3-1
  }
};

// Integration rules (MFEM IntRules): tensor GL with (order|1)/2+1 points per axis,
// x fastest; triangles: collapsed GL rule exact for the requested degree (any
// exact rule gives MFEM's values up to rounding since integrands are polynomial).
struct Rule {
  int n = 0, dim = 1;
  std::vector<real> pts, w;
};
Rule seg_rule(int order) {
  Rule r;
  int n = (order | 1) / 2 + 1;
  r.n = n;
  r.dim = 1;
  r.pts.resize(n);
  r.w.resize(n);
  gauss_legendre(n, r.pts.data(), r.w.data());
  return r;
}
Rule geom_rule(int geom, int order) {
  Rule s = seg_rule(order);
  Rule r;
  if (geom == REF_GEOM_QUAD) {
    r.dim = 2;
    for (int j = 0; j < s.n; j++)
      for (int i = 0; i < s.n; i++) {
        r.pts.push_back(s.pts[i]);
        r.pts.push_back(s.pts[j]);
        r.w.push_back(s.w[i] * s.w[j]);
      }
  } else if (geom == REF_GEOM_HEX) {
    r.dim = 3;
    for (int k = 0; k < s.n; k++)
      for (int j = 0; j < s.n; j++)
        for (int i = 0; i < s.n; i++) {
          r.pts.push_back(s.pts[i]);
          r.pts.push_back(s.pts[j]);
          r.pts.push_back(s.pts[k]);
          r.w.push_back(s.w[i] * s.w[j] * s.w[k]);
        }
  } else {
    r.dim = 2;
    int n = (order + 3) / 2;
    std::vector<real> x(n), w(n);
    gauss_legendre(n, x.data(), w.data());
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        r.pts.push_back(x[i]);
        r.pts.push_back((1.0 - x[i]) * x[j]);
        r.w.push_back(w[i] * w[j] * (1.0 - x[i]));
      }
  }
  r.n = (int)r.w.size();
  return r;
}
// face rule: segment (2D) or square (3D hex faces)
Rule face_rule(int geom, int order) {
  if (geom == REF_GEOM_HEX) return geom_rule(REF_GEOM_QUAD, order);
  return seg_rule(order);
}

// ----------------------------------------------------------------------------
// Element transformation for linear-geometry elements (bilinear quad, trilinear
// hex, affine triangle). J[i*dim + r] = dx_i / dxi_r.
struct Trans {
  const RefFE* fe;
  const real* X;  // nv*dim vertex coordinates
  void jacobian(const real* xi, real* J) const {
    int d = fe->dim;
    if (fe->geom == REF_GEOM_QUAD) {
      // the bilinear map's derivatives written on edge vectors,
      // dx/dxi = (1-eta)(X1-X0) + eta(X2-X3), dx/deta = (1-xi)(X3-X0) + xi(X2-X1):
      // the same map as MFEM's PointMat * dshape, without the rounding that
      // sum leaves when |X| is much larger than the element (|X|/h ulps in J)
      for (int i = 0; i < 2; i++) {
        J[i * 2 + 0] = (1 - xi[1]) * (X[2 + i] - X[i]) + xi[1] * (X[4 + i] - X[6 + i]);
        J[i * 2 + 1] = (1 - xi[0]) * (X[6 + i] - X[i]) + xi[0] * (X[4 + i] - X[2 + i]);
      }
      return;
    }
    if (fe->geom == REF_GEOM_HEX) {
      // the trilinear map's derivatives on edge vectors: dx/dxi_r is the
      // bilinear blend (in the other two coordinates) of the four edges along
      // axis r; same map as PointMat * dshape, without its rounding on
      // elements small against their coordinates
      for (int r = 0; r < 3; r++) {
        const int s1 = r == 0 ? 1 : 0, s2 = r == 2 ? 1 : 2;
        for (int i = 0; i < 3; i++) J[i * 3 + r] = 0.0;
        for (int b2 = 0; b2 < 2; b2++)
          for (int b1 = 0; b1 < 2; b1++) {
            const real w = (b1 ? xi[s1] : 1 - xi[s1]) * (b2 ? xi[s2] : 1 - xi[s2]);
            int i0[3] = {0, 0, 0}, i1[3];
            i0[s1] = b1;
            i0[s2] = b2;
            for (int q = 0; q < 3; q++) i1[q] = i0[q];
            i1[r] = 1;
            const int v0 = (i0[2] ? 4 : 0) + (i0[1] ? (i0[0] ? 2 : 3) : (i0[0] ? 1 : 0));
            const int v1 = (i1[2] ? 4 : 0) + (i1[1] ? (i1[0] ? 2 : 3) : (i1[0] ? 1 : 0));
            for (int i = 0; i < 3; i++) J[i * 3 + r] += w * (X[v1 * 3 + i] - X[v0 * 3 + i]);
          }
      }
      return;
    }
    std::fill(J, J + d * d, 0.0);
    for (int v = 0; v < fe->nv; v++) {
      real dN[3];
      if (fe->geom == REF_GEOM_TRI) {
        if (v == 0) { dN[0] = -1; dN[1] = -1; }
        else if (v == 1) { dN[0] = 1; dN[1] = 0; }
        else { dN[0] = 0; dN[1] = 1; }
      } else {
        const real* R = fe->vert(v);
        for (int r = 0; r < d; r++) {
          real t = R[r] > 0.5 ? 1.0 : -1.0;
          for (int s = 0; s < d; s++)
            if (s != r) t *= R[s] > 0.5 ? xi[s] : 1.0 - xi[s];
          dN[r] = t;
        }
      }
      for (int i = 0; i < d; i++)
        for (int r = 0; r < d; r++) J[i * d + r] += X[v * d + i] * dN[r];
    }
  }
  void transform(const real* xi, real* x) const {
    int d = fe->dim;
    std::fill(x, x + d, 0.0);
    for (int v = 0; v < fe->nv; v++) {
      real N;
      if (fe->geom == REF_GEOM_TRI) N = v == 0 ? 1 - xi[0] - xi[1] : xi[v - 1];
      else {
        const real* R = fe->vert(v);
        N = 1.0;
        for (int s = 0; s < d; s++) N *= R[s] > 0.5 ? xi[s] : 1.0 - xi[s];
      }
      for (int i = 0; i < d; i++) x[i] += N * X[v * d + i];
    }
  }
};
real det(int d, const real* J) {
  if (d == 2) return J[0] * J[3] - J[1] * J[2];
  return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
         J[2] * (J[3] * J[7] - J[4] * J[6]);
}
void adjugate(int d, const real* J, real* A);
// x = J^-1 b for d = 2, 3 (adjugate over determinant)
void solve_small(int d, const real* J, const real* b, real* x) {
  real A[9];
  adjugate(d, J, A);
  const real dj = det(d, J);
  for (int i = 0; i < d; i++) {
    real s = 0;
    for (int k = 0; k < d; k++) s += A[i * d + k] * b[k];
    x[i] = s / dj;
  }
}
// adj(J): J * adj(J) = det(J) I   (MFEM CalcAdjugate)
void adjugate(int d, const real* J, real* A) {
  if (d == 2) {
    A[0] = J[3];
    A[1] = -J[1];
    A[2] = -J[2];
    A[3] = J[0];
    return;
  }
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

// ----------------------------------------------------------------------------
// Physics (src/physics/physics.cpp). State layout: rho, rho*u[dim], rho*E.
struct Gas {
  real gamma, R, b;
};
// physics.cpp:101-112
real pressure(const real* s, int dim, const Gas& g) {
  real v2 = 0;
  for (int d = 0; d < dim; d++) v2 += s[1 + d] * s[1 + d];
  v2 /= s[0];
  return (g.gamma - 1.0) * (s[1 + dim] - 0.5 * v2) / (1.0 - s[0] * g.b);
}
// physics.cpp:134-140
real sound_speed(const real* s, int dim, const Gas& g) {
  real p = pressure(s, dim, g);
  return std::sqrt(g.gamma * p / s[0] / (1.0 - s[0] * g.b));
}
// physics.cpp:6-50
bool state_is_physical(const real* s, int dim, const Gas& g) {
  real den = s[0], E = s[1 + dim];
  if (den < 0 || den != den) return false;
  if (E <= 0 || E != E) return false;
  real p = pressure(s, dim, g);
  if (p <= 0 || p != p) return false;
  return true;
}
// physics.cpp:53-97 (pressure check ignores the covolume, :84)
bool state_is_physical_say(const real* s, int dim, const Gas& g) {
  real den = s[0], E = s[1 + dim];
  if (den < 0 || den != den) return false;
  if (E <= 0 || E != E) return false;
  real v2 = 0;
  for (int i = 0; i < dim; i++) v2 += s[1 + i] * s[1 + i];
  v2 /= den;
  real p = (g.gamma - 1.0) * (E - 0.5 * v2);
  if (p <= 0 || p != p) return false;
  return true;
}
// physics.cpp:143-168; f[k*dim + d]
void flux_tensor(const real* s, int dim, const Gas& g, real* f) {
  real den = s[0];
  real p = pressure(s, dim, g);
  for (int d = 0; d < dim; d++) {
    f[0 * dim + d] = s[1 + d];
    for (int i = 0; i < dim; i++) f[(1 + i) * dim + d] = s[1 + i] * s[1 + d] / den;
    f[(1 + d) * dim + d] += p;
  }
  real H = (s[1 + dim] + p) / den;
  for (int d = 0; d < dim; d++) f[(1 + dim) * dim + d] = s[1 + d] * H;
}
// physics.cpp:171-195
void flux_dot_n(const real* s, const real* nor, int dim, const Gas& g, real* fn) {
  real den = s[0];
  real p = pressure(s, dim, g);
  real vn = 0;
  for (int d = 0; d < dim; d++) vn += s[1 + d] * nor[d];
  fn[0] = vn;
  for (int d = 0; d < dim; d++) fn[1 + d] = vn * s[1 + d] / den + p * nor[d];
  real H = (s[1 + dim] + p) / den;
  fn[1 + dim] = vn * H;
}
// physics.cpp:198-231 (x-direction flux of a rotated state)
void flux_f(const real* s, int dim, const Gas& g, real* f) {
  real p = pressure(s, dim, g);
  f[0] = s[1];
  f[1] = s[1] * s[1] / s[0] + p;
  for (int d = 1; d < dim; d++) f[d + 1] = s[1] * s[d + 1] / s[0];
  real H = (s[1 + dim] + p) / s[0];
  f[1 + dim] = s[1] * H;
}
// physics.cpp:252-266
real max_char_speed(const real* s, int dim, const Gas& g) {
  real v2 = 0;
  for (int d = 0; d < dim; d++) v2 += s[1 + d] * s[1 + d];
  v2 /= s[0];
  real c = sound_speed(s, dim, g);
  real vel = std::sqrt(v2 / s[0]);
  return vel + c;
}
// physics.cpp:235-249
real mach(const real* s, int dim, const Gas& g) {
  real v2 = 0;
  for (int d = 0; d < dim; d++) v2 += s[1 + d] * s[1 + d];
  v2 /= s[0];
  real c = sound_speed(s, dim, g);
  return std::sqrt(v2 / s[0]) / c;
}
// physics.cpp:309-350 (Toro two-rarefaction estimate)
void toro_speeds(const real* s1, const real* s2, int dim, const Gas& g, real* lam) {
  real cL = sound_speed(s1, dim, g), cR = sound_speed(s2, dim, g);
  real uL = s1[1] / s1[0], uR = s2[1] / s2[0];
  real pL = pressure(s1, dim, g), pR = pressure(s2, dim, g);
  real gm = g.gamma;
  real z = 0.5 * (gm - 1.0) / gm;
  real pStar = std::pow((cL + cR - 0.5 * (gm - 1.0) * (uR - uL)) /
                              (cL / std::pow(pL, z) + cR / std::pow(pR, z)),
                          1.0 / z);
  real qL = pStar > pL ? std::sqrt(1.0 + 0.5 * (gm + 1.0) * (pStar / pL - 1.0) / gm) : 1.0;
  real qR = pStar > pR ? std::sqrt(1.0 + 0.5 * (gm + 1.0) * (pStar / pR - 1.0) / gm) : 1.0;
  lam[0] = uL - qL * cL;
  lam[dim + 1] = uR + qR * cR;
  for (int i = 1; i < dim + 1; i++) lam[i] = uL;
}
// physics.cpp:269-296 (Einfeldt averaged speeds; unused by the reference solvers)
void einfeldt_speeds(const real* s1, const real* s2, int dim, const Gas& g, real* lam) {
  real cL = sound_speed(s1, dim, g), cR = sound_speed(s2, dim, g);
  real ruL = s1[1], ruR = s2[1];
  real uL = ruL / s1[0], uR = ruR / s2[0];
  real sL = std::sqrt(s1[0]), sR = std::sqrt(s2[0]);
  real sum = sL + sR;
  real eta2 = 0.5 * sL * sR / ((sL + sR) * (sL + sR));
  real uav = (ruL / sL + ruR / sR) / sum;
  real cav = std::sqrt((sL * (cL * cL) + sR * (cR * cR)) / sum + eta2 * ((uR - uL) * (uR - uL)));
  lam[0] = uav - cav;
  lam[dim + 1] = uav + cav;
  for (int i = 1; i < dim + 1; i++) lam[i] = uav;
}

// rs_basic.cpp:11-38
void rotate(real* s, const real* nor, int dim) {
  real* m = s + 1;
  if (dim == 2) {
    real mx = nor[0] * m[0] + nor[1] * m[1];
    real my = -nor[1] * m[0] + nor[0] * m[1];
    m[0] = mx;
    m[1] = my;
  } else {
    real sp = nor[2];
    real cp = std::sqrt(1 - sp * sp);
    real st = std::fabs(cp) > 1e-6 ? nor[1] / cp : 1.0;
    real ct = std::fabs(cp) > 1e-6 ? nor[0] / cp : 0.0;
    real mx = cp * ct * m[0] + cp * st * m[1] + sp * m[2];
    real my = -st * m[0] + ct * m[1];
    real mz = -sp * ct * m[0] - sp * st * m[1] + cp * m[2];
    m[0] = mx;
    m[1] = my;
    m[2] = mz;
  }
}
// rs_basic.cpp:40-67
void inverse_rotate(real* s, const real* nor, int dim) {
  real* m = s + 1;
  if (dim == 2) {
    real mx = nor[0] * m[0] - nor[1] * m[1];
    real my = nor[1] * m[0] + nor[0] * m[1];
    m[0] = mx;
    m[1] = my;
  } else {
    real sp = nor[2];
    real cp = std::sqrt(1 - sp * sp);
    real st = std::fabs(cp) > 1e-6 ? nor[1] / cp : 1.0;
    real ct = std::fabs(cp) > 1e-6 ? nor[0] / cp : 0.0;
    real mx = cp * ct * m[0] - st * m[1] - sp * ct * m[2];
    real my = cp * st * m[0] + ct * m[1] - sp * st * m[2];
    real mz = sp * m[0] + cp * m[2];
    m[0] = mx;
    m[1] = my;
    m[2] = mz;
  }
}

struct Scheme {
  ref_scheme s;
  Gas gas() const { return Gas{s.gamma, s.gas_constant, s.covolume}; }
};

// Riemann solvers. The inputs are rotated in place as the reference does
// (rs_llf.cpp:18-19, rs_hll.cpp:14-15, rs_hllc.cpp:79-80). Returns -1 on an
// unphysical input state.
real riemann(const Scheme& sc, int dim, real* s1, real* s2, const real* nor, real* flux) {
  const Gas g = sc.gas();
  const int ne = dim + 2;
  if (!state_is_physical_say(s1, dim, g)) return -1;
  if (!state_is_physical_say(s2, dim, g)) return -1;
  real f1[5], f2[5], lam[5];
  switch (sc.s.riemann) {
    case REF_RS_RUSANOV: {  // rs_rusanov.cpp:6-40
      real m1 = max_char_speed(s1, dim, g), m2 = max_char_speed(s2, dim, g);
      real mE = std::max(m1, m2);
      flux_dot_n(s1, nor, dim, g, f1);
      flux_dot_n(s2, nor, dim, g, f2);
      real nm = 0;
      for (int i = 0; i < dim; i++) nm += nor[i] * nor[i];
      nm = std::sqrt(nm);
      for (int i = 0; i < ne; i++) flux[i] = 0.5 * (f1[i] + f2[i]) / nm - 0.5 * mE * (s2[i] - s1[i]);
      return mE;
    }
    case REF_RS_LLF: {  // rs_llf.cpp:6-37
      rotate(s1, nor, dim);
      rotate(s2, nor, dim);
      flux_f(s1, dim, g, f1);
      flux_f(s2, dim, g, f2);
      if (sc.s.wave_speeds == REF_WS_EINFELDT) einfeldt_speeds(s1, s2, dim, g, lam);
      else toro_speeds(s1, s2, dim, g, lam);
      real mE = std::max(std::fabs(lam[0]), std::fabs(lam[dim + 1]));
      for (int i = 0; i < ne; i++) flux[i] = 0.5 * (f1[i] + f2[i]) - 0.5 * mE * (s2[i] - s1[i]);
      inverse_rotate(flux, nor, dim);
      return mE;
    }
    case REF_RS_HLL: {  // rs_hll.cpp:6-50
      rotate(s1, nor, dim);
      rotate(s2, nor, dim);
      flux_f(s1, dim, g, f1);
      flux_f(s2, dim, g, f2);
      if (sc.s.wave_speeds == REF_WS_EINFELDT) einfeldt_speeds(s1, s2, dim, g, lam);
      else toro_speeds(s1, s2, dim, g, lam);
      real mE = std::max(std::fabs(lam[0]), std::fabs(lam[dim + 1]));
      real SL = lam[0], SR = lam[dim + 1];
      if (SL >= 0)
        for (int i = 0; i < ne; i++) flux[i] = f1[i];
      else if (SR <= 0)
        for (int i = 0; i < ne; i++) flux[i] = f2[i];
      else
        for (int i = 0; i < ne; i++)
          flux[i] = (SR * f1[i] - SL * f2[i] + SL * SR * (s2[i] - s1[i])) / (SR - SL);
      inverse_rotate(flux, nor, dim);
      return mE;
    }
    default: {  // HLLC rs_hllc.cpp:71-161, getFStar :39-67
      rotate(s1, nor, dim);
      rotate(s2, nor, dim);
      flux_f(s1, dim, g, f1);
      flux_f(s2, dim, g, f2);
      if (sc.s.wave_speeds == REF_WS_EINFELDT) einfeldt_speeds(s1, s2, dim, g, lam);
      else toro_speeds(s1, s2, dim, g, lam);
      real SL = lam[0], SR = lam[dim + 1];
      real mE = sc.s.hllc_abs ? std::max(std::fabs(SL), std::fabs(SR)) : std::max(SL, SR);
      if (SL >= 0)
        for (int i = 0; i < ne; i++) flux[i] = f1[i];
      else if (SR <= 0)
        for (int i = 0; i < ne; i++) flux[i] = f2[i];
      else {
        real pL = pressure(s1, dim, g), pR = pressure(s2, dim, g);
        real cL = SL - s1[1] / s1[0];
        real cR = SR - s2[1] / s2[0];
        real sStar = (pR - pL + s1[1] * cL - s2[1] * cR) / (s1[0] * cL - s2[0] * cR);
        const real* st = sStar > 0.0 ? s1 : s2;
        const real* fK = sStar > 0.0 ? f1 : f2;
        real pK = sStar > 0.0 ? pL : pR;
        real SK = sStar > 0.0 ? SL : SR;
        real cK = sStar > 0.0 ? cL : cR;
        real rhoK = st[0];
        real D[5] = {0, 0, 0, 0, 0};
        D[1] = 1.0;
        D[dim + 1] = sStar;
        for (int i = 0; i < ne; i++)
          flux[i] = (sStar * (st[i] * SK - fK[i]) + D[i] * SK * (pK + rhoK * cK * (sStar - st[1] / st[0]))) *
                    (1.0 / (SK - sStar));
      }
      inverse_rotate(flux, nor, dim);
      return mE;
    }
  }
}

// Ghost states: src/bc/boundary_integrator_*.cpp computeRightState.
void bc_state(const Scheme& sc, int dim, int bi, const real* s1, const real* nor, real* s2) {
  const Gas g = sc.gas();
  const int ne = dim + 2;
  real P[8];
  for (int k = 0; k < 8; k++) P[k] = sc.s.bc_param[bi][k];
  const real gm = g.gamma, R = g.R;
  for (int i = 0; i < ne; i++) s2[i] = s1[i];
  switch (sc.s.bc_type[bi]) {
    case REF_BC_SLIP:  // boundary_integrator_slip.cpp:7-17
      rotate(s2, nor, dim);
      s2[1] *= -1;
      inverse_rotate(s2, nor, dim);
      break;
    case REF_BC_OUTLET:  // boundary_integrator_open.cpp:9-12
      break;
    case REF_BC_SUPERSONIC_INLET: {  // boundary_integrator_supersonic_inlet.cpp:8-19;
      // state from case_manager.cpp:736-750 (energy omits w^2, :749)
      real rho = P[0], u = P[1], v = P[2], w = P[3], p = P[4];
      real st[5];
      st[0] = rho;
      st[1] = rho * u;
      st[2] = rho * v;
      if (dim == 3) st[3] = rho * w;
      st[ne - 1] = p / (gm - 1.) + 0.5 * rho * (u * u + v * v);
      for (int i = 0; i < ne; i++) s2[i] = st[i];
      break;
    }
    case REF_BC_SUBSONIC_INLET_TOTAL_PRESSURE:  // ..._subsonic_inlet_total_pressure.cpp:8-38
    case REF_BC_SUBSONIC_INLET_FIXED_PRESSURE: {  // ..._subsonic_inlet_fixed_pressure.cpp:8-38
      real pTot = P[0], TTot = P[1];
      real rhoU2 = s2[1] * s2[1] + s2[2] * s2[2];
      if (dim == 3) rhoU2 += s2[3] * s2[3];
      rhoU2 /= s2[0];
      real MTTotSq = rhoU2 / s2[0] / gm / R / TTot;
      real MSq = MTTotSq / (1.0 - 0.5 * (gm - 1.0) * MTTotSq);
      real p = sc.s.bc_type[bi] == REF_BC_SUBSONIC_INLET_TOTAL_PRESSURE
                     ? pTot * std::pow(1.0 + 0.5 * (gm - 1.0) * MSq, -gm / (gm - 1.0))
                     : pTot;
      real T = TTot / (1.0 + 0.5 * (gm - 1.0) * MSq);
      real rho = p / R / T;
      s2[ne - 1] = p / (gm - 1.0) + 0.5 * rhoU2 / s2[0] * rho;
      s2[1] = s2[1] / s2[0] * rho;
      s2[2] = s2[2] / s2[0] * rho;
      if (dim == 3) s2[3] = s2[3] / s2[0] * rho;
      s2[0] = rho;
      break;
    }
    case REF_BC_FIXED_PRESSURE_OUTLET: {  // boundary_integrator_open_fixed_pressure.cpp:9-32
      real pFix = P[0];
      real rhoU2 = s2[1] * s2[1] + s2[2] * s2[2];
      if (dim == 3) rhoU2 += s2[3] * s2[3];
      real U2 = rhoU2 / s2[0] / s2[0];
      real epsIn = (s2[ne - 1] - 0.5 * rhoU2) / s2[0];
      real rho = pFix / (gm - 1.0) / epsIn;
      real rhoE = pFix / (gm - 1.0) + 0.5 * rho * U2;
      s2[ne - 1] = rhoE;
      s2[1] = s1[1] / s1[0] * rho;
      s2[2] = s1[2] / s1[0] * rho;
      if (dim == 3) s2[3] = s2[3] / s1[0] * rho;
      s2[0] = rho;
      break;
    }
    case REF_BC_TOTAL_PRESSURE_OUTLET: {  // ..._open_total_pressure.cpp:9-42, normal index fixed (nor[0..dim-1])
      real pTotal = P[0], TFix = P[1];
      real rho2U2 = s2[1] * s2[1] + s2[2] * s2[2];
      if (dim == 3) rho2U2 += s2[3] * s2[3];
      real rhoN = s2[1] * nor[0] + s2[2] * nor[1];
      if (dim == 3) rhoN += s2[3] * nor[2];
      real U2 = rho2U2 / s2[0] / s2[0];
      real epsIn = rhoN > 0 ? (s2[ne - 1] - 0.5 * rho2U2) / s2[0] : R / (gm - 1.0) * TFix;
      real M = mach(s2, dim, g);
      real pr = M > 1 ? pTotal * std::pow(1.0 + 0.5 * (gm - 1.0) * M * M, -gm / (gm - 1.0)) : pTotal;
      real rho = M > 1 ? pr / (gm - 1.0) / epsIn : pr / (0.5 * U2 + (gm - 1.0) * epsIn);
      pr = M > 1 ? pr : pr - 0.5 * U2 * rho;
      real rhoE = pr / (gm - 1.0) + 0.5 * rho * U2;
      s2[ne - 1] = rhoE;
      s2[1] *= rho / s2[0];
      s2[2] *= rho / s2[0];
      if (dim == 3) s2[3] *= rho / s2[0];
      s2[0] = rho;
      break;
    }
    case REF_BC_CHAR_OUTLET: {  // boundary_integrator_char_outlet.cpp:14-51 with its ctor :6-12
      // the fixed state as case_manager.cpp:781-799 builds it (energy without w^2)
      real fs[5] = {0, 0, 0, 0, 0};
      fs[0] = P[0];
      fs[1] = P[0] * P[1];
      fs[2] = P[0] * P[2];
      if (dim == 3) fs[3] = P[0] * P[3];
      fs[ne - 1] = P[4] / (gm - 1.) + 0.5 * P[0] * (P[1] * P[1] + P[2] * P[2]);
      const real cOut = sound_speed(fs, dim, g);
      const real UOut = fs[1] / fs[0];
      const real cIn = sound_speed(s1, dim, g);
      // state2 starts as state1 (the reference rotates it uninitialised, Q10)
      rotate(s2, nor, dim);
      const real UIn = s2[1] / s2[0];
      const real UB = 0.5 * (UIn + UOut) - (cOut - cIn) / (gm - 1.0);
      const real cB = -0.25 * (gm - 1.0) * (UOut - UIn) + 0.5 * (cOut + cIn);
      real sB;
      if (s2[1] < 0)  // inflow
        sB = cOut * cOut / std::pow(fs[0], gm - 1.0) / gm;
      else  // outflow
        sB = cIn * cIn / std::pow(s2[0], gm - 1.0) / gm;
      const real rhoB = std::pow(cB * cB / sB / gm, 1.0 / (gm - 1.0));
      const real pB = rhoB * cB * cB / gm;
      s2[0] = rhoB;
      // the normal momentum rhoB UB (the reference stores the velocity UB in
      // the momentum slot); the tangential momenta stay, so the tangential
      // velocities in the energy are those momenta over rhoB, as there
      s2[1] = rhoB * UB;
      const real vt = s2[2] / s2[0], wt = dim == 3 ? s2[3] / s2[0] : 0.0;
      s2[ne - 1] = pB / (gm - 1.0) * (1.0 - rhoB * g.b) + 0.5 * rhoB * (UB * UB + vt * vt + wt * wt);
      inverse_rotate(s2, nor, dim);
      break;
    }
    default:
      break;
  }
}

// boundary_integrator_char_outlet.cpp:153-189: the characteristic outlet's flux
// is the physical flux of its ghost state (no Riemann solver); returns |u|+c of
// the ghost
real char_outlet_flux(const Scheme& sc, int dim, const real* s2, const real* nor, real* fl) {
  const Gas g = sc.gas();
  flux_dot_n(s2, nor, dim, g, fl);
  return max_char_speed(s2, dim, g);
}

// ----------------------------------------------------------------------------
struct Face {
  int64_t e1 = -1, e2 = -1;
  int lf1 = -1, lf2 = -1;
  int attr = 0;       // boundary attribute (bdr faces)
  int fv[4];          // topological vertices in Elem1's local face order
  int64_t fvt[4];
  // non-conforming faces (MFEM NCMesh): nc = 1 on a slave face, whose Elem1 is
  // the fine element and Elem2 the coarse (master) element; pm[m] holds the
  // master's reference coordinates of the slave face's corner m (the point
  // matrix of Loc2). ncm = 1 on a master face: not integrated (its slave faces
  // are), but an interior face of the master for the limiter.
  int nc = 0, ncm = 0;
  real pm[4][3];
};

struct CSR {
  std::vector<int64_t> I;
  std::vector<int64_t> J;
  std::vector<real> A;
};

struct Ref {
  RefFE fe;
  Scheme sc;
  int dim = 2, nv = 4, nd = 4;
  int64_t NE = 0, NDofs = 0;
  std::vector<real> X;       // NE*nv*dim
  std::vector<int64_t> topo;   // NE*nv
  std::vector<Face> faces;     // all faces in face-id order
  std::vector<int64_t> interior;  // interior face ids (ascending)
  std::vector<int64_t> bdr_faces; // face id per bdr element (in bdr order)
  std::vector<int> bdr_attr;
  std::vector<std::vector<int64_t>> elem_faces;  // local face -> face id
  std::vector<std::vector<int64_t>> el2el;       // face neighbours
  CSR Aflux;
  std::vector<real> Me_inv;  // NE*nd*nd (row-major)
  std::vector<real> avgs;    // nv_eq*NE  (byNODES: [v*NE + e])
  // per-element face-contribution order for deterministic parallel accumulation
  std::vector<std::vector<std::pair<int64_t, int>>> contrib;  // (slot, side)
  std::vector<real> face_buf;  // slot * 2 * nd * ne
  int64_t n_slots = 0;

  std::vector<char> affine;    // element map is affine (parallelogram/parallelepiped/triangle)

  int ne() const { return dim + 2; }
  Trans trans(int64_t e) const { return Trans{&fe, X.data() + e * nv * dim}; }
  // exact Jacobian of an affine element: J(:,r) = X(v_r) - X(v_0) with v_r the
  // vertex at the unit point of reference axis r
  void affine_jacobian(int64_t e, real* J) const {
    const real* Xe = X.data() + e * nv * dim;
    int cols[3] = {1, fe.geom == REF_GEOM_TRI ? 2 : 3, 4};
    for (int i = 0; i < dim; i++)
      for (int r = 0; r < dim; r++) J[i * dim + r] = Xe[cols[r] * dim + i] - Xe[i];
  }
  void detect_affine() {
    affine.assign(NE, 1);
    if (fe.geom == REF_GEOM_TRI) return;
    // quadrilaterals are mapped bilinearly everywhere, as MFEM maps them; the
    // edge-vector Jacobian above is exact on parallelograms already
    if (fe.geom == REF_GEOM_QUAD) {
      affine.assign(NE, 0);
      return;
    }
    for (int64_t e = 0; e < NE; e++) {
      real J[9];
      affine_jacobian(e, J);
      const real* Xe = X.data() + e * nv * dim;
      real scale = 0, big = 0;
      for (int i = 0; i < dim * dim; i++) scale = std::max(scale, std::fabs(J[i]));
      for (int k = 0; k < nv * dim; k++) big = std::max(big, std::fabs(Xe[k]));
      // affine to within rounding of the coordinates (1e-13 of the size or 16
      // ulps of the largest double coordinate); beyond that the element is
      // mapped exactly, as MFEM maps every element
      const real tol = std::max(real(1e-13) * scale, real(16.0 * 2.220446049250313e-16) * big);
      for (int v = 0; v < nv; v++) {
        const real* R = fe.vert(v);
        for (int i = 0; i < dim; i++) {
          real xv = Xe[i];
          for (int r = 0; r < dim; r++) xv += J[i * dim + r] * R[r];
          if (std::fabs(xv - Xe[v * dim + i]) > tol) affine[e] = 0;
        }
      }
    }
  }

  // reference coordinates of face point s (face param) in element e's local face
  void face_to_elem(const Face& f, int side, const real* s, real* xi) const {
    real N[4];
    if (dim == 2) { N[0] = 1 - s[0]; N[1] = s[0]; }
    else { N[0] = (1 - s[0]) * (1 - s[1]); N[1] = s[0] * (1 - s[1]); N[2] = s[0] * s[1]; N[3] = (1 - s[0]) * s[1]; }
    int nvf = fe.nvf;
    std::fill(xi, xi + dim, 0.0);
    int64_t e = side == 0 ? f.e1 : f.e2;
    if (side == 1 && f.nc) {
      for (int m = 0; m < nvf; m++)
        for (int d = 0; d < dim; d++) xi[d] += N[m] * f.pm[m][d];
      return;
    }
    for (int m = 0; m < nvf; m++) {
      int lv;
      if (side == 0) lv = fe.face_vert(f.lf1, m);
      else {
        lv = -1;
        for (int v = 0; v < nv; v++)
          if (topo[e * nv + v] == f.fvt[m]) {
            // must be one of e2's local face vertices
            for (int mm = 0; mm < nvf; mm++)
              if (fe.face_vert(f.lf2, mm) == v) lv = v;
          }
      }
      const real* R = fe.vert(lv);
      for (int d = 0; d < dim; d++) xi[d] += N[m] * R[d];
    }
  }
  // CalcOrtho of the face Jacobian at face point s (normal out of Elem1, |n| = measure density)
  // The reference faces are affine images of the unit square/segment, so the
  // face-map derivative D is exact (edge vectors of the reference face); for
  // affine elements the element Jacobian is the exact vertex-difference matrix.
  // Exactness matters: the reference 3D rotation (rs_basic.cpp:21-28) loses
  // orthogonality when n_z = 1 - ulp, so last-bit noise in an axis-aligned
  // unit normal would perturb rotated fluxes at the 1e-8 level.
  // Non-affine elements: the Jacobian of the face's own vertex map (MFEM's
  // FaceElementTransformations: the face's linear element over its vertices in
  // Elem1's local order, PointMat * dshape), so a planar axis-aligned face has
  // exact zero tangent components, as in the reference; the element Jacobian at
  // the embedded point would carry the rounding of the embedding (a coordinate
  // of 1 - ulp for 1) into them.
  void face_normal(const Face& f, const real* s, real* nor) const {
    if (!affine[f.e1]) {
      const real* Xe = X.data() + f.e1 * nv * dim;
      const real* P[4];
      for (int m = 0; m < fe.nvf; m++) P[m] = Xe + fe.face_vert(f.lf1, m) * dim;
      if (dim == 2) {
        // segment: dshape = (-1, 1)
        nor[0] = -P[0][1] + P[1][1];
        nor[1] = -(-P[0][0] + P[1][0]);
      } else {
        // bilinear square, vertices (0,0), (1,0), (1,1), (0,1)
        const real ds[4] = {-(1 - s[1]), 1 - s[1], s[1], -s[1]};
        const real dt[4] = {-(1 - s[0]), -s[0], s[0], 1 - s[0]};
        real a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
        for (int m = 0; m < 4; m++)
          for (int i = 0; i < 3; i++) {
            a[i] += P[m][i] * ds[m];
            b[i] += P[m][i] * dt[m];
          }
        nor[0] = a[1] * b[2] - a[2] * b[1];
        nor[1] = a[2] * b[0] - a[0] * b[2];
        nor[2] = a[0] * b[1] - a[1] * b[0];
      }
      return;
    }
    real xi[3];
    face_to_elem(f, 0, s, xi);
    real J[9];
    affine_jacobian(f.e1, J);
    const real* R[4];
    for (int m = 0; m < fe.nvf; m++) R[m] = fe.vert(fe.face_vert(f.lf1, m));
    if (dim == 2) {
      real D[2] = {R[1][0] - R[0][0], R[1][1] - R[0][1]};
      real Jf[2] = {J[0] * D[0] + J[1] * D[1], J[2] * D[0] + J[3] * D[1]};
      nor[0] = Jf[1];
      nor[1] = -Jf[0];
    } else {
      real Ds[3], Dt[3];
      for (int d = 0; d < 3; d++) {
        Ds[d] = R[1][d] - R[0][d];
        Dt[d] = R[3][d] - R[0][d];
      }
      real a[3], b[3];
      for (int i = 0; i < 3; i++) {
        a[i] = J[i * 3 + 0] * Ds[0] + J[i * 3 + 1] * Ds[1] + J[i * 3 + 2] * Ds[2];
        b[i] = J[i * 3 + 0] * Dt[0] + J[i * 3 + 1] * Dt[1] + J[i * 3 + 2] * Dt[2];
      }
      nor[0] = a[1] * b[2] - a[2] * b[1];
      nor[1] = a[2] * b[0] - a[0] * b[2];
      nor[2] = a[0] * b[1] - a[1] * b[0];
    }
  }
  int face_intorder() const {
    // face_integrator.cpp:45-58 (all elements share type and order)
    int o = fe.order_w() + 2 * fe.p;
    if (fe.geom == REF_GEOM_TRI) o++;
    return o;
  }

  bool build(std::string& err);
  bool attach_nc(int64_t n, const int64_t* slave, const int* slave_face, const int64_t* master,
                 const int* master_face, const double* param, std::string& err);
  void assemble_aflux();
  void assemble_mass();
  int face_mult(const real* x, real* z, real& mcs, int64_t* bad);
  int mult(const real* x, real* y, real& mcs, int64_t* bad);
  void compute_means(const real* x);
  void limiter_update(real* x, real* ind_values, real* min_values);
  void transform_back(int64_t k, const real* x, real* xi) const;
  void extrap_averages(int64_t t, const real* x, const std::vector<int64_t>& cells, real* out) const;
  real hmin() const;
};

bool Ref::build(std::string& err) {
  std::map<std::vector<int64_t>, int64_t> key2face;
  elem_faces.assign(NE, std::vector<int64_t>(fe.nf, -1));
  for (int64_t e = 0; e < NE; e++) {
    for (int lf = 0; lf < fe.nf; lf++) {
      std::vector<int64_t> key(fe.nvf);
      for (int m = 0; m < fe.nvf; m++) key[m] = topo[e * nv + fe.face_vert(lf, m)];
      std::sort(key.begin(), key.end());
      auto it = key2face.find(key);
      if (it == key2face.end()) {
        Face f;
        f.e1 = e;
        f.lf1 = lf;
        for (int m = 0; m < fe.nvf; m++) {
          f.fv[m] = fe.face_vert(lf, m);
          f.fvt[m] = topo[e * nv + f.fv[m]];
        }
        key2face[key] = (int64_t)faces.size();
        elem_faces[e][lf] = (int64_t)faces.size();
        faces.push_back(f);
      } else {
        Face& f = faces[it->second];
        if (f.e2 >= 0) { err = "face shared by more than two elements (periodic direction shorter than 3 elements?)"; return false; }
        if (f.e1 == e) { err = "element adjacent to itself (periodic direction too short)"; return false; }
        f.e2 = e;
        f.lf2 = lf;
        elem_faces[e][lf] = it->second;
      }
    }
  }
  for (int64_t f = 0; f < (int64_t)faces.size(); f++)
    if (faces[f].e2 >= 0) interior.push_back(f);
  el2el.assign(NE, {});
  for (int64_t f : interior) {
    el2el[faces[f].e1].push_back(faces[f].e2);
    el2el[faces[f].e2].push_back(faces[f].e1);
  }
  return true;
}

// Non-conforming faces (MFEM's NCMesh master/slave lists): face slave_face of
// the fine element `slave` covers part of face master_face of the coarser
// element `master`; param[(k * nvf + c) * (dim - 1) + j] is coordinate j, in
// the master face's own parametrisation (face vertex 0 at the origin, as
// face_to_elem parametrises faces), of the slave face's corner c (in the
// slave's local face order). The slave face becomes an interior face with
// Elem1 = the slave and Elem2 = the master, whose Loc2 point matrix holds the
// master's reference coordinates of those corners (MFEM's slave point
// matrix). The master face is integrated through its slaves only
// (face_integrator.cpp runs over interior faces) and enters the limiter's
// stencil as an interior face (limiter.cpp:166-217, the infoNC branch); the
// element-to-element table links master and slave.
bool Ref::attach_nc(int64_t n, const int64_t* slave, const int* slave_face, const int64_t* master,
                    const int* master_face, const double* param, std::string& err) {
  if (n <= 0) return true;
  const int fd = dim - 1;
  for (int64_t k = 0; k < n; k++) {
    const int64_t s = slave[k], m = master[k];
    const int ls = slave_face[k], lm = master_face[k];
    if (s < 0 || s >= NE || m < 0 || m >= NE || s == m || ls < 0 || ls >= fe.nf || lm < 0 || lm >= fe.nf) {
      err = "non-conforming face record " + std::to_string(k) + " out of range";
      return false;
    }
    Face& fs = faces[elem_faces[s][ls]];
    Face& fm = faces[elem_faces[m][lm]];
    if (fs.e2 >= 0 || fs.e1 != s || fs.nc) {
      err = "non-conforming face record " + std::to_string(k) + ": the slave face is not an open face";
      return false;
    }
    if (fm.e2 >= 0 || fm.e1 != m) {
      err = "non-conforming face record " + std::to_string(k) + ": the master face is not an open face";
      return false;
    }
    fs.e2 = m;
    fs.lf2 = lm;
    fs.nc = 1;
    fm.ncm = 1;
    for (int c = 0; c < fe.nvf; c++) {
      real sp[2] = {0, 0};
      for (int j = 0; j < fd; j++) {
        sp[j] = param[(k * fe.nvf + c) * fd + j];
        if (!(sp[j] >= 0.0 && sp[j] <= 1.0)) {
          err = "non-conforming face record " + std::to_string(k) + ": face parameter outside [0, 1]";
          return false;
        }
      }
      // the master face's bilinear (linear) map of the parameter
      real N[4];
      if (dim == 2) { N[0] = 1 - sp[0]; N[1] = sp[0]; }
      else { N[0] = (1 - sp[0]) * (1 - sp[1]); N[1] = sp[0] * (1 - sp[1]); N[2] = sp[0] * sp[1]; N[3] = (1 - sp[0]) * sp[1]; }
      for (int d = 0; d < dim; d++) fs.pm[c][d] = 0;
      for (int q = 0; q < fe.nvf; q++) {
        const real* R = fe.vert(fe.face_vert(lm, q));
        for (int d = 0; d < dim; d++) fs.pm[c][d] += N[q] * R[d];
      }
    }
  }
  interior.clear();
  for (int64_t f = 0; f < (int64_t)faces.size(); f++)
    if (faces[f].e2 >= 0) interior.push_back(f);
  el2el.assign(NE, {});
  for (int64_t f : interior) {
    el2el[faces[f].e1].push_back(faces[f].e2);
    el2el[faces[f].e2].push_back(faces[f].e1);
  }
  return true;
}

// domain_integrator.cpp:8-58, assembled into the CSR Aflux (main.cpp:148-151)
void Ref::assemble_aflux() {
  Rule ir = geom_rule(fe.geom, 2 * fe.p);
  int64_t rows = NE * nd;
  int rowlen = dim * nd;
  Aflux.I.resize(rows + 1);
  Aflux.J.resize(rows * rowlen);
  Aflux.A.resize(rows * rowlen);
  for (int64_t r = 0; r <= rows; r++) Aflux.I[r] = r * rowlen;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < NE; e++) {
    std::vector<real> elmat(nd * rowlen, 0.0), shape(nd), dshr(nd * dim), dshx(nd * dim);
    Trans T = trans(e);
    for (int q = 0; q < ir.n; q++) {
      const real* ip = &ir.pts[q * dim];
      fe.shape(ip, shape.data());
      for (int k = 0; k < nd; k++) shape[k] *= ir.w[q];
      real J[9], A[9];
      T.jacobian(ip, J);
      adjugate(dim, J, A);
      fe.dshape(ip, dshr.data());
      for (int j = 0; j < nd; j++)
        for (int d = 0; d < dim; d++) {
          real s = 0;
          for (int r = 0; r < dim; r++) s += dshr[j * dim + r] * A[r * dim + d];
          dshx[j * dim + d] = s;
        }
      for (int d = 0; d < dim; d++)
        for (int j = 0; j < nd; j++)
          for (int k = 0; k < nd; k++) elmat[j * rowlen + k + d * nd] += shape[k] * dshx[j * dim + d];
    }
    for (int j = 0; j < nd; j++) {
      int64_t row = e * nd + j;
      int64_t off = Aflux.I[row];
      for (int d = 0; d < dim; d++)
        for (int k = 0; k < nd; k++) {
          Aflux.J[off + d * nd + k] = d * NDofs + e * nd + k;
          Aflux.A[off + d * nd + k] = elmat[j * rowlen + k + d * nd];
        }
    }
  }
}

// fe_evolution.cpp:42-69 (MassIntegrator order 2p+OrderW, then inverse)
void Ref::assemble_mass() {
  Rule ir = geom_rule(fe.geom, 2 * fe.p + fe.order_w());
  Me_inv.assign(NE * nd * nd, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < NE; e++) {
    std::vector<real> M(nd * nd, 0.0), shape(nd);
    Trans T = trans(e);
    for (int q = 0; q < ir.n; q++) {
      const real* ip = &ir.pts[q * dim];
      real J[9];
      T.jacobian(ip, J);
      real w = det(dim, J) * ir.w[q];
      fe.shape(ip, shape.data());
      for (int i = 0; i < nd; i++)
        for (int j = 0; j < nd; j++) M[i * nd + j] += w * shape[i] * shape[j];
    }
    invert_dense(nd, M.data(), &Me_inv[e * nd * nd]);
  }
}

// NonlinearForm::Mult face/boundary loops with FaceIntegrator::AssembleFaceVector
// (face_integrator.cpp:11-276) and BoundaryIntegrator::AssembleFaceVector
// (boundary_integrator.cpp:12-194). Per-face element vectors are computed in
// parallel; each element then adds them in face-id order (interior faces, then
// boundary elements), which is the reference's serial accumulation order.
int Ref::face_mult(const real* x, real* z, real& mcs, int64_t* bad) {
  const int NEQ = ne();
  const int stride = 2 * nd * NEQ;
  Rule ir = face_rule(fe.geom, face_intorder());
  int nslots = (int)(interior.size() + bdr_faces.size());
  face_buf.assign((size_t)nslots * stride, 0.0);
  int err = 0;
  int64_t bad1 = -1, bad2 = -1;
  real gmax = 0.0;
#pragma omp parallel
  {
    std::vector<real> sh1(nd), sh2(nd);
    real lmax = 0.0;
#pragma omp for schedule(static)
    for (int64_t slot = 0; slot < nslots; slot++) {
      bool isb = slot >= (int64_t)interior.size();
      int64_t fid = isb ? bdr_faces[slot - interior.size()] : interior[slot];
      const Face& f = faces[fid];
      real* ev1 = &face_buf[(size_t)slot * stride];
      real* ev2 = ev1 + nd * NEQ;
      int bi = -1;
      if (isb) {
        for (int b = 0; b < sc.s.num_bc; b++)
          if (sc.s.bc_attr[b] == f.attr) bi = b;
        if (bi < 0) continue;  // unmarked boundary attribute: no integrator
      }
      for (int q = 0; q < ir.n; q++) {
        const real* s = &ir.pts[q * ir.dim];
        real xi1[3], xi2[3], nor[3], u1[5], u2[5], fl[5];
        face_to_elem(f, 0, s, xi1);
        fe.shape(xi1, sh1.data());
        for (int k = 0; k < NEQ; k++) {
          real v = 0;
          for (int i = 0; i < nd; i++) v += x[k * NDofs + f.e1 * nd + i] * sh1[i];
          u1[k] = v;
        }
        if (!isb) {
          face_to_elem(f, 1, s, xi2);
          fe.shape(xi2, sh2.data());
          for (int k = 0; k < NEQ; k++) {
            real v = 0;
            for (int i = 0; i < nd; i++) v += x[k * NDofs + f.e2 * nd + i] * sh2[i];
            u2[k] = v;
          }
        }
        face_normal(f, s, nor);
        real normag = 0;
        for (int i = 0; i < dim; i++) normag += nor[i] * nor[i];
        normag = std::sqrt(normag);
        // Vector::Norml2 (scaled hypot-style) then nor *= 1/norm
        real scale = 0.0, sum = 0.0;
        for (int i = 0; i < dim; i++) {
          if (nor[i] != 0.0) {
            real a = std::fabs(nor[i]);
            if (scale <= a) {
              real r = scale / a;
              sum = 1.0 + sum * (r * r);
              scale = a;
            } else {
              real r = a / scale;
              sum += r * r;
            }
          }
        }
        real n2 = scale * std::sqrt(sum);
        real inv = 1.0 / n2;
        for (int i = 0; i < dim; i++) nor[i] *= inv;
        if (isb) bc_state(sc, dim, bi, u1, nor, u2);
        real m = (isb && sc.s.bc_type[bi] == REF_BC_CHAR_OUTLET) ? char_outlet_flux(sc, dim, u2, nor, fl)
                                                                   : riemann(sc, dim, u1, u2, nor, fl);
        if (m < 0) {
#pragma omp critical
          {
            if (!err || fid < bad1) { err = 1; bad1 = f.e1; bad2 = f.e2; }
          }
          continue;
        }
        if (m > lmax) lmax = m;
        for (int k = 0; k < NEQ; k++) fl[k] *= ir.w[q];
        for (int k = 0; k < NEQ; k++) {
          for (int i = 0; i < nd; i++) ev1[k * nd + i] -= fl[k] * sh1[i] * normag;
          if (!isb)
            for (int i = 0; i < nd; i++) ev2[k * nd + i] += fl[k] * sh2[i] * normag;
        }
      }
    }
#pragma omp critical
    { if (lmax > gmax) gmax = lmax; }
  }
  if (gmax > mcs) mcs = gmax;
  // z = 0; AddElementVector per face in order
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < NE; e++) {
    for (int k = 0; k < NEQ; k++)
      for (int i = 0; i < nd; i++) z[k * NDofs + e * nd + i] = 0.0;
    for (auto& c : contrib[e]) {
      const real* ev = &face_buf[(size_t)c.first * stride] + c.second * nd * NEQ;
      for (int k = 0; k < NEQ; k++)
        for (int i = 0; i < nd; i++) z[k * NDofs + e * nd + i] += ev[k * nd + i];
    }
  }
  if (err) {
    if (bad) { bad[0] = bad1; bad[1] = bad2; }
    return -1;
  }
  return 0;
}

// fe_evolution.cpp:102-216 (Mult) and :219-245 (GetFlux)
int Ref::mult(const real* x, real* y, real& mcs, int64_t* bad) {
  const int NEQ = ne();
  mcs = 0.0;  // :105
  std::vector<real> z(NEQ * NDofs);
  int rc = face_mult(x, z.data(), mcs, bad);
  // GetFlux: flux(i, d, k) stored at i + NDofs*(d + dim*k)
  std::vector<real> flux((size_t)NDofs * dim * NEQ);
  real gmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : gmax)
  for (int64_t i = 0; i < NDofs; i++) {
    real st[5], f[15];
    for (int k = 0; k < NEQ; k++) st[k] = x[k * NDofs + i];
    Gas g = sc.gas();
    flux_tensor(st, dim, g, f);
    for (int d = 0; d < dim; d++)
      for (int k = 0; k < NEQ; k++) flux[i + NDofs * (d + dim * k)] = f[k * dim + d];
    real m = max_char_speed(st, dim, g);
    if (m > gmax) gmax = m;
  }
  if (gmax > mcs) mcs = gmax;
  // Aflux->AddMult(fk, zk) per equation (SparseMatrix::AddMult: row sum then add)
  for (int k = 0; k < NEQ; k++) {
    const real* fk = &flux[(size_t)k * dim * NDofs];
    real* zk = &z[(size_t)k * NDofs];
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < NDofs; r++) {
      real d = 0.0;
      for (int64_t j = Aflux.I[r]; j < Aflux.I[r + 1]; j++) d += Aflux.A[j] * fk[Aflux.J[j]];
      zk[r] += d;
    }
  }
  // y_e = Me_inv[e] * z_e  (mfem::Mult dense)
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < NE; e++) {
    const real* Mi = &Me_inv[e * nd * nd];
    for (int k = 0; k < NEQ; k++)
      for (int i = 0; i < nd; i++) {
        real s = 0.0;
        for (int j = 0; j < nd; j++) s += Mi[i * nd + j] * z[k * NDofs + e * nd + j];
        y[k * NDofs + e * nd + i] = s;
      }
  }
  return rc;
}

// averager.cpp:84-109 -> GridFunction::GetElementAverages: (sum_k m_k u_k)/(sum_k m_k),
// m_k = integral of phi_k with MassIntegrator order p + OrderW.
void Ref::compute_means(const real* x) {
  const int NEQ = ne();
  Rule ir = geom_rule(fe.geom, fe.p + fe.order_w());
  avgs.assign(NEQ * NE, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < NE; e++) {
    std::vector<real> m(nd, 0.0), shape(nd);
    Trans T = trans(e);
    for (int q = 0; q < ir.n; q++) {
      const real* ip = &ir.pts[q * dim];
      real J[9];
      T.jacobian(ip, J);
      real w = det(dim, J) * ir.w[q];
      fe.shape(ip, shape.data());
      for (int k = 0; k < nd; k++) m[k] += w * 1.0 * shape[k];
    }
    real ipsi = 0;
    for (int k = 0; k < nd; k++) ipsi += m[k] * 1.0;
    for (int v = 0; v < NEQ; v++) {
      real s = 0;
      for (int k = 0; k < nd; k++) s += m[k] * x[v * NDofs + e * nd + k];
      avgs[v * NE + e] = s / ipsi;
    }
  }
}

// Reference coordinates in element k of the physical point x (TransformBack):
// the affine inverse, or Newton's method on the bilinear map. The point may lie
// outside element k (the Shu indicator extrapolates the neighbour's polynomial
// over the troubled cell); the iteration is not clipped to the element.
void Ref::transform_back(int64_t k, const real* x, real* xi) const {
  real J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (affine[k] || fe.geom == REF_GEOM_TRI) {
    affine_jacobian(k, J);
    const real* X0 = X.data() + k * nv * dim;
    real b[3];
    for (int i = 0; i < dim; i++) b[i] = x[i] - X0[i];
    solve_small(dim, J, b, xi);
    return;
  }
  Trans T = trans(k);
  for (int d = 0; d < dim; d++) xi[d] = 0.5;
  for (int it = 0; it < 60; it++) {
    real xc[3], r[3], dxi[3];
    T.transform(xi, xc);
    for (int i = 0; i < dim; i++) r[i] = x[i] - xc[i];
    T.jacobian(xi, J);
    solve_small(dim, J, r, dxi);
    real step = 0;
    for (int d = 0; d < dim; d++) {
      xi[d] += dxi[d];
      step = std::max(step, std::fabs(dxi[d]));
    }
    if (step < 1e-15) break;
  }
}

// Averager::computeStencilExtrapAverages for the stencil of cell t
// (averager.cpp:152-273): for every cell k of the stencil (t itself first, then
// its face neighbours), the neighbour's polynomial integrated with the weights
// of the troubled cell's reference rule (order p + OrderW) at the points
// T_k^-1(T_t(xi_q)), times Trans_k.Weight() at the last point of the mass rule
// (assembleShiftedElementMatrix reads the weight without setting a point,
// :262-266), divided by |k| (the integral of the constant test function).
// out[m * NEQ + v] for the m-th stencil cell.
void Ref::extrap_averages(int64_t t, const real* x, const std::vector<int64_t>& cells, real* out) const {
  const int NEQ = ne();
  Rule ir = geom_rule(fe.geom, fe.p + fe.order_w());
  std::vector<real> shape(nd);
  Trans Tt = trans(t);
  for (size_t m = 0; m < cells.size(); m++) {
    const int64_t k = cells[m];
    Trans Tk = trans(k);
    real vol = 0, wlast = 0;
    for (int q = 0; q < ir.n; q++) {
      real J[9];
      Tk.jacobian(&ir.pts[q * dim], J);
      const real dJ = det(dim, J);
      vol += dJ * ir.w[q];
      wlast = dJ;
    }
    real acc[5] = {};
    for (int q = 0; q < ir.n; q++) {
      real xp[3], xi[3];
      Tt.transform(&ir.pts[q * dim], xp);
      if (k == t) {
        for (int d = 0; d < dim; d++) xi[d] = ir.pts[q * dim + d];
      } else {
        transform_back(k, xp, xi);
      }
      fe.shape(xi, shape.data());
      const real w = wlast * ir.w[q];
      for (int v = 0; v < NEQ; v++) {
        real u = 0;
        for (int i = 0; i < nd; i++) u += shape[i] * x[v * NDofs + k * nd + i];
        acc[v] += w * u;
      }
    }
    for (int v = 0; v < NEQ; v++) out[m * NEQ + v] = acc[v] / vol;
  }
}

// Limiter::update (limiter.cpp:22-94), IndicatorBJ::checkDiscontinuity
// (indicator_bj.cpp:88-200) with updateYMin (:15-86), LimiterMultiplier::limit
// (limiter_multiplier.cpp:12-99), LimiterFinDiff::limit (limiter_findiff.cpp:3-33).
void Ref::limiter_update(real* x, real* ind_values, real* min_values) {
  const int NEQ = ne();
  const Gas g = sc.gas();
  std::vector<real> xNew(x, x + NEQ * NDofs);
  compute_means(x);
  Rule fr = face_rule(fe.geom, 2);  // indicator_bj.cpp:150 (intorder 2)
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < NE; c++) {
    real yMin[5], mI[5], MI[5], mean[5];
    std::vector<real> el(nd * NEQ), shape(nd);
    for (int k = 0; k < NEQ; k++)
      for (int i = 0; i < nd; i++) el[k * nd + i] = x[k * NDofs + c * nd + i];
    for (int v = 0; v < NEQ; v++) mean[v] = avgs[v * NE + c];
    // --- indicator
    if (sc.s.indicator == REF_IND_NOWHERE || sc.s.indicator == REF_IND_EVERYWHERE) {
      real val = sc.s.indicator == REF_IND_NOWHERE ? 1.0 : 0.0;
      for (int v = 0; v < NEQ; v++) yMin[v] = val;
    } else if (sc.s.indicator == REF_IND_SHU) {
      // IndicatorShu::checkDiscontinuity (indicator_shu.cpp:13-74): maxFabsPj
      // takes the troubled cell's own mean at every stencil cell (:40-48)
      std::vector<int64_t> cells;
      cells.push_back(c);
      for (int64_t nb : el2el[c]) cells.push_back(nb);
      std::vector<real> ex(cells.size() * NEQ);
      extrap_averages(c, x, cells, ex.data());
      const real Ck = 0.03, eps = 1e-6;
      for (int v = 0; v < NEQ; v++) {
        real mx = -1e9, sum = 0;
        for (size_t m = 0; m < cells.size(); m++) {
          mx = std::fabs(mean[v]) > mx ? std::fabs(mean[v]) : mx;
          sum += std::fabs(mean[v] - ex[m * NEQ + v]);
        }
        yMin[v] = (std::fabs(mx) < eps && std::fabs(sum) < eps) ? real(1) : std::min(real(1), Ck * mx / (sum + eps));
      }
    } else {
      for (int v = 0; v < NEQ; v++) { mI[v] = 1e9; MI[v] = -1e9; yMin[v] = 1e9; }
      std::vector<int64_t> cells;
      cells.push_back(c);
      for (int64_t nb : el2el[c]) cells.push_back(nb);
      for (int64_t k : cells)
        for (int v = 0; v < NEQ; v++) {
          real a = avgs[v * NE + k];
          mI[v] = a < mI[v] ? a : mI[v];
          MI[v] = a > MI[v] ? a : MI[v];
        }
      real hK = 0.0;
      if (sc.s.phi == REF_PHI_VENKAT && sc.s.venkat_k > 0) {
        // element size (volume^(1/dim)) for eps^2 = (K h)^3
        Rule vr2 = geom_rule(fe.geom, 2);
        real vol = 0;
        Trans T = trans(c);
        for (int q = 0; q < vr2.n; q++) {
          real J[9];
          T.jacobian(&vr2.pts[q * dim], J);
          vol += det(dim, J) * vr2.w[q];
        }
        hK = std::pow(vol, 1.0 / dim);
      }
      for (int lf = 0; lf < fe.nf; lf++) {
        int64_t fid = elem_faces[c][lf];
        const Face& f = faces[fid];
        if (f.e2 < 0 && !f.ncm) continue;  // boundary faces are not in the stencil
        int side = f.e1 == c ? 0 : 1;
        for (int q = 0; q < fr.n; q++) {
          real xi[3], u[5];
          face_to_elem(f, side, &fr.pts[q * fr.dim], xi);
          fe.shape(xi, shape.data());
          for (int k = 0; k < NEQ; k++) {
            real s = 0;
            for (int i = 0; i < nd; i++) s += el[k * nd + i] * shape[i];
            u[k] = s;
          }
          for (int i = 0; i < NEQ; i++) {
            real diff = u[i] - mean[i];
            real y;
            if (diff > 1e-3) y = (MI[i] - mean[i]) / diff;
            else if (diff < -1e-3) y = (mI[i] - mean[i]) / diff;
            else y = 1.0;
            if (sc.s.phi == REF_PHI_BJ) {  // indicator_bj.cpp:66-68 (commented in reference)
              yMin[i] = y < yMin[i] ? y : yMin[i];
              yMin[i] = yMin[i] > 1.0 ? 1.0 : yMin[i];
            } else if (sc.s.phi == REF_PHI_VENKAT) {  // :70-72 (commented in reference)
              real yCur;
              if (sc.s.venkat_k > 0) {
                real eps2 = std::pow(sc.s.venkat_k * hK, 3.0);
                real d2 = diff * diff;
                yCur = (d2 * (y * y + 2.0 * y) + eps2) / (d2 * (y * y + y + 2.0) + eps2);
              } else {
                yCur = (y * y + 2.0 * y) / (y * y + y + 2.0);
              }
              yMin[i] = yCur < yMin[i] ? yCur : yMin[i];
            } else {  // michalak, :75-79 (active in the reference)
              real yStar = 1.5;
              real yRel = y / yStar;
              real yCur = y < 1.0 ? y + (3.0 - 2.0 * yStar) * yRel * yRel + (yStar - 2.0) * yRel * yRel * yRel : 1.0;
              yMin[i] = yCur < yMin[i] ? yCur : yMin[i];
            }
          }
        }
      }
    }
    for (int v = 0; v < NEQ; v++) if (ind_values) ind_values[v * NE + c] = yMin[v];
    real minV = yMin[0];
    if (min_values) min_values[c] = minV;
    // --- limit
    if (sc.s.limiter == REF_LIM_MULTI) {
      if (sc.s.linearize && fe.geom != REF_GEOM_TRI) {  // triangles are skipped (limiter_multiplier.cpp:29)
        // LimiterMultiplier::linearize (limiter_multiplier.cpp:101-178) for p = 1:
        // u_i = mean - dF . (xi_c - xi_i) with dF the reference gradient at the
        // centre xi_c (the image of the element centre, transformed back) and
        // xi_i the points of the order-2 rule, which are the nodes
        real xc[3] = {0.5, 0.5, 0.5};
        std::vector<real> dsh(nd * dim);
        fe.dshape(xc, dsh.data());
        Rule lr = geom_rule(fe.geom, fe.p + 1);
        for (int k = 0; k < NEQ; k++) {
          real dF[3] = {};
          for (int d = 0; d < dim; d++)
            for (int i = 0; i < nd; i++) dF[d] += el[k * nd + i] * dsh[i * dim + d];
          std::vector<real> nv_(nd);
          for (int i = 0; i < nd; i++) {
            real u = mean[k];
            for (int d = 0; d < dim; d++) u -= dF[d] * (xc[d] - lr.pts[i * dim + d]);
            nv_[i] = u;
          }
          for (int i = 0; i < nd; i++) el[k * nd + i] = nv_[i];
        }
      }
      int reps = sc.s.pow_nd ? nd : 1;
      for (int k = 0; k < NEQ; k++)
        for (int r = 0; r < reps; r++)
          for (int i = 0; i < nd; i++) el[k * nd + i] = mean[k] + minV * (el[k * nd + i] - mean[k]);
      if (sc.s.last_hope) {
        for (int v = 0; v < fe.nv; v++) {
          fe.shape(fe.vert(v), shape.data());
          real u[5] = {};
          for (int k = 0; k < NEQ; k++) {
            real s = 0;
            for (int i = 0; i < nd; i++) s += el[k * nd + i] * shape[i];
            u[k] = s;
          }
          if (!state_is_physical(u, dim, g)) {
            for (int k = 0; k < NEQ; k++)
              for (int i = 0; i < nd; i++) el[k * nd + i] = mean[k];
          }
        }
      }
    } else if (sc.s.limiter == REF_LIM_FINDIFF) {
      bool total = false;
      for (int k = 0; k < NEQ; k++) total = yMin[k] < 0.999 ? true : false;
      if (total)
        for (int k = 0; k < NEQ; k++)
          for (int i = 0; i < nd; i++) el[k * nd + i] = mean[k];
    }
    for (int k = 0; k < NEQ; k++)
      for (int i = 0; i < nd; i++) xNew[k * NDofs + c * nd + i] = el[k * nd + i];
  }
  std::memcpy(x, xNew.data(), sizeof(real) * NEQ * NDofs);
}

// Mesh::GetElementSize(i, 1): smallest singular value of the (perfect-geometry)
// Jacobian at the element centre; main.cpp:191-201 takes the minimum.
real Ref::hmin() const {
  real hm = 1e300;
  for (int64_t e = 0; e < NE; e++) {
    real xc[3] = {0.5, 0.5, 0.5};
    if (fe.geom == REF_GEOM_TRI) { xc[0] = 1.0 / 3.0; xc[1] = 1.0 / 3.0; }
    real J[9];
    trans(e).jacobian(xc, J);
    if (fe.geom == REF_GEOM_TRI) {
      // multiply by the Jacobian mapping the equilateral triangle onto the reference one
      real P[4] = {1.0, -1.0 / std::sqrt(3.0), 0.0, 2.0 / std::sqrt(3.0)};
      real Q[4];
      Q[0] = J[0] * P[0] + J[1] * P[2];
      Q[1] = J[0] * P[1] + J[1] * P[3];
      Q[2] = J[2] * P[0] + J[3] * P[2];
      Q[3] = J[2] * P[1] + J[3] * P[3];
      std::memcpy(J, Q, sizeof(Q));
    }
    int d = dim;
    // smallest singular value: sqrt of smallest eigenvalue of J^T J
    real A[9] = {0};
    for (int i = 0; i < d; i++)
      for (int j = 0; j < d; j++) {
        real s = 0;
        for (int k = 0; k < d; k++) s += J[k * d + i] * J[k * d + j];
        A[i * d + j] = s;
      }
    real lmin;
    if (d == 2) {
      real tr = A[0] + A[3], dt = A[0] * A[3] - A[1] * A[2];
      real disc = std::sqrt(std::max(real(0), tr * tr / 4 - dt));
      lmin = tr / 2 - disc;
    } else {
      // symmetric 3x3 eigenvalues (trigonometric method)
      real p1 = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
      real q = (A[0] + A[4] + A[8]) / 3;
      if (p1 == 0) lmin = std::min(A[0], std::min(A[4], A[8]));
      else {
        real p2 = (A[0] - q) * (A[0] - q) + (A[4] - q) * (A[4] - q) + (A[8] - q) * (A[8] - q) + 2 * p1;
        real p = std::sqrt(p2 / 6);
        real B[9];
        for (int i = 0; i < 9; i++) B[i] = (A[i] - (i % 4 == 0 ? q : 0)) / p;
        real r = det(3, B) / 2;
        real phi = r <= -1 ? M_PI / 3 : (r >= 1 ? 0 : std::acos(r) / 3);
        real e1 = q + 2 * p * std::cos(phi);
        real e3 = q + 2 * p * std::cos(phi + 2 * M_PI / 3);
        lmin = e3;
        (void)e1;
      }
    }
    real h = std::sqrt(std::max(real(0), lmin));
    if (h < hm) hm = h;
  }
  return hm;
}

}  // namespace

// ============================================================================
// The C interface is double; when the restatement is built wider (REF_REAL),
// inputs are widened on the way in and results rounded on the way out. In the
// double build these are pass-throughs without copies.
namespace {
const real* widen(const double* x, size_t n, std::vector<real>& tmp) {
  if constexpr (std::is_same<real, double>::value) {
    (void)n;
    (void)tmp;
    return reinterpret_cast<const real*>(x);
  } else {
    tmp.assign(x, x + n);
    return tmp.data();
  }
}
real* widen_out(double* y, size_t n, std::vector<real>& tmp) {
  if constexpr (std::is_same<real, double>::value) {
    (void)n;
    (void)tmp;
    return reinterpret_cast<real*>(y);
  } else {
    tmp.assign(n, real(0));
    return tmp.data();
  }
}
void narrow(const std::vector<real>& tmp, double* y) {
  if constexpr (!std::is_same<real, double>::value) {
    for (size_t i = 0; i < tmp.size(); i++) y[i] = (double)tmp[i];
  } else {
    (void)tmp;
    (void)y;
  }
}
}  // namespace

extern "C" {

void* ref_create_nc(int geom, int order, int64_t num_elements, const double* elem_coords,
                    const int64_t* elem_verts, int64_t num_bdr, const int64_t* bdr_verts,
                    const int* bdr_attr, const ref_scheme* scheme, int64_t num_nc,
                    const int64_t* nc_slave, const int* nc_slave_face, const int64_t* nc_master,
                    const int* nc_master_face, const double* nc_param, char* err, int errlen) {
  Ref* r = new Ref();
  r->fe.init(geom, order);
  r->sc.s = *scheme;
  r->dim = r->fe.dim;
  r->nv = r->fe.nv;
  r->nd = r->fe.nd;
  r->NE = num_elements;
  r->NDofs = num_elements * r->nd;
  r->X.assign(elem_coords, elem_coords + num_elements * r->nv * r->dim);  // widened element-wise
  r->topo.assign(elem_verts, elem_verts + num_elements * r->nv);
  std::string e;
  if (scheme->indicator == REF_IND_SHU && scheme->limiter == REF_LIM_MULTI)
    e = "the Shu indicator writes no minValues, which the multiplier reads: pair it with FinDiff";
  else if (scheme->linearize && scheme->limiter == REF_LIM_MULTI && order != 1 && geom != REF_GEOM_TRI)
    e = "linearize reads nd points of an order p+1 rule: p = 1 quadrilaterals and hexahedra only";
  if (!e.empty() || !r->build(e) ||
      !r->attach_nc(num_nc, nc_slave, nc_slave_face, nc_master, nc_master_face, nc_param, e)) {
    if (err) snprintf(err, errlen, "%s", e.c_str());
    delete r;
    return nullptr;
  }
  // boundary elements -> faces
  std::map<std::vector<int64_t>, int64_t> key2face;
  for (int64_t f = 0; f < (int64_t)r->faces.size(); f++) {
    if (r->faces[f].e2 >= 0) continue;
    std::vector<int64_t> key(r->faces[f].fvt, r->faces[f].fvt + r->fe.nvf);
    std::sort(key.begin(), key.end());
    key2face[key] = f;
  }
  for (int64_t b = 0; b < num_bdr; b++) {
    std::vector<int64_t> key(bdr_verts + b * r->fe.nvf, bdr_verts + (b + 1) * r->fe.nvf);
    std::sort(key.begin(), key.end());
    auto it = key2face.find(key);
    if (it == key2face.end()) {
      if (err) snprintf(err, errlen, "boundary element %ld matches no boundary face", (long)b);
      delete r;
      return nullptr;
    }
    r->faces[it->second].attr = bdr_attr[b];
    r->bdr_faces.push_back(it->second);
    r->bdr_attr.push_back(bdr_attr[b]);
  }
  // accumulation order per element
  r->contrib.assign(r->NE, {});
  int64_t slot = 0;
  for (int64_t f : r->interior) {
    r->contrib[r->faces[f].e1].push_back({slot, 0});
    r->contrib[r->faces[f].e2].push_back({slot, 1});
    slot++;
  }
  for (int64_t f : r->bdr_faces) {
    r->contrib[r->faces[f].e1].push_back({slot, 0});
    slot++;
  }
  r->detect_affine();
  r->assemble_aflux();
  r->assemble_mass();
  return r;
}

void* ref_create(int geom, int order, int64_t num_elements, const double* elem_coords,
                 const int64_t* elem_verts, int64_t num_bdr, const int64_t* bdr_verts,
                 const int* bdr_attr, const ref_scheme* scheme, char* err, int errlen) {
  return ref_create_nc(geom, order, num_elements, elem_coords, elem_verts, num_bdr, bdr_verts, bdr_attr, scheme, 0,
                       nullptr, nullptr, nullptr, nullptr, nullptr, err, errlen);
}

void ref_destroy(void* h) { delete (Ref*)h; }
int64_t ref_num_dofs(void* h) { return ((Ref*)h)->NDofs; }
int ref_nodes_per_element(void* h) { return ((Ref*)h)->nd; }
int64_t ref_num_interior_faces(void* h) { return (int64_t)((Ref*)h)->interior.size(); }

void ref_node_coords(void* h, double* xyz) {
  Ref* r = (Ref*)h;
  for (int64_t e = 0; e < r->NE; e++) {
    Trans T = r->trans(e);
    for (int i = 0; i < r->nd; i++) {
      real x[3];
      T.transform(&r->fe.nodes[i * r->dim], x);
      for (int d = 0; d < r->dim; d++) xyz[(e * r->nd + i) * r->dim + d] = (double)x[d];
    }
  }
}

int ref_mult(void* h, const double* x, double* y, double* mcs, int64_t* bad) {
  Ref* r = (Ref*)h;
  const size_t n = (size_t)r->ne() * r->NDofs;
  std::vector<real> xt, yt;
  real m = 0;
  int rc = r->mult(widen(x, n, xt), widen_out(y, n, yt), m, bad);
  narrow(yt, y);
  if (mcs) *mcs = (double)m;
  return rc;
}

int ref_face_mult(void* h, const double* x, double* z, double* mcs, int64_t* bad) {
  Ref* r = (Ref*)h;
  const size_t n = (size_t)r->ne() * r->NDofs;
  std::vector<real> xt, zt;
  real m = 0;
  int rc = r->face_mult(widen(x, n, xt), widen_out(z, n, zt), m, bad);
  narrow(zt, z);
  if (mcs) *mcs = (double)m;
  return rc;
}

void ref_means(void* h, const double* x, double* means) {
  Ref* r = (Ref*)h;
  std::vector<real> xt;
  r->compute_means(widen(x, (size_t)r->ne() * r->NDofs, xt));
  for (size_t i = 0; i < r->avgs.size(); i++) means[i] = (double)r->avgs[i];
}

void ref_limiter_update(void* h, double* x, double* ind, double* mv) {
  Ref* r = (Ref*)h;
  const size_t n = (size_t)r->ne() * r->NDofs;
  std::vector<real> xt, it, mt;
  real* xr = widen_out(x, n, xt);
  if constexpr (!std::is_same<real, double>::value) std::copy(x, x + n, xr);
  real* ir = ind ? widen_out(ind, (size_t)r->ne() * r->NE, it) : nullptr;
  real* mr = mv ? widen_out(mv, (size_t)r->NE, mt) : nullptr;
  r->limiter_update(xr, ir, mr);
  narrow(xt, x);
  if (ind) narrow(it, ind);
  if (mv) narrow(mt, mv);
}

// ExplicitRKLimitedSolver::Step (rk_explicit_limited.cpp:30-82): packed a[l++]
// consumption, MFEM add(x, a*dt, k0, y) and y.Add(a*dt, kj) association.
int ref_step(void* h, double* x, double t, double dt, int s, const double* a, const double* b,
             const double* c, double* mcs, int64_t* bad) {
  (void)t;
  (void)c;
  Ref* r = (Ref*)h;
  int64_t n = r->ne() * r->NDofs;
  std::vector<std::vector<real>> k(s, std::vector<real>(n));
  std::vector<real> y(n, 0.0);
  std::vector<real> xt;
  real* xr = widen_out(x, (size_t)n, xt);
  if constexpr (!std::is_same<real, double>::value) std::copy(x, x + n, xr);
  real m = 0;
  int rc = r->mult(xr, k[0].data(), m, bad);
  const real dtr = dt;
  for (int l = 0, i = 1; i < s; i++) {
    real a0 = a[l++] * dtr;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; j++) y[j] = xr[j] + a0 * k[0][j];
    for (int jj = 1; jj < i; jj++) {
      real aj = a[l++] * dtr;
#pragma omp parallel for schedule(static)
      for (int64_t j = 0; j < n; j++) y[j] += aj * k[jj][j];
    }
    r->limiter_update(y.data(), nullptr, nullptr);
    int rc2 = r->mult(y.data(), k[i].data(), m, bad);
    if (rc2) rc = rc2;
  }
  for (int i = 0; i < s; i++) {
    real bi = b[i] * dtr;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; j++) xr[j] += bi * k[i][j];
  }
  r->limiter_update(xr, nullptr, nullptr);
  narrow(xt, x);
  if (mcs) *mcs = (double)m;
  return rc;
}

double ref_hmin(void* h) { return (double)((Ref*)h)->hmin(); }

void ref_element_mass_inverse(void* h, int64_t e, double* out) {
  Ref* r = (Ref*)h;
  for (int i = 0; i < r->nd * r->nd; i++) out[i] = (double)r->Me_inv[e * r->nd * r->nd + i];
}

void ref_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

double ref_riemann_eval(const ref_scheme* s, int dim, const double* s1, const double* s2,
                        const double* nor, double* flux) {
  Scheme sc;
  sc.s = *s;
  real a[5] = {}, b[5] = {}, n[3] = {}, f[5] = {};
  for (int i = 0; i < dim + 2; i++) {
    a[i] = s1[i];
    b[i] = s2[i];
  }
  for (int i = 0; i < dim; i++) n[i] = nor[i];
  const real lam = riemann(sc, dim, a, b, n, f);
  for (int i = 0; i < dim + 2; i++) flux[i] = (double)f[i];
  return (double)lam;
}

void ref_bc_state(const ref_scheme* s, int dim, int bi, const double* s1, const double* nor, double* s2) {
  Scheme sc;
  sc.s = *s;
  real a[5] = {}, n[3] = {}, g[5] = {};
  for (int i = 0; i < dim + 2; i++) a[i] = s1[i];
  for (int i = 0; i < dim; i++) n[i] = nor[i];
  bc_state(sc, dim, bi, a, n, g);
  for (int i = 0; i < dim + 2; i++) s2[i] = (double)g[i];
}

void ref_gl_rule(int n, double* x, double* w) {
  real xr[64], wr[64];
  gauss_legendre(n, xr, wr);
  for (int i = 0; i < n; i++) {
    x[i] = (double)xr[i];
    w[i] = (double)wr[i];
  }
}

void ref_shape(int geom, int order, const double* xi, double* phi, double* dphi) {
  RefFE fe;
  fe.init(geom, order);
  real xr[3] = {xi[0], xi[1], fe.dim == 3 ? xi[2] : 0.0};
  std::vector<real> ph(fe.nd), dph((size_t)fe.nd * fe.dim);
  if (phi) {
    fe.shape(xr, ph.data());
    for (int i = 0; i < fe.nd; i++) phi[i] = (double)ph[i];
  }
  if (dphi) {
    fe.dshape(xr, dph.data());
    for (size_t i = 0; i < dph.size(); i++) dphi[i] = (double)dph[i];
  }
}

}  // extern "C"
