// fr_host_ops.cpp -- operator data of FRPSpace1D / FRPSpace2D built natively on the host, the
// error plumbing of the C ABI, and the fr_* operator entry points of fr.h.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fr_common.h"

namespace fr {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  g_error = std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" + std::to_string(line) + ")";
  return FR_E_HIP;
}

// Legendre recurrence returning P_n(x) and P_{n-1}(x).
static void legendre_pair(int n, double x, double &pn, double &pnm1) {
  double a = 1.0, b = x;  // P_0, P_1
  if (n == 0) {
    pn = 1.0;
    pnm1 = 0.0;
    return;
  }
  for (int k = 1; k < n; ++k) {
    const double c = ((2.0 * k + 1.0) * x * b - k * a) / (k + 1.0);
    a = b;
    b = c;
  }
  pn = b;
  pnm1 = a;
}

// Gauss-Legendre rule on [-1, 1], nodes ascending (src/Polynomial/poly_legendre.jl:6):
// Newton's method on P_n from cos(pi (i + 3/4) / (n + 1/2)); the rule is made exactly symmetric.
void gauss_legendre(int n, double *x, double *w) {
  if (n == 1) {
    x[0] = 0.0;
    w[0] = 2.0;
    return;
  }
  const int half = (n + 1) / 2;
  for (int i = 0; i < half; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double pn, pm;
      legendre_pair(n, z, pn, pm);
      const double dp = n * (z * pn - pm) / (z * z - 1.0);
      const double dz = pn / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double pn, pm;
    legendre_pair(n, z, pn, pm);
    const double dp = n * (z * pn - pm) / (z * z - 1.0);
    const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    if ((n & 1) && i == half - 1) z = 0.0;
    x[n - 1 - i] = z;
    x[i] = -z;
    w[n - 1 - i] = wi;
    w[i] = wi;
  }
}

// P_n'(x) through P'_{k+1} = x P'_k + (k+1) P_k (GSL sf_legendre_Pl_deriv_array's last entry,
// src/Polynomial/poly_legendre.jl:13).
double dlegendre(int n, double x) {
  if (n <= 0) return 0.0;
  double p = 1.0, pprev = 0.0, dp = 0.0;
  for (int k = 0; k < n; ++k) {
    const double dnext = x * dp + (k + 1.0) * p;
    const double pnext = (k == 0) ? x : ((2.0 * k + 1.0) * x * p - k * pprev) / (k + 1.0);
    pprev = p;
    p = pnext;
    dp = dnext;
  }
  return dp;
}

// Correction-function derivatives (src/Polynomial/poly_legendre.jl:29-71): Radau (nodal DG),
// spectral difference, Huynh g2.
void dcorrection(int kind, int deg, const double *x, int n, double *dgl, double *dgr) {
  const double sgn = (deg & 1) ? -1.0 : 1.0;
  for (int i = 0; i < n; ++i) {
    const double D = dlegendre(deg, x[i]);
    const double Dp = dlegendre(deg + 1, x[i]);
    double y = Dp;
    if (kind != FR_CORR_RADAU) {
      const double Dm = dlegendre(deg - 1, x[i]);
      y = (kind == FR_CORR_SD) ? (deg * Dm + (deg + 1) * Dp) / (2.0 * deg + 1.0)
                               : ((deg + 1) * Dm + deg * Dp) / (2.0 * deg + 1.0);
    }
    dgl[i] = sgn * 0.5 * (D - y);
    dgr[i] = 0.5 * (D + y);
  }
}

// src/Polynomial/poly_lagrange.jl:6-21
void lagrange_point(const double *sp, int nsp, double x, double *l) {
  for (int k = 0; k < nsp; ++k) {
    double prod = 1.0;
    for (int j = 0; j < nsp; ++j)
      if (j != k) prod *= (x - sp[j]) / (sp[k] - sp[j]);
    l[k] = prod;
  }
}

// src/Polynomial/poly_lagrange.jl:38-59, lpdm(m, k) = l_k'(sp_m) column-major
void dlagrange(const double *sp, int nsp, double *lpdm) {
  for (int k = 0; k < nsp; ++k)
    for (int m = 0; m < nsp; ++m) {
      double acc = 0.0;
      for (int l = 0; l < nsp; ++l) {
        if (l == k) continue;
        double prod = 1.0;
        for (int j = 0; j < nsp; ++j)
          if (j != k && j != l) prod *= (sp[m] - sp[j]) / (sp[k] - sp[j]);
        acc += prod / (sp[k] - sp[l]);
      }
      lpdm[m + nsp * k] = acc;
    }
}

// Orthonormal Jacobi polynomial by the symmetric three-term recurrence
// (src/Polynomial/poly_jacobi.jl:7-42).
double jacobi_p(double x, double alpha, double beta, int N) {
  const double ab = alpha + beta;
  const double g0 = std::pow(2.0, ab + 1.0) / (ab + 1.0) * std::tgamma(alpha + 1.0) *
                    std::tgamma(beta + 1.0) / std::tgamma(ab + 1.0);
  double pm = 1.0 / std::sqrt(g0);
  if (N == 0) return pm;
  const double g1 = (alpha + 1.0) * (beta + 1.0) / (ab + 3.0) * g0;
  double pc = ((ab + 2.0) * x / 2.0 + (alpha - beta) / 2.0) / std::sqrt(g1);
  if (N == 1) return pc;
  double aold = 2.0 / (2.0 + ab) * std::sqrt((alpha + 1.0) * (beta + 1.0) / (ab + 3.0));
  for (int i = 1; i <= N - 1; ++i) {
    const double h1 = 2.0 * i + ab;
    const double anew = 2.0 / (h1 + 2.0) *
                        std::sqrt((i + 1.0) * (i + 1.0 + ab) * (i + 1.0 + alpha) * (i + 1.0 + beta) /
                                  (h1 + 1.0) / (h1 + 3.0));
    const double bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2.0);
    const double pn = 1.0 / anew * (-aold * pm + (x - bnew) * pc);
    aold = anew;
    pm = pc;
    pc = pn;
  }
  return pc;
}

// src/Polynomial/poly_jacobi.jl:88-95
double djacobi_p(double x, double alpha, double beta, int N) {
  if (N == 0) return 0.0;
  return std::sqrt(N * (N + alpha + beta + 1.0)) * jacobi_p(x, alpha + 1.0, beta + 1.0, N - 1);
}

// Gauss-Jordan elimination with partial pivoting; column-major in and out.
bool invert(int n, const double *A, double *Ainv) {
  std::vector<double> M((size_t)n * 2 * n);
  auto at = [&](int r, int c) -> double & { return M[(size_t)r * 2 * n + c]; };
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      at(r, c) = A[r + (size_t)n * c];
      at(r, n + c) = (r == c) ? 1.0 : 0.0;
    }
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(at(r, c)) > std::fabs(at(piv, c))) piv = r;
    if (at(piv, c) == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < 2 * n; ++k) std::swap(at(c, k), at(piv, k));
    const double d = at(c, c);
    for (int k = 0; k < 2 * n; ++k) at(c, k) /= d;
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = at(r, c);
      if (f == 0.0) continue;
      for (int k = 0; k < 2 * n; ++k) at(r, k) -= f * at(c, k);
    }
  }
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) Ainv[r + (size_t)n * c] = at(r, n + c);
  return true;
}

// FRPSpace1D / FRPSpace2D operator data (src/struct.jl:40-85, 130-227).
bool build_ops(int deg, int correction, HostOps &o) {
  if (deg < 1 || deg > 8) return false;
  const int n = deg + 1;
  o.deg = deg;
  o.nsp = n;
  gauss_legendre(n, o.r, o.w);
  lagrange_point(o.r, n, -1.0, o.ll);
  lagrange_point(o.r, n, 1.0, o.lr);
  dlagrange(o.r, n, o.dl);
  dcorrection(correction, deg, o.r, n, o.dhl, o.dhr);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) o.V1[i + n * j] = jacobi_p(o.r[i], 0.0, 0.0, j);
  if (!invert(n, o.V1, o.iV1)) return false;
  for (int l = 0; l < n; ++l)
    for (int k = 0; k < n; ++k) o.wp[k + n * l] = o.w[k] * o.w[l];
  // nodes of the quad: node q = k + n*l sits at (r_k, r_l); mode sk = a*n + b is P_a(r) P_b(s)
  const int np = n * n;
  if (np > kMaxNp2D) {  // the 2D spaces stop at degree 5
    std::memset(o.V, 0, sizeof(o.V));
    std::memset(o.iV, 0, sizeof(o.iV));
    return true;
  }
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b)
      for (int l = 0; l < n; ++l)
        for (int k = 0; k < n; ++k)
          o.V[(k + n * l) + np * (a * n + b)] = jacobi_p(o.r[k], 0.0, 0.0, a) * jacobi_p(o.r[l], 0.0, 0.0, b);
  return invert(np, o.V, o.iV);
}

}  // namespace fr

// ====================================================================== C ABI: operator entry points
using namespace fr;

extern "C" {

const char *fr_last_error(void) { return g_error.c_str(); }

const char *fr_version(void) { return "frhip 0.1.0 (gfx950)"; }

int fr_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int fr_legendre_point(int deg, double *r, double *w) {
  FR_REQUIRE(deg >= 0 && deg < 64 && r, FR_E_INVALID, "fr_legendre_point: bad arguments");
  std::vector<double> ww(deg + 1);
  gauss_legendre(deg + 1, r, ww.data());
  if (w) std::memcpy(w, ww.data(), sizeof(double) * (deg + 1));
  return FR_OK;
}

int fr_lagrange_point(const double *sp, int nsp, double x, double *l) {
  FR_REQUIRE(sp && l && nsp > 0, FR_E_INVALID, "fr_lagrange_point: bad arguments");
  lagrange_point(sp, nsp, x, l);
  return FR_OK;
}

int fr_dlagrange(const double *sp, int nsp, double *lpdm) {
  FR_REQUIRE(sp && lpdm && nsp > 0, FR_E_INVALID, "fr_dlagrange: bad arguments");
  dlagrange(sp, nsp, lpdm);
  return FR_OK;
}

int fr_dcorrection(int kind, int deg, const double *x, int n, double *dgl, double *dgr) {
  FR_REQUIRE(kind >= FR_CORR_RADAU && kind <= FR_CORR_HUYNH, FR_E_INVALID, "fr_dcorrection: unknown kind");
  FR_REQUIRE(deg >= 1 && x && dgl && dgr && n > 0, FR_E_INVALID, "fr_dcorrection: bad arguments");
  dcorrection(kind, deg, x, n, dgl, dgr);
  return FR_OK;
}

int fr_jacobi_p(const double *x, int n, double alpha, double beta, int N, double *out) {
  FR_REQUIRE(x && out && n >= 0 && N >= 0, FR_E_INVALID, "fr_jacobi_p: bad arguments");
  for (int i = 0; i < n; ++i) out[i] = jacobi_p(x[i], alpha, beta, N);
  return FR_OK;
}

int fr_djacobi_p(const double *x, int n, double alpha, double beta, int N, double *out) {
  FR_REQUIRE(x && out && n >= 0 && N >= 0, FR_E_INVALID, "fr_djacobi_p: bad arguments");
  for (int i = 0; i < n; ++i) out[i] = djacobi_p(x[i], alpha, beta, N);
  return FR_OK;
}

int fr_vandermonde_1d(int N, const double *r, int n, double *V) {
  FR_REQUIRE(r && V && N >= 0 && n > 0, FR_E_INVALID, "fr_vandermonde_1d: bad arguments");
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i < n; ++i) V[i + (size_t)n * j] = jacobi_p(r[i], 0.0, 0.0, j);
  return FR_OK;
}

int fr_dvandermonde_1d(int N, const double *r, int n, double *Vr) {
  FR_REQUIRE(r && Vr && N >= 0 && n > 0, FR_E_INVALID, "fr_dvandermonde_1d: bad arguments");
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i < n; ++i) Vr[i + (size_t)n * j] = djacobi_p(r[i], 0.0, 0.0, j);
  return FR_OK;
}

int fr_vandermonde_quad(int N, const double *r, const double *s, int n, double *V) {
  FR_REQUIRE(r && s && V && N >= 0 && n > 0, FR_E_INVALID, "fr_vandermonde_quad: bad arguments");
  int sk = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N; ++j, ++sk)
      for (int q = 0; q < n; ++q) V[q + (size_t)n * sk] = jacobi_p(r[q], 0.0, 0.0, i) * jacobi_p(s[q], 0.0, 0.0, j);
  return FR_OK;
}

int fr_inv(int n, const double *A, double *Ainv) {
  FR_REQUIRE(A && Ainv && n > 0, FR_E_INVALID, "fr_inv: bad arguments");
  FR_REQUIRE(invert(n, A, Ainv), FR_E_INVALID, "fr_inv: singular matrix");
  return FR_OK;
}

int fr_rs_jacobi_quad(double r, double s, const double *vert, double *J) {
  FR_REQUIRE(vert && J, FR_E_INVALID, "fr_rs_jacobi_quad: bad arguments");
  for (int d = 0; d < 2; ++d) {
    const double v1 = vert[0 + 4 * d], v2 = vert[1 + 4 * d], v3 = vert[2 + 4 * d], v4 = vert[3 + 4 * d];
    J[d] = (s - 1.0) * v1 / 4 + (1.0 - s) * v2 / 4 + (s + 1.0) * v3 / 4 - (s + 1.0) * v4 / 4;
    J[d + 2] = (r - 1.0) * v1 / 4 - (r + 1.0) * v2 / 4 + (r + 1.0) * v3 / 4 + (1.0 - r) * v4 / 4;
  }
  return FR_OK;
}

int fr_rs_jacobi_mesh(int ncx, int ncy, const double *vert, int n, const double *r, double *J, double *iJ) {
  FR_REQUIRE(ncx > 0 && ncy > 0 && n > 0 && vert && r && J, FR_E_INVALID, "fr_rs_jacobi_mesh: bad arguments");
  const size_t cells = (size_t)ncx * ncy, pts = cells * n * n;
  for (int l = 0; l < n; ++l)
    for (int k = 0; k < n; ++k)
      for (size_t c = 0; c < cells; ++c) {
        double v[8];
        for (int q = 0; q < 8; ++q) v[q] = vert[c + cells * q];  // vertex q % 4, coordinate q / 4
        double m[4];
        fr_rs_jacobi_quad(r[k], r[l], v, m);
        const size_t at = c + cells * ((size_t)k + n * l);
        for (int e = 0; e < 4; ++e) J[at + pts * e] = m[e];
        if (iJ) {  // inverse of [a b; c d] = [d -b; -c a] / (ad - bc)
          const double a = m[0], cc = m[1], b = m[2], d = m[3], det = a * d - b * cc;
          iJ[at] = d / det;
          iJ[at + pts] = -cc / det;
          iJ[at + 2 * pts] = -b / det;
          iJ[at + 3 * pts] = a / det;
        }
      }
  return FR_OK;
}

int fr_shock_detector(double Se, int deg, double S0, double kappa) {
  (void)deg;
  double sigma;
  if (Se < S0 - kappa)
    sigma = 1.0;
  else if (Se < S0 + kappa)
    sigma = 0.5 * (1.0 - std::sin(0.5 * M_PI * (Se - S0) / kappa));
  else
    sigma = 0.0;
  return sigma < 0.99 ? 1 : 0;
}

}  // extern "C"
