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

// ------------------------------------------------------------------ one-element positivity limiter
// The reference's per-element host routine positive_limiter (src/dissipation.jl:37-215), for
// callers that limit one element at a time; the device passes (fr1d_positivity, fr2d_positivity)
// run the same steps over whole states.
namespace {

// largest t in [0, 1] with rho(t) (p(t) - eps) = 0, where the state moves on the segment from the
// mean um (t = 0) to ut (t = 1): with rho, m, E linear in t, rho p = (g - 1)(rho E - |m|^2 / 2)
// is a quadratic in t (the root the reference's Newton solve of tj_equation, dissipation.jl:208-215,
// finds from t = 1; 0 when there is none)
double squeeze_root(const double *ut, const double *um, int nv, double g, double eps) {
  const int ne = nv - 1;  // energy slot; momenta 1 .. ne-1
  const double r0 = um[0], r1 = ut[0] - um[0];
  const double e0 = um[ne], e1 = ut[ne] - um[ne];
  double mm00 = 0.0, mm01 = 0.0, mm11 = 0.0;
  for (int v = 1; v < ne; ++v) {
    const double a = um[v], b = ut[v] - um[v];
    mm00 += a * a;
    mm01 += a * b;
    mm11 += b * b;
  }
  const double A = (g - 1.0) * (r1 * e1 - 0.5 * mm11);
  const double B = (g - 1.0) * (r0 * e1 + r1 * e0 - mm01) - eps * r1;
  const double C = (g - 1.0) * (r0 * e0 - 0.5 * mm00) - eps * r0;
  double best = 0.0;
  if (std::fabs(A) < 1e-300) {
    if (B != 0.0) {
      const double t = -C / B;
      if (t >= 0.0 && t <= 1.0) best = t;
    }
    return best;
  }
  const double disc = B * B - 4.0 * A * C;
  if (disc < 0.0) return 0.0;
  const double q = -0.5 * (B + (B >= 0.0 ? std::sqrt(disc) : -std::sqrt(disc)));
  const double ta = q / A, tb = q != 0.0 ? C / q : ta;
  if (ta >= 0.0 && ta <= 1.0 && ta > best) best = ta;
  if (tb >= 0.0 && tb <= 1.0 && tb > best) best = tb;
  return best;
}

// KitBase conserve_prim's last entry, lambda = rho / (2 p), of a state (rho, m..., E)
double lambda_of(const double *w, int nv, double g) {
  double mm = 0.0;
  for (int v = 1; v < nv - 1; ++v) mm += w[v] * w[v];
  const double p = (g - 1.0) * (w[nv - 1] - 0.5 * mm / w[0]);
  return 0.5 * w[0] / p;
}

// the Euler forms (1D: np points, nv = 3; 2D: np = nsp^2 points, nv = 4) once the traces
// tr[nt][nv] of the element's faces are formed: density squeeze, then the pressure squeeze
int limit_euler(double *u, int np, int nv, double g, const double *w, const double *tr, int nt, double t0,
                const char *who) {
  double um[4] = {0.0, 0.0, 0.0, 0.0};
  for (int v = 0; v < nv; ++v)
    for (int q = 0; q < np; ++q) um[v] += u[q + np * v] * w[q];
  const double pm = 0.5 * um[0] / lambda_of(um, nv, g);
  const double eps = std::fmin(1e-13, std::fmin(um[0], pm));
  double rmin = INFINITY;
  for (int f = 0; f < nt; ++f) rmin = std::fmin(rmin, tr[f * nv]);
  for (int q = 0; q < np; ++q) rmin = std::fmin(rmin, u[q]);
  const double t1 = std::fmin((um[0] - eps) / (um[0] - rmin + 1e-8), 1.0);
  if (!(t1 > 0.0 && t1 <= 1.0)) return fail(FR_E_LIMITER, std::string(who) + ": density limiter parameter out of (0, 1]");
  for (int q = 0; q < np; ++q) u[q] = t1 * (u[q] - um[0]) + um[0];
  // pressure squeeze over the face traces (taken before the density squeeze, as the reference
  // forms them once) and the solution points
  bool any = false;
  double t2 = INFINITY;
  for (int f = 0; f < nt; ++f)
    if (lambda_of(tr + f * nv, nv, g) < eps) {
      any = true;
      t2 = std::fmin(t2, squeeze_root(tr + f * nv, um, nv, g, eps));
    }
  for (int q = 0; q < np; ++q) {
    double s[4];
    for (int v = 0; v < nv; ++v) s[v] = u[q + np * v];
    if (lambda_of(s, nv, g) < eps) {
      any = true;
      t2 = std::fmin(t2, squeeze_root(s, um, nv, g, eps));
    }
  }
  if (any) {
    // min(minimum(tj), t0): the reference's minimum(tj, t0) (dissipation.jl:116,199) is a
    // MethodError in Julia
    t2 = std::fmin(t2, t0);
    for (int v = 0; v < nv; ++v)
      for (int q = 0; q < np; ++q) u[q + np * v] = t2 * (u[q + np * v] - um[v]) + um[v];
  }
  return FR_OK;
}

}  // namespace

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

int fr_dlegendre(int p, const double *x, int n, double *out) {
  FR_REQUIRE(p >= 0 && n >= 0 && (n == 0 || (x && out)), FR_E_INVALID, "fr_dlegendre: bad arguments");
  for (int i = 0; i < n; ++i) out[i] = dlegendre(p, x[i]);
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

int fr_modal_filter_l2(double *u, int n, double lambda) {
  FR_REQUIRE(u && n >= 0 && lambda >= 0.0 && std::isfinite(lambda), FR_E_INVALID,
             "fr_modal_filter_l2: bad arguments");
  // mode k (0-based) of the modal vector divided by 1 + lambda (k+1)^2 k^2; mode 0 is kept
  for (int k = 1; k < n; ++k) {
    const double kk = (double)k * (double)(k + 1);
    u[k] /= 1.0 + lambda * kk * kk;
  }
  return FR_OK;
}

int fr_positive_limiter_scalar(double *u, int n, const double *weights, double wsum, const double *ll,
                               const double *lr) {
  FR_REQUIRE(u && weights && ll && lr && n >= 1 && wsum != 0.0, FR_E_INVALID, "fr_positive_limiter_scalar: bad arguments");
  double um = 0.0, bl = 0.0, br = 0.0, umin = INFINITY;
  for (int q = 0; q < n; ++q) {
    um += u[q] * weights[q];
    bl += u[q] * ll[q];
    br += u[q] * lr[q];
    umin = std::fmin(umin, u[q]);
  }
  um /= wsum;
  const double eps = std::fmin(1e-13, um);
  const double theta = std::fmin(std::fmin(bl, br), umin);
  const double t = std::fmin((um - eps) / (um - theta + 1e-8), 1.0);
  FR_REQUIRE(t > 0.0 && t <= 1.0, FR_E_LIMITER, "fr_positive_limiter_scalar: limiter parameter out of (0, 1]");
  for (int q = 0; q < n; ++q) u[q] = t * (u[q] - um) + um;
  return FR_OK;
}

int fr_positive_limiter_euler1d(double *u, int nsp, double gamma, const double *weights, const double *ll,
                                const double *lr, double t0) {
  FR_REQUIRE(u && weights && ll && lr && nsp >= 1 && nsp <= kMaxNp2D && gamma > 1.0, FR_E_INVALID,
             "fr_positive_limiter_euler1d: bad arguments");
  double tr[2][3];
  for (int v = 0; v < 3; ++v) {
    double a = 0.0, b = 0.0;
    for (int q = 0; q < nsp; ++q) {
      a += u[q + nsp * v] * ll[q];
      b += u[q + nsp * v] * lr[q];
    }
    tr[0][v] = a;
    tr[1][v] = b;
  }
  return limit_euler(u, nsp, 3, gamma, weights, &tr[0][0], 2, t0, "fr_positive_limiter_euler1d");
}

int fr_positive_limiter_euler2d(double *u, int nsp, double gamma, const double *weights, const double *ll,
                                const double *lr, double t0) {
  FR_REQUIRE(u && weights && ll && lr && nsp >= 1 && nsp <= 6 && gamma > 1.0, FR_E_INVALID,
             "fr_positive_limiter_euler2d: bad arguments");
  const int np = nsp * nsp;
  // traces of face f (1 bottom, 2 right, 3 top, 4 left) at face point j, every variable
  std::vector<double> tr(4 * nsp * 4);
  for (int j = 0; j < nsp; ++j)
    for (int v = 0; v < 4; ++v) {
      double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
      for (int q = 0; q < nsp; ++q) {
        s1 += u[j + nsp * q + np * v] * ll[q];
        s2 += u[q + nsp * j + np * v] * lr[q];
        s3 += u[j + nsp * q + np * v] * lr[q];
        s4 += u[q + nsp * j + np * v] * ll[q];
      }
      tr[(0 * nsp + j) * 4 + v] = s1;
      tr[(1 * nsp + j) * 4 + v] = s2;
      tr[(2 * nsp + j) * 4 + v] = s3;
      tr[(3 * nsp + j) * 4 + v] = s4;
    }
  return limit_euler(u, np, 4, gamma, weights, tr.data(), 4 * nsp, t0, "fr_positive_limiter_euler2d");
}

}  // extern "C"
