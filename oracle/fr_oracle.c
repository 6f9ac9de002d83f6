/*
 * fr_oracle.c -- CPU FP64 oracle (TEST INFRASTRUCTURE ONLY, see fr_oracle.h).
 *
 * Each function cites the reference file:line it restates.  Paths are relative
 * to the root of vavrines/FluxReconstruction.jl.  KitBase routines (not in the
 * reference tree) are restated from their published algorithms and marked so.
 */
#include "fr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* =====================================================================
 * Polynomials
 * ===================================================================== */

/* Legendre P_n(x) by the three-term recurrence. */
double or_legendre_p(int n, double x) {
  if (n == 0) return 1.0;
  double p0 = 1.0, p1 = x;
  for (int k = 1; k < n; ++k) {
    double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

/* P_n'(x) -- restates src/Polynomial/poly_legendre.jl:13 (GSL sf_legendre_Pl_deriv_array, last
 * entry) with the identity P'_{k+1} = x P'_k + (k+1) P_k, valid on the closed interval.
 * Pinned by tests: numpy.polynomial.legendre.legder/legval. */
double or_dlegendre(int n, double x) {
  if (n <= 0) return 0.0;
  double p = 1.0, dp = 0.0; /* P_0, P_0' */
  double pm1 = 0.0;
  for (int k = 0; k < n; ++k) {
    double dnext = x * dp + (k + 1.0) * p;
    double pnext = (k == 0) ? x : ((2.0 * k + 1.0) * x * p - k * pm1) / (k + 1.0);
    pm1 = p;
    p = pnext;
    dp = dnext;
  }
  return dp;
}

/* Gauss-Legendre nodes (ascending) and weights: src/Polynomial/poly_legendre.jl:6 and
 * src/struct.jl:47,178 (FastGaussQuadrature.gausslegendre).  Newton iteration on P_n from the
 * Chebyshev-like initial guess; the positive half is computed and mirrored so the node set is
 * exactly symmetric.  Pinned by tests: numpy.polynomial.legendre.leggauss + closed forms. */
void or_gauss_legendre(int n, double *x, double *w) {
  if (n == 1) {
    x[0] = 0.0;
    w[0] = 2.0;
    return;
  }
  int half = (n + 1) / 2;
  for (int i = 0; i < half; ++i) {
    double z = cos(PI * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 1; k < n; ++k) {
        double p2 = ((2.0 * k + 1.0) * z * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    {
      double p0 = 1.0, p1 = z;
      for (int k = 1; k < n; ++k) {
        double p2 = ((2.0 * k + 1.0) * z * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
    }
    double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    if ((n % 2 == 1) && i == half - 1) z = 0.0;
    x[n - 1 - i] = z;
    x[i] = -z;
    w[n - 1 - i] = wi;
    w[i] = wi;
  }
}

/* Correction-function derivatives at x: src/Polynomial/poly_legendre.jl:29-37 (radau),
 * 44-54 (sd), 61-71 (huynh).  Radau pinned by tests: g_R(1)=1, g_R(-1)=0 through Gauss
 * quadrature of dgr, mirror symmetry, and the p=3 values. */
void or_dcorrection(int kind, int p, const double *x, int nx, double *dgl, double *dgr) {
  double sgn = (p % 2 == 0) ? 1.0 : -1.0; /* (-1)^p */
  for (int i = 0; i < nx; ++i) {
    double D = or_dlegendre(p, x[i]);
    double Dp = or_dlegendre(p + 1, x[i]);
    double y;
    if (kind == OR_CORR_RADAU) {
      y = Dp;
    } else {
      double Dm = or_dlegendre(p - 1, x[i]);
      if (kind == OR_CORR_SD)
        y = (p * Dm + (p + 1) * Dp) / (2.0 * p + 1.0);
      else
        y = ((p + 1) * Dm + p * Dp) / (2.0 * p + 1.0);
    }
    dgl[i] = sgn * 0.5 * (D - y);
    dgr[i] = 0.5 * (D + y);
  }
}

/* Lagrange basis at x: src/Polynomial/poly_lagrange.jl:6-21. */
void or_lagrange_point(const double *sp, int nsp, double x, double *l) {
  for (int k = 0; k < nsp; ++k) {
    double tmp = 1.0;
    for (int j = 0; j < nsp; ++j)
      if (j != k) tmp *= (x - sp[j]) / (sp[k] - sp[j]);
    l[k] = tmp;
  }
}

/* lpdm(m,k) = l_k'(sp_m), column-major: src/Polynomial/poly_lagrange.jl:38-59. */
void or_dlagrange(const double *sp, int nsp, double *lpdm) {
  for (int k = 0; k < nsp; ++k)
    for (int m = 0; m < nsp; ++m) {
      double lsum = 0.0;
      for (int l = 0; l < nsp; ++l) {
        double tmp = 1.0;
        for (int j = 0; j < nsp; ++j)
          if (j != k && j != l) tmp *= (sp[m] - sp[j]) / (sp[k] - sp[j]);
        if (l != k) lsum += tmp / (sp[k] - sp[l]);
      }
      lpdm[m + nsp * k] = lsum;
    }
}

/* Orthonormal Jacobi polynomial: src/Polynomial/poly_jacobi.jl:7-42. */
double or_jacobi_p(double x, double alpha, double beta, int N) {
  double PL[64];
  double gamma0 = pow(2.0, alpha + beta + 1.0) / (alpha + beta + 1.0) * tgamma(alpha + 1.0) *
                  tgamma(beta + 1.0) / tgamma(alpha + beta + 1.0);
  PL[0] = 1.0 / sqrt(gamma0);
  if (N == 0) return PL[0];
  double gamma1 = (alpha + 1.0) * (beta + 1.0) / (alpha + beta + 3.0) * gamma0;
  PL[1] = ((alpha + beta + 2.0) * x / 2.0 + (alpha - beta) / 2.0) / sqrt(gamma1);
  if (N == 1) return PL[1];
  double aold = 2.0 / (2.0 + alpha + beta) * sqrt((alpha + 1.0) * (beta + 1.0) / (alpha + beta + 3.0));
  for (int i = 1; i <= N - 1; ++i) {
    double h1 = 2.0 * i + alpha + beta;
    double anew = 2.0 / (h1 + 2.0) *
                  sqrt((i + 1.0) * (i + 1.0 + alpha + beta) * (i + 1.0 + alpha) * (i + 1.0 + beta) /
                       (h1 + 1.0) / (h1 + 3.0));
    double bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2.0);
    PL[i + 1] = 1.0 / anew * (-aold * PL[i - 1] + (x - bnew) * PL[i]);
    aold = anew;
  }
  return PL[N];
}

/* src/Polynomial/poly_jacobi.jl:88-95. */
double or_djacobi_p(double x, double alpha, double beta, int N) {
  if (N == 0) return 0.0;
  return sqrt(N * (N + alpha + beta + 1.0)) * or_jacobi_p(x, alpha + 1.0, beta + 1.0, N - 1);
}

/* src/Transform/transform.jl:18-26; V(i, j) column-major nr x (N+1). */
void or_vandermonde_1d(int N, const double *r, int nr, double *V) {
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i < nr; ++i) V[i + nr * j] = or_jacobi_p(r[i], 0.0, 0.0, j);
}

/* src/Transform/transform.jl:76-84. */
void or_dvandermonde_1d(int N, const double *r, int nr, double *Vr) {
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i < nr; ++i) Vr[i + nr * j] = or_djacobi_p(r[i], 0.0, 0.0, j);
}

/* src/Transform/transform.jl:55-68: mode sk = i*(N+1) + j holds P_i(r) P_j(s). */
void or_vandermonde_quad(int N, const double *r, const double *s, int n, double *V) {
  int sk = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N; ++j, ++sk)
      for (int q = 0; q < n; ++q)
        V[q + n * sk] = or_jacobi_p(r[q], 0.0, 0.0, i) * or_jacobi_p(s[q], 0.0, 0.0, j);
}

/* Dense inverse by Gauss-Jordan with partial pivoting (stands in for LinearAlgebra.inv,
 * src/struct.jl:54,140,205).  Column-major. */
int or_inv(int n, const double *A, double *Ainv) {
  double *M = (double *)malloc(sizeof(double) * n * 2 * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      M[i * 2 * n + j] = A[i + n * j];
      M[i * 2 * n + n + j] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(M[r * 2 * n + c]) > fabs(M[piv * 2 * n + c])) piv = r;
    if (M[piv * 2 * n + c] == 0.0) {
      free(M);
      return -1;
    }
    if (piv != c)
      for (int j = 0; j < 2 * n; ++j) {
        double t = M[c * 2 * n + j];
        M[c * 2 * n + j] = M[piv * 2 * n + j];
        M[piv * 2 * n + j] = t;
      }
    double d = M[c * 2 * n + c];
    for (int j = 0; j < 2 * n; ++j) M[c * 2 * n + j] /= d;
    for (int r = 0; r < n; ++r)
      if (r != c) {
        double f = M[r * 2 * n + c];
        if (f != 0.0)
          for (int j = 0; j < 2 * n; ++j) M[r * 2 * n + j] -= f * M[c * 2 * n + j];
      }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Ainv[i + n * j] = M[i * 2 * n + n + j];
  free(M);
  return 0;
}

/* x = A \ b (column-major A). */
int or_solve(int n, const double *A, const double *b, double *x) {
  double *Ai = (double *)malloc(sizeof(double) * n * n);
  if (or_inv(n, A, Ai)) {
    free(Ai);
    return -1;
  }
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += Ai[i + n * j] * b[j];
    x[i] = s;
  }
  free(Ai);
  return 0;
}

/* Bilinear quadrilateral Jacobian [xr xs; yr ys]: src/Geometry/geo_jacobi.jl:77-88.
 * vert(v, d) column-major 4x2, J column-major 2x2. */
void or_rs_jacobi_quad(double r, double s, const double *vert, double *J) {
  for (int d = 0; d < 2; ++d) {
    const double v1 = vert[0 + 4 * d], v2 = vert[1 + 4 * d], v3 = vert[2 + 4 * d], v4 = vert[3 + 4 * d];
    double dr = (s - 1.0) * v1 / 4 + (1.0 - s) * v2 / 4 + (s + 1.0) * v3 / 4 - (s + 1.0) * v4 / 4;
    double ds = (r - 1.0) * v1 / 4 - (r + 1.0) * v2 / 4 + (r + 1.0) * v3 / 4 + (1.0 - r) * v4 / 4;
    J[d + 2 * 0] = dr; /* xr (d=0), yr (d=1) */
    J[d + 2 * 1] = ds; /* xs, ys */
  }
}

/* =====================================================================
 * KitBase physics (parity unpinned against KitBase itself; restated from the
 * published KitBase 0.9 formulas; pinned by consistency / rotation / upwind tests)
 * ===================================================================== */

/* prim = [rho, U, V, lambda], lambda = rho / (2 p) */
void or_conserve_prim2(const double *w, double g, double *prim) {
  prim[0] = w[0];
  prim[1] = w[1] / w[0];
  prim[2] = w[2] / w[0];
  prim[3] = 0.5 * w[0] / (g - 1.0) / (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]);
}

void or_prim_conserve2(const double *prim, double g, double *w) {
  w[0] = prim[0];
  w[1] = prim[0] * prim[1];
  w[2] = prim[0] * prim[2];
  w[3] = 0.5 * prim[0] / prim[3] / (g - 1.0) + 0.5 * prim[0] * (prim[1] * prim[1] + prim[2] * prim[2]);
}

void or_conserve_prim1(const double *w, double g, double *prim) {
  prim[0] = w[0];
  prim[1] = w[1] / w[0];
  prim[2] = 0.5 * w[0] / (g - 1.0) / (w[2] - 0.5 * w[1] * w[1] / w[0]);
}

void or_prim_conserve1(const double *prim, double g, double *w) {
  w[0] = prim[0];
  w[1] = prim[0] * prim[1];
  w[2] = 0.5 * prim[0] / prim[2] / (g - 1.0) + 0.5 * prim[0] * prim[1] * prim[1];
}

void or_euler_flux2(const double *w, double g, double *F, double *G) {
  double prim[4];
  or_conserve_prim2(w, g, prim);
  double p = 0.5 * w[0] / prim[3];
  F[0] = w[1];
  F[1] = w[1] * w[1] / w[0] + p;
  F[2] = w[1] * w[2] / w[0];
  F[3] = (w[3] + p) * w[1] / w[0];
  G[0] = w[2];
  G[1] = w[2] * w[1] / w[0];
  G[2] = w[2] * w[2] / w[0] + p;
  G[3] = (w[3] + p) * w[2] / w[0];
}

void or_euler_flux1(const double *w, double g, double *F) {
  double prim[3];
  or_conserve_prim1(w, g, prim);
  double p = 0.5 * w[0] / prim[2];
  F[0] = w[1];
  F[1] = w[1] * w[1] / w[0] + p;
  F[2] = (w[2] + p) * w[1] / w[0];
}

/* local_frame / global_frame (KitBase), 2D state [rho, mx, my, E], n = (c, s). */
void or_local_frame2(const double *w, double c, double s, double *out) {
  out[0] = w[0];
  out[1] = w[1] * c + w[2] * s;
  out[2] = w[2] * c - w[1] * s;
  out[3] = w[3];
}
void or_global_frame2(const double *w, double c, double s, double *out) {
  out[0] = w[0];
  out[1] = w[1] * c - w[2] * s;
  out[2] = w[1] * s + w[2] * c;
  out[3] = w[3];
}

/* Generic x-direction Riemann flux on nv-variable Euler states (nv = 3 in 1D, 4 in 2D).
 * HLL with Davis wave-speed bounds (KitBase flux_hll!, called with dt = len = 1 at
 * example/shock-vortex.jl:78,86 and src/Equation/eq_euler.jl:53); Rusanov (local Lax-Friedrichs);
 * Roe (Toro ch.11, Roe-averaged eigen-decomposition, no entropy fix). */
static void or_riemann_nv(int nv, int kind, const double *wL, const double *wR, double g, double *fw) {
  double FL[4], FR[4], GL[4], GR[4];
  double uL, uR, cL, cR, pL, pR;
  if (nv == 4) {
    double pl[4], pr[4];
    or_conserve_prim2(wL, g, pl);
    or_conserve_prim2(wR, g, pr);
    or_euler_flux2(wL, g, FL, GL);
    or_euler_flux2(wR, g, FR, GR);
    uL = pl[1];
    uR = pr[1];
    cL = sqrt(0.5 * g / pl[3]);
    cR = sqrt(0.5 * g / pr[3]);
    pL = 0.5 * pl[0] / pl[3];
    pR = 0.5 * pr[0] / pr[3];
  } else {
    double pl[3], pr[3];
    or_conserve_prim1(wL, g, pl);
    or_conserve_prim1(wR, g, pr);
    or_euler_flux1(wL, g, FL);
    or_euler_flux1(wR, g, FR);
    uL = pl[1];
    uR = pr[1];
    cL = sqrt(0.5 * g / pl[2]);
    cR = sqrt(0.5 * g / pr[2]);
    pL = 0.5 * pl[0] / pl[2];
    pR = 0.5 * pr[0] / pr[2];
  }
  if (kind == OR_FLUX_HLL) {
    double lmin = fmin(uL - cL, uR - cR);
    double lmax = fmax(uL + cL, uR + cR);
    if (lmin >= 0.0) {
      for (int m = 0; m < nv; ++m) fw[m] = FL[m];
    } else if (lmax <= 0.0) {
      for (int m = 0; m < nv; ++m) fw[m] = FR[m];
    } else {
      double factor = 1.0 / (lmax - lmin);
      for (int m = 0; m < nv; ++m)
        fw[m] = factor * (lmax * FL[m] - lmin * FR[m] + lmax * lmin * (wR[m] - wL[m]));
    }
  } else if (kind == OR_FLUX_RUSANOV) {
    double smax = fmax(fabs(uL) + cL, fabs(uR) + cR);
    for (int m = 0; m < nv; ++m) fw[m] = 0.5 * (FL[m] + FR[m]) - 0.5 * smax * (wR[m] - wL[m]);
  } else {
    /* Roe */
    double sL = sqrt(wL[0]), sR = sqrt(wR[0]);
    double vL = (nv == 4) ? wL[2] / wL[0] : 0.0, vR = (nv == 4) ? wR[2] / wR[0] : 0.0;
    double EL = wL[nv - 1], ER = wR[nv - 1];
    double HL = (EL + pL) / wL[0], HR = (ER + pR) / wR[0];
    double ws = 1.0 / (sL + sR);
    double ut = (sL * uL + sR * uR) * ws;
    double vt = (sL * vL + sR * vR) * ws;
    double Ht = (sL * HL + sR * HR) * ws;
    double q2 = ut * ut + vt * vt;
    double a2 = (g - 1.0) * (Ht - 0.5 * q2);
    double at = sqrt(a2);
    double dr = wR[0] - wL[0], dp = pR - pL, du = uR - uL, dv = vR - vL;
    double rt = sL * sR;
    double al1 = (dp - rt * at * du) / (2.0 * a2);
    double al5 = (dp + rt * at * du) / (2.0 * a2);
    double al2 = dr - dp / a2;
    double al3 = rt * dv;
    double l1 = fabs(ut - at), l2 = fabs(ut), l5 = fabs(ut + at);
    double diss[4];
    if (nv == 4) {
      diss[0] = l1 * al1 + l2 * al2 + l5 * al5;
      diss[1] = l1 * al1 * (ut - at) + l2 * al2 * ut + l5 * al5 * (ut + at);
      diss[2] = l1 * al1 * vt + l2 * (al2 * vt + al3) + l5 * al5 * vt;
      diss[3] = l1 * al1 * (Ht - ut * at) + l2 * (al2 * 0.5 * q2 + al3 * vt) + l5 * al5 * (Ht + ut * at);
    } else {
      diss[0] = l1 * al1 + l2 * al2 + l5 * al5;
      diss[1] = l1 * al1 * (ut - at) + l2 * al2 * ut + l5 * al5 * (ut + at);
      diss[2] = l1 * al1 * (Ht - ut * at) + l2 * al2 * 0.5 * q2 + l5 * al5 * (Ht + ut * at);
    }
    for (int m = 0; m < nv; ++m) fw[m] = 0.5 * (FL[m] + FR[m]) - 0.5 * diss[m];
  }
}

void or_riemann2(int kind, const double *wL, const double *wR, double g, double *fw) {
  or_riemann_nv(4, kind, wL, wR, g, fw);
}
void or_riemann1(int kind, const double *wL, const double *wR, double g, double *fw) {
  or_riemann_nv(3, kind, wL, wR, g, fw);
}

/* =====================================================================
 * 2D Euler FR residual -- restates example/shock-vortex.jl:26-118 (loop structure, face
 * numbering 1=bottom 2=right 3=top 4=left, f contravariant through iJ) with
 * example/euler2d_wave.jl:35-107 as the non-threaded twin.  OR_BC_PERIODIC first fills the
 * ghost layer from the opposite interior (example/euler2d_wave.jl:125-135 without the
 * momentum sign flips) and then runs the same ghost-reading residual.
 * ===================================================================== */

size_t or_fr2d_len(const or_fr2d *d) {
  return (size_t)(d->nx + 2) * (d->ny + 2) * d->nsp * d->nsp * 4;
}

#define U5(a, i, j, k, l, m) \
  (a)[(size_t)(i) + NX2 * ((size_t)(j) + NY2 * ((size_t)(k) + NSP * ((size_t)(l) + NSP * (size_t)(m))))]

void or_fr2d_fill_periodic(const or_fr2d *d, double *u) {
  const size_t NX2 = d->nx + 2, NY2 = d->ny + 2, NSP = d->nsp;
  const int nx = d->nx, ny = d->ny;
  for (int m = 0; m < 4; ++m)
    for (int l = 0; l < (int)NSP; ++l)
      for (int k = 0; k < (int)NSP; ++k) {
        for (int j = 1; j <= ny; ++j) {
          U5(u, 0, j, k, l, m) = U5(u, nx, j, k, l, m);
          U5(u, nx + 1, j, k, l, m) = U5(u, 1, j, k, l, m);
        }
        for (int i = 0; i <= nx + 1; ++i) {
          U5(u, i, 0, k, l, m) = U5(u, i, ny, k, l, m);
          U5(u, i, ny + 1, k, l, m) = U5(u, i, 1, k, l, m);
        }
      }
}

static void or_fr2d_rhs_ghost(const or_fr2d *d, const double *u, double *du) {
  const int nx = d->nx, ny = d->ny, nsp = d->nsp;
  const size_t NX2 = nx + 2, NY2 = ny + 2, NSP = nsp;
  const double g = d->gamma;
  const double *ll = d->ll, *lr = d->lr, *lpdm = d->dl, *dhl = d->dhl, *dhr = d->dhr;
  const double iJ11 = d->iJ[0], iJ21 = d->iJ[1], iJ12 = d->iJ[2], iJ22 = d->iJ[3];
  const size_t ncell = NX2 * NY2;
  const size_t n = ncell * NSP * NSP * 4;

  /* f(i,j,k,l,m,n) ; u_face(i,j,face,l,m) ; f_face(i,j,face,l,m,n) */
  double *f = (double *)malloc(sizeof(double) * n * 2);
  double *u_face = (double *)malloc(sizeof(double) * ncell * 4 * NSP * 4);
  double *f_face = (double *)malloc(sizeof(double) * ncell * 4 * NSP * 4 * 2);
  double *fx = (double *)malloc(sizeof(double) * (nx + 1) * ny * NSP * 4);
  double *fy = (double *)malloc(sizeof(double) * nx * (ny + 1) * NSP * 4);
  double *rhs1 = (double *)malloc(sizeof(double) * nx * ny * NSP * NSP * 4);
  double *rhs2 = (double *)malloc(sizeof(double) * nx * ny * NSP * NSP * 4);

#define F6(i, j, k, l, m, c) f[(size_t)(i) + NX2 * ((size_t)(j) + NY2 * ((size_t)(k) + NSP * ((size_t)(l) + NSP * ((size_t)(m) + 4 * (size_t)(c)))))]
#define UF(i, j, fc, l, m) u_face[(size_t)(i) + NX2 * ((size_t)(j) + NY2 * ((size_t)(fc) + 4 * ((size_t)(l) + NSP * (size_t)(m))))]
#define FF(i, j, fc, l, m, c) f_face[(size_t)(i) + NX2 * ((size_t)(j) + NY2 * ((size_t)(fc) + 4 * ((size_t)(l) + NSP * ((size_t)(m) + 4 * (size_t)(c)))))]
#define FX(i, j, k, m) fx[(size_t)((i)-1) + (size_t)(nx + 1) * ((size_t)((j)-1) + (size_t)ny * ((size_t)(k) + NSP * (size_t)(m)))]
#define FY(i, j, k, m) fy[(size_t)((i)-1) + (size_t)nx * ((size_t)((j)-1) + (size_t)(ny + 1) * ((size_t)(k) + NSP * (size_t)(m)))]
#define R5(a, i, j, k, l, m) (a)[(size_t)((i)-1) + (size_t)nx * ((size_t)((j)-1) + (size_t)ny * ((size_t)(k) + NSP * ((size_t)(l) + NSP * (size_t)(m))))]

  /* du .= 0.0  (:27) */
  memset(du, 0, sizeof(double) * n);

  /* discontinuous flux, all cells incl. ghosts (:48-55) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int l = 0; l < nsp; ++l)
    for (int k = 0; k < nsp; ++k)
      for (int j = 0; j <= ny + 1; ++j)
        for (int i = 0; i <= nx + 1; ++i) {
          double w[4], Fg[4], Gg[4];
          for (int m = 0; m < 4; ++m) w[m] = U5(u, i, j, k, l, m);
          or_euler_flux2(w, g, Fg, Gg);
          /* f = iJ[i, j][k, l] * [fg, gg] (:50-53) */
          double a11 = iJ11, a21 = iJ21, a12 = iJ12, a22 = iJ22;
          if (d->iJp) {
            a11 = U5(d->iJp, i, j, k, l, 0);
            a21 = U5(d->iJp, i, j, k, l, 1);
            a12 = U5(d->iJp, i, j, k, l, 2);
            a22 = U5(d->iJp, i, j, k, l, 3);
          }
          for (int s = 0; s < 4; ++s) {
            F6(i, j, k, l, s, 0) = a11 * Fg[s] + a12 * Gg[s];
            F6(i, j, k, l, s, 1) = a21 * Fg[s] + a22 * Gg[s];
          }
        }

  /* face interpolation (:57-71) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < 4; ++m)
    for (int l = 0; l < nsp; ++l)
      for (int j = 0; j <= ny + 1; ++j)
        for (int i = 0; i <= nx + 1; ++i) {
          double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
          for (int q = 0; q < nsp; ++q) {
            s1 += U5(u, i, j, l, q, m) * ll[q];
            s2 += U5(u, i, j, q, l, m) * lr[q];
            s3 += U5(u, i, j, l, q, m) * lr[q];
            s4 += U5(u, i, j, q, l, m) * ll[q];
          }
          UF(i, j, 0, l, m) = s1;
          UF(i, j, 1, l, m) = s2;
          UF(i, j, 2, l, m) = s3;
          UF(i, j, 3, l, m) = s4;
          for (int c = 0; c < 2; ++c) {
            double t1 = 0, t2 = 0, t3 = 0, t4 = 0;
            for (int q = 0; q < nsp; ++q) {
              t1 += F6(i, j, l, q, m, c) * ll[q];
              t2 += F6(i, j, q, l, m, c) * lr[q];
              t3 += F6(i, j, l, q, m, c) * lr[q];
              t4 += F6(i, j, q, l, m, c) * ll[q];
            }
            FF(i, j, 0, l, m, c) = t1;
            FF(i, j, 1, l, m, c) = t2;
            FF(i, j, 2, l, m, c) = t3;
            FF(i, j, 3, l, m, c) = t4;
          }
        }

  /* face normals and the contravariant scale of the common fluxes: shock-vortex.jl uses
   * n1 = (1, 0), n2 = (0, 1) with iJ11 / iJ22 (:73-115); parallelogram.jl rotates both face
   * families and scales by (iJ n1)_1, (iJ n2)_2 (dev/parallelogram.jl:116-162) */
  const double c1 = d->geom ? d->n1[0] : 1.0, s1 = d->geom ? d->n1[1] : 0.0;
  const double c2 = d->geom ? d->n2[0] : 0.0, s2 = d->geom ? d->n2[1] : 1.0;
  const double q1 = d->geom ? iJ11 * c1 + iJ12 * s1 : iJ11;
  const double q2 = d->geom ? iJ21 * c2 + iJ22 * s2 : iJ22;

  /* x-faces: flux between (i-1, j) right trace and (i, j) left trace (:73-80) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nsp; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx + 1; ++i) {
        double a[4], b[4], wL[4], wR[4], fw[4], fg[4];
        for (int m = 0; m < 4; ++m) {
          a[m] = UF(i - 1, j, 1, k, m);
          b[m] = UF(i, j, 3, k, m);
        }
        if (d->geom) {
          or_local_frame2(a, c1, s1, wL);
          or_local_frame2(b, c1, s1, wR);
          or_riemann2(d->flux, wL, wR, g, fw);
          or_global_frame2(fw, c1, s1, fg);
        } else {
          or_riemann2(d->flux, a, b, g, fg);
        }
        for (int m = 0; m < 4; ++m) FX(i, j, k, m) = fg[m];
      }
  /* y-faces in the face-normal frame n2 (:81-89) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nsp; ++k)
    for (int j = 1; j <= ny + 1; ++j)
      for (int i = 1; i <= nx; ++i) {
        double a[4], b[4], wL[4], wR[4], fw[4], fg[4];
        for (int m = 0; m < 4; ++m) {
          a[m] = UF(i, j - 1, 2, k, m);
          b[m] = UF(i, j, 0, k, m);
        }
        or_local_frame2(a, c2, s2, wL);
        or_local_frame2(b, c2, s2, wR);
        or_riemann2(d->flux, wL, wR, g, fw);
        or_global_frame2(fw, c2, s2, fg);
        for (int m = 0; m < 4; ++m) FY(i, j, k, m) = fg[m];
      }

  /* derivative-matrix application (:91-100) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < 4; ++m)
    for (int l = 0; l < nsp; ++l)
      for (int k = 0; k < nsp; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i) {
            double s1 = 0.0, s2 = 0.0;
            for (int q = 0; q < nsp; ++q) {
              s1 += F6(i, j, q, l, m, 0) * lpdm[k + NSP * q];
              s2 += F6(i, j, k, q, m, 1) * lpdm[l + NSP * q];
            }
            R5(rhs1, i, j, k, l, m) = s1;
            R5(rhs2, i, j, k, l, m) = s2;
          }

  /* correction assembly (:102-115); the common fluxes scaled by q1 = iJ11, q2 = iJ22 in the
   * shock-vortex form, by iJ[i, j][k, l][1, 1] and [2, 2] of the point with a point metric */
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < 4; ++m)
    for (int l = 0; l < nsp; ++l)
      for (int k = 0; k < nsp; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i) {
            const double p1 = d->iJp ? U5(d->iJp, i, j, k, l, 0) : q1;
            const double p2 = d->iJp ? U5(d->iJp, i, j, k, l, 3) : q2;
            U5(du, i, j, k, l, m) =
                -(R5(rhs1, i, j, k, l, m) + R5(rhs2, i, j, k, l, m) +
                  (FX(i, j, l, m) * p1 - FF(i, j, 3, l, m, 0)) * dhl[k] +
                  (FX(i + 1, j, l, m) * p1 - FF(i, j, 1, l, m, 0)) * dhr[k] +
                  (FY(i, j, k, m) * p2 - FF(i, j, 0, k, m, 1)) * dhl[l] +
                  (FY(i, j + 1, k, m) * p2 - FF(i, j, 2, k, m, 1)) * dhr[l]);
          }

  free(f);
  free(u_face);
  free(f_face);
  free(fx);
  free(fy);
  free(rhs1);
  free(rhs2);
#undef F6
#undef UF
#undef FF
#undef FX
#undef FY
#undef R5
}

void or_fr2d_rhs(const or_fr2d *d, const double *u, double *du) {
  if (d->bc == OR_BC_PERIODIC) {
    size_t n = or_fr2d_len(d);
    double *uc = (double *)malloc(sizeof(double) * n);
    memcpy(uc, u, sizeof(double) * n);
    or_fr2d_fill_periodic(d, uc);
    or_fr2d_rhs_ghost(d, uc, du);
    free(uc);
  } else {
    or_fr2d_rhs_ghost(d, u, du);
  }
}

/* One explicit stage out = a0*u0 + a1*us + b*L(us) on interior cells; ghost cells are copied
 * from us (L vanishes there, shock-vortex.jl:27).  This fixed expression is the stage
 * convention shared with the GPU stepper (Shu-Osher SSP-RK3, classical RK4, midpoint). */
void or_fr2d_stage(const or_fr2d *d, const double *u0, const double *us, double *out, double a0,
                   double a1, double b) {
  const int nx = d->nx, ny = d->ny;
  const size_t NX2 = nx + 2, NY2 = ny + 2, NSP = d->nsp;
  size_t n = or_fr2d_len(d);
  double *du = (double *)malloc(sizeof(double) * n);
  or_fr2d_rhs(d, us, du);
  for (size_t q = 0; q < n; ++q) {
    size_t e = q % (NX2 * NY2);
    int i = (int)(e % NX2), j = (int)(e / NX2);
    if (i >= 1 && i <= nx && j >= 1 && j <= ny)
      out[q] = (a0 != 0.0 ? a0 * u0[q] : 0.0) + a1 * us[q] + b * du[q];
    else
      out[q] = us[q];
  }
  (void)NSP;
  free(du);
}

/* Shu-Osher SSP-RK3:  u1 = u + dt L(u);  u2 = 3/4 u + 1/4 u1 + dt/4 L(u1);
 * u^{n+1} = 1/3 u + 2/3 u2 + 2 dt/3 L(u2).  Pinned by tests through the stability polynomial
 * 1 + z + z^2/2 + z^3/6 on the linear advection operator. */
void or_fr2d_ssprk3(const or_fr2d *d, double *u, double dt, int nsteps) {
  size_t n = or_fr2d_len(d);
  double *u1 = (double *)malloc(sizeof(double) * n), *u2 = (double *)malloc(sizeof(double) * n);
  for (int s = 0; s < nsteps; ++s) {
    or_fr2d_stage(d, u, u, u1, 0.0, 1.0, dt);
    or_fr2d_stage(d, u, u1, u2, 0.75, 0.25, 0.25 * dt);
    or_fr2d_stage(d, u, u2, u, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt);
  }
  if (d->bc == OR_BC_PERIODIC) or_fr2d_fill_periodic(d, u);
  free(u1);
  free(u2);
}

/* Classical RK4 with an accumulator: y2 = u + dt/2 k1, y3 = u + dt/2 k2, y4 = u + dt k3,
 * u^{n+1} = acc + dt/6 k4 where acc = u + dt/6 k1 + dt/3 k2 + dt/3 k3. */
void or_fr2d_rk4(const or_fr2d *d, double *u, double dt, int nsteps) {
  const int nx = d->nx, ny = d->ny;
  const size_t NX2 = nx + 2, NY2 = ny + 2;
  size_t n = or_fr2d_len(d);
  double *y = (double *)malloc(sizeof(double) * n), *y2 = (double *)malloc(sizeof(double) * n);
  double *acc = (double *)malloc(sizeof(double) * n), *k = (double *)malloc(sizeof(double) * n);
  const double cs[4] = {0.5, 0.5, 1.0, 0.0};
  const double bs[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  for (int s = 0; s < nsteps; ++s) {
    memcpy(y, u, sizeof(double) * n);
    memcpy(acc, u, sizeof(double) * n);
    for (int st = 0; st < 4; ++st) {
      or_fr2d_rhs(d, y, k);
      for (size_t q = 0; q < n; ++q) {
        size_t e = q % (NX2 * NY2);
        int i = (int)(e % NX2), j = (int)(e / NX2);
        int inner = (i >= 1 && i <= nx && j >= 1 && j <= ny);
        if (st < 3) {
          y2[q] = inner ? u[q] + (cs[st] * dt) * k[q] : u[q];
          acc[q] = inner ? acc[q] + (bs[st] * dt) * k[q] : acc[q];
        } else {
          y2[q] = inner ? acc[q] + (bs[st] * dt) * k[q] : u[q];
        }
      }
      double *t = y;
      y = y2;
      y2 = t;
    }
    memcpy(u, y, sizeof(double) * n);
  }
  if (d->bc == OR_BC_PERIODIC) or_fr2d_fill_periodic(d, u);
  free(y);
  free(y2);
  free(acc);
  free(k);
}

/* Explicit midpoint (the reference scripts' Midpoint(), shock-vortex.jl:294):
 * y = u + dt/2 L(u);  u^{n+1} = u + dt L(y). */
void or_fr2d_midpoint(const or_fr2d *d, double *u, double dt, int nsteps) {
  size_t n = or_fr2d_len(d);
  double *y = (double *)malloc(sizeof(double) * n);
  for (int s = 0; s < nsteps; ++s) {
    or_fr2d_stage(d, u, u, y, 0.0, 1.0, 0.5 * dt);
    or_fr2d_stage(d, u, y, u, 1.0, 0.0, dt);
  }
  if (d->bc == OR_BC_PERIODIC) or_fr2d_fill_periodic(d, u);
  free(y);
}

/* =====================================================================
 * Dissipation
 * ===================================================================== */

/* src/dissipation.jl:13-23.  Pinned by the known answers of test/runtests.jl:12-14. */
int or_shock_detector(double Se, int deg, double S0, double kappa) {
  (void)deg;
  double sigma;
  if (Se < S0 - kappa)
    sigma = 1.0;
  else if (S0 - kappa <= Se && Se < S0 + kappa)
    sigma = 0.5 * (1.0 - sin(0.5 * PI * (Se - S0) / kappa));
  else
    sigma = 0.0;
  return sigma < 0.99 ? 1 : 0;
}

/* KitBase filter_l2! on a vector (called as modal_filter!(u, lambda; filter=:l2) at
 * example/shock-vortex.jl:316 and example/euler_highlevel.jl:45): coefficient n (0-based)
 * is divided by 1 + lambda (n+1)^2 n^2 for n >= 1.  Parity unpinned (KitBase absent). */
void or_modal_filter_l2(double *u, int n, double lambda) {
  for (int i = 1; i < n; ++i) u[i] /= (1.0 + lambda * (double)(i + 1) * (i + 1) * (double)i * i);
}

/* KitBase filter_exp1d (src/Polynomial/poly_filter.jl:14): sigma_i = exp(-alpha ((i-Nc)/(N-Nc))^s)
 * for i >= Nc, alpha = -log(eps).  Parity unpinned (KitBase absent). */
void or_filter_exp1d(int N, int s, int Nc, double *diag) {
  double alpha = -log(2.220446049250313e-16);
  for (int i = 0; i <= N; ++i) diag[i] = 1.0;
  for (int i = Nc; i <= N; ++i) diag[i] = exp(-alpha * pow((double)(i - Nc) / (N - Nc), s));
}

/* filter_exp2d (src/Polynomial/poly_filter.jl:32-47): modes i = 0..N outer, j = 0..N-i inner */
void or_filter_exp2d(int N, int s, int Nc, double *diag) {
  double alpha = -log(2.220446049250313e-16);
  int sk = 0;
  for (int i = 0; i <= N; ++i)
    for (int j = 0; j <= N - i; ++j) {
      diag[sk] = 1.0;
      if (i + j >= Nc) diag[sk] = exp(-alpha * pow((double)(i + j - Nc) / (N - Nc), s));
      sk += 1;
    }
}

/* filter_exp (src/Polynomial/poly_filter.jl:11-24): F = V * diagm(filterdiag) * invV; returns -1
 * when nv is neither N+1 nor (N+1)(N+2)/2 (filterdiag undefined in the reference) */
int or_filter_exp(int N, int s, int Nc, const double *V, int nv, const double *invV, double *F) {
  double diag[1024];
  if (nv > 1024) return -1;
  if (nv == N + 1)
    or_filter_exp1d(N, s, Nc, diag);
  else if (nv == (N + 1) * (N + 2) / 2)
    or_filter_exp2d(N, s, Nc, diag);
  else
    return -1;
  for (int j = 0; j < nv; ++j)
    for (int i = 0; i < nv; ++i) {
      double a = 0.0;
      for (int k = 0; k < nv; ++k) a += V[i + nv * k] * diag[k] * invV[k + nv * j];
      F[i + nv * j] = a;
    }
  return 0;
}

/* Modal sensor + filter, example/shock-vortex.jl:308-321 generalised from the hard-coded 3x3
 * (deg = 2) slices to nsp x nsp: u_hat = iV * rho(:), su = u_hat[end]^2 / sum(u_hat^2),
 * shock_detector(log10(su), deg, S0, kappa); on shock every variable is filtered in modal
 * space and mapped back by V.  Loops over every cell including ghosts, as the reference. */
int or_fr2d_sensor_filter(const or_fr2d *d, const double *V, const double *iV, double *u,
                          double lambda, double S0, double kappa, int *flags) {
  const int nx = d->nx, ny = d->ny, nsp = d->nsp;
  const size_t NX2 = nx + 2, NY2 = ny + 2, NSP = nsp;
  const int np = nsp * nsp;
  int count = 0;
#pragma omp parallel for collapse(2) reduction(+ : count) schedule(static)
  for (int i = 0; i <= nx + 1; ++i)
    for (int j = 0; j <= ny + 1; ++j) {
      double nod[64], mod[64];
      for (int l = 0; l < nsp; ++l)
        for (int k = 0; k < nsp; ++k) nod[k + nsp * l] = U5(u, i, j, k, l, 0);
      double ss = 0.0;
      for (int a = 0; a < np; ++a) {
        double s = 0.0;
        for (int b = 0; b < np; ++b) s += iV[a + np * b] * nod[b];
        mod[a] = s;
      }
      for (int a = 0; a < np; ++a) ss += mod[a] * mod[a];
      double su = mod[np - 1] * mod[np - 1] / ss;
      int shock = or_shock_detector(log10(su), nsp - 1, S0, kappa);
      if (flags) flags[i + NX2 * j] = shock;
      if (shock) {
        ++count;
        for (int m = 0; m < 4; ++m) {
          for (int l = 0; l < nsp; ++l)
            for (int k = 0; k < nsp; ++k) nod[k + nsp * l] = U5(u, i, j, k, l, m);
          for (int a = 0; a < np; ++a) {
            double s = 0.0;
            for (int b = 0; b < np; ++b) s += iV[a + np * b] * nod[b];
            mod[a] = s;
          }
          or_modal_filter_l2(mod, np, lambda);
          for (int a = 0; a < np; ++a) {
            double s = 0.0;
            for (int b = 0; b < np; ++b) s += V[a + np * b] * mod[b];
            nod[a] = s;
          }
          for (int l = 0; l < nsp; ++l)
            for (int k = 0; k < nsp; ++k) U5(u, i, j, k, l, m) = nod[k + nsp * l];
        }
      }
    }
  return count;
}

/* The same L2 modal filter with a per-cell strength lam[i + (nx+2) j] (cells with lam <= 0 are
 * left unchanged): the filter step of the learned-sensor path. */
int or_fr2d_filter_cells(const or_fr2d *d, const double *V, const double *iV, double *u, const double *lam) {
  const int nx = d->nx, ny = d->ny, nsp = d->nsp;
  const size_t NX2 = nx + 2, NY2 = ny + 2, NSP = nsp;
  const int np = nsp * nsp;
  int count = 0;
  for (int i = 0; i <= nx + 1; ++i)
    for (int j = 0; j <= ny + 1; ++j) {
      const double l = lam[i + NX2 * j];
      if (!(l > 0.0)) continue;
      ++count;
      double nod[64], mod[64];
      for (int m = 0; m < 4; ++m) {
        for (int ll = 0; ll < nsp; ++ll)
          for (int k = 0; k < nsp; ++k) nod[k + nsp * ll] = U5(u, i, j, k, ll, m);
        for (int a = 0; a < np; ++a) {
          double s = 0.0;
          for (int b = 0; b < np; ++b) s += iV[a + np * b] * nod[b];
          mod[a] = s;
        }
        or_modal_filter_l2(mod, np, l);
        for (int a = 0; a < np; ++a) {
          double s = 0.0;
          for (int b = 0; b < np; ++b) s += V[a + np * b] * mod[b];
          nod[a] = s;
        }
        for (int ll = 0; ll < nsp; ++ll)
          for (int k = 0; k < nsp; ++k) U5(u, i, j, k, ll, m) = nod[k + nsp * ll];
      }
    }
  (void)NY2;
  return count;
}

/* tj_equation (src/dissipation.jl:208-215): p(t) - eps of the state t (u - ubar) + ubar.
 * The reference solves it by NewtonRaphson from t = 1 with tol 1e-9 (dissipation.jl:184,193);
 * here the root is taken in closed form: rho(t), m(t), E(t) are affine in t, so
 * rho(t) (p(t) - eps) = (g-1) (rho E - |m|^2/2) - eps rho is a quadratic in t.  The root
 * returned is the largest one in [0, 1] (the one Newton reaches from t = 1 when p(1) < eps
 * and p(0) > eps).  Parity vs NonlinearSolve unpinned (<= 1e-9 by its tolerance); pinned by
 * substitution back into tj_equation in the tests. */
static double or_tj_root(const double *ut, const double *um, double g, double eps) {
  double r0 = um[0], r1 = ut[0] - um[0];
  double x0 = um[1], x1 = ut[1] - um[1];
  double y0 = um[2], y1 = ut[2] - um[2];
  double e0 = um[3], e1 = ut[3] - um[3];
  /* q(t) = (g-1)(rho E - (mx^2+my^2)/2) - eps rho  = A t^2 + B t + C */
  double A = (g - 1.0) * (r1 * e1 - 0.5 * (x1 * x1 + y1 * y1));
  double B = (g - 1.0) * (r0 * e1 + r1 * e0 - (x0 * x1 + y0 * y1)) - eps * r1;
  double C = (g - 1.0) * (r0 * e0 - 0.5 * (x0 * x0 + y0 * y0)) - eps * r0;
  double best = 0.0;
  if (fabs(A) < 1e-300) {
    if (B != 0.0) {
      double t = -C / B;
      if (t >= 0.0 && t <= 1.0) best = t;
    }
    return best;
  }
  double disc = B * B - 4.0 * A * C;
  if (disc < 0.0) return 0.0;
  double sq = sqrt(disc);
  double qq = -0.5 * (B + (B >= 0 ? sq : -sq));
  double t1 = qq / A, t2 = (qq != 0.0) ? C / qq : t1;
  if (t1 >= 0.0 && t1 <= 1.0 && t1 > best) best = t1;
  if (t2 >= 0.0 && t2 <= 1.0 && t2 > best) best = t2;
  return best;
}

/* src/dissipation.jl:125-206 on one element ue(k, l, m) (column-major nsp x nsp x 4).
 * The reference's t2 = minimum(tj, t0) (line 199) raises a MethodError in Julia whenever the
 * pressure corrector triggers; the intended min(minimum(tj), t0) is used here. */
int or_positive_limiter2d(double *ue, int nsp, double g, const double *weights, const double *ll,
                          const double *lr) {
  const int np = nsp * nsp;
#define UE(k, l, m) ue[(k) + nsp * ((l) + nsp * (m))]
  double um[4];
  for (int m = 0; m < 4; ++m) {
    double s = 0.0;
    for (int l = 0; l < nsp; ++l)
      for (int k = 0; k < nsp; ++k) s += UE(k, l, m) * weights[k + nsp * l];
    um[m] = s;
  }
  double pm_prim[4];
  or_conserve_prim2(um, g, pm_prim);
  double t_mean = 1.0 / pm_prim[3];
  double p_mean = 0.5 * um[0] * t_mean;
  /* boundary values b(face, j, var) */
  double bv[4][16][4];
  for (int j = 0; j < nsp; ++j)
    for (int m = 0; m < 4; ++m) {
      double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
      for (int q = 0; q < nsp; ++q) {
        s1 += UE(j, q, m) * ll[q];
        s2 += UE(q, j, m) * lr[q];
        s3 += UE(j, q, m) * lr[q];
        s4 += UE(q, j, m) * ll[q];
      }
      bv[0][j][m] = s1;
      bv[1][j][m] = s2;
      bv[2][j][m] = s3;
      bv[3][j][m] = s4;
    }
  double eps = fmin(1e-13, fmin(um[0], p_mean));
  double rmin = INFINITY;
  for (int f = 0; f < 4; ++f)
    for (int j = 0; j < nsp; ++j) rmin = fmin(rmin, bv[f][j][0]);
  for (int q = 0; q < np; ++q) rmin = fmin(rmin, ue[q]);
  double t1 = fmin((um[0] - eps) / (um[0] - rmin + 1e-8), 1.0);
  if (!(t1 > 0.0 && t1 <= 1.0)) return -1;
  for (int l = 0; l < nsp; ++l)
    for (int k = 0; k < nsp; ++k) UE(k, l, 0) = t1 * (UE(k, l, 0) - um[0]) + um[0];
  /* pressure corrector: boundary values are the pre-density-fix ones, as in the reference */
  int have = 0;
  double tmin = INFINITY;
  for (int j = 0; j < nsp; ++j)
    for (int f = 0; f < 4; ++f) {
      double prim[4];
      or_conserve_prim2(bv[f][j], g, prim);
      if (prim[3] < eps) {
        double t = or_tj_root(bv[f][j], um, g, eps);
        have = 1;
        tmin = fmin(tmin, t);
      }
    }
  for (int l = 0; l < nsp; ++l)
    for (int k = 0; k < nsp; ++k) {
      double w[4], prim[4];
      for (int m = 0; m < 4; ++m) w[m] = UE(k, l, m);
      or_conserve_prim2(w, g, prim);
      if (prim[3] < eps) {
        double t = or_tj_root(w, um, g, eps);
        have = 1;
        tmin = fmin(tmin, t);
      }
    }
  if (have) {
    double t2 = fmin(tmin, 1.0);
    for (int m = 0; m < 4; ++m)
      for (int l = 0; l < nsp; ++l)
        for (int k = 0; k < nsp; ++k) UE(k, l, m) = t2 * (UE(k, l, m) - um[m]) + um[m];
  }
#undef UE
  return 0;
}

/* Limiter over the interior cells (example/shock-vortex.jl:298-303); returns the number of
 * elements whose density assert failed (left untouched). */
int or_fr2d_positivity(const or_fr2d *d, const double *wp4, double *u) {
  const int nx = d->nx, ny = d->ny, nsp = d->nsp;
  const size_t NX2 = nx + 2, NY2 = ny + 2, NSP = nsp;
  int bad = 0;
#pragma omp parallel for collapse(2) reduction(+ : bad) schedule(static)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      double ue[4 * 64];
      for (int m = 0; m < 4; ++m)
        for (int l = 0; l < nsp; ++l)
          for (int k = 0; k < nsp; ++k) ue[k + nsp * (l + nsp * m)] = U5(u, i, j, k, l, m);
      if (or_positive_limiter2d(ue, nsp, d->gamma, wp4, d->ll, d->lr)) {
        ++bad;
        continue;
      }
      for (int m = 0; m < 4; ++m)
        for (int l = 0; l < nsp; ++l)
          for (int k = 0; k < nsp; ++k) U5(u, i, j, k, l, m) = ue[k + nsp * (l + nsp * m)];
    }
  return bad;
}

/* src/dissipation.jl:38-59 (scalar vector version). */
int or_positive_limiter1d_scalar(double *u, int n, const double *weights, double wsum, const double *ll,
                                 const double *lr) {
  double s = 0.0, b0 = 0.0, b1 = 0.0, mn = INFINITY;
  for (int i = 0; i < n; ++i) s += u[i] * weights[i];
  double um = s / wsum;
  for (int i = 0; i < n; ++i) {
    b0 += u[i] * ll[i];
    b1 += u[i] * lr[i];
    mn = fmin(mn, u[i]);
  }
  double eps = fmin(1e-13, um);
  double theta = fmin(fmin(b0, b1), mn);
  double t = fmin((um - eps) / (um - theta + 1e-8), 1.0);
  if (!(t > 0.0 && t <= 1.0)) return -1;
  for (int i = 0; i < n; ++i) u[i] = t * (u[i] - um) + um;
  return 0;
}

/* src/dissipation.jl:61-123 on one 1D Euler element ue(p, k) (column-major nsp x 3): density
 * squeeze toward the mean (t1), then, where lambda < eps at the two flux points (pre-squeeze
 * traces, as in the reference) or at a solution point, the pressure squeeze t2 = min(tj, t0)
 * with tj the closed-form root of tj_equation (or_tj_root with the y momentum zero).  The
 * weights sum to one.  Returns -1 (element untouched) when the density assert fails. */
int or_positive_limiter1d_euler(double *ue, int nsp, double g, const double *weights, const double *ll,
                                const double *lr) {
#define UE(p, k) ue[(p) + nsp * (k)]
  double um[3];
  for (int k = 0; k < 3; ++k) {
    double s = 0.0;
    for (int p = 0; p < nsp; ++p) s += UE(p, k) * weights[p];
    um[k] = s;
  }
  double pm[3];
  or_conserve_prim1(um, g, pm);
  const double p_mean = 0.5 * um[0] / pm[2];
  double bv[2][3];
  for (int k = 0; k < 3; ++k) {
    double a = 0.0, b = 0.0;
    for (int p = 0; p < nsp; ++p) {
      a += UE(p, k) * ll[p];
      b += UE(p, k) * lr[p];
    }
    bv[0][k] = a;
    bv[1][k] = b;
  }
  const double eps = fmin(1e-13, fmin(um[0], p_mean));
  double rmin = fmin(bv[0][0], bv[1][0]);
  for (int p = 0; p < nsp; ++p) rmin = fmin(rmin, UE(p, 0));
  const double t1 = fmin((um[0] - eps) / (um[0] - rmin + 1e-8), 1.0);
  if (!(t1 > 0.0 && t1 <= 1.0)) return -1;
  for (int p = 0; p < nsp; ++p) UE(p, 0) = t1 * (UE(p, 0) - um[0]) + um[0];
  const double um4[4] = {um[0], um[1], 0.0, um[2]};
  int have = 0;
  double tmin = INFINITY;
  for (int f = 0; f < 2; ++f) {
    double prim[3];
    or_conserve_prim1(bv[f], g, prim);
    if (prim[2] < eps) {
      const double w4[4] = {bv[f][0], bv[f][1], 0.0, bv[f][2]};
      tmin = fmin(tmin, or_tj_root(w4, um4, g, eps));
      have = 1;
    }
  }
  for (int p = 0; p < nsp; ++p) {
    const double w[3] = {UE(p, 0), UE(p, 1), UE(p, 2)};
    double prim[3];
    or_conserve_prim1(w, g, prim);
    if (prim[2] < eps) {
      const double w4[4] = {w[0], w[1], 0.0, w[2]};
      tmin = fmin(tmin, or_tj_root(w4, um4, g, eps));
      have = 1;
    }
  }
  if (have) {
    const double t2 = fmin(tmin, 1.0);
    for (int k = 0; k < 3; ++k)
      for (int p = 0; p < nsp; ++p) UE(p, k) = t2 * (UE(p, k) - um[k]) + um[k];
  }
#undef UE
  return 0;
}

/* 1D modal sensor + filter of example/euler_highlevel.jl:37-52 on u(i, p, k): per cell
 * u_hat = iV rho, su = u_hat[end]^2 / (sum(u_hat^2) + 1e-6), shock_detector(log10(su), deg) with
 * S0 = -3 log10(deg), kappa = 4; on a shock every variable is filtered in modal space with
 * lambda = dt exp(0.875 deg) 0.15 (modal_filter!, :l2).  Returns the number of flagged cells. */
int or_fr1d_sensor_filter(int ncell, int nsp, const double *V, const double *iV, double dt, double *u) {
  const int deg = nsp - 1;
  const double S0 = -3.0 * log10((double)deg), kappa = 4.0;
  const double lambda = dt * exp(0.875 * deg) * 0.15;
  int count = 0;
  for (int i = 0; i < ncell; ++i) {
    double mod[16], nod[16];
    double num = 0.0, den = 0.0;
    for (int a = 0; a < nsp; ++a) {
      double s = 0.0;
      for (int b = 0; b < nsp; ++b) s += iV[a + nsp * b] * u[i + (size_t)ncell * b];
      mod[a] = s;
      den += s * s;
    }
    num = mod[nsp - 1] * mod[nsp - 1];
    const double su = num / (den + 1e-6);
    if (!or_shock_detector(log10(su), deg, S0, kappa)) continue;
    ++count;
    for (int k = 0; k < 3; ++k) {
      for (int a = 0; a < nsp; ++a) {
        double s = 0.0;
        for (int b = 0; b < nsp; ++b) s += iV[a + nsp * b] * u[i + (size_t)ncell * (b + nsp * k)];
        mod[a] = s;
      }
      or_modal_filter_l2(mod, nsp, lambda);
      for (int a = 0; a < nsp; ++a) {
        double s = 0.0;
        for (int b = 0; b < nsp; ++b) s += V[a + nsp * b] * mod[b];
        nod[a] = s;
      }
      for (int a = 0; a < nsp; ++a) u[i + (size_t)ncell * (a + nsp * k)] = nod[a];
    }
  }
  return count;
}

/* The limiter over every cell of a 1D Euler state u(i, p, k); returns the failed count. */
int or_fr1d_positivity(int ncell, int nsp, double g, const double *weights, const double *ll, const double *lr,
                       double *u) {
  int bad = 0;
  for (int i = 0; i < ncell; ++i) {
    double ue[3 * 16];
    for (int k = 0; k < 3; ++k)
      for (int p = 0; p < nsp; ++p) ue[p + nsp * k] = u[i + (size_t)ncell * (p + nsp * k)];
    if (or_positive_limiter1d_euler(ue, nsp, g, weights, ll, lr)) {
      ++bad;
      continue;
    }
    for (int k = 0; k < 3; ++k)
      for (int p = 0; p < nsp; ++p) u[i + (size_t)ncell * (p + nsp * k)] = ue[p + nsp * k];
  }
  return bad;
}

#undef U5

/* =====================================================================
 * 1D residuals
 * ===================================================================== */

/* src/Equation/eq_euler.jl:29-98 with layout u(i, p, k) = u[i + ncell*(p + nsp*k)].
 * dirichlet: du of the first and last cell zeroed (:75-78); period: both end faces from the
 * (ncell right, 1 left) pair (:80-98). */
void or_fr1d_euler_rhs(int ncell, int nsp, const double *J, const double *ll, const double *lr,
                       const double *dl, const double *dgl, const double *dgr, double g, int flux,
                       int bc, const double *u, double *du) {
#define A3(a, i, p, k) (a)[(size_t)(i) + (size_t)ncell * ((size_t)(p) + (size_t)nsp * (size_t)(k))]
  size_t n = (size_t)ncell * nsp * 3;
  double *f = (double *)malloc(sizeof(double) * n);
  double *uf = (double *)malloc(sizeof(double) * ncell * 2 * 3);
  double *ff = (double *)malloc(sizeof(double) * ncell * 2 * 3);
  double *fi = (double *)calloc((size_t)(ncell + 1) * 3, sizeof(double));
  double *rhs1 = (double *)malloc(sizeof(double) * n);
#define FC(a, i, s, k) (a)[(size_t)(i) + (size_t)ncell * ((size_t)(s) + 2 * (size_t)(k))]
  for (int j = 0; j < nsp; ++j)
    for (int i = 0; i < ncell; ++i) {
      double w[3], F[3];
      for (int k = 0; k < 3; ++k) w[k] = A3(u, i, j, k);
      or_euler_flux1(w, g, F);
      for (int k = 0; k < 3; ++k) A3(f, i, j, k) = F[k] / J[i];
    }
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < ncell; ++i) {
      double a = 0, b = 0, c = 0, e = 0;
      for (int q = 0; q < nsp; ++q) {
        a += A3(u, i, q, k) * ll[q];
        b += A3(f, i, q, k) * ll[q];
        c += A3(u, i, q, k) * lr[q];
        e += A3(f, i, q, k) * lr[q];
      }
      FC(uf, i, 0, k) = a;
      FC(ff, i, 0, k) = b;
      FC(uf, i, 1, k) = c;
      FC(ff, i, 1, k) = e;
    }
  for (int i = 1; i < ncell; ++i) {
    double wL[3], wR[3], fw[3];
    for (int k = 0; k < 3; ++k) {
      wL[k] = FC(uf, i - 1, 1, k);
      wR[k] = FC(uf, i, 0, k);
    }
    or_riemann1(flux, wL, wR, g, fw);
    for (int k = 0; k < 3; ++k) fi[i + (size_t)(ncell + 1) * k] = fw[k];
  }
  for (int k = 0; k < 3; ++k)
    for (int p = 0; p < nsp; ++p)
      for (int i = 0; i < ncell; ++i) {
        double s = 0.0;
        for (int q = 0; q < nsp; ++q) s += A3(f, i, q, k) * dl[p + (size_t)nsp * q];
        A3(rhs1, i, p, k) = s;
      }
  for (int i = 1; i < ncell - 1; ++i)
    for (int p = 0; p < nsp; ++p)
      for (int k = 0; k < 3; ++k)
        A3(du, i, p, k) = -(A3(rhs1, i, p, k) +
                            (fi[i + (size_t)(ncell + 1) * k] / J[i] - FC(ff, i, 0, k)) * dgl[p] +
                            (fi[i + 1 + (size_t)(ncell + 1) * k] / J[i] - FC(ff, i, 1, k)) * dgr[p]);
  if (bc == OR_BC1_DIRICHLET) {
    for (int p = 0; p < nsp; ++p)
      for (int k = 0; k < 3; ++k) {
        A3(du, 0, p, k) = 0.0;
        A3(du, ncell - 1, p, k) = 0.0;
      }
  } else {
    double wL[3], wR[3], fw[3];
    for (int k = 0; k < 3; ++k) {
      wL[k] = FC(uf, ncell - 1, 1, k);
      wR[k] = FC(uf, 0, 0, k);
    }
    or_riemann1(flux, wL, wR, g, fw);
    for (int k = 0; k < 3; ++k) {
      fi[0 + (size_t)(ncell + 1) * k] = fw[k];
      fi[ncell + (size_t)(ncell + 1) * k] = fw[k];
    }
    int ends[2] = {0, ncell - 1};
    for (int e = 0; e < 2; ++e) {
      int i = ends[e];
      for (int p = 0; p < nsp; ++p)
        for (int k = 0; k < 3; ++k)
          A3(du, i, p, k) = -(A3(rhs1, i, p, k) +
                              (fi[i + (size_t)(ncell + 1) * k] / J[i] - FC(ff, i, 0, k)) * dgl[p] +
                              (fi[i + 1 + (size_t)(ncell + 1) * k] / J[i] - FC(ff, i, 1, k)) * dgr[p]);
    }
  }
  free(f);
  free(uf);
  free(ff);
  free(fi);
  free(rhs1);
#undef A3
#undef FC
}

/* src/Equation/eq_advection.jl:153-175, 201-273 and src/Equation/eq_scalar.jl:1-12,
 * layout u(i, p) = u[i + ncell*p]; the regularisations +1e-8 (interior) and +1e-6 (periodic
 * face) are kept. */
void or_fr1d_advection_rhs(int ncell, int nsp, const double *J, const double *ll, const double *lr,
                           const double *dl, const double *dgl, const double *dgr, double a, int bc,
                           const double *u, double *du) {
  size_t n = (size_t)ncell * nsp;
  double *f = (double *)malloc(sizeof(double) * n);
  double *uf = (double *)malloc(sizeof(double) * ncell * 2);
  double *ff = (double *)malloc(sizeof(double) * ncell * 2);
  double *fi = (double *)calloc((size_t)ncell + 1, sizeof(double));
  double *rhs1 = (double *)malloc(sizeof(double) * n);
  for (int j = 0; j < nsp; ++j)
    for (int i = 0; i < ncell; ++i) f[i + (size_t)ncell * j] = a * u[i + (size_t)ncell * j] / J[i];
  for (int i = 0; i < ncell; ++i) {
    double s0 = 0, s1 = 0, t0 = 0, t1 = 0;
    for (int q = 0; q < nsp; ++q) {
      s0 += u[i + (size_t)ncell * q] * ll[q];
      s1 += u[i + (size_t)ncell * q] * lr[q];
      t0 += f[i + (size_t)ncell * q] * ll[q];
      t1 += f[i + (size_t)ncell * q] * lr[q];
    }
    uf[i] = s0;
    uf[i + ncell] = s1;
    ff[i] = t0;
    ff[i + ncell] = t1;
  }
  for (int i = 1; i < ncell; ++i) {
    double au = (ff[i] - ff[i - 1 + ncell]) / (uf[i] - uf[i - 1 + ncell] + 1e-8);
    fi[i] = 0.5 * (ff[i] + ff[i - 1 + ncell]) - 0.5 * fabs(au) * (uf[i] - uf[i - 1 + ncell]);
  }
  for (int p = 0; p < nsp; ++p)
    for (int i = 0; i < ncell; ++i) {
      double s = 0.0;
      for (int q = 0; q < nsp; ++q) s += f[i + (size_t)ncell * q] * dl[p + (size_t)nsp * q];
      rhs1[i + (size_t)ncell * p] = s;
    }
  for (int p = 0; p < nsp; ++p)
    for (int i = 1; i < ncell - 1; ++i)
      du[i + (size_t)ncell * p] = -(rhs1[i + (size_t)ncell * p] + (fi[i] - ff[i]) * dgl[p] +
                                    (fi[i + 1] - ff[i + ncell]) * dgr[p]);
  if (bc == OR_BC1_DIRICHLET) {
    for (int p = 0; p < nsp; ++p) {
      du[0 + (size_t)ncell * p] = 0.0;
      du[ncell - 1 + (size_t)ncell * p] = 0.0;
    }
  } else {
    double au = (ff[0] - ff[ncell - 1 + ncell]) / (uf[0] - uf[ncell - 1 + ncell] + 1e-6);
    fi[0] = 0.5 * (ff[ncell - 1 + ncell] + ff[0]) - 0.5 * fabs(au) * (uf[0] - uf[ncell - 1 + ncell]);
    fi[ncell] = fi[0];
    for (int p = 0; p < nsp; ++p) {
      int ends[2] = {0, ncell - 1};
      for (int e = 0; e < 2; ++e) {
        int i = ends[e];
        du[i + (size_t)ncell * p] = -(rhs1[i + (size_t)ncell * p] + (fi[i] - ff[i]) * dgl[p] +
                                      (fi[i + 1] - ff[i + ncell]) * dgr[p]);
      }
    }
  }
  free(f);
  free(uf);
  free(ff);
  free(fi);
  free(rhs1);
}

/* =====================================================================
 * 2D discrete-velocity BGK -- example/bgk_wave.jl:69-129 tensorised to quads: moments and
 * Maxwellian per solution point (:77-81, KitBase moments_conserve / conserve_prim / maxwellian in
 * 2D), f_flux = v f / J per direction (:85-88), face interpolation (:99-101), upwind interface
 * flux through the Heaviside blend (:103-107), derivative (:114-116) and correction plus the
 * relaxation (M - f) / tau (:122-128).  Periodic in both directions.
 * ===================================================================== */
size_t or_bgk2d_len(const or_bgk2d *d) {
  return (size_t)d->nx * d->ny * d->nsp * d->nsp * d->nu * d->nv;
}

void or_bgk2d_moments(const or_bgk2d *d, const double *f, double *w) {
  const size_t npts = (size_t)d->nx * d->ny * d->nsp * d->nsp;
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < npts; ++p) {
    double m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    for (int b = 0; b < d->nv; ++b)
      for (int a = 0; a < d->nu; ++a) {
        const double fv = f[p + npts * ((size_t)a + (size_t)d->nu * b)];
        const double wt = d->wu[a] * d->wv[b];
        m0 += wt * fv;
        m1 += wt * d->uv[a] * fv;
        m2 += wt * d->vv[b] * fv;
        m3 += wt * 0.5 * (d->uv[a] * d->uv[a] + d->vv[b] * d->vv[b]) * fv;
      }
    w[p] = m0;
    w[p + npts] = m1;
    w[p + 2 * npts] = m2;
    w[p + 3 * npts] = m3;
  }
}

void or_bgk2d_rhs(const or_bgk2d *d, const double *f, double *df) {
  const int nx = d->nx, ny = d->ny, n = d->nsp;
  const size_t npts = (size_t)nx * ny * n * n;
  const size_t ncell = (size_t)nx * ny;
  double *w = (double *)malloc(sizeof(double) * npts * 4);
  or_bgk2d_moments(d, f, w);
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < d->nv; ++b)
    for (int a = 0; a < d->nu; ++a) {
      const double ua = d->uv[a], vb = d->vv[b];
      const double *fv = f + npts * ((size_t)a + (size_t)d->nu * b);
      double *dv = df + npts * ((size_t)a + (size_t)d->nu * b);
#define FV(arr, i, j, k, l) (arr)[(size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)ny * ((size_t)(k) + (size_t)n * (size_t)(l)))]
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const int il = (i + nx - 1) % nx, ir = (i + 1) % nx, jb = (j + ny - 1) % ny, jt = (j + 1) % ny;
          for (int l = 0; l < n; ++l)
            for (int k = 0; k < n; ++k) {
              /* x direction, row l: flux traces of this cell and of its neighbours */
              double dfx = 0, fL = 0, fR = 0, nbR = 0, nbL = 0;
              for (int q = 0; q < n; ++q) {
                const double g = ua * FV(fv, i, j, q, l) * d->iJx;
                dfx += g * d->dl[k + n * q];
                fL += g * d->ll[q];
                fR += g * d->lr[q];
                nbR += ua * FV(fv, il, j, q, l) * d->iJx * d->lr[q]; /* right trace of the left cell */
                nbL += ua * FV(fv, ir, j, q, l) * d->iJx * d->ll[q]; /* left trace of the right cell */
              }
              const double dx_pos = ua > 0.0 ? 1.0 : 0.0; /* heaviside */
              const double FsL = fL * (1.0 - dx_pos) + nbR * dx_pos;
              const double FsR = nbL * (1.0 - dx_pos) + fR * dx_pos;
              double dfy = 0, gB = 0, gT = 0, nbT = 0, nbB = 0;
              for (int q = 0; q < n; ++q) {
                const double g = vb * FV(fv, i, j, k, q) * d->iJy;
                dfy += g * d->dl[l + n * q];
                gB += g * d->ll[q];
                gT += g * d->lr[q];
                nbT += vb * FV(fv, i, jb, k, q) * d->iJy * d->lr[q];
                nbB += vb * FV(fv, i, jt, k, q) * d->iJy * d->ll[q];
              }
              const double dy_pos = vb > 0.0 ? 1.0 : 0.0;
              const double GsB = gB * (1.0 - dy_pos) + nbT * dy_pos;
              const double GsT = nbB * (1.0 - dy_pos) + gT * dy_pos;
              const size_t pt = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * ((size_t)k + (size_t)n * (size_t)l));
              double wp[4] = {w[pt], w[pt + npts], w[pt + 2 * npts], w[pt + 3 * npts]}, prim[4];
              or_conserve_prim2(wp, d->gamma, prim);
              const double du = ua - prim[1], dvv = vb - prim[2];
              const double M = prim[0] * (prim[3] / PI) * exp(-prim[3] * (du * du + dvv * dvv));
              const double fc = FV(fv, i, j, k, l);
              FV(dv, i, j, k, l) = -(dfx + (FsL - fL) * d->dhl[k] + (FsR - fR) * d->dhr[k] + dfy +
                                     (GsB - gB) * d->dhl[l] + (GsT - gT) * d->dhr[l]) +
                                   (M - fc) / d->tau;
            }
        }
#undef FV
    }
  (void)ncell;
  free(w);
}

/* =====================================================================
 * 1D discrete-velocity BGK -- example/bgk_wave.jl:69-129 as the script writes it: per cell and
 * point the moments over the velocities (KitBase moments_conserve: sum w f, sum w u f,
 * 1/2 sum w u^2 f), the primitive variables at gamma (conserve_prim: rho, U,
 * lambda = rho / (2 (gamma - 1) (E - rho U^2 / 2))) and the Maxwellian
 * rho sqrt(lambda / pi) exp(-lambda (u - U)^2) (:76-81); f = u_j u / J (:84-88); the face traces
 * f_face[., ., 1] = f . ll, f_face[., ., 2] = f . lr (interp_face!, :99-101); the interface value
 * of face f between its left cell L and right cell R, f_face[R, 1] (1 - delta) + f_face[L, 2]
 * delta with delta = heaviside(u_j) (:103-107, faces 1 and nface both join cell nx to cell 1);
 * the derivative (:114-116) and the correction plus (M - u) / tau (:122-128).
 * ===================================================================== */
void or_bgk1d_moments(const or_bgk1d *d, const double *f, double *w) {
  const size_t nx = d->nx, nu = d->nu, npts = nx * d->nsp;
  for (size_t k = 0; k < (size_t)d->nsp; ++k)
    for (size_t i = 0; i < nx; ++i) {
      double m0 = 0.0, m1 = 0.0, m2 = 0.0;
      for (size_t j = 0; j < nu; ++j) {
        const double fv = f[i + nx * (j + nu * k)];
        m0 += d->wv[j] * fv;
        m1 += d->wv[j] * d->uv[j] * fv;
        m2 += d->wv[j] * d->uv[j] * d->uv[j] * fv;
      }
      w[i + nx * k] = m0;
      w[i + nx * k + npts] = m1;
      w[i + nx * k + 2 * npts] = 0.5 * m2;
    }
}

void or_bgk1d_rhs(const or_bgk1d *d, const double *f, double *df) {
  const int nx = d->nx, nu = d->nu, n = d->nsp;
  const size_t npts = (size_t)nx * n, len = npts * nu;
  double *w = (double *)malloc(sizeof(double) * npts * 3);
  double *M = (double *)malloc(sizeof(double) * len);
  double *fl = (double *)malloc(sizeof(double) * len);   /* f = u_j u / J */
  double *face = (double *)malloc(sizeof(double) * (size_t)nx * nu * 2);
  double *fint = (double *)malloc(sizeof(double) * (size_t)(nx + 1) * nu);
#define A3(a, i, j, k) (a)[(size_t)(i) + (size_t)nx * ((size_t)(j) + (size_t)nu * (size_t)(k))]
  or_bgk1d_moments(d, f, w);
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < nx; ++i) {
      const size_t p = (size_t)i + (size_t)nx * k;
      const double rho = w[p], U = w[p + npts] / w[p];
      const double lam = 0.5 * w[p] / (d->gamma - 1.0) / (w[p + 2 * npts] - 0.5 * w[p + npts] * w[p + npts] / w[p]);
      for (int j = 0; j < nu; ++j) {
        const double c = d->uv[j] - U;
        A3(M, i, j, k) = rho * sqrt(lam / PI) * exp(-lam * (c * c));
      }
    }
  const double J = 0.5 * d->dx;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < nu; ++j)
      for (int i = 0; i < nx; ++i) A3(fl, i, j, k) = d->uv[j] * A3(f, i, j, k) / J;
  for (int j = 0; j < nu; ++j)
    for (int i = 0; i < nx; ++i) {
      double a = 0.0, b = 0.0;
      for (int k = 0; k < n; ++k) {
        a += A3(fl, i, j, k) * d->ll[k];
        b += A3(fl, i, j, k) * d->lr[k];
      }
      face[(size_t)i + (size_t)nx * ((size_t)j + (size_t)nu * 0)] = a;
      face[(size_t)i + (size_t)nx * ((size_t)j + (size_t)nu * 1)] = b;
    }
  /* faces 0..nx (0-based): right cell R = fi (fi < nx) or 0, left cell L = fi - 1 (fi > 0) or nx - 1 */
  for (int fi = 0; fi <= nx; ++fi) {
    const int R = fi < nx ? fi : 0, L = fi > 0 ? fi - 1 : nx - 1;
    for (int j = 0; j < nu; ++j) {
      const double delta = d->uv[j] > 0.0 ? 1.0 : 0.0;
      fint[(size_t)fi + (size_t)(nx + 1) * j] = face[(size_t)R + (size_t)nx * (size_t)j] * (1.0 - delta) +
                                                face[(size_t)L + (size_t)nx * ((size_t)j + (size_t)nu)] * delta;
    }
  }
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < nu; ++j)
      for (int i = 0; i < nx; ++i) {
        double r1 = 0.0;
        for (int q = 0; q < n; ++q) r1 += A3(fl, i, j, q) * d->dl[k + n * q];
        /* left face of cell i: face i (cell 0: face nx); right face: i + 1 (cell nx - 1: face 0) */
        const int fL = i > 0 ? i : nx, fR = i < nx - 1 ? i + 1 : 0;
        const double iL = fint[(size_t)fL + (size_t)(nx + 1) * j], iR = fint[(size_t)fR + (size_t)(nx + 1) * j];
        const double tl = face[(size_t)i + (size_t)nx * (size_t)j];
        const double tr = face[(size_t)i + (size_t)nx * ((size_t)j + (size_t)nu)];
        A3(df, i, j, k) = -(r1 + (iL - tl) * d->dhl[k] + (iR - tr) * d->dhr[k]) + (A3(M, i, j, k) - A3(f, i, j, k)) / d->tau;
      }
#undef A3
  free(w);
  free(M);
  free(fl);
  free(face);
  free(fint);
}

/* =====================================================================
 * Metrics: src/tools.jl:43-53
 * ===================================================================== */
double or_L1_error(const double *u, const double *ue, size_t n, double dx) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += fabs(u[i] - ue[i]) * dx;
  return s;
}
double or_L2_error(const double *u, const double *ue, size_t n, double dx) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double e = fabs(u[i] - ue[i]) * dx;
    s += e * e;
  }
  return sqrt(s);
}
double or_Linf_error(const double *u, const double *ue, size_t n, double dx) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s = fmax(s, fabs(u[i] - ue[i]) * dx);
  return s;
}

/* counter-based splitmix64 stream, identical in numpy (tests/np_reference.py) */
uint64_t or_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double or_uniform(uint64_t *state) { return (double)(or_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0); }
