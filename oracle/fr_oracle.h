/*
 * fr_oracle.h -- CPU FP64 oracle for the flux-reconstruction residual path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Julia package (vavrines/FluxReconstruction.jl) and of the KitBase routines it
 * calls.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the reference results / the timed CPU baseline.  The product
 * library (libfrhip.so) never links, loads or calls anything in this folder.
 *
 * Array conventions follow the reference: every multi-dimensional array is
 * Julia column-major.  The 2D state u(i, j, k, l, m) has axes
 * (0:nx+1, 0:ny+1, 1:nsp, 1:nsp, 1:nvar) (example/shock-vortex.jl:160) and is
 * stored at   i + (nx+2)*(j + (ny+2)*(k + nsp*(l + nsp*m)))   (0-based).
 * Matrices such as dl(m, k) sit at m + nsp*k.
 *
 * Parity status of each function (how it is pinned) is recorded next to it.
 * "parity unpinned" marks routines whose exact reference arithmetic lives in
 * KitBase / OrdinaryDiffEq / NonlinearSolve, which are absent from the
 * reference tree; those are restated from the published algorithm.
 */
#ifndef FR_ORACLE_H
#define FR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_FLUX_HLL = 0, OR_FLUX_RUSANOV = 1, OR_FLUX_ROE = 2 };
enum { OR_BC_GHOST = 0, OR_BC_PERIODIC = 1 };
enum { OR_CORR_RADAU = 0, OR_CORR_SD = 1, OR_CORR_HUYNH = 2 };

/* ---------------- polynomial / operator data (setup) ---------------- */
void or_gauss_legendre(int n, double *x, double *w);
double or_legendre_p(int n, double x);
double or_dlegendre(int n, double x);
void or_dcorrection(int kind, int p, const double *x, int nx, double *dgl, double *dgr);
void or_lagrange_point(const double *sp, int nsp, double x, double *l);
void or_dlagrange(const double *sp, int nsp, double *lpdm);
double or_jacobi_p(double x, double alpha, double beta, int N);
double or_djacobi_p(double x, double alpha, double beta, int N);
void or_vandermonde_1d(int N, const double *r, int nr, double *V);
void or_dvandermonde_1d(int N, const double *r, int nr, double *Vr);
void or_vandermonde_quad(int N, const double *r, const double *s, int n, double *V);
int or_inv(int n, const double *A, double *Ainv);
int or_solve(int n, const double *A, const double *b, double *x);
void or_rs_jacobi_quad(double r, double s, const double *vert /* 4x2 col-major */, double *J /*2x2 col-major*/);

/* ---------------- KitBase physics restated (2D: 4 vars, 1D: 3 vars) ---------------- */
void or_conserve_prim2(const double *w, double g, double *prim);
void or_prim_conserve2(const double *prim, double g, double *w);
void or_conserve_prim1(const double *w, double g, double *prim);
void or_prim_conserve1(const double *prim, double g, double *w);
void or_euler_flux2(const double *w, double g, double *F, double *G);
void or_euler_flux1(const double *w, double g, double *F);
void or_riemann2(int kind, const double *wL, const double *wR, double g, double *fw);
void or_riemann1(int kind, const double *wL, const double *wR, double g, double *fw);
void or_local_frame2(const double *w, double c, double s, double *out);
void or_global_frame2(const double *w, double c, double s, double *out);

/* ---------------- 2D Euler FR residual (example/shock-vortex.jl:26-118) ---------------- */
typedef struct {
  int nx, ny, nsp;
  double gamma;
  int flux; /* OR_FLUX_* */
  int bc;   /* OR_BC_* */
  double iJ[4]; /* uniform inverse Jacobian, 2x2 column-major: iJ11, iJ21, iJ12, iJ22 */
  const double *ll, *lr, *dl, *dhl, *dhr;
  /* geom 0: interface terms of example/shock-vortex.jl:73-115 (x faces unrotated, y faces in the
   * frame n = (0, 1), scaled by iJ11 / iJ22).  geom 1: those of dev/parallelogram.jl:116-162 (x
   * faces in the frame n1, y faces in the frame n2, scaled by (iJ n1)_1 / (iJ n2)_2); with
   * n1 = (1, 0), n2 = (0, 1) the two agree bitwise. */
  int geom;
  double n1[2], n2[2];
  /* optional point-varying metric (geom 0 only): iJ[i, j][k, l] of every solution point in the
   * state layout with the 2x2 entry e = a + 2b in place of the variable (r_x, s_x, r_y, s_y); NULL
   * takes the uniform iJ above.  src/struct.jl:134-142 builds it per point for any mesh, and
   * example/shock-vortex.jl:48-55,102-115 reads it per point. */
  const double *iJp;
} or_fr2d;

size_t or_fr2d_len(const or_fr2d *d);
void or_fr2d_fill_periodic(const or_fr2d *d, double *u);
void or_fr2d_rhs(const or_fr2d *d, const double *u, double *du);
/* generic stage: out = a0*u0 + a1*us + b*L(us) on interior cells, ghosts copied from us */
void or_fr2d_stage(const or_fr2d *d, const double *u0, const double *us, double *out, double a0,
                   double a1, double b);
void or_fr2d_ssprk3(const or_fr2d *d, double *u, double dt, int nsteps);
void or_fr2d_rk4(const or_fr2d *d, double *u, double dt, int nsteps);
void or_fr2d_midpoint(const or_fr2d *d, double *u, double dt, int nsteps);

/* ---------------- dissipation (src/dissipation.jl, shock-vortex.jl:296-331) ---------------- */
int or_shock_detector(double Se, int deg, double S0, double kappa);
void or_modal_filter_l2(double *u, int n, double lambda);
void or_filter_exp1d(int N, int s, int Nc, double *diag);
void or_filter_exp2d(int N, int s, int Nc, double *diag);
int or_filter_exp(int N, int s, int Nc, const double *V, int nv, const double *invV, double *F);
/* per-element modal sensor + L2 filter over every cell (ghosts included, as the reference loop);
   returns the number of filtered cells, writes the per-cell flag if flags != NULL */
int or_fr2d_sensor_filter(const or_fr2d *d, const double *V, const double *iV, double *u,
                          double lambda, double S0, double kappa, int *flags);
int or_fr2d_filter_cells(const or_fr2d *d, const double *V, const double *iV, double *u, const double *lam);
/* 2D tensor positivity limiter on one element u(k,l,m) (dissipation.jl:125-206).
   weights(k,l) as passed by the caller (the reference passes wp/4).  Returns 0, or -1 if the
   reference's @assert on t1 fails, with u untouched in that case. */
int or_positive_limiter2d(double *ue, int nsp, double g, const double *weights, const double *ll,
                          const double *lr);
int or_fr2d_positivity(const or_fr2d *d, const double *wp4, double *u);
int or_positive_limiter1d_euler(double *ue, int nsp, double g, const double *weights, const double *ll,
                                const double *lr);
int or_fr1d_sensor_filter(int ncell, int nsp, const double *V, const double *iV, double dt, double *u);
int or_fr1d_positivity(int ncell, int nsp, double g, const double *weights, const double *ll, const double *lr,
                       double *u);
/* src/dissipation.jl:38-59 on one scalar element; wsum = sum(weights) as Julia forms it (the sum of
 * the array, or the weight itself when the caller passes one scalar weight) */
int or_positive_limiter1d_scalar(double *u, int n, const double *weights, double wsum, const double *ll,
                                 const double *lr);

/* ---------------- 1D residuals (src/Equation/eq_euler.jl, eq_advection.jl, eq_scalar.jl) ------ */
enum { OR_BC1_DIRICHLET = 0, OR_BC1_PERIOD = 1 };
void or_fr1d_euler_rhs(int ncell, int nsp, const double *J, const double *ll, const double *lr,
                       const double *dl, const double *dgl, const double *dgr, double g, int flux,
                       int bc, const double *u, double *du);
void or_fr1d_advection_rhs(int ncell, int nsp, const double *J, const double *ll, const double *lr,
                           const double *dl, const double *dgl, const double *dgr, double a, int bc,
                           const double *u, double *du);

/* ---------------- 2D discrete-velocity BGK (example/bgk_wave.jl:69-129, tensorised) -----------
 * f(i, j, k, l, a, b): periodic nx x ny cells (no ghost layer), nsp x nsp points, nu x nv
 * velocities; column-major, f[i + nx*(j + ny*(k + nsp*(l + nsp*(a + nu*b))))]. */
typedef struct {
  int nx, ny, nsp, nu, nv;
  double iJx, iJy, tau, gamma;            /* 2/dx, 2/dy; relaxation time; gamma of conserve_prim */
  const double *uv, *vv, *wu, *wv;        /* abscissae and quadrature weights per direction */
  const double *ll, *lr, *dl, *dhl, *dhr;
} or_bgk2d;
size_t or_bgk2d_len(const or_bgk2d *d);
/* moments w = (rho, rho U, rho V, E) at every solution point: w[pt + npts*c] */
void or_bgk2d_moments(const or_bgk2d *d, const double *f, double *w);
void or_bgk2d_rhs(const or_bgk2d *d, const double *f, double *df);

/* ---------------- 1D discrete-velocity BGK (example/bgk_wave.jl:69-129 as written) ----------
 * u(i, j, k): periodic nx cells, nu velocities, nsp points; u[i + nx*(j + nu*k)]. */
typedef struct {
  int nx, nu, nsp;
  double dx, tau, gamma;                  /* cell size (J = dx/2); relaxation time; gamma (3 there) */
  const double *uv, *wv;                  /* velocities and weights (KitBase VSpace1D u, weights) */
  const double *ll, *lr, *dl, *dhl, *dhr; /* standard_lagrange, dradau */
} or_bgk1d;
/* conserved moments (rho, rho U, E) at every point p = i + nx*k: w[p + nx*nsp*c] */
void or_bgk1d_moments(const or_bgk1d *d, const double *f, double *w);
void or_bgk1d_rhs(const or_bgk1d *d, const double *f, double *df);

/* ---------------- metrics (src/tools.jl:43-53) ---------------- */
double or_L1_error(const double *u, const double *ue, size_t n, double dx);
double or_L2_error(const double *u, const double *ue, size_t n, double dx);
double or_Linf_error(const double *u, const double *ue, size_t n, double dx);

/* ---------------- workload generators used by parity tests ---------------- */
uint64_t or_splitmix64(uint64_t *state);
double or_uniform(uint64_t *state); /* [0,1) from the top 53 bits */

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
