/*
 * fr.h -- C ABI of libfrhip.so, the MI355X (gfx950) flux-reconstruction solver core.
 *
 * Plain pointers and sizes only.  Every array uses the reference's Julia column-major layout, so
 * a Julia caller passes its arrays (or their OffsetArray parents) unchanged:
 *   2D state u(i, j, k, l, m), axes (0:nx+1, 0:ny+1, 1:nsp, 1:nsp, 1:4)
 *       -> u[i + (nx+2)*(j + (ny+2)*(k + nsp*(l + nsp*m)))]      (example/shock-vortex.jl:160)
 *   1D Euler state u(i, p, k), axes (1:ncell, 1:nsp, 1:3)        (src/Equation/eq_euler.jl:1)
 *   1D advection state u(i, p), axes (1:ncell, 1:nsp)            (src/Equation/eq_advection.jl:1)
 *   matrices such as dl(m, k) = l_k'(r_m)  -> dl[m + nsp*k]      (src/Polynomial/poly_lagrange.jl:109)
 *
 * Every call returns FR_OK (0) or an FR_E_* code; fr_last_error() gives the message of the
 * calling thread's last failure.  Functions with a _dev suffix take device pointers and run
 * asynchronously on the context's stream; all others take host pointers and return once done.
 * Calls on one context are serialised by the caller; distinct contexts are independent.
 */
#ifndef FR_H
#define FR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status / enums */
enum {
  FR_OK = 0,
  FR_E_INVALID = 1,        /* bad argument (null pointer, size, enum value) */
  FR_E_HIP = 2,            /* HIP runtime error */
  FR_E_NONFINITE = 3,      /* NaN/Inf found by a checked call */
  FR_E_NEGATIVE_STATE = 4, /* density or pressure <= 0 found by a checked call */
  FR_E_UNSUPPORTED = 5,    /* configuration not implemented (e.g. degree out of range) */
  FR_E_NOMEM = 6,
  FR_E_LIMITER = 7         /* positivity limiter could not bring an element in range */
};
enum { FR_FLUX_HLL = 0, FR_FLUX_RUSANOV = 1, FR_FLUX_ROE = 2 };
/* FR_BC_GHOST: the ghost layer of u is data (reference semantics: the host refreshes ghosts,
 * example/shock-vortex.jl:324-326).  FR_BC_PERIODIC: neighbours wrap inside the kernels. */
enum { FR_BC_GHOST = 0, FR_BC_PERIODIC = 1 };
enum { FR_CORR_RADAU = 0, FR_CORR_SD = 1, FR_CORR_HUYNH = 2 };
enum { FR_RK_SSPRK3 = 0, FR_RK_RK4 = 1, FR_RK_MIDPOINT = 2, FR_RK_EULER = 3, FR_RK_TSIT5 = 4 };
/* ghost refresh rules for fr2d_fill_ghost (host loops of the reference scripts) */
enum {
  FR_GHOST_PERIODIC = 0, /* copy the opposite interior row/column (euler2d_wave.jl:127-131, no sign flips) */
  FR_GHOST_OUTFLOW = 1,  /* copy the adjacent interior row/column on every side */
  FR_GHOST_INFLOW_LEFT = 2, /* shock-vortex.jl:322-325: bottom, top and right copy the adjacent row/column,
                               the left (supersonic inflow) column keeps its values */
  FR_GHOST_WAVE = 3, /* euler2d_wave.jl:127-134 (wave along x): FR_GHOST_PERIODIC with the y-momentum
                        (variable 3) of the bottom and top ghost rows, corners included, negated */
  FR_GHOST_WAVE_Y = 4 /* euler2d_wave.jl:153-160 (wave along y): FR_GHOST_PERIODIC with the x-momentum
                         (variable 2) of the left and right ghost columns, corners included, negated */
};
/* sensor modes for fr2d_sensor_filter */
enum { FR_SENSOR_OFF = 0, FR_SENSOR_MODAL = 1, FR_SENSOR_LEARNED = 2 };

const char *fr_last_error(void);
const char *fr_version(void);
/* number of visible HIP devices (0 on a machine without a GPU; never fails).  fr_last_error,
 * fr_version and fr_device_count have no reference counterpart: the status plumbing of a C ABI
 * in place of Julia exceptions. */
int fr_device_count(void);

/* ------------------------------------------------------------------ operator data (host, native)
 * These replace the reference's setup routines; outputs are caller-owned host arrays. */
/* Gauss-Legendre points/weights, ascending (src/Polynomial/poly_legendre.jl:6; struct.jl:47,178) */
int fr_legendre_point(int deg, double *r, double *w);
/* l_k(x) for solution points sp (src/Polynomial/poly_lagrange.jl:6-21) */
int fr_lagrange_point(const double *sp, int nsp, double x, double *l);
/* lpdm(m,k) = l_k'(sp_m) (src/Polynomial/poly_lagrange.jl:38-59) */
int fr_dlagrange(const double *sp, int nsp, double *lpdm);
/* the derivative of the Legendre polynomial P_p at each of the n points x (∂legendre,
 * src/Polynomial/poly_legendre.jl:13-22; the last entry of GSL's sf_legendre_Pl_deriv_array) */
int fr_dlegendre(int p, const double *x, int n, double *out);
/* correction-function derivatives dgl, dgr at x (src/Polynomial/poly_legendre.jl:29-71) */
int fr_dcorrection(int kind, int deg, const double *x, int n, double *dgl, double *dgr);
/* orthonormal Jacobi P_N^(a,b) at n points (src/Polynomial/poly_jacobi.jl:7-83,88-107) */
int fr_jacobi_p(const double *x, int n, double alpha, double beta, int N, double *out);
int fr_djacobi_p(const double *x, int n, double alpha, double beta, int N, double *out);
/* 1D Vandermonde n x (N+1) and its r-derivative (src/Transform/transform.jl:18-31,76-89) */
int fr_vandermonde_1d(int N, const double *r, int n, double *V);
int fr_dvandermonde_1d(int N, const double *r, int n, double *Vr);
/* quad tensor Vandermonde n x (N+1)^2, modes i-outer / j-inner (src/Transform/transform.jl:55-68) */
int fr_vandermonde_quad(int N, const double *r, const double *s, int n, double *V);
/* dense inverse (column-major): LinearAlgebra.inv as used for iV = inv(V) (src/struct.jl:54,205) */
int fr_inv(int n, const double *A, double *Ainv);
/* bilinear quad Jacobian [xr xs; yr ys] at (r,s), vertices 4x2 (src/Geometry/geo_jacobi.jl:77-88) */
int fr_rs_jacobi_quad(double r, double s, const double *vert, double *J);
/* the per-point Jacobians of an FRPSpace2D over any quad mesh (src/struct.jl:134-142,
 * J[i, j][k, l] = rs_jacobi(r[k], r[l], vertices[i, j]), iJ its inverse; geo_jacobi.jl:90-98):
 * vert is the mesh's (ncx, ncy, 4, 2) vertex array, J and iJ (NULL: not formed) are
 * (ncx, ncy, n, n, 2, 2), all column-major.  The 2x2 inverse is the closed form
 * [d -b; -c a] / (ad - bc). */
int fr_rs_jacobi_mesh(int ncx, int ncy, const double *vert, int n, const double *r, double *J, double *iJ);
/* Persson-Peraire switch; returns 1 (shock) / 0 (src/dissipation.jl:13-23) */
int fr_shock_detector(double Se, int deg, double S0, double kappa);
/* positive_limiter of src/dissipation.jl on one element, in place, on the host (the reference's
 * per-element routine; fr1d_positivity / fr2d_positivity run it over device states).  Returns
 * FR_E_LIMITER, leaving u unchanged, where the reference's @assert 0 < t <= 1 fails.
 * scalar (:37-58): u(n), um = sum(u w) / wsum (wsum = sum(weights), or the weight itself when
 * the reference passes one number);
 * euler1d (:61-123): u(nsp, 3) column-major, weights summing to one;
 * euler2d (:125-206): u(nsp, nsp, 4) column-major, weights(nsp, nsp) summing to one (wp / 4);
 * the pressure squeeze takes min(minimum(tj), t0) (the reference's minimum(tj, t0) raises a
 * MethodError) and the closed-form root of tj_equation (:208-215) in place of its Newton solve. */
int fr_positive_limiter_scalar(double *u, int n, const double *weights, double wsum, const double *ll,
                               const double *lr);
/* KitBase's modal_filter!(u, lambda; filter = :l2) on one modal vector u(n), in place, on the host
 * (called by example/shock-vortex.jl:316 and example/euler_highlevel.jl:45 on the modes iV * u of
 * an element): mode k (0-based) is divided by 1 + lambda (k+1)^2 k^2.  The same damping as the
 * device filters of fr2d_sensor_filter / fr1d_sensor_filter.  KitBase is not part of the
 * reference tree, so the coefficients follow its published form, not its code. */
int fr_modal_filter_l2(double *u, int n, double lambda);
int fr_positive_limiter_euler1d(double *u, int nsp, double gamma, const double *weights, const double *ll,
                                const double *lr, double t0);
int fr_positive_limiter_euler2d(double *u, int nsp, double gamma, const double *weights, const double *ll,
                                const double *lr, double t0);
/* filter_exp(N, s, V, Nc, invV) (src/Polynomial/poly_filter.jl:11-24): the nodal exponential
 * filter F = V diag(sigma) invV (nv x nv, column-major) with sigma of KitBase filter_exp1d for a
 * line basis (nv = N+1) or filter_exp2d (poly_filter.jl:32-47) for a triangle basis
 * (nv = (N+1)(N+2)/2); invV NULL: inverted here.  Requires 0 <= Nc < N, s >= 1.  Other nv return
 * FR_E_UNSUPPORTED (the reference leaves the filter undefined there).  Applied to device states
 * with fr_poly_derivative_dev (pdm = F). */
int fr_filter_exp(int N, int s, int Nc, const double *V, int nv, const double *invV, double *F);

/* ------------------------------------------------------------------ row contractions (device)
 * f(row, q, b) is a column-major (nrow, n, nbatch) device array, 1 <= n <= 16; the operators are
 * host arrays copied into the launch.  Asynchronous on `stream` (a hipStream_t, NULL for HIP's null
 * stream); input and output must not overlap. */
/* interp_face!(fδ, f, ll, lr) of src/interpolate.jl:1-10 on every row: fδ(row, 1, b) = f(row, :, b)·ll,
 * fδ(row, 2, b) = f(row, :, b)·lr, fδ column-major (nrow, 2, nbatch) */
int fr_interp_face_dev(const double *d_f, long long nrow, int n, int nbatch, const double *ll, const double *lr,
                       double *d_ff, void *stream);
/* poly_derivative!(df, f, pdm) of src/derivative.jl:1-10 on every row: df(row, i, b) =
 * f(row, :, b)·pdm(i, :), pdm n x n column-major, df column-major (nrow, n, nbatch) */
int fr_poly_derivative_dev(const double *d_f, long long nrow, int n, int nbatch, const double *pdm, double *d_df,
                           void *stream);

/* ------------------------------------------------------------------ 2D Euler on quads */
typedef struct fr2d_ctx fr2d_ctx;

typedef struct {
  int nx, ny;       /* interior cells; the state carries one ghost layer (ngx = ngy = 1) */
  int deg;          /* polynomial degree p, 1 <= p <= 5 */
  double gamma;     /* ratio of specific heats */
  int flux;         /* FR_FLUX_* */
  int bc;           /* FR_BC_* */
  double iJ11;      /* iJ[i,j][k,l][1,1] = 2/dx of FRPSpace2D (src/struct.jl:137-142) */
  double iJ22;      /* iJ[i,j][k,l][2,2] = 2/dy */
  /* Operator data as FRPSpace2D stores it (ps.ll, ps.lr, ps.dl, ps.dhl, ps.dhr, ps.V, ps.iV).
   * Any of them may be NULL, in which case the library builds the Gauss/Radau operators itself
   * (src/struct.jl:130-227). */
  const double *ll, *lr, *dl, *dhl, *dhr, *V, *iV;
} fr2d_desc;

/* Replaces FRPSpace2D + the scratch tuple p of example/shock-vortex.jl:120-153 (the context owns
 * all scratch and a device copy of the state for the stepping entry points). */
int fr2d_create(fr2d_ctx **out, const fr2d_desc *desc);
int fr2d_destroy(fr2d_ctx *ctx);
/* Run on a caller-provided hipStream_t (NULL is HIP's null stream), or on the context's own
 * non-blocking stream again with FR_OWN_STREAM (the default after fr2d_create).  The stream and
 * size queries have no reference counterpart (device plumbing of the C ABI). */
#define FR_OWN_STREAM ((void *)~(uintptr_t)0)
int fr2d_set_stream(fr2d_ctx *ctx, void *hip_stream);
void *fr2d_get_stream(fr2d_ctx *ctx);
/* number of doubles in one state array: (nx+2)(ny+2) nsp^2 4 */
size_t fr2d_state_len(const fr2d_ctx *ctx);

/* The drop-in dudt!(du, u, p, t) of example/shock-vortex.jl:26-118 (also euler2d_wave.jl:35-107):
 * du = L(u) on interior cells, du = 0 on the ghost layer (shock-vortex.jl:27). */
int fr2d_euler_rhs(fr2d_ctx *ctx, const double *u, double *du, double t);          /* host ptrs */
int fr2d_euler_rhs_dev(fr2d_ctx *ctx, const double *d_u, double *d_du, double t);  /* device ptrs */

/* Device-resident state for the integrator path (replaces ODEProblem/init/step!,
 * shock-vortex.jl:293-305). */
int fr2d_set_state(fr2d_ctx *ctx, const double *u);  /* host -> device */
int fr2d_get_state(fr2d_ctx *ctx, double *u);        /* device -> host (synchronises) */
int fr2d_set_state_dev(fr2d_ctx *ctx, const double *d_u);
int fr2d_get_state_dev(fr2d_ctx *ctx, double *d_u);
/* Device pointer of the state (or NULL).  The device-resident arrays keep the reference layout
 * with padded rows: element (i, j) of plane p is at i + pitch*j + plane*p, with pitch = nx+2
 * rounded up to a multiple of 16 and plane = pitch*(ny+2) (fr2d_state_pitch), which aligns the
 * 16-element row segments the kernels move to 128-byte lines.  fr2d_set_state(_dev) and
 * fr2d_get_state(_dev) convert from / to the unpadded reference layout.  No reference
 * counterpart (device interop). */
double *fr2d_state_dev(fr2d_ctx *ctx);
int fr2d_state_pitch(const fr2d_ctx *ctx, int *pitch, long long *plane);
/* nsteps explicit steps of size dt with scheme FR_RK_*, replacing step!(itg) of OrdinaryDiffEq
 * (shock-vortex.jl:294,305; Midpoint there, SSP-RK3 / RK4 / Euler / fixed-step Tsit5 here too);
 * each stage is one fused residual+update launch (Tsit5: a residual launch and an elementwise
 * combination launch).  In FR_BC_GHOST mode the ghost layer is held fixed (L = 0 there). */
int fr2d_step(fr2d_ctx *ctx, int scheme, double dt, int nsteps);
/* the launches of nsteps steps captured into a hipGraph and replayed nrep times on the context's
 * stream (asynchronous, like fr2d_step); the context keeps the instantiated graph and reuses it
 * while scheme, dt and nsteps are unchanged.  No reference counterpart (a replay of fr2d_step). */
int fr2d_step_graph(fr2d_ctx *ctx, int scheme, double dt, int nsteps, int nrep);
/* No reference counterpart (launch tuning of this library): the staggered-start delay the
 * context's first stepping call picked on the device for its RK stage launches, in units of 512
 * clocks per workgroup slot of a CU (0: off, -1: not tuned yet).  Grids whose tiles are all
 * resident at once start the workgroups sharing a CU one after another, so that their load phases
 * do not coincide; the delay changes the timing only, never the results. */
int fr2d_get_stagger(const fr2d_ctx *ctx, int *stagger);
/* No reference counterpart (launch tuning of this library): fixes the staggered-start delay (0..16
 * units; 0 off) instead of the device timing of the first stepping call, or (-1) asks the next
 * stepping call to time the candidates again.  For runs whose timing of the candidates is not
 * representative, such as runs under a profiler; only the timing changes, never the results. */
int fr2d_set_stagger(fr2d_ctx *ctx, int stagger);
/* No counterpart in the reference: fr2d_step for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  nsteps steps
 * of FR_RK_SSPRK3, FR_RK_MIDPOINT or FR_RK_EULER on the caller's device state d_u (reference
 * layout, read and written in place; the same arithmetic as fr2d_step, bitwise) with the stage
 * buffers taken from a fixed scratch array of the library: 2 x 6,291,456 doubles (96 MiB) per
 * GPU, held until the process ends.  No allocation and no host synchronisation; asynchronous on
 * the context's stream.  A state above 6,291,456 doubles, FR_RK_RK4 (a third array) or
 * FR_RK_TSIT5 (six more) returns FR_E_UNSUPPORTED.  Every call on one GPU shares that GPU's
 * scratch (which keeps the last call's contents and needs no initial values): calls that overlap
 * on one GPU race, whatever their stream, thread or context, and must be serialised by the
 * caller. */
int fr2d_step_dev_noalloc(fr2d_ctx *ctx, double *d_u, int scheme, double dt, int nsteps);
/* refresh the ghost layer of the device state (FR_GHOST_*): the ghost copies of the reference
 * host loops (shock-vortex.jl:322-325, euler2d_wave.jl:125-135) */
int fr2d_fill_ghost(fr2d_ctx *ctx, int rule);

/* Per-step dissipation of example/shock-vortex.jl:296-321 on the device state:
 * modal sensor (u_hat = iV * rho, su = u_hat[end]^2 / sum u_hat^2, shock_detector(log10 su, p,
 * S0, kappa)) and KitBase's L2 modal filter with strength lambda on shocked cells.
 * mode FR_SENSOR_LEARNED uses the MLP loaded by fr2d_set_sensor_mlp (strength = its output).
 * If n_flagged is non-NULL it receives the number of filtered cells (synchronises). */
int fr2d_sensor_filter(fr2d_ctx *ctx, int mode, double lambda, double S0, double kappa,
                       int *n_flagged);
/* Per-cell strength multiplier of the sensor without filtering (d_out: float per cell of the
 * (nx+2)(ny+2) layout, device memory): modal 1 / 0 (the shock_detector decision of
 * shock-vortex.jl:308-317), learned max(MLP, 0). */
int fr2d_sensor_eval(fr2d_ctx *ctx, int mode, double S0, double kappa, float *d_out);
/* Positivity limiter of src/dissipation.jl:125-206 on every interior cell of the device state,
 * weights wp/4 (shock-vortex.jl:298-303).  Returns FR_E_LIMITER if an element fails the
 * reference's t1 assertion (that element is left unchanged). */
int fr2d_positivity(fr2d_ctx *ctx);
/* The end of one iteration of the shock-vortex host loop and the start of the next in one pass
 * over the state (example/shock-vortex.jl:306-325 followed by :298-303): sensor + filter of every
 * cell (mode FR_SENSOR_*, as fr2d_sensor_filter), the ghost refresh `rule` (FR_GHOST_*) written
 * from the filtered values, and the positivity limiter on every interior cell.  Equal to
 * fr2d_sensor_filter, fr2d_fill_ghost(rule), fr2d_positivity in sequence, reading and writing
 * each cell once.  n_filtered (optional): filtered cells whose values are kept.  Returns
 * FR_E_LIMITER like fr2d_positivity. */
int fr2d_filter_refresh_limit(fr2d_ctx *ctx, int mode, double lambda, double S0, double kappa, int rule,
                              int *n_filtered);
/* nsteps iterations of the host loop of example/shock-vortex.jl:296-326 on the device state,
 * enqueued without host round trips: the positivity limiter, then per iteration one step of
 * `scheme`, the sensor + filter (`mode`), the ghost refresh `rule` and (but for the last
 * iteration) the next positivity limiter, the last three fused as in fr2d_filter_refresh_limit.
 * Equal to those calls in sequence; limiter failures are summed on the device and reported once
 * at the end (FR_E_LIMITER) instead of at the iteration that met them.  n_filtered (optional):
 * filtered cells summed over the loop. */
int fr2d_run_host_loop(fr2d_ctx *ctx, int scheme, double dt, int nsteps, int mode, double lambda, double S0,
                       double kappa, int rule, int *n_filtered);
/* Scan the device state for NaN/Inf (FR_E_NONFINITE) or rho <= 0 / p <= 0 (FR_E_NEGATIVE_STATE).
 * No reference counterpart: the reference only asserts inside positive_limiter
 * (src/dissipation.jl:143-145). */
int fr2d_check_state(fr2d_ctx *ctx);
/* Uniform parallelogram cells (dev/parallelogram.jl:78-162): iJ is the 2x2 inverse Jacobian of
 * every cell, column-major like Julia's Matrix (iJ[0] = r_x, iJ[1] = s_x, iJ[2] = r_y, iJ[3] = s_y),
 * with a positive determinant; n1 / n2 are the normals of the x faces (r = const) and y faces
 * (s = const), normalised here, or NULL for grad r / grad s.  The volume flux becomes
 * iJ [F, G], the face fluxes are solved in the frames of n1 / n2 (KitBase local_frame /
 * global_frame) and scaled by (iJ n1)_1 / (iJ n2)_2.  Without this call the cells are the
 * rectangles of example/shock-vortex.jl (iJ11, iJ22 of fr2d_desc). */
int fr2d_set_geometry(fr2d_ctx *ctx, const double *iJ, const double *n1, const double *n2);
/* Cells with point-varying metrics: the FRPSpace2D of any quad mesh (src/struct.jl:130-134,
 * J[i, j][k, l] = rs_jacobi(r[k], r[l], vertices[i, j]) and iJ its inverse, per solution point;
 * src/Geometry/geo_jacobi.jl:77-98).  iJ holds iJ[i, j][k, l] of every cell in the reference state
 * layout with the 2x2 entry in place of the variable: element (i, j, k, l, e) at
 * i + (nx+2) (j + (ny+2) (k + nsp (l + nsp e))), e = 0 r_x, 1 s_x, 2 r_y, 3 s_y (column-major),
 * ghost cells included but not read.  The residual is then example/shock-vortex.jl:26-118 as
 * written, with the metric of each point: volume flux iJ[k, l] [F, G] (:48-55), interface fluxes
 * as on rectangles (:73-89), the x-face flux of point (k, l) scaled by its own iJ[k, l][1, 1] and
 * the y-face flux by iJ[k, l][2, 2] (:102-115).  Interior entries must be finite with a positive
 * determinant.  The library keeps two device copies (both layouts); replaces fr2d_set_geometry's
 * cells. */
int fr2d_set_metrics(fr2d_ctx *ctx, const double *iJ);
/* Primitive variables of every solution point on the global point grid, the post-processing block
 * of example/shock-vortex.jl:333-352 (its `sol`): column-major (nx*nsp, ny*nsp, 4) with
 * sol[(i-1)*nsp + k, (j-1)*nsp + l, :] = (rho, U, V, 1/lambda), lambda = rho/(2p) as in KitBase
 * conserve_prim.  fr2d_primitives copies to a host array (synchronises); fr2d_primitives_dev
 * writes a device array on the context stream. */
int fr2d_primitives(fr2d_ctx *ctx, double *sol);
int fr2d_primitives_dev(fr2d_ctx *ctx, double *d_sol);

/* ------------------------------------------------------------------ learned troubled-cell sensor
 * No reference counterpart (a learned replacement of shock_detector): MLP 16 -> 32 -> 32 -> 1 (ReLU), BF16 weights
 * and inputs on MFMA with FP32 accumulation, FP32 master parameters and Adam.  Parameter vector
 * (fr_mlp_nparam() floats): W1[32][16] b1[32] W2[32][32] b2[32] W3[32] b3.  Inputs: per cell the
 * 16 density modal coefficients iV*rho normalised to unit 2-norm, as BF16 (fr2d_sensor_features). */
typedef struct fr_mlp fr_mlp;
int fr_mlp_nparam(void);
int fr_mlp_create(fr_mlp **out);
int fr_mlp_destroy(fr_mlp *m);
int fr_mlp_set_params(fr_mlp *m, const float *params);  /* host; resets the Adam state */
int fr_mlp_get_params(fr_mlp *m, float *params);        /* host */
/* s = max(MLP(x), 0) for n rows of d_x (BF16 bits, 16 per row); device pointers */
int fr_mlp_forward(fr_mlp *m, const uint16_t *d_x, float *d_out, long long n);
/* gradients of L = mean (MLP(x) - y)^2 over n rows into host grads[nparam]; *loss = L */
int fr_mlp_grad(fr_mlp *m, const uint16_t *d_x, const float *d_y, long long n, float *grads, float *loss);
/* nsteps Adam steps on consecutive batches of `batch` rows; *loss = mean of the last batch losses */
int fr_mlp_train(fr_mlp *m, const uint16_t *d_x, const float *d_y, long long n, int batch, int nsteps, float lr,
                 float *loss);
/* checks the MFMA operand/result lane layout the kernels assume; *max_err must be 0 */
int fr_mfma_selftest(float *max_err);
/* use m as the learned sensor of ctx (NULL detaches; the context keeps a pointer, m must outlive it) */
int fr2d_set_sensor_mlp(fr2d_ctx *ctx, fr_mlp *m);
/* features of every cell of the device state into d_x ((nx+2)(ny+2) rows of 16 BF16) */
int fr2d_sensor_features(fr2d_ctx *ctx, uint16_t *d_x);

/* ------------------------------------------------------------------ 1D (FREulerProblem / FRAdvectionProblem) */
typedef struct fr1d_ctx fr1d_ctx;
enum { FR_EQ_EULER = 0, FR_EQ_ADVECTION = 1 };
enum { FR_BC1_DIRICHLET = 0, FR_BC1_PERIOD = 1 };
typedef struct {
  int equation;      /* FR_EQ_* */
  int ncell, deg;    /* 1 <= deg <= 8 (example/euler_highlevel.jl runs degree 7) */
  double dx;         /* uniform cell size; J = dx/2 (src/struct.jl:42) */
  double gamma;      /* Euler */
  double a;          /* advection speed */
  int flux;          /* Euler Riemann flux, FR_FLUX_* */
  int bc;            /* FR_BC1_* (src/Equation/eq_euler.jl:68-98, eq_advection.jl:251-273) */
  int correction;    /* FR_CORR_* (src/struct.jl:63) */
} fr1d_desc;
/* Replaces FREulerProblem / FRAdvectionProblem (src/Equation/eq_euler.jl:1-27,
 * eq_advection.jl:1-26): the scratch tuple p and the operator data, owned by the context. */
int fr1d_create(fr1d_ctx **out, const fr1d_desc *desc);
int fr1d_destroy(fr1d_ctx *ctx);
size_t fr1d_state_len(const fr1d_ctx *ctx);
/* frode_euler!(du, u, p, t) (src/Equation/eq_euler.jl:29-73) or
 * frode_advection!(du, u, p, t) (src/Equation/eq_advection.jl:153-175) */
int fr1d_rhs(fr1d_ctx *ctx, const double *u, double *du, double t);
int fr1d_rhs_dev(fr1d_ctx *ctx, const double *d_u, double *d_du, double t);
/* the stream of every later call on the context, with fr2d_set_stream's meaning (NULL is HIP's
 * null stream, FR_OWN_STREAM the context's own non-blocking stream, the default) */
int fr1d_set_stream(fr1d_ctx *ctx, void *hip_stream);
/* device-resident state and nsteps steps of FR_RK_*, replacing init / step!(itg) on the
 * ODEProblem (example/euler_highlevel.jl:34,54).  FR_RK_TSIT5 is OrdinaryDiffEq's Tsit5 with a
 * fixed dt (adaptive = false: example/advection_highlevel.jl:26, euler1d_convergence.jl:133),
 * six residual evaluations per step and six more state-sized arrays (fr2d_step takes it too;
 * fr_bgk1d_step and fr2d_step_dev_noalloc return FR_E_UNSUPPORTED for it). */
int fr1d_set_state(fr1d_ctx *ctx, const double *u);
int fr1d_get_state(fr1d_ctx *ctx, double *u);
int fr1d_step(fr1d_ctx *ctx, int scheme, double dt, int nsteps);
/* Execution path of fr1d_step, fr_ens_step and fr_ens_run (same bits on every path; no
 * counterpart in the reference, the path is this library's execution choice):
 * FR1D_PATH_BATCHED: a launch per stage and per dissipation pass over all members' work items;
 * FR1D_PATH_MEMBER: a workgroup per member runs the whole call's schedule (every stage,
 * dissipation pass and copy, a workgroup barrier after each) in launches of up to 3072 operations;
 * FR1D_PATH_AUTO (the default): the member path when no member has more than two work items of
 * the context's workgroup size (up to 508 cells at 256 threads), else batched.  Tsit5,
 * fr_ens_step's mode 1 and the graph entry points always take the batched launches. */
#define FR1D_PATH_AUTO 0
#define FR1D_PATH_BATCHED 1
#define FR1D_PATH_MEMBER 2
int fr1d_set_path(fr1d_ctx *ctx, int path);
/* the path fr1d_step takes for this scheme under the current setting */
int fr1d_get_path(const fr1d_ctx *ctx, int scheme, int *path);
/* The per-step host loop of example/euler_highlevel.jl:37-52 on the device state (Euler only):
 * modal sensor on rho (su = u_hat[end]^2 / (sum u_hat^2 + 1e-6), shock_detector of
 * src/dissipation.jl:13 with S0 = -3 log10(deg), kappa = 4) and, on a shock, modal_filter!
 * (:l2) of every variable with lambda = dt exp(0.875 deg) 0.15.  n_flagged (optional). */
int fr1d_sensor_filter(fr1d_ctx *ctx, double dt, int *n_flagged);
/* positive_limiter(u::AbstractMatrix, ...) of src/dissipation.jl:61-123 on every cell (Gauss
 * weights / 2).  Returns FR_E_LIMITER if a cell has a non-positive mean density (left as is);
 * n_limited (optional) counts the cells changed. */
int fr1d_positivity(fr1d_ctx *ctx, int *n_limited);

/* ------------------------------------------------------------------ batched 1D Euler ensembles */
typedef struct fr_ens_ctx fr_ens_ctx;
/* nprob independent 1D Euler problems (layout of each member as fr1d); member b has ncell[b]
 * cells of degree deg[b] (1..8) and cell size dx[b].  No reference counterpart: the reference
 * runs one FREulerProblem per script; an ensemble advances many in one launch per stage. */
int fr_ens_create(fr_ens_ctx **out, int nprob, const int *ncell, const int *deg, const double *dx,
                  double gamma, int flux, int bc);
int fr_ens_destroy(fr_ens_ctx *ctx);
size_t fr_ens_state_len(const fr_ens_ctx *ctx);          /* total doubles over all members */
size_t fr_ens_offset(const fr_ens_ctx *ctx, int member); /* first double of a member */
int fr_ens_set_state(fr_ens_ctx *ctx, const double *u);  /* concatenated member states */
int fr_ens_get_state(fr_ens_ctx *ctx, double *u);
/* mode 0: one batched launch per stage; 1: one launch per member per stage (looped baseline) */
int fr_ens_step(fr_ens_ctx *ctx, int mode, double dt, int nsteps);
/* the execution path of fr_ens_step / fr_ens_run (FR1D_PATH_*; see fr1d_set_path) */
int fr_ens_set_path(fr_ens_ctx *ctx, int path);
/* capture nsteps batched steps in a hipGraph and replay nrep times (fr_ens_run_graph, no options) */
int fr_ens_step_graph(fr_ens_ctx *ctx, double dt, int nsteps, int nrep);
/* The per-step host loop of example/euler_highlevel.jl:37-54 over every member, with SSP-RK3 steps
 * or, with FR_ENS_MIDPOINT, the script's own Midpoint steps: FR_ENS_FILTER runs
 * fr1d_sensor_filter's sensor + filter (filter time step dt) on the state before every step;
 * FR_ENS_LIMIT runs positive_limiter (src/dissipation.jl:61-123) on the input of every stage.
 * mode 0: one batched launch per pass; 1: one launch per member per pass (looped baseline).
 * n_flagged / n_limited (optional) sum the filtered / limited cells over the loop.  A cell whose
 * mean density is not positive is left unchanged and reported as FR_E_LIMITER after the loop. */
enum { FR_ENS_FILTER = 1, FR_ENS_LIMIT = 2, FR_ENS_MIDPOINT = 4 };
int fr_ens_run(fr_ens_ctx *ctx, int mode, int opts, double dt, int nsteps, int *n_flagged, int *n_limited);
/* the batched loop of nsteps steps captured in a hipGraph (kept while opts, dt, nsteps and the
 * work mapping are unchanged, so later calls only replay it) and replayed nrep times */
int fr_ens_run_graph(fr_ens_ctx *ctx, int opts, double dt, int nsteps, int nrep, int *n_flagged, int *n_limited);
/* time candidate work mappings on the device and keep the fastest; returns its index */
int fr_ens_autotune(fr_ens_ctx *ctx, double dt, int *chosen);
/* fr1d_sensor_filter / fr1d_positivity over every member (each with its own degree) */
int fr_ens_sensor_filter(fr_ens_ctx *ctx, double dt, int *n_flagged);
int fr_ens_positivity(fr_ens_ctx *ctx, int *n_limited);

/* ------------------------------------------------------------------ 1D discrete-velocity BGK
 * example/bgk_wave.jl:69-129 (mol!) as the script writes it: nx periodic cells of size dx, nu
 * velocities u with quadrature weights w (KitBase VSpace1D u, weights), Radau FR of degree deg;
 * per point the moments (moments_conserve), conserve_prim at gamma (the script: 3.0) and the
 * Maxwellian rho sqrt(lambda / pi) exp(-lambda (u - U)^2); f = u_j u / J, the upwind interface value
 * of the Heaviside blend (:103-107), derivative and corrections, plus (M - u) / tau (the script:
 * tau = 1e-2).  State u(i, j, k) column-major (cell, velocity, point) per member, members
 * concatenated (nmem > 1: a batched ensemble whose members differ in tau). */
typedef struct fr_bgk1d_ctx fr_bgk1d_ctx;
typedef struct {
  int nx;            /* periodic cells */
  int deg;           /* 1..8 */
  int nu;            /* velocities */
  int nmem;          /* ensemble members */
  double dx;         /* cell size (J = dx/2) */
  double gamma;      /* of conserve_prim (3 in bgk_wave.jl:79) */
  const double *u, *w;  /* velocities and quadrature weights */
  const double *tau;    /* nmem relaxation times */
} fr_bgk1d_desc;
int fr_bgk1d_create(fr_bgk1d_ctx **out, const fr_bgk1d_desc *desc);
int fr_bgk1d_destroy(fr_bgk1d_ctx *ctx);
size_t fr_bgk1d_state_len(const fr_bgk1d_ctx *ctx);
/* mol!(du, u, p, t) of every member (bgk_wave.jl:69-129): host pointers, or device pointers on a
 * caller stream (NULL: HIP's null stream) */
int fr_bgk1d_rhs(fr_bgk1d_ctx *ctx, const double *f, double *df);
int fr_bgk1d_rhs_dev(fr_bgk1d_ctx *ctx, const double *d_f, double *d_df, void *stream);
int fr_bgk1d_set_state(fr_bgk1d_ctx *ctx, const double *f);
int fr_bgk1d_get_state(fr_bgk1d_ctx *ctx, double *f);
/* nsteps steps of the device state: FR_RK_SSPRK3, FR_RK_MIDPOINT (the script's Midpoint()) or
 * FR_RK_EULER */
int fr_bgk1d_step(fr_bgk1d_ctx *ctx, int scheme, double dt, int nsteps);
/* nrep replays of a hipGraph of nsteps such steps, captured on the first call and kept while
 * scheme, dt and nsteps are unchanged: the bits of fr_bgk1d_step over nrep * nsteps steps, for the
 * launch-bound small cases (the script's 20 cells x 100 velocities) */
int fr_bgk1d_step_graph(fr_bgk1d_ctx *ctx, int scheme, double dt, int nsteps, int nrep);
/* Execution path of fr_bgk1d_step / fr_bgk1d_step_graph (same bits on every path):
 * FR_BGK1D_PATH_STREAM: two launches per stage (velocity sums, transport/relaxation) over the
 * whole state in HBM; FR_BGK1D_PATH_RESIDENT: one launch per up to 256 steps, a workgroup per
 * member holding the member's state, stage arrays and point data in LDS (members of up to about
 * 6,800 values under SSP-RK3, 10,000 under midpoint / Euler, at 160 KiB of LDS; a step on a larger
 * member returns FR_E_UNSUPPORTED); FR_BGK1D_PATH_AUTO (the default): resident whenever it fits.
 * No counterpart in the reference: the path is this library's execution choice. */
#define FR_BGK1D_PATH_AUTO 0
#define FR_BGK1D_PATH_STREAM 1
#define FR_BGK1D_PATH_RESIDENT 2
int fr_bgk1d_set_path(fr_bgk1d_ctx *ctx, int path);
/* the path fr_bgk1d_step takes for this scheme under the current setting */
int fr_bgk1d_get_path(const fr_bgk1d_ctx *ctx, int scheme, int *path);
/* (rho, rho U, E) of the device state at every point: w[member][c][i + nx k] */
int fr_bgk1d_moments(fr_bgk1d_ctx *ctx, double *w);

/* ------------------------------------------------------------------ 2D discrete-velocity BGK
 * example/bgk_wave.jl:69-129 tensorised to periodic uniform quads (Radau FR, upwind interface flux
 * per velocity, relaxation (M - f)/tau to the 2D Maxwellian of the local moments).  State
 * f(i, j, k, l, a, b) column-major per member, members concatenated (nmem > 1: a batched ensemble
 * whose members differ in tau). */
typedef struct fr_bgk_ctx fr_bgk_ctx;
typedef struct {
  int nx, ny;        /* periodic cells, no ghost layer */
  int deg;           /* 1..4 */
  int nu, nv;        /* velocities per direction */
  int nmem;          /* ensemble members */
  double dx, dy;     /* cell sizes (J = dx/2, dy/2) */
  double gamma;      /* of KitBase conserve_prim for the moments (2 for a 2D monatomic gas) */
  const double *u, *v, *wu, *wv;  /* abscissae and quadrature weights per velocity direction */
  const double *tau;              /* nmem relaxation times */
} fr_bgk_desc;
int fr_bgk_create(fr_bgk_ctx **out, const fr_bgk_desc *desc);
int fr_bgk_destroy(fr_bgk_ctx *ctx);
size_t fr_bgk_state_len(const fr_bgk_ctx *ctx);
/* mol!(du, u, p, t) for every member (host pointers) */
int fr_bgk_rhs(fr_bgk_ctx *ctx, const double *f, double *df);
int fr_bgk_set_state(fr_bgk_ctx *ctx, const double *f);
int fr_bgk_set_state_dev(fr_bgk_ctx *ctx, const double *d_f);  /* device pointer, synchronises */
int fr_bgk_get_state(fr_bgk_ctx *ctx, double *f);
/* nsteps SSP-RK3 steps of all members, one moments + one transport launch per stage */
int fr_bgk_step(fr_bgk_ctx *ctx, double dt, int nsteps);
/* conserved moments of the device state, w[member][c][point] (host) */
int fr_bgk_moments(fr_bgk_ctx *ctx, double *w);
/* Gauss-Hermite nodes / weights for the weight exp(-x^2) (host, native): the velocity quadrature
 * of the BGK state, in place of KitBase's VSpace1D of bgk_wave.jl:25 */
int fr_gauss_hermite(int n, double *x, double *w);

#ifdef __cplusplus
}
#endif
#endif
