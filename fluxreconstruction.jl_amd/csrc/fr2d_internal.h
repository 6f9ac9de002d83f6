// fr2d_internal.h -- context of the 2D Euler solver and its kernel launchers.
#pragma once

#include "fr_common.h"

namespace fr {

struct Fr2dCtx {
  int nx = 0, ny = 0, deg = 0, nsp = 0, flux = 0, bc = 0;
  double gamma = 1.4, iJ11 = 1.0, iJ22 = 1.0;
  // cell geometry: 0 rectangles (iJ11, iJ22); 1 uniform parallelograms with the full inverse
  // Jacobian and the unit normals of the x faces (n1) and y faces (n2); 2 point-varying metrics
  // (d_mj_ref / d_mj_pad: the inverse Jacobian of every solution point in the reference and in
  // the padded layout, four planes per point like a state's variables)
  int geom = 0;
  double iJ12 = 0.0, iJ21 = 0.0;
  double n1[2] = {1.0, 0.0}, n2[2] = {0.0, 1.0};
  double *d_mj_ref = nullptr, *d_mj_pad = nullptr, *raw_mj_pad = nullptr;
  HostOps ops;
  size_t len = 0;  // doubles per state array in the reference layout, (nx+2)(ny+2) nsp^2 4
  // The device-resident arrays (state, stage buffers, RK4 accumulator) use a padded form of the
  // reference layout: rows of `pitch` elements (nx+2 rounded up to a multiple of 16) and planes
  // of ldpi = pitch (ny+2) elements, each array starting 15 doubles into its allocation.  Element
  // (1, j) of every plane then starts a 128-byte line, so the 16-element row segments the kernels
  // read and write are aligned (the reference pitch leaves every segment across two lines).
  int pitch = 0;
  long long ldpi = 0;
  size_t lenp = 0;  // doubles per padded array, ldpi nsp^2 4
  static constexpr int kPadFront = 15;
  int device = 0;
  int num_cu = 0;  // compute units of the device
  // Staggered start of the RK stage launches whose tiles are all resident at once (at most four
  // workgroups per CU): workgroup slot s of a CU (s = block / num_cu) first sleeps s * stagger
  // units of 64 x 8 clocks, so the load phases of the workgroups sharing a CU do not coincide.
  // -1: not tuned yet (fr2d_tune_stagger picks it on the device), 0: off.  Only the timing
  // changes, never the arithmetic.
  int stagger = -1;
  // the caller's stream (caller_stream: any hipStream_t, NULL being HIP's null stream) or the
  // context's own non-blocking stream
  bool caller_stream = false;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  // device state and stage buffers of the integrator path (padded layout; raw_* are the
  // allocations, d_* = raw_* + kPadFront)
  double *d_u = nullptr, *d_s1 = nullptr, *d_s2 = nullptr, *d_acc = nullptr;
  double *raw_u = nullptr, *raw_s1 = nullptr, *raw_s2 = nullptr, *raw_acc = nullptr;
  // Tsit5: six stage derivatives in the padded layout, kstride doubles apart (one zeroed
  // allocation; the kernels never write their ghost cells or row padding, which stay zero)
  double *d_k = nullptr, *raw_k = nullptr;
  size_t kstride = 0;
  // staging buffers of the host-pointer residual
  double *d_in = nullptr, *d_out = nullptr;
  // per-cell flags / counters of the dissipation kernels
  int *d_count = nullptr;
  // learned sensor: BF16 MLP parameters, per-cell features and strengths in device memory
  void *d_mlp = nullptr;
  void *d_feat = nullptr;
  void *d_strength = nullptr;
  bool mlp_loaded = false;
  // device copies of V, iV, wp/4, ll, lr used by the dissipation kernels
  double *d_dops = nullptr;
  // the last step sequence captured by fr2d_step_graph, replayed while its parameters match
  hipGraphExec_t gexec = nullptr;
  int g_scheme = -1, g_nsteps = 0;
  double g_dt = 0.0;
};

inline hipStream_t fr2d_stream(const Fr2dCtx &c) { return c.caller_stream ? c.stream : c.own_stream; }

// Every launcher below works on the padded arrays of the context (pitch, ldpi), except
// fr2d_launch_residual, whose u / du are caller arrays in the reference layout.
int fr2d_launch_residual(const Fr2dCtx &c, const double *d_u, double *d_du, hipStream_t st);
// ghost_copy0/1 (optional, ignored on periodic grids): buffers that receive d_us's ghost layer
int fr2d_launch_stage(const Fr2dCtx &c, const double *d_u0, const double *d_us, double *d_out, double a0,
                      double a1, double b, const double *acc_in, double *acc_out, double b2, hipStream_t st,
                      double *ghost_copy0 = nullptr, double *ghost_copy1 = nullptr);
int fr2d_launch_fill_ghost(const Fr2dCtx &c, double *d_u, int rule, hipStream_t st);
// whether the RK stage launches of c have every tile resident at once (the staggered start
// applies), and the timing of the stagger candidates on the device (keeps the fastest in c.stagger)
bool fr2d_single_round(const Fr2dCtx &c);
int fr2d_tune_stagger(Fr2dCtx &c, hipStream_t st);
// eval_out != nullptr: write the per-cell strength multiplier (modal: 0/1, learned: MLP output)
// instead of filtering
int fr2d_launch_sensor_filter(const Fr2dCtx &c, double *d_u, int mode, double lambda, double S0, double kappa,
                              int *d_count, float *eval_out, hipStream_t st);
int fr2d_launch_positivity(const Fr2dCtx &c, double *d_u, int *d_count, hipStream_t st);
// sensor + filter of every cell, ghost refresh `rule` from the filtered values, positivity of the
// interior cells, in one pass (d_count[0] filtered, d_count[1] limiter failures)
int fr2d_launch_filter_refresh_limit(const Fr2dCtx &c, double *d_u, int mode, double lambda, double S0,
                                     double kappa, int rule, int *d_count, hipStream_t st);
int fr2d_launch_check(const Fr2dCtx &c, const double *d_u, int *d_count, hipStream_t st);
// primitive variables of every solution point on the global point grid (fr2d_post.hip)
int fr2d_launch_primitives(const Fr2dCtx &c, const double *d_u, double *d_sol, hipStream_t st);
// learned sensor: per-cell strengths from the MLP into c.d_strength (fr_mlp.hip)
int fr2d_launch_mlp_sensor(const Fr2dCtx &c, const double *d_u, hipStream_t st);

// device copies of the dissipation operators, carved out of Fr2dCtx::d_dops
struct DissOps {
  const double *V, *iV, *wts, *ll, *lr;
};
DissOps fr2d_diss_ops(const Fr2dCtx &c);

}  // namespace fr

// opaque handle of the C ABI (include/fr.h)
struct fr2d_ctx : fr::Fr2dCtx {};
