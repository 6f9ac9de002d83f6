// fr2d_dissipation.hip -- per-step dissipation kernels of the 2D solver (gfx950):
// modal shock sensor + L2 modal filter (example/shock-vortex.jl:308-321, src/dissipation.jl:13-23),
// positivity limiter (src/dissipation.jl:125-215), state checks and ghost-layer copies.
// The filter and the limiter run four lanes per element (one per conserved variable); the sensor
// evaluation and the state check one thread per element.  Element planes are read straight from
// the SoA state (consecutive elements on consecutive lanes, so accesses stay contiguous).
#include <algorithm>

#include "fr2d_internal.h"

namespace fr {

__device__ __forceinline__ int shock_switch(double Se, double S0, double kappa) {
  double sigma;
  if (Se < S0 - kappa)
    sigma = 1.0;
  else if (Se < S0 + kappa)
    sigma = 0.5 * (1.0 - sin(0.5 * M_PI * (Se - S0) / kappa));
  else
    sigma = 0.0;
  return sigma < 0.99 ? 1 : 0;
}

// Sensor + filter over every cell, ghosts included (the reference loops over axes(itg.u, 1:2)).
// V, iV: nodal<->modal matrices (np x np column-major, node q = k + NSP*l, mode a*NSP + b).
// mode FR_SENSOR_MODAL: strength lambda on cells flagged by the Persson-Peraire switch;
// FR_SENSOR_LEARNED: strength = lambda * s[e], s produced per cell by the MLP sensor kernels.
template <int NSP>
__global__ __launch_bounds__(256) void fr2d_sensor_filter_kernel(double *u, int ncell, int nxg, int pitch,
                                                                 long long ldp,
                                                                 const double *__restrict__ V,
                                                                 const double *__restrict__ iV, int mode,
                                                                 double lambda, double S0, double kappa,
                                                                 const float *__restrict__ strength, int *count,
                                                                 float *eval_out) {
  constexpr int NP = NSP * NSP;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;  // cell, reference order
  if (e >= ncell) return;
  const long long ep = e % nxg + (long long)pitch * (e / nxg);             // its place in a plane
  double nod[NP], mod[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) nod[q] = u[ep + ldp * q];
#pragma unroll
  for (int a = 0; a < NP; ++a) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < NP; ++b) s += iV[a + NP * b] * nod[b];
    mod[a] = s;
  }
  double lam = 0.0;
  if (mode == FR_SENSOR_MODAL) {
    double ss = 0.0;
#pragma unroll
    for (int a = 0; a < NP; ++a) ss += mod[a] * mod[a];
    const double su = mod[NP - 1] * mod[NP - 1] / ss;
    lam = shock_switch(log10(su), S0, kappa) ? lambda : 0.0;
  } else {
    lam = lambda * (double)strength[e];
  }
  if (eval_out) {  // evaluation only: per-cell strength multiplier, state untouched
    eval_out[e] = (lambda > 0.0) ? (float)(lam / lambda) : (lam > 0.0 ? 1.0f : 0.0f);
    return;
  }
  if (!(lam > 0.0)) return;
  if (count) atomicAdd(count, 1);
  for (int m = 0; m < 4; ++m) {
    if (m > 0) {
#pragma unroll
      for (int q = 0; q < NP; ++q) nod[q] = u[ep + ldp * (q + NP * m)];
#pragma unroll
      for (int a = 0; a < NP; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NP; ++b) s += iV[a + NP * b] * nod[b];
        mod[a] = s;
      }
    }
    // KitBase filter_l2! on the flattened mode vector: mode n /= 1 + lambda (n+1)^2 n^2
#pragma unroll
    for (int a = 1; a < NP; ++a) mod[a] /= (1.0 + lam * (double)((a + 1) * (a + 1)) * (double)(a * a));
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < NP; ++b) s += V[q + NP * b] * mod[b];
      u[ep + ldp * (q + NP * m)] = s;
    }
  }
}


// Source cell of ghost cell (gi, gj) under a ghost refresh rule (as fr2d_fill_ghost_kernel);
// keep = the rule leaves the cell as it is (the inflow column of FR_GHOST_INFLOW_LEFT); negv =
// the variable copied negated (2: the y-momentum of the y ghost rows of FR_GHOST_WAVE, 1: the
// x-momentum of the x ghost columns of FR_GHOST_WAVE_Y), -1 none.
__device__ __forceinline__ void ghost_source(int gi, int gj, int nx, int ny, int rule, int &si, int &sj, bool &keep,
                                             int &negv) {
  keep = false;
  negv = -1;
  if (rule == FR_GHOST_PERIODIC || rule == FR_GHOST_WAVE || rule == FR_GHOST_WAVE_Y) {
    si = (gi == 0) ? nx : (gi == nx + 1 ? 1 : gi);
    sj = (gj == 0) ? ny : (gj == ny + 1 ? 1 : gj);
    if (rule == FR_GHOST_WAVE && (gj == 0 || gj == ny + 1)) negv = 2;
    if (rule == FR_GHOST_WAVE_Y && (gi == 0 || gi == nx + 1)) negv = 1;
  } else if (rule == FR_GHOST_OUTFLOW) {
    si = (gi == 0) ? 1 : (gi == nx + 1 ? nx : gi);
    sj = (gj == 0) ? 1 : (gj == ny + 1 ? ny : gj);
  } else {
    keep = gi == 0 && gj >= 1 && gj <= ny;
    si = (gi == nx + 1) ? nx : gi;
    sj = (gj == 0) ? 1 : (gj == ny + 1 ? ny : gj);
  }
}

// The end of one iteration of the shock-vortex host loop and the start of the next in one pass
// (example/shock-vortex.jl:306-325, then :298-303): sensor + filter of every cell, the ghost
// refresh `rule` written from the filtered values, and the positivity limiter of every interior
// cell.  Each cell is read once; a ghost cell that the rule overwrites is written only by the
// threads of its source cell, so no cell has two writers.  nfilt counts the filtered cells whose
// values are kept (interior and kept ghosts), nfail the cells whose density assert failed (left
// as filtered); each is touched only when its event occurs.
//
// Four lanes per cell: lane r of a quad holds variable r of the cell at every solution point, so
// the modal transform and the filter of the four variables run side by side.  The density
// lane's sensor value is broadcast over the quad; the limiter's test points are split over the
// quad (lane r: the points of face r and the nodes q = r (mod 4)), their four variables gathered
// through LDS, and the corrector minimum reduced over the quad.  Every value is formed with the
// operations, in the order, of one thread per cell running src/dissipation.jl:13-23 and
// :125-215, so the result does not depend on how the cells or the points are split.
//
// The same kernel serves the separate entry points: rule -1 skips the ghost refresh (every cell
// is then filtered where it lies, ghosts included, as shock-vortex.jl:308-321), limit 0 skips the
// limiter, mode FR_SENSOR_OFF the filter.
//
// Launch bound (256, 4): four wavefronts per SIMD; at nsp = 4 the compiler then keeps the kernel
// in 128 VGPRs with a few spilled values, measured 3-4 % faster than 142 VGPRs at three.
template <int NSP>
__global__ __launch_bounds__(256, 4) void fr2d_filter_refresh_limit_kernel(
    double *u, int nx, int ny, int pitch, long long ldp, const double *__restrict__ V,
    const double *__restrict__ iV, const double *__restrict__ wts, const double *__restrict__ ll,
    const double *__restrict__ lr, int mode, double lambda, double S0, double kappa,
    const float *__restrict__ strength, int rule, int limit, double g, int *nfilt, int *nfail) {
  constexpr int NP = NSP * NSP;
  constexpr int NS = NP > 4 * NSP ? NP : 4 * NSP;  // exchange slots per cell and variable
  constexpr int SS = 64 * 4 + 4;                   // doubles per slot: 64 cells x 4 variables (+ pad)
  __shared__ double sx[NS * SS];
  __shared__ double smean[64 * 4];
  // the rows of iV, then the rows of V: every lane reads the same entry at a time (LDS broadcast),
  // so the matrices occupy neither scalar nor vector registers
  __shared__ double smat[2 * NP * NP];
  for (int t = threadIdx.x; t < NP * NP; t += 256) {
    const int a = t / NP, b = t % NP;
    smat[t] = iV[a + NP * b];
    smat[NP * NP + t] = V[a + NP * b];
  }
  __syncthreads();
  const int r = threadIdx.x & 3;   // variable held by this lane
  const int c = threadIdx.x >> 2;  // cell slot of the workgroup
  const int nxg = nx + 2;
  // The workgroups walk the padded plane in place order, shifted by the front offset of the
  // array (Fr2dCtx::kPadFront), so that the 16 cells of a wavefront are one aligned 128-byte segment of every
  // plane; padding cells and the places before the plane leave at once (whole quads together).
  const long long ep = blockIdx.x * 64LL + c - Fr2dCtx::kPadFront;  // place in a padded plane
  if (ep < 0 || ep >= ldp) return;
  const int i = (int)(ep % pitch), j = (int)(ep / pitch);
  if (i >= nxg) return;
  const long long e = i + (long long)nxg * j;  // the cell in the reference order (strength index)
  const bool ghost = i == 0 || i == nx + 1 || j == 0 || j == ny + 1;
  if (ghost) {
    if (rule < 0) {
      if (mode == FR_SENSOR_OFF) return;  // neither filtered nor limited
    } else {
      // a kept ghost stays: it may be the source of other ghosts (the corners of
      // FR_GHOST_INFLOW_LEFT copy the inflow column)
      int si, sj;
      bool keep;
      int negv;
      ghost_source(i, j, nx, ny, rule, si, sj, keep, negv);
      if (!keep) return;  // refreshed by the threads of its source cell
    }
  }
  auto X = [&](int q, int m) -> double & { return sx[q * SS + c * 4 + m]; };
  double *const ur = u + ep + ldp * (NP * r);
  double x[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) x[q] = ur[ldp * q];

  // sensor + filter.  The modal coefficients of the density lane give the switch; they are
  // formed again for a flagged cell, so that no coefficient is live across the switch's sine.
  bool changed = false, flagged = false;
  if (mode != FR_SENSOR_OFF) {
    auto modal = [&](double *mod) {
#pragma unroll
      for (int a = 0; a < NP; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NP; ++b) s += smat[a * NP + b] * x[b];
        mod[a] = s;
      }
    };
    double lam = 0.0;
    if (mode == FR_SENSOR_MODAL) {
      double su;
      {
        double mod[NP];
        modal(mod);
        double ss = 0.0;
#pragma unroll
        for (int a = 0; a < NP; ++a) ss += mod[a] * mod[a];
        su = __shfl(mod[NP - 1] * mod[NP - 1] / ss, 0, 4);  // the density lane's value
      }
      lam = shock_switch(log10(su), S0, kappa) ? lambda : 0.0;
    } else {
      lam = lambda * (double)strength[e];
    }
    if (lam > 0.0) {
      changed = flagged = true;
      double mod[NP];
      modal(mod);
#pragma unroll
      for (int a = 1; a < NP; ++a) mod[a] /= (1.0 + lam * (double)((a + 1) * (a + 1)) * (double)(a * a));
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NP; ++b) s += smat[NP * NP + q * NP + b] * mod[b];
        x[q] = s;
      }
    }
  }
  {  // one count update per wavefront
    const unsigned long long fb = __ballot(flagged && r == 0);
    if (fb && (int)(threadIdx.x & 63) == __ffsll((long long)fb) - 1) atomicAdd(nfilt, __popcll(fb));
  }

  // ghost refresh: copy the filtered cell into every ghost cell whose source it is
  if (rule >= 0 && !(i >= 2 && i <= nx - 1 && j >= 2 && j <= ny - 1)) {
    const int ci[5] = {i - 1, i, i + 1, 0, nx + 1}, cj[5] = {j - 1, j, j + 1, 0, ny + 1};
    for (int a = 0; a < 5; ++a) {
      const int gi = ci[a];
      if (gi < 0 || gi > nx + 1) continue;
      bool dup = false;
      for (int b = 0; b < a; ++b) dup = dup || ci[b] == gi;
      if (dup) continue;
      for (int cc = 0; cc < 5; ++cc) {
        const int gj = cj[cc];
        if (gj < 0 || gj > ny + 1) continue;
        bool dupj = false;
        for (int d = 0; d < cc; ++d) dupj = dupj || cj[d] == gj;
        if (dupj || (gi == i && gj == j)) continue;
        if (!(gi == 0 || gi == nx + 1 || gj == 0 || gj == ny + 1)) continue;
        int si, sj;
        bool keep;
        int negv;
        ghost_source(gi, gj, nx, ny, rule, si, sj, keep, negv);
        if (keep || si != i || sj != j) continue;
        double *const gr = u + (long long)gi + (long long)pitch * gj + ldp * (NP * r);
        const double sg = r == negv ? -1.0 : 1.0;
#pragma unroll
        for (int q = 0; q < NP; ++q) gr[ldp * q] = sg * x[q];
      }
    }
  }

  // positivity limiter of src/dissipation.jl:125-206 on the interior cells (weights wp/4)
  bool failed = false;
  if (limit && !ghost) {
    double um_own = 0.0;
#pragma unroll
    for (int l = 0; l < NSP; ++l)
#pragma unroll
      for (int k = 0; k < NSP; ++k) um_own += x[k + NSP * l] * wts[k + NSP * l];
    smean[c * 4 + r] = um_own;
#pragma unroll
    for (int q = 0; q < NP; ++q) X(q, r) = x[q];
    wave_lds_sync();
    double um[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) um[m] = smean[c * 4 + m];
    const double p_mean = 0.5 * um[0] * (1.0 / prim_lambda(um, g));
    const double eps = fmin(1e-13, fmin(um[0], p_mean));
    // density minimum over the boundary points (faces 1 bottom, 2 right, 3 top, 4 left) and the
    // nodes, taken by the density lane from its registers and broadcast over the quad
    double rmin = INFINITY;
#pragma unroll
    for (int jj = 0; jj < NSP; ++jj)
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < NSP; ++q) {
          const double wq = (f == 0 || f == 3) ? ll[q] : lr[q];
          s += x[(f == 0 || f == 2) ? jj + NSP * q : q + NSP * jj] * wq;
        }
        rmin = fmin(rmin, s);
      }
#pragma unroll
    for (int q = 0; q < NP; ++q) rmin = fmin(rmin, x[q]);
    rmin = __shfl(rmin, 0, 4);
    const double t1 = fmin((um[0] - eps) / (um[0] - rmin + 1e-8), 1.0);
    if (!(t1 > 0.0 && t1 <= 1.0)) {
      failed = true;
    } else {
      double tmin = INFINITY;
      bool have = false;
      // the density-fixed nodes q = r (mod 4)
#pragma unroll
      for (int t = 0; t < (NP + 3) / 4; ++t) {
        const int q = 4 * t + r;
        if (q < NP) {
          double w[4] = {t1 * (X(q, 0) - um[0]) + um[0], X(q, 1), X(q, 2), X(q, 3)};
          if (prim_lambda_below(w, g, eps)) {
            have = true;
            tmin = fmin(tmin, tj_root(w, um, g, eps));
          }
        }
      }
      wave_lds_sync();  // the nodes are read: the slots take this lane's variable at every
                        // boundary point (x is still the state before the density fix)
#pragma unroll
      for (int jj = 0; jj < NSP; ++jj)
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < NSP; ++q) {
            const double wq = (f == 0 || f == 3) ? ll[q] : lr[q];
            s += x[(f == 0 || f == 2) ? jj + NSP * q : q + NSP * jj] * wq;
          }
          X(4 * jj + f, r) = s;
        }
      wave_lds_sync();
      // the boundary points of face r
#pragma unroll
      for (int jj = 0; jj < NSP; ++jj) {
        double w[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) w[m] = X(4 * jj + r, m);
        if (prim_lambda_below(w, g, eps)) {
          have = true;
          tmin = fmin(tmin, tj_root(w, um, g, eps));
        }
      }
      tmin = fmin(tmin, __shfl_xor(tmin, 1, 4));
      tmin = fmin(tmin, __shfl_xor(tmin, 2, 4));
      int hv = have ? 1 : 0;
      hv |= __shfl_xor(hv, 1, 4);
      hv |= __shfl_xor(hv, 2, 4);
      if (t1 < 1.0) changed = true;
      if (r == 0) {
#pragma unroll
        for (int q = 0; q < NP; ++q) x[q] = t1 * (x[q] - um[0]) + um[0];
      }
      if (hv) {
        changed = true;
        const double t2 = fmin(tmin, 1.0);
#pragma unroll
        for (int q = 0; q < NP; ++q) x[q] = t2 * (x[q] - um_own) + um_own;
      }
    }
  }
  {
    const unsigned long long fb = __ballot(failed && r == 0);
    if (fb && (int)(threadIdx.x & 63) == __ffsll((long long)fb) - 1) atomicAdd(nfail, __popcll(fb));
  }
  if (changed) {
#pragma unroll
    for (int q = 0; q < NP; ++q) ur[ldp * q] = x[q];
  }
}

// counts non-finite values (count[0]) and non-positive density / pressure points (count[1])
template <int NSP>
__global__ __launch_bounds__(256) void fr2d_check_kernel(const double *u, int nx, int ny, int pitch, long long ldp,
                                                         double g, int *count) {
  constexpr int NP = NSP * NSP;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)nx * ny) return;
  const long long e = (t % nx + 1) + (long long)pitch * (t / nx + 1);
  int nonfinite = 0, negative = 0;
  for (int q = 0; q < NP; ++q) {
    double w[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      w[m] = u[e + ldp * (q + NP * m)];
      if (!isfinite(w[m])) ++nonfinite;
    }
    const double p = (g - 1.0) * (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]);
    if (!(w[0] > 0.0) || !(p > 0.0)) ++negative;
  }
  if (nonfinite) atomicAdd(count, nonfinite);
  if (negative) atomicAdd(count + 1, negative);
}


// ------------------------------------------------------------------ launchers


// One launch of the four-lanes-per-cell pass: the filter (mode), the ghost refresh (rule, -1 for
// none: every cell is then processed where it lies) and the limiter (limit) as selected.
template <int NSP>
static int quad_nsp(const Fr2dCtx &c, double *d_u, int mode, double lambda, double S0, double kappa, int rule,
                    int limit, int *nfilt, int *nfail, hipStream_t st) {
  DissOps o = fr2d_diss_ops(c);
  // four lanes per cell, 64 places of the padded plane per workgroup (shifted by the front offset)
  hipLaunchKernelGGL(fr2d_filter_refresh_limit_kernel<NSP>, dim3((unsigned)((c.ldpi + Fr2dCtx::kPadFront + 63) / 64)), dim3(256), 0, st,
                     d_u, c.nx, c.ny, c.pitch, c.ldpi, o.V, o.iV, o.wts, o.ll, o.lr, mode, lambda, S0, kappa,
                     (const float *)c.d_strength, rule, limit, c.gamma, nfilt, nfail);
  FR_LAUNCHED("fr2d_filter_refresh_limit_kernel");
  return FR_OK;
}

static int launch_quad(const Fr2dCtx &c, double *d_u, int mode, double lambda, double S0, double kappa, int rule,
                       int limit, int *nfilt, int *nfail, hipStream_t st) {
  switch (c.nsp) {
    case 2: return quad_nsp<2>(c, d_u, mode, lambda, S0, kappa, rule, limit, nfilt, nfail, st);
    case 3: return quad_nsp<3>(c, d_u, mode, lambda, S0, kappa, rule, limit, nfilt, nfail, st);
    case 4: return quad_nsp<4>(c, d_u, mode, lambda, S0, kappa, rule, limit, nfilt, nfail, st);
    case 5: return quad_nsp<5>(c, d_u, mode, lambda, S0, kappa, rule, limit, nfilt, nfail, st);
    case 6: return quad_nsp<6>(c, d_u, mode, lambda, S0, kappa, rule, limit, nfilt, nfail, st);
    default: return fail(FR_E_UNSUPPORTED, "dissipation: degree out of range");
  }
}

// per-cell strength evaluation without touching the state (the sensor evaluation entry point)
template <int NSP>
static int sensor_eval_nsp(const Fr2dCtx &c, double *d_u, int mode, double lambda, double S0, double kappa,
                           float *eval_out, hipStream_t st) {
  const int ncell = (c.nx + 2) * (c.ny + 2);
  DissOps o = fr2d_diss_ops(c);
  hipLaunchKernelGGL(fr2d_sensor_filter_kernel<NSP>, dim3((ncell + 255) / 256), dim3(256), 0, st, d_u, ncell,
                     c.nx + 2, c.pitch, c.ldpi,
                     o.V, o.iV, mode, lambda, S0, kappa, (const float *)c.d_strength, nullptr, eval_out);
  FR_LAUNCHED("fr2d_sensor_filter_kernel");
  return FR_OK;
}

int fr2d_launch_sensor_filter(const Fr2dCtx &c, double *d_u, int mode, double lambda, double S0, double kappa,
                              int *d_count, float *eval_out, hipStream_t st) {
  if (!eval_out) return launch_quad(c, d_u, mode, lambda, S0, kappa, -1, 0, d_count, nullptr, st);
  switch (c.nsp) {
    case 2: return sensor_eval_nsp<2>(c, d_u, mode, lambda, S0, kappa, eval_out, st);
    case 3: return sensor_eval_nsp<3>(c, d_u, mode, lambda, S0, kappa, eval_out, st);
    case 4: return sensor_eval_nsp<4>(c, d_u, mode, lambda, S0, kappa, eval_out, st);
    case 5: return sensor_eval_nsp<5>(c, d_u, mode, lambda, S0, kappa, eval_out, st);
    case 6: return sensor_eval_nsp<6>(c, d_u, mode, lambda, S0, kappa, eval_out, st);
    default: return fail(FR_E_UNSUPPORTED, "sensor: degree out of range");
  }
}

int fr2d_launch_positivity(const Fr2dCtx &c, double *d_u, int *d_count, hipStream_t st) {
  return launch_quad(c, d_u, FR_SENSOR_OFF, 0.0, 0.0, 1.0, -1, 1, nullptr, d_count, st);
}

int fr2d_launch_filter_refresh_limit(const Fr2dCtx &c, double *d_u, int mode, double lambda, double S0,
                                     double kappa, int rule, int *d_count, hipStream_t st) {
  return launch_quad(c, d_u, mode, lambda, S0, kappa, rule, 1, d_count, d_count + 1, st);
}

template <int NSP>
static int check_nsp(const Fr2dCtx &c, const double *d_u, int *cnt, hipStream_t st) {
  const long long n = (long long)c.nx * c.ny;
  hipLaunchKernelGGL(fr2d_check_kernel<NSP>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_u, c.nx, c.ny,
                     c.pitch, c.ldpi, c.gamma, cnt);
  FR_LAUNCHED("fr2d_check_kernel");
  return FR_OK;
}

int fr2d_launch_check(const Fr2dCtx &c, const double *d_u, int *d_count, hipStream_t st) {
  switch (c.nsp) {
    case 2: return check_nsp<2>(c, d_u, d_count, st);
    case 3: return check_nsp<3>(c, d_u, d_count, st);
    case 4: return check_nsp<4>(c, d_u, d_count, st);
    case 5: return check_nsp<5>(c, d_u, d_count, st);
    case 6: return check_nsp<6>(c, d_u, d_count, st);
    default: return fail(FR_E_UNSUPPORTED, "check: degree out of range");
  }
}

}  // namespace fr
