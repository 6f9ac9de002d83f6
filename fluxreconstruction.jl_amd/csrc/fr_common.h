// fr_common.h -- shared host/device helpers of libfrhip (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/fr.h"

namespace fr {

// ------------------------------------------------------------------ error plumbing
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define FR_HIP(call)                                                   \
  do {                                                                 \
    hipError_t fr_e_ = (call);                                         \
    if (fr_e_ != hipSuccess) return fr::hip_fail(fr_e_, #call, __FILE__, __LINE__); \
  } while (0)

#define FR_REQUIRE(cond, code, msg)          \
  do {                                       \
    if (!(cond)) return fr::fail(code, msg); \
  } while (0)

// A kernel launch is checked with hipGetLastError right after it.
#define FR_LAUNCHED(name) FR_HIP(hipGetLastError())

// ------------------------------------------------------------------ host operator builders
// (fr_host_ops.cpp; the same routines back the fr_* operator entry points of fr.h)
void gauss_legendre(int n, double *x, double *w);
double dlegendre(int n, double x);
void dcorrection(int kind, int deg, const double *x, int n, double *dgl, double *dgr);
void lagrange_point(const double *sp, int nsp, double x, double *l);
void dlagrange(const double *sp, int nsp, double *lpdm);
double jacobi_p(double x, double alpha, double beta, int N);
double djacobi_p(double x, double alpha, double beta, int N);
bool invert(int n, const double *A, double *Ainv);

// Operators of one degree: Gauss points, ll/lr, dl(m,k) column-major, correction derivatives,
// tensor weights wp(k,l) and the 1D Vandermonde pair for degrees 1..8; the quad Vandermonde
// (nsp^2 x nsp^2) and its inverse for degrees 1..5 only (the 2D spaces; zero above).
constexpr int kMaxNp2D = 36;  // nsp^2 of the highest 2D degree (5)
struct HostOps {
  int deg = 0, nsp = 0;
  double r[16], w[16], ll[16], lr[16], dl[256], dhl[16], dhr[16];
  double V[kMaxNp2D * kMaxNp2D], iV[kMaxNp2D * kMaxNp2D], wp[256];
  double V1[256], iV1[256];
};
bool build_ops(int deg, int correction, HostOps &o);

// XCD-aware tile order: consecutive tiles land on the same XCD (workgroups are dispatched to the
// 8 XCDs round-robin), so neighbouring tiles share that XCD's L2 for their halo reads.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int per = nb >> 3, rem = nb & 7;
  const int x = bid & 7, loc = bid >> 3;
  return (x < rem) ? x * (per + 1) + loc : rem * (per + 1) + (x - rem) * per + loc;
}

// ------------------------------------------------------------------ buffer access to one plane
// A scalar buffer resource spans one plane of `bytes` bytes starting at `base`; lanes address it
// with a 32-bit byte offset, so no per-lane 64-bit address arithmetic is spent, and the hardware
// range check turns an access outside the plane into a zero load / dropped store.
typedef unsigned int fr_u2 __attribute__((ext_vector_type(2)));
constexpr int kBufferWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const double *base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), 0, (int)bytes, kBufferWord3);
}
__device__ __forceinline__ double bload(__amdgpu_buffer_rsrc_t r, unsigned byte_off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
}
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, unsigned byte_off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(fr_u2, v), r, byte_off, 0, 0);
}

// Cache policy of buffer stores whose data only a later kernel reads: agent scope (sc1), which
// writes the line through the XCD's L2.  Each XCD has its own L2; at a kernel boundary the dirty
// lines of every L2 are written back and the next kernel's acquire invalidates them, so lines
// written through at store time take that write-back off the critical path between dependent
// launches.  Measured on the 256^2 p = 3 RK stage (aligned 128-byte row segments): 19.0 ->
// 18.1 us per launch; the streaming hint (nt) instead: 21.0 us.  Rows that straddle lines (the
// reference layout of the drop-in residual) measured slower with it (20.4 -> 21.7 us).
constexpr int kStoreAgent = 16;

// Value of v held by lane src of the same wavefront (ds_bpermute: a crossbar read, no LDS
// storage).  Every lane of the wavefront must execute the call with the same control flow; src is
// taken modulo 64.
__device__ __forceinline__ double lane_fetch(double v, int src) {
  const long long b = __builtin_bit_cast(long long, v);
  const int a = src << 2;
  const int lo = __builtin_amdgcn_ds_bpermute(a, (int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_ds_bpermute(a, (int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// Orders LDS traffic between the lanes of one wavefront: LDS operations are not moved across it
// by the compiler and the wavefront's own outstanding LDS operations complete before it (the
// LDS executes one wavefront's operations in order), with no wait for the other wavefronts.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0) (vmcnt, expcnt untouched)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Marks v as produced at this point: the instructions that compute it are not sunk by the
// compiler towards a later use (past a barrier, into the next phase's register peak).
__device__ __forceinline__ void fr_pin(double &v) { asm volatile("" : "+v"(v)); }

// ------------------------------------------------------------------ FP64 reciprocal / square root
// The hardware estimates v_rcp_f64 / v_rsq_f64 refined by two Newton-Raphson (reciprocal) and one
// Goldschmidt + Newton (square root) iterations in FMA arithmetic: about one ulp from the
// correctly rounded result at roughly half the instructions of the IEEE division / sqrt
// sequences.  FR_IEEE_DIV selects the correctly rounded operations instead.
__device__ __forceinline__ double fr_rcp(double x) {
#ifndef FR_IEEE_DIV
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / x;
#endif
}

__device__ __forceinline__ double fr_sqrt(double x) {
#ifndef FR_IEEE_DIV
  const double y = __builtin_amdgcn_rsq(x);
  double s = x * y, h = 0.5 * y;
  const double r = fma(-s, h, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  const double d = fma(-s, s, x);
  return fma(d, h, s);
#else
  return sqrt(x);
#endif
}

// ------------------------------------------------------------------ Euler physics (device)
// Conserved 2D state w = [rho, rho u, rho v, E]; p = (g-1)(E - |m|^2 / (2 rho)).

// x-direction physical flux and the primitive quantities the Riemann solvers need.
struct XState {
  double F[4];
  double u, p, rinv;
};

__device__ __forceinline__ void xstate(const double *w, double gm1, XState &s) {
  s.rinv = fr_rcp(w[0]);
  s.u = w[1] * s.rinv;
  const double v = w[2] * s.rinv;
  s.p = gm1 * (w[3] - 0.5 * (w[1] * s.u + w[2] * v));
  s.F[0] = w[1];
  s.F[1] = w[1] * s.u + s.p;
  s.F[2] = w[2] * s.u;
  s.F[3] = (w[3] + s.p) * s.u;
}

// Interface flux in the x-direction for the normal-frame states wL | wR (4 variables).
// FLUX: FR_FLUX_HLL (Davis wave-speed bounds, KitBase flux_hll!), FR_FLUX_RUSANOV, FR_FLUX_ROE
// (Roe-averaged eigen-decomposition, no entropy fix).
template <int FLUX>
__device__ __forceinline__ void riemann_x4(const double *wL, const double *wR, double g, double gm1,
                                           double *fw) {
  XState L, R;
  xstate(wL, gm1, L);
  xstate(wR, gm1, R);
  if constexpr (FLUX == FR_FLUX_HLL) {
    const double cL = fr_sqrt(g * L.p * L.rinv), cR = fr_sqrt(g * R.p * R.rinv);
    const double lmin = fmin(L.u - cL, R.u - cR);
    const double lmax = fmax(L.u + cL, R.u + cR);
    const double factor = fr_rcp(lmax - lmin);
    const bool left = lmin >= 0.0, right = lmax <= 0.0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double h = factor * (lmax * L.F[m] - lmin * R.F[m] + lmax * lmin * (wR[m] - wL[m]));
      fw[m] = left ? L.F[m] : (right ? R.F[m] : h);
    }
  } else if constexpr (FLUX == FR_FLUX_RUSANOV) {
    const double cL = fr_sqrt(g * L.p * L.rinv), cR = fr_sqrt(g * R.p * R.rinv);
    const double smax = fmax(fabs(L.u) + cL, fabs(R.u) + cR);
#pragma unroll
    for (int m = 0; m < 4; ++m) fw[m] = 0.5 * (L.F[m] + R.F[m]) - 0.5 * smax * (wR[m] - wL[m]);
  } else {
    const double sL = sqrt(wL[0]), sR = sqrt(wR[0]);
    const double vL = wL[2] * L.rinv, vR = wR[2] * R.rinv;
    const double HL = (wL[3] + L.p) * L.rinv, HR = (wR[3] + R.p) * R.rinv;
    const double ws = 1.0 / (sL + sR);
    const double ut = (sL * L.u + sR * R.u) * ws;
    const double vt = (sL * vL + sR * vR) * ws;
    const double Ht = (sL * HL + sR * HR) * ws;
    const double q2 = ut * ut + vt * vt;
    const double a2 = gm1 * (Ht - 0.5 * q2);
    const double at = sqrt(a2);
    const double dr = wR[0] - wL[0], dp = R.p - L.p, du = R.u - L.u, dv = vR - vL;
    const double rt = sL * sR;
    const double ia2 = 1.0 / a2;
    const double al1 = (dp - rt * at * du) * 0.5 * ia2;
    const double al5 = (dp + rt * at * du) * 0.5 * ia2;
    const double al2 = dr - dp * ia2;
    const double al3 = rt * dv;
    const double l1 = fabs(ut - at), l2 = fabs(ut), l5 = fabs(ut + at);
    const double d0 = l1 * al1 + l2 * al2 + l5 * al5;
    const double d1 = l1 * al1 * (ut - at) + l2 * al2 * ut + l5 * al5 * (ut + at);
    const double d2 = l1 * al1 * vt + l2 * (al2 * vt + al3) + l5 * al5 * vt;
    const double d3 = l1 * al1 * (Ht - ut * at) + l2 * (al2 * 0.5 * q2 + al3 * vt) + l5 * al5 * (Ht + ut * at);
    fw[0] = 0.5 * (L.F[0] + R.F[0]) - 0.5 * d0;
    fw[1] = 0.5 * (L.F[1] + R.F[1]) - 0.5 * d1;
    fw[2] = 0.5 * (L.F[2] + R.F[2]) - 0.5 * d2;
    fw[3] = 0.5 * (L.F[3] + R.F[3]) - 0.5 * d3;
  }
}

// Root of tj_equation (src/dissipation.jl:208-215) in closed form: rho, m and E are affine in t,
// so rho(t) (p(t) - eps) = (g-1)(rho E - |m|^2/2) - eps rho is a quadratic A t^2 + B t + C; the
// largest root in [0, 1] is returned (the one NewtonRaphson reaches from t = 1).
__device__ __forceinline__ double tj_root(const double *ut, const double *um, double g, double eps) {
  const double r0 = um[0], r1 = ut[0] - um[0];
  const double x0 = um[1], x1 = ut[1] - um[1];
  const double y0 = um[2], y1 = ut[2] - um[2];
  const double e0 = um[3], e1 = ut[3] - um[3];
  const double A = (g - 1.0) * (r1 * e1 - 0.5 * (x1 * x1 + y1 * y1));
  const double B = (g - 1.0) * (r0 * e1 + r1 * e0 - (x0 * x1 + y0 * y1)) - eps * r1;
  const double C = (g - 1.0) * (r0 * e0 - 0.5 * (x0 * x0 + y0 * y0)) - eps * r0;
  double best = 0.0;
  if (fabs(A) < 1e-300) {
    if (B != 0.0) {
      const double t = -C / B;
      if (t >= 0.0 && t <= 1.0) best = t;
    }
    return best;
  }
  const double disc = B * B - 4.0 * A * C;
  if (disc < 0.0) return 0.0;
  const double sq = sqrt(disc);
  const double qq = -0.5 * (B + (B >= 0.0 ? sq : -sq));
  const double t1 = qq / A, t2 = (qq != 0.0) ? C / qq : t1;
  if (t1 >= 0.0 && t1 <= 1.0 && t1 > best) best = t1;
  if (t2 >= 0.0 && t2 <= 1.0 && t2 > best) best = t2;
  return best;
}

// lambda = rho / (2 p) of KitBase conserve_prim
__device__ __forceinline__ double prim_lambda(const double *w, double g) {
  return 0.5 * w[0] / (g - 1.0) / (w[3] - 0.5 * (w[1] * w[1] + w[2] * w[2]) / w[0]);
}

// prim_lambda(w, g) < eps with the same outcome as the three-division form, at the cost of one
// refined reciprocal on all but borderline states.  With w0 > 0 the approximate
// d = E - |m|^2/(2 rho) is within 2^-49 (|E| + |m|^2/(2 rho)) of the exactly rounded one, so
// d > 2^-24 (|E| + |m|^2/(2 rho)) gives d > 0 and fixes d to a relative 2^-25.  Then
// lambda = (rho/(2(g-1))) / d > eps for certain when d eps < rho/(2(g-1)) (1 - 2^-20), a
// margin far above every rounding error of the two sides.  Every other state (non-positive or
// non-finite values, d near 0, lambda near eps) takes the exact form.
__device__ __forceinline__ bool prim_lambda_below(const double *w, double g, double eps) {
  if (w[0] > 0.0) {
    const double ri = fr_rcp(w[0]);
    const double k = 0.5 * (w[1] * w[1] + w[2] * w[2]) * ri;
    const double d = w[3] - k;
    const double a = 0.5 * w[0] * fr_rcp(g - 1.0);
    if (d > 0x1p-24 * (fabs(w[3]) + k) && d * eps < a * (1.0 - 0x1p-20)) return false;
  }
  return prim_lambda(w, g) < eps;
}

// y-direction interface flux: rotate into the face-normal frame n = (0, 1), solve, rotate back
// (KitBase local_frame / global_frame with cos = 0, sin = 1; example/shock-vortex.jl:84-87).
// Riemann flux across a face with unit normal n = (c, s): the states go to the face frame
// (rho, m.n, m.t, E) with t = (-s, c), the x-direction solver runs there and the flux comes back
// as (f0, c f1 - s f2, s f1 + c f2, f3) (KitBase local_frame / global_frame)
template <int FLUX>
__device__ __forceinline__ void riemann_n4(const double *wL, const double *wR, double c, double s, double g,
                                           double gm1, double *fw) {
  const double a[4] = {wL[0], c * wL[1] + s * wL[2], c * wL[2] - s * wL[1], wL[3]};
  const double b[4] = {wR[0], c * wR[1] + s * wR[2], c * wR[2] - s * wR[1], wR[3]};
  double f[4];
  riemann_x4<FLUX>(a, b, g, gm1, f);
  fw[0] = f[0];
  fw[1] = c * f[1] - s * f[2];
  fw[2] = s * f[1] + c * f[2];
  fw[3] = f[3];
}

template <int FLUX>
__device__ __forceinline__ void riemann_y4(const double *wB, const double *wT, double g, double gm1,
                                           double *fw) {
  const double a[4] = {wB[0], wB[2], -wB[1], wB[3]};
  const double b[4] = {wT[0], wT[2], -wT[1], wT[3]};
  double f[4];
  riemann_x4<FLUX>(a, b, g, gm1, f);
  fw[0] = f[0];
  fw[1] = -f[2];
  fw[2] = f[1];
  fw[3] = f[3];
}

// 1D Euler (3 variables) interface flux.
template <int FLUX>
__device__ __forceinline__ void riemann_x3(const double *wL, const double *wR, double g, double gm1,
                                           double *fw) {
  const double a[4] = {wL[0], wL[1], 0.0, wL[2]};
  const double b[4] = {wR[0], wR[1], 0.0, wR[2]};
  double f[4];
  riemann_x4<FLUX>(a, b, g, gm1, f);
  fw[0] = f[0];
  fw[1] = f[1];
  fw[2] = f[3];
}

}  // namespace fr
