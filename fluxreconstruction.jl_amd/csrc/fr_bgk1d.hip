// fr_bgk1d.hip -- the 1D discrete-velocity BGK model of example/bgk_wave.jl:69-129 (mol!) as the
// script writes it, and batched ensembles of such problems that differ in the relaxation time,
// for gfx950.
//
// State u(i, j, k, member): nx periodic cells, nu velocities, nsp solution points, column-major,
// the cell fastest, so a wavefront of consecutive cells reads contiguous memory for a fixed
// (velocity, point).  Two kernels per stage:
// - moments: one or kMomSlices threads per (member, cell, point), the sums over the velocities of
//   w f, w u f and w u^2 f / 2 (KitBase moments_conserve), then the primitive variables at gamma (conserve_prim:
//   rho, U, lambda) of the point and the Maxwellian's prefactor rho sqrt(lambda / pi), so that the
//   transport kernel spends one exp per value on M;
// - transport + relaxation + RK stage: one thread per (member, cell, 8 velocities).  f = u_j u / J, the
//   traces f . ll and f . lr, the upwind interface value (the Heaviside blend of the script picks
//   the left cell's right trace for u_j > 0, the right cell's left trace otherwise; the other term
//   of the blend is the other trace times zero), the derivative, the Radau corrections and
//   (M - u) / tau with M = rho sqrt(lambda / pi) exp(-lambda (u_j - U)^2) of the point.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "fr_common.h"

namespace fr {

constexpr int kBgk1dMaxNsp = 9;  // degrees 1..8, as the 1D Euler kernels

struct Bgk1dGeo {
  int nx, nu, nsp, nmem;
  long long npts;     // nx * nsp: points of one member
  long long mstride;  // nx * nu * nsp: doubles per member
  double iJ, gamma;   // 2 / dx; gamma of conserve_prim
};

struct Bgk1dOps {
  double ll[kBgk1dMaxNsp], lr[kBgk1dMaxNsp], dl[kBgk1dMaxNsp * kBgk1dMaxNsp], dhl[kBgk1dMaxNsp],
      dhr[kBgk1dMaxNsp];
};

struct Bgk1dStage {
  const double *u, *u0;
  double *out;
  double a0, a1, b;
  int residual;
};

// prim[member][c][p], p = i + nx k: rho, U, lambda and rho sqrt(lambda / pi) of every point.
// The velocity sums run in kMomSlices interleaved slices (slice s: j = s, s + kMomSlices, ..),
// added in slice order at the end, so every launch shape gives the same bits.  A workgroup takes
// P consecutive points (of the member-major point list), lanes on consecutive cells, and S
// threads per point:
// - S = kMomSlices (P = kMomSplitPts): thread group s sums slice s, the partial sums meet in LDS
//   and group 0 adds them; for small states (a one-member state of a few ten thousand points:
//   one thread per point left three quarters of the chip idle, each with a serial chain of nu
//   loads);
// - S = 1 (P = 64): each thread sums all slices of its point itself; for large states, where one
//   thread per point already fills the chip.
constexpr int kMomSlices = 8;
constexpr int kMomSplitPts = 32;
constexpr long long kMomSplitBelow = 262144;  // points: below, S = kMomSlices

// U, lambda and rho sqrt(lambda / pi) of a point from its moments (rho, rho U, 2E) (conserve_prim
// at gamma; the Maxwellian's prefactor); shared by the streaming and the LDS-resident kernels
__device__ __forceinline__ void bgk1d_close(double m0, double m1, double m2, double gamma, double &U, double &lam,
                                            double &c0) {
  const double w2 = 0.5 * m2;
  const double l = 0.5 * m0 / (gamma - 1.0) / (w2 - 0.5 * m1 * m1 / m0);
  U = m1 / m0;
  lam = l;
  c0 = m0 * sqrt(l / M_PI);
}

template <int S, int P>
__global__ __launch_bounds__(S * P) void bgk1d_moments_kernel(Bgk1dGeo g, const double *__restrict__ f,
                                                              const double *__restrict__ uv,
                                                              const double *__restrict__ wv, double *prim) {
  static_assert(S == 1 || S == kMomSlices, "one thread per point or one per slice");
  __shared__ double part[S > 1 ? S - 1 : 1][3][P];
  const int lane = threadIdx.x % P, s = threadIdx.x / P;
  const long long t = blockIdx.x * (long long)P + lane;
  const bool act = t < g.npts * g.nmem;
  const long long mem = act ? t / g.npts : 0, p = act ? t % g.npts : 0;
  double m0 = 0.0, m1 = 0.0, m2 = 0.0;
  if (act) {
    const long long i = p % g.nx, k = p / g.nx;
    const double *fb = f + mem * g.mstride + i + (long long)g.nx * g.nu * k;
    if constexpr (S == 1) {
      double a0[kMomSlices], a1[kMomSlices], a2[kMomSlices];
#pragma unroll
      for (int q = 0; q < kMomSlices; ++q) a0[q] = a1[q] = a2[q] = 0.0;
      for (int j0 = 0; j0 < g.nu; j0 += kMomSlices) {
#pragma unroll
        for (int q = 0; q < kMomSlices; ++q) {
          const int j = j0 + q;
          if (j < g.nu) {
            const double fv = fb[(long long)g.nx * j];
            const double wj = wv[j], uj = uv[j];
            a0[q] += wj * fv;
            a1[q] += wj * uj * fv;
            a2[q] += wj * uj * uj * fv;
          }
        }
      }
      m0 = a0[0];
      m1 = a1[0];
      m2 = a2[0];
#pragma unroll
      for (int q = 1; q < kMomSlices; ++q) {
        m0 += a0[q];
        m1 += a1[q];
        m2 += a2[q];
      }
    } else {
#pragma unroll 8
      for (int j = s; j < g.nu; j += kMomSlices) {
        const double fv = fb[(long long)g.nx * j];
        const double wj = wv[j], uj = uv[j];
        m0 += wj * fv;
        m1 += wj * uj * fv;
        m2 += wj * uj * uj * fv;
      }
    }
  }
  if constexpr (S > 1) {
    if (s > 0) {
      part[s - 1][0][lane] = m0;
      part[s - 1][1][lane] = m1;
      part[s - 1][2][lane] = m2;
    }
    __syncthreads();
    if (s != 0) return;
#pragma unroll
    for (int q = 0; q < S - 1; ++q) {
      m0 += part[q][0][lane];
      m1 += part[q][1][lane];
      m2 += part[q][2][lane];
    }
  }
  if (!act) return;
  double *pr = prim + mem * 4 * g.npts + p;
  pr[0] = m0;
  bgk1d_close(m0, m1, m2, g.gamma, pr[g.npts], pr[2 * g.npts], pr[3 * g.npts]);
}

// res[k] of the transport/relaxation operator (bgk_wave.jl:84-128) at one cell and velocity vel,
// every point k: f = vel u / J, the upwind interface value (the neighbour's trace: the left cell's
// right trace for vel > 0, the right cell's left trace otherwise), derivative and Radau
// corrections, plus (M - u) / tau with M = c0 exp(-lambda (vel - U)^2).  u holds the state values
// at base + kst k (the cell) and nbase + kst k (the upwind neighbour); w receives the cell's own.
// Shared by the streaming and the LDS-resident kernels, so both give the same bits.
template <int NSP>
__device__ __forceinline__ void bgk1d_cell_velocity(const Bgk1dOps &ops, double iJ, const double *__restrict__ u,
                                                    long long base, long long nbase, long long kst, double vel,
                                                    bool pos, const double *U, const double *lam, const double *c0,
                                                    double tv, double *w, double *res) {
  double f[NSP];
#pragma unroll
  for (int k = 0; k < NSP; ++k) {
    w[k] = u[base + kst * k];
    f[k] = vel * w[k] * iJ;
  }
  double tl = 0.0, tr = 0.0;
#pragma unroll
  for (int k = 0; k < NSP; ++k) {
    tl += f[k] * ops.ll[k];
    tr += f[k] * ops.lr[k];
  }
  double nb = 0.0;
#pragma unroll
  for (int k = 0; k < NSP; ++k) nb += vel * u[nbase + kst * k] * iJ * (pos ? ops.lr[k] : ops.ll[k]);
  const double iL = pos ? nb : tl, iR = pos ? tr : nb;
#pragma unroll
  for (int k = 0; k < NSP; ++k) {
    double r1 = 0.0;
#pragma unroll
    for (int q = 0; q < NSP; ++q) r1 += f[q] * ops.dl[k + NSP * q];
    const double c = vel - U[k];
    const double M = c0[k] * exp(-lam[k] * (c * c));
    res[k] = -(r1 + (iL - tl) * ops.dhl[k] + (iR - tr) * ops.dhr[k]) + (M - w[k]) / tv;
  }
}

// a0 u0 + a1 u + b res, the stage combination (u0 read only when a0 != 0)
__device__ __forceinline__ double bgk1d_combine(double a0, double a1, double b, double w, const double *u0,
                                                long long o, double res) {
  double v = a1 * w;
  if (a0 != 0.0) v = a0 * u0[o] + v;
  return v + b * res;
}

// A thread takes one cell and kVelPerThread consecutive velocities: the point data of the cell
// (U, lambda, rho sqrt(lambda / pi)) is loaded once for all of them.
constexpr int kVelPerThread = 8;

template <int NSP>
__global__ __launch_bounds__(256) void bgk1d_rhs_kernel(Bgk1dGeo g, Bgk1dOps ops, const double *__restrict__ uv,
                                                        const double *__restrict__ tau,
                                                        const double *__restrict__ prim, Bgk1dStage sa) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int nvg = (g.nu + kVelPerThread - 1) / kVelPerThread;  // velocity groups
  const long long per = (long long)g.nx * nvg;
  if (t >= per * g.nmem) return;
  const long long mem = t / per, r = t % per;
  const int i = (int)(r % g.nx), j0 = (int)(r / g.nx) * kVelPerThread;
  const long long kst = (long long)g.nx * g.nu;  // stride of the point index
  const double *__restrict__ u = sa.u;
  const double tv = tau[mem];
  const double *pr = prim + mem * 4 * g.npts + i;
  double U[NSP], lam[NSP], c0[NSP];
#pragma unroll
  for (int k = 0; k < NSP; ++k) {
    U[k] = pr[g.npts + (long long)g.nx * k];
    lam[k] = pr[2 * g.npts + (long long)g.nx * k];
    c0[k] = pr[3 * g.npts + (long long)g.nx * k];
  }
  const int jend = min(g.nu, j0 + kVelPerThread);
  for (int j = j0; j < jend; ++j) {
    const double vel = uv[j];
    const bool pos = vel > 0.0;  // heaviside(u_j)
    const long long base = mem * g.mstride + (long long)i + (long long)g.nx * j;
    // the upwind neighbour: the left cell (u_j > 0) or the right cell, periodic
    const int in = pos ? (i == 0 ? g.nx - 1 : i - 1) : (i == g.nx - 1 ? 0 : i + 1);
    double w[NSP], res[NSP];
    bgk1d_cell_velocity<NSP>(ops, g.iJ, u, base, base + (long long)(in - i), kst, vel, pos, U, lam, c0, tv, w, res);
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      const long long o = base + kst * k;
      sa.out[o] = sa.residual ? res[k] : bgk1d_combine(sa.a0, sa.a1, sa.b, w[k], sa.u0, o, res[k]);
    }
  }
}

// ------------------------------------------------------------------ LDS-resident steps
// Small members (the script's 20 cells x 100 velocities: 6,000 values) are bound by the latency of
// two dependent launches per stage, not by HBM: a workgroup per member keeps the member's state,
// its stage arrays, the slice sums and the point data in LDS for up to kResidentSteps steps per
// launch, with a workgroup barrier after the velocity sums, the point data and the transport of
// every stage.  The arithmetic is the streaming kernels' (the same kMomSlices slices added in slice
// order, bgk1d_close, bgk1d_cell_velocity, bgk1d_combine with run-time coefficients), so both paths
// give the same bits.
constexpr int kResidentThreads = 1024;
constexpr int kResidentSteps = 256;

struct Bgk1dSteps {
  int scheme, nstage;
  double a0[3], a1[3], b[3];  // stage s: a0[s] u^n + a1[s] u_s + b[s] L(u_s)
};

__host__ __device__ inline int resident_arrays(int scheme) { return scheme == FR_RK_SSPRK3 ? 3 : 2; }

// LDS of one member: the state and stage arrays, 3 kMomSlices slice sums and U, lambda, c0 per
// point, the velocities and weights
inline size_t resident_lds_bytes(const Bgk1dGeo &g, int scheme) {
  return ((size_t)resident_arrays(scheme) * (size_t)g.mstride + (size_t)(3 * kMomSlices + 3) * (size_t)g.npts +
          2 * (size_t)g.nu) *
         sizeof(double);
}

// One stage C = a0 B + a1 A + b L(A) of the member held in LDS; every thread of the workgroup
// calls it (it ends with a barrier).  C may be B (written at the value each thread read), never A.
template <int NSP>
__device__ __forceinline__ void resident_stage(const Bgk1dGeo &g, const Bgk1dOps &ops, const double *uv,
                                               const double *wv, double tv, const double *A,
                                               const double *B, double *C, double a0, double a1, double b,
                                               double *part, double *pt) {
  const int tid = threadIdx.x;
  const int npts = (int)g.npts;
  const long long kst = (long long)g.nx * g.nu;
  for (int idx = tid; idx < kMomSlices * npts; idx += kResidentThreads) {
    const int sl = idx / npts, p = idx - sl * npts;
    const double *fb = A + (p % g.nx) + kst * (p / g.nx);
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
#pragma unroll 4
    for (int j = sl; j < g.nu; j += kMomSlices) {
      const double fv = fb[(long long)g.nx * j];
      const double wj = wv[j], uj = uv[j];
      m0 += wj * fv;
      m1 += wj * uj * fv;
      m2 += wj * uj * uj * fv;
    }
    part[(3 * sl) * npts + p] = m0;
    part[(3 * sl + 1) * npts + p] = m1;
    part[(3 * sl + 2) * npts + p] = m2;
  }
  __syncthreads();
  for (int p = tid; p < npts; p += kResidentThreads) {
    double m0 = part[p], m1 = part[npts + p], m2 = part[2 * npts + p];
#pragma unroll
    for (int q = 1; q < kMomSlices; ++q) {
      m0 += part[(3 * q) * npts + p];
      m1 += part[(3 * q + 1) * npts + p];
      m2 += part[(3 * q + 2) * npts + p];
    }
    bgk1d_close(m0, m1, m2, g.gamma, pt[p], pt[npts + p], pt[2 * npts + p]);
  }
  __syncthreads();
  for (int r = tid; r < g.nx * g.nu; r += kResidentThreads) {
    const int i = r % g.nx, j = r / g.nx;
    const double vel = uv[j];
    const bool pos = vel > 0.0;
    double U[NSP], lam[NSP], c0[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      U[k] = pt[i + g.nx * k];
      lam[k] = pt[npts + i + g.nx * k];
      c0[k] = pt[2 * npts + i + g.nx * k];
    }
    const long long base = (long long)i + (long long)g.nx * j;
    const int in = pos ? (i == 0 ? g.nx - 1 : i - 1) : (i == g.nx - 1 ? 0 : i + 1);
    double w[NSP], res[NSP];
    bgk1d_cell_velocity<NSP>(ops, g.iJ, A, base, base + (long long)(in - i), kst, vel, pos, U, lam, c0, tv, w, res);
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      const long long o = base + kst * k;
      C[o] = bgk1d_combine(a0, a1, b, w[k], B, o, res[k]);
    }
  }
  __syncthreads();
}

// nsteps steps of member blockIdx.x of `state` (read once, written once), the member held in LDS
template <int NSP>
__global__ __launch_bounds__(kResidentThreads) void bgk1d_resident_kernel(Bgk1dGeo g, Bgk1dOps ops,
                                                                          const double *__restrict__ uvg,
                                                                          const double *__restrict__ wvg,
                                                                          const double *__restrict__ tau,
                                                                          double *state, Bgk1dSteps sc, int nsteps) {
  extern __shared__ double lds[];
  const long long ms = g.mstride;
  double *X = lds, *Y = lds + ms, *Z = lds + 2 * ms;
  double *part = lds + (long long)resident_arrays(sc.scheme) * ms;
  double *pt = part + (long long)3 * kMomSlices * g.npts;
  double *uv = pt + 3 * g.npts, *wv = uv + g.nu;
  double *gs = state + (long long)blockIdx.x * ms;
  const double tv = tau[blockIdx.x];
  for (long long v = threadIdx.x; v < ms; v += kResidentThreads) X[v] = gs[v];
  for (int j = threadIdx.x; j < g.nu; j += kResidentThreads) {
    uv[j] = uvg[j];
    wv[j] = wvg[j];
  }
  __syncthreads();
  for (int n = 0; n < nsteps; ++n) {
    if (sc.scheme == FR_RK_SSPRK3) {
      resident_stage<NSP>(g, ops, uv, wv, tv, X, X, Y, sc.a0[0], sc.a1[0], sc.b[0], part, pt);
      resident_stage<NSP>(g, ops, uv, wv, tv, Y, X, Z, sc.a0[1], sc.a1[1], sc.b[1], part, pt);
      resident_stage<NSP>(g, ops, uv, wv, tv, Z, X, X, sc.a0[2], sc.a1[2], sc.b[2], part, pt);
    } else if (sc.scheme == FR_RK_MIDPOINT) {
      resident_stage<NSP>(g, ops, uv, wv, tv, X, X, Y, sc.a0[0], sc.a1[0], sc.b[0], part, pt);
      resident_stage<NSP>(g, ops, uv, wv, tv, Y, X, X, sc.a0[1], sc.a1[1], sc.b[1], part, pt);
    } else {
      resident_stage<NSP>(g, ops, uv, wv, tv, X, X, Y, sc.a0[0], sc.a1[0], sc.b[0], part, pt);
      double *t = X;
      X = Y;
      Y = t;
    }
  }
  for (long long v = threadIdx.x; v < ms; v += kResidentThreads) gs[v] = X[v];
}

// dst = src over n values (the Euler step's final copy; a kernel, so a captured graph holds no
// copy node)
static __global__ __launch_bounds__(256) void bgk1d_copy_kernel(double *__restrict__ dst,
                                                               const double *__restrict__ src, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    dst[i] = src[i];
}

struct Bgk1dCtx {
  int nx = 0, nsp = 0, nu = 0, nmem = 1;
  double dx = 1.0, gamma = 3.0;
  HostOps ops;
  size_t len = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;  // orders work on a caller stream against the context's stream
  double *d_vel = nullptr;  // u | w
  double *d_tau = nullptr, *d_prim = nullptr;
  double *d_u = nullptr, *d_s1 = nullptr, *d_s2 = nullptr, *d_in = nullptr, *d_out = nullptr;
  hipGraphExec_t gexec = nullptr;  // captured steps of fr_bgk1d_step_graph
  int path = FR_BGK1D_PATH_AUTO;
  size_t lds_max = 0;
  int g_scheme = -1, g_nsteps = 0;
  double g_dt = 0.0;

  Bgk1dGeo geo() const {
    Bgk1dGeo g;
    g.nx = nx;
    g.nu = nu;
    g.nsp = nsp;
    g.nmem = nmem;
    g.npts = (long long)nx * nsp;
    g.mstride = g.npts * nu;
    g.iJ = 2.0 / dx;
    g.gamma = gamma;
    return g;
  }

  Bgk1dOps dev_ops() const {
    Bgk1dOps o{};
    for (int q = 0; q < nsp; ++q) {
      o.ll[q] = ops.ll[q];
      o.lr[q] = ops.lr[q];
      o.dhl[q] = ops.dhl[q];
      o.dhr[q] = ops.dhr[q];
    }
    for (int q = 0; q < nsp * nsp; ++q) o.dl[q] = ops.dl[q];
    return o;
  }

  // the LDS-resident path: FR_BGK1D_PATH_AUTO takes it whenever a member's arrays fit the
  // workgroup's LDS (lds_max: 0 when the device or the kernel attribute refuses)
  bool use_resident(int scheme) const {
    return path != FR_BGK1D_PATH_STREAM && resident_lds_bytes(geo(), scheme) <= lds_max;
  }

  template <int NSP>
  int launch_resident_n(const Bgk1dSteps &sc, int n, hipStream_t st) const {
    const Bgk1dGeo g = geo();
    hipLaunchKernelGGL(bgk1d_resident_kernel<NSP>, dim3((unsigned)nmem), dim3(kResidentThreads),
                       (unsigned)resident_lds_bytes(g, sc.scheme), st, g, dev_ops(), d_vel, d_vel + nu, d_tau, d_u,
                       sc, n);
    FR_LAUNCHED("bgk1d_resident_kernel");
    return FR_OK;
  }

  int launch_resident(const Bgk1dSteps &sc, int n, hipStream_t st) const {
    switch (nsp) {
      case 2: return launch_resident_n<2>(sc, n, st);
      case 3: return launch_resident_n<3>(sc, n, st);
      case 4: return launch_resident_n<4>(sc, n, st);
      case 5: return launch_resident_n<5>(sc, n, st);
      case 6: return launch_resident_n<6>(sc, n, st);
      case 7: return launch_resident_n<7>(sc, n, st);
      case 8: return launch_resident_n<8>(sc, n, st);
      default: return launch_resident_n<9>(sc, n, st);
    }
  }

  // largest dynamic LDS the resident kernel of this degree may take on the current device
  template <int NSP>
  size_t resident_limit_n() const {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) return 0;
    if (hipFuncSetAttribute((const void *)bgk1d_resident_kernel<NSP>, hipFuncAttributeMaxDynamicSharedMemorySize, v) !=
        hipSuccess) {
      (void)hipGetLastError();
      return 0;
    }
    return (size_t)v;
  }

  size_t resident_limit() const {
    switch (nsp) {
      case 2: return resident_limit_n<2>();
      case 3: return resident_limit_n<3>();
      case 4: return resident_limit_n<4>();
      case 5: return resident_limit_n<5>();
      case 6: return resident_limit_n<6>();
      case 7: return resident_limit_n<7>();
      case 8: return resident_limit_n<8>();
      default: return resident_limit_n<9>();
    }
  }

  template <int NSP>
  int launch_rhs(const Bgk1dStage &sa, hipStream_t st) const {
    const Bgk1dOps o = dev_ops();
    const Bgk1dGeo g = geo();
    const long long n = (long long)nx * ((nu + kVelPerThread - 1) / kVelPerThread) * nmem;
    hipLaunchKernelGGL(bgk1d_rhs_kernel<NSP>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, o, d_vel,
                       d_tau, d_prim, sa);
    FR_LAUNCHED("bgk1d_rhs_kernel");
    return FR_OK;
  }

  int launch_moments(const double *d_f, hipStream_t st) const {
    const Bgk1dGeo g = geo();
    const long long n = g.npts * nmem;
    if (n < kMomSplitBelow)
      hipLaunchKernelGGL((bgk1d_moments_kernel<kMomSlices, kMomSplitPts>),
                         dim3((unsigned)((n + kMomSplitPts - 1) / kMomSplitPts)), dim3(kMomSlices * kMomSplitPts), 0,
                         st, g, d_f, d_vel, d_vel + nu, d_prim);
    else
      hipLaunchKernelGGL((bgk1d_moments_kernel<1, 64>), dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, g, d_f,
                         d_vel, d_vel + nu, d_prim);
    FR_LAUNCHED("bgk1d_moments_kernel");
    return FR_OK;
  }

  int stage(const Bgk1dStage &sa, hipStream_t st) const {
    int rc = launch_moments(sa.u, st);
    if (rc) return rc;
    switch (nsp) {
      case 2: return launch_rhs<2>(sa, st);
      case 3: return launch_rhs<3>(sa, st);
      case 4: return launch_rhs<4>(sa, st);
      case 5: return launch_rhs<5>(sa, st);
      case 6: return launch_rhs<6>(sa, st);
      case 7: return launch_rhs<7>(sa, st);
      case 8: return launch_rhs<8>(sa, st);
      default: return launch_rhs<9>(sa, st);
    }
  }

  // nsteps explicit steps of every member, enqueued on st: FR_RK_SSPRK3 (Shu-Osher),
  // FR_RK_MIDPOINT (OrdinaryDiffEq Midpoint, bgk_wave.jl:136) or FR_RK_EULER; each stage =
  // moments + transport/relaxation.  Consecutive steps alternate the state between d_u and d_s1
  // (every stage writes every value), so no stage writes an array whose neighbours it reads: the
  // last stage of a step reads u^n only at its own value.  The last step writes d_u (Euler: copied
  // back when it ends in d_s1).  Stage buffers must exist.
  int enqueue_steps(int scheme, double dt, int nsteps, hipStream_t st) const {
    if (use_resident(scheme)) {
      Bgk1dSteps sc{};
      sc.scheme = scheme;
      if (scheme == FR_RK_SSPRK3) {
        sc.nstage = 3;
        const double a0[3] = {0.0, 0.75, 1.0 / 3.0}, a1[3] = {1.0, 0.25, 2.0 / 3.0},
                     b[3] = {dt, 0.25 * dt, 2.0 / 3.0 * dt};
        for (int q = 0; q < 3; ++q) sc.a0[q] = a0[q], sc.a1[q] = a1[q], sc.b[q] = b[q];
      } else if (scheme == FR_RK_MIDPOINT) {
        sc.nstage = 2;
        sc.a0[0] = 0.0, sc.a1[0] = 1.0, sc.b[0] = 0.5 * dt;
        sc.a0[1] = 1.0, sc.a1[1] = 0.0, sc.b[1] = dt;
      } else {
        sc.nstage = 1;
        sc.a0[0] = 0.0, sc.a1[0] = 1.0, sc.b[0] = dt;
      }
      for (int done = 0; done < nsteps;) {
        const int n = std::min(nsteps - done, kResidentSteps);
        int rc = launch_resident(sc, n, st);
        if (rc) return rc;
        done += n;
      }
      return FR_OK;
    }
    for (int n = 0; n < nsteps; ++n) {
      int rc = FR_OK;
      double *const X = (n & 1) ? d_s1 : d_u, *const Y = (n & 1) ? d_u : d_s1;
      double *const out = n == nsteps - 1 ? d_u : Y;
      if (scheme == FR_RK_SSPRK3) {
        if ((rc = stage(Bgk1dStage{X, X, Y, 0.0, 1.0, dt, 0}, st))) return rc;
        if ((rc = stage(Bgk1dStage{Y, X, d_s2, 0.75, 0.25, 0.25 * dt, 0}, st))) return rc;
        if ((rc = stage(Bgk1dStage{d_s2, X, out, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt, 0}, st))) return rc;
      } else if (scheme == FR_RK_MIDPOINT) {
        if ((rc = stage(Bgk1dStage{X, X, d_s2, 0.0, 1.0, 0.5 * dt, 0}, st))) return rc;
        if ((rc = stage(Bgk1dStage{d_s2, X, out, 1.0, 0.0, dt, 0}, st))) return rc;
      } else {
        if ((rc = stage(Bgk1dStage{X, X, Y, 0.0, 1.0, dt, 0}, st))) return rc;
        if (n == nsteps - 1 && Y != d_u) {
          const long long blocks = std::min<long long>(((long long)len + 255) / 256, 4096);
          hipLaunchKernelGGL(bgk1d_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, st, d_u, Y, (long long)len);
          FR_LAUNCHED("bgk1d_copy_kernel");
        }
      }
    }
    return FR_OK;
  }

  int ensure_stage_buffers() {
    if (!d_s1) FR_HIP(hipMalloc(&d_s1, len * sizeof(double)));
    if (!d_s2) FR_HIP(hipMalloc(&d_s2, len * sizeof(double)));
    return FR_OK;
  }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    double *bufs[] = {d_vel, d_tau, d_prim, d_u, d_s1, d_s2, d_in, d_out};
    for (double *b : bufs)
      if (b) (void)hipFree(b);
    if (ev) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace fr

using namespace fr;

struct fr_bgk1d_ctx : Bgk1dCtx {};

extern "C" {

int fr_bgk1d_create(fr_bgk1d_ctx **out, const fr_bgk1d_desc *d) {
  FR_REQUIRE(out && d, FR_E_INVALID, "fr_bgk1d_create: null argument");
  *out = nullptr;
  FR_REQUIRE(d->nx >= 1 && d->nu >= 1 && d->nmem >= 1, FR_E_INVALID, "fr_bgk1d_create: sizes must be >= 1");
  FR_REQUIRE(d->deg >= 1 && d->deg <= kBgk1dMaxNsp - 1, FR_E_UNSUPPORTED, "fr_bgk1d_create: degree must be in 1..8");
  FR_REQUIRE(d->dx > 0.0 && std::isfinite(d->dx) && d->gamma > 1.0, FR_E_INVALID,
             "fr_bgk1d_create: dx > 0, gamma > 1");
  FR_REQUIRE(d->u && d->w && d->tau, FR_E_INVALID, "fr_bgk1d_create: null velocity grid or tau");
  for (int m = 0; m < d->nmem; ++m) FR_REQUIRE(d->tau[m] > 0.0, FR_E_INVALID, "fr_bgk1d_create: tau must be > 0");
  FR_REQUIRE((double)d->nx * d->nu * d->nmem < 2147483647.0 * 256.0, FR_E_INVALID, "fr_bgk1d_create: too large");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FR_E_HIP, "fr_bgk1d_create: no HIP device visible");
  fr_bgk1d_ctx *c = new (std::nothrow) fr_bgk1d_ctx();
  FR_REQUIRE(c, FR_E_NOMEM, "fr_bgk1d_create: out of host memory");
  c->nx = d->nx;
  c->nsp = d->deg + 1;
  c->nu = d->nu;
  c->nmem = d->nmem;
  c->dx = d->dx;
  c->gamma = d->gamma;
  if (!build_ops(d->deg, FR_CORR_RADAU, c->ops)) {
    delete c;
    return fail(FR_E_INVALID, "fr_bgk1d_create: operator construction failed");
  }
  const Bgk1dGeo g = c->geo();
  c->len = (size_t)g.mstride * d->nmem;
  std::vector<double> vel(d->u, d->u + d->nu);
  vel.insert(vel.end(), d->w, d->w + d->nu);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&c->d_vel, vel.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(c->d_vel, vel.data(), vel.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->d_tau, d->nmem * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(c->d_tau, d->tau, d->nmem * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->d_prim, (size_t)g.npts * 4 * d->nmem * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->d_u, c->len * sizeof(double));
  if (e != hipSuccess) {
    int rc = hip_fail(e, "fr_bgk1d_create", __FILE__, __LINE__);
    c->release();
    delete c;
    return rc;
  }
  c->lds_max = c->resident_limit();
  *out = c;
  return FR_OK;
}

int fr_bgk1d_destroy(fr_bgk1d_ctx *c) {
  if (!c) return FR_OK;
  c->release();
  delete c;
  return FR_OK;
}

size_t fr_bgk1d_state_len(const fr_bgk1d_ctx *c) { return c ? c->len : 0; }

int fr_bgk1d_rhs(fr_bgk1d_ctx *c, const double *f, double *df) {
  FR_REQUIRE(c && f && df, FR_E_INVALID, "fr_bgk1d_rhs: null argument");
  const size_t bytes = c->len * sizeof(double);
  if (!c->d_in) FR_HIP(hipMalloc(&c->d_in, bytes));
  if (!c->d_out) FR_HIP(hipMalloc(&c->d_out, bytes));
  FR_HIP(hipMemcpyAsync(c->d_in, f, bytes, hipMemcpyHostToDevice, c->stream));
  int rc = c->stage(Bgk1dStage{c->d_in, nullptr, c->d_out, 0.0, 0.0, 0.0, 1}, c->stream);
  if (rc) return rc;
  FR_HIP(hipMemcpyAsync(df, c->d_out, bytes, hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr_bgk1d_rhs_dev(fr_bgk1d_ctx *c, const double *d_f, double *d_df, void *stream) {
  FR_REQUIRE(c && d_f && d_df, FR_E_INVALID, "fr_bgk1d_rhs_dev: null argument");
  FR_REQUIRE(d_f != d_df, FR_E_INVALID, "fr_bgk1d_rhs_dev: u and du must not alias");
  // the moments buffer belongs to the context: the caller's stream waits for the context's
  // queued work, and the context's later work waits for this call
  hipStream_t st = (hipStream_t)stream;
  FR_HIP(hipEventRecord(c->ev, c->stream));
  FR_HIP(hipStreamWaitEvent(st, c->ev, 0));
  int rc = c->stage(Bgk1dStage{d_f, nullptr, d_df, 0.0, 0.0, 0.0, 1}, st);
  if (rc) return rc;
  FR_HIP(hipEventRecord(c->ev, st));
  FR_HIP(hipStreamWaitEvent(c->stream, c->ev, 0));
  return FR_OK;
}

int fr_bgk1d_set_state(fr_bgk1d_ctx *c, const double *f) {
  FR_REQUIRE(c && f, FR_E_INVALID, "fr_bgk1d_set_state: null argument");
  FR_HIP(hipMemcpyAsync(c->d_u, f, c->len * sizeof(double), hipMemcpyHostToDevice, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr_bgk1d_get_state(fr_bgk1d_ctx *c, double *f) {
  FR_REQUIRE(c && f, FR_E_INVALID, "fr_bgk1d_get_state: null argument");
  FR_HIP(hipMemcpyAsync(f, c->d_u, c->len * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr_bgk1d_set_path(fr_bgk1d_ctx *c, int path) {
  FR_REQUIRE(c, FR_E_INVALID, "fr_bgk1d_set_path: null context");
  FR_REQUIRE(path >= FR_BGK1D_PATH_AUTO && path <= FR_BGK1D_PATH_RESIDENT, FR_E_INVALID,
             "fr_bgk1d_set_path: path must be FR_BGK1D_PATH_AUTO, _STREAM or _RESIDENT");
  if (path != c->path && c->gexec) {  // a captured graph holds the launches of its path
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  c->path = path;
  return FR_OK;
}

int fr_bgk1d_get_path(const fr_bgk1d_ctx *c, int scheme, int *path) {
  FR_REQUIRE(c && path, FR_E_INVALID, "fr_bgk1d_get_path: null argument");
  *path = c->use_resident(scheme) ? FR_BGK1D_PATH_RESIDENT : FR_BGK1D_PATH_STREAM;
  return FR_OK;
}

static int check_step_args(const fr_bgk1d_ctx *c, int scheme, double dt, const char *who) {
  FR_REQUIRE(c && std::isfinite(dt), FR_E_INVALID, who);
  FR_REQUIRE(scheme == FR_RK_SSPRK3 || scheme == FR_RK_MIDPOINT || scheme == FR_RK_EULER, FR_E_UNSUPPORTED,
             "fr_bgk1d_step: schemes are SSP-RK3, midpoint and Euler");
  FR_REQUIRE(c->path != FR_BGK1D_PATH_RESIDENT || c->use_resident(scheme), FR_E_UNSUPPORTED,
             "fr_bgk1d_step: a member's arrays exceed the workgroup LDS of the resident path");
  return FR_OK;
}

int fr_bgk1d_step(fr_bgk1d_ctx *c, int scheme, double dt, int nsteps) {
  int rc = check_step_args(c, scheme, dt, "fr_bgk1d_step: bad arguments");
  if (rc) return rc;
  FR_REQUIRE(nsteps >= 0, FR_E_INVALID, "fr_bgk1d_step: bad arguments");
  if ((rc = c->ensure_stage_buffers())) return rc;
  if ((rc = c->enqueue_steps(scheme, dt, nsteps, c->stream))) return rc;
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

// nrep replays of a graph of nsteps steps, captured on the first call and kept while scheme, dt
// and nsteps are unchanged (it holds the context's own buffers, which never move once
// allocated).  Each replay ends with the state in d_u, so nrep replays give the bits of
// fr_bgk1d_step over nrep * nsteps steps.
int fr_bgk1d_step_graph(fr_bgk1d_ctx *c, int scheme, double dt, int nsteps, int nrep) {
  int rc = check_step_args(c, scheme, dt, "fr_bgk1d_step_graph: bad arguments");
  if (rc) return rc;
  FR_REQUIRE(nsteps >= 1 && nrep >= 0, FR_E_INVALID, "fr_bgk1d_step_graph: nsteps >= 1, nrep >= 0");
  if ((rc = c->ensure_stage_buffers())) return rc;
  hipStream_t st = c->stream;
  if (!(c->gexec && c->g_scheme == scheme && c->g_dt == dt && c->g_nsteps == nsteps)) {
    if (c->gexec) {
      (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
    hipGraph_t graph = nullptr;
    FR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = c->enqueue_steps(scheme, dt, nsteps, st);
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    FR_HIP(e);
    e = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      c->gexec = nullptr;
      return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    c->g_scheme = scheme;
    c->g_dt = dt;
    c->g_nsteps = nsteps;
  }
  for (int r = 0; r < nrep; ++r) FR_HIP(hipGraphLaunch(c->gexec, st));
  FR_HIP(hipStreamSynchronize(st));
  return FR_OK;
}

// conserved moments (rho, rho U, E) of the device state at every point, host output
// w[member][c][p], p = i + nx k
int fr_bgk1d_moments(fr_bgk1d_ctx *c, double *w) {
  FR_REQUIRE(c && w, FR_E_INVALID, "fr_bgk1d_moments: null argument");
  const Bgk1dGeo g = c->geo();
  int rc = c->launch_moments(c->d_u, c->stream);
  if (rc) return rc;
  std::vector<double> prim((size_t)g.npts * 4 * c->nmem);
  FR_HIP(hipMemcpyAsync(prim.data(), c->d_prim, prim.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  for (long long mem = 0; mem < c->nmem; ++mem)
    for (long long p = 0; p < g.npts; ++p) {
      const double *pr = prim.data() + mem * 4 * g.npts + p;
      const double rho = pr[0], U = pr[g.npts], lam = pr[2 * g.npts];
      double *o = w + mem * 3 * g.npts + p;
      o[0] = rho;
      o[g.npts] = rho * U;
      o[2 * g.npts] = 0.5 * rho / lam / (c->gamma - 1.0) + 0.5 * rho * U * U;
    }
  return FR_OK;
}

}  // extern "C"
