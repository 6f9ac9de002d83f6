// fr1d.hip -- 1D FR residuals (src/Equation/eq_euler.jl:29-98, eq_advection.jl:153-273,
// eq_scalar.jl:1-12) and batched ensembles of independent 1D Euler problems, for gfx950.
//
// One kernel body handles a chunk of consecutive elements of one problem: lane t of a workgroup
// of T threads owns element first-1+t, so lanes 0 and T-1 are halo elements that only provide
// their face traces; the T-2 interior lanes produce du.  Traces and interface fluxes go through
// LDS.  A single problem is a work list of its chunks; an ensemble is the concatenated work list
// of all members (any degree 1..4 per member) processed by a persistent grid.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fr_common.h"
#include "fr_tsit5.h"

namespace fr {

constexpr int NSP_MAX = 9;  // degrees 1..8 (example/euler_highlevel.jl runs degree 7)
// device operator table: entry nsp (2..NSP_MAX) at OPS_STRIDE*(nsp-2):
// ll lr dhl dhr [NSP_MAX each] dl [NSP_MAX^2] | w/2 [NSP_MAX] (Gauss weights summing to one)
// V1 iV1 [NSP_MAX^2 each]
constexpr int OFF_W = 4 * NSP_MAX + NSP_MAX * NSP_MAX;
constexpr int OFF_V = OFF_W + NSP_MAX, OFF_IV = OFF_V + NSP_MAX * NSP_MAX;
constexpr int OPS_STRIDE = OFF_IV + NSP_MAX * NSP_MAX;
struct Prob1D {
  long long off;  // first double of the member in the concatenated state
  int ncell, nsp;
  double iJ;      // 2/dx
};

struct Work1D {
  int prob, first;  // member, first interior element of the chunk
};

struct Stage1D {
  const double *u;
  double *out;
  const double *u0;
  const double *acc_in;
  double *acc_out;
  double a0, a1, b, b2;
  int residual;
};

struct Phys1D {
  int eq;       // FR_EQ_EULER / FR_EQ_ADVECTION
  int bc;       // FR_BC1_*
  double g, gm1, a;
};

template <int NSP, int EQ, int FLUX>
__device__ __forceinline__ void chunk1d(const Prob1D &pb, int first, const double *__restrict__ ops,
                                        const Phys1D &ph, const Stage1D &sa, double *sh) {
  constexpr int NV = (EQ == FR_EQ_EULER) ? 3 : 1;
  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int n = pb.ncell;
  const double *ll = ops, *lr = ops + NSP_MAX, *dhl = ops + 2 * NSP_MAX, *dhr = ops + 3 * NSP_MAX,
               *dl = ops + 4 * NSP_MAX;
  // shared layout: uR[NV][T] | fR[NV][T] | Fs[NV][T]
  double *sUR = sh, *sFR = sh + NV * T, *sFS = sh + 2 * NV * T;
  int e = first - 1 + tid;
  bool valid = true;
  if (e < 0 || e >= n) {
    // only the two neighbours of the domain wrap (periodic); lanes further out are unused
    if (ph.bc == FR_BC1_PERIOD && (e == -1 || e == n))
      e = (e < 0) ? e + n : e - n;
    else
      valid = false;
  }
  const bool interior = tid >= 1 && tid <= T - 2 && (first - 1 + tid) < n;
  const double *__restrict__ u = sa.u + pb.off;
  double w[NSP][NV];
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int p = 0; p < NSP; ++p) w[p][k] = valid ? u[e + (long long)n * (p + NSP * k)] : (k == 0 ? 1.0 : (k == 2 ? 2.5 : 0.0));
  // discontinuous flux f = F(u) / J, traces of u and f
  double f[NSP][NV], uL[NV], uR[NV], fL[NV], fR[NV];
#pragma unroll
  for (int p = 0; p < NSP; ++p) {
    if constexpr (EQ == FR_EQ_EULER) {
      const double r = 1.0 / w[p][0];
      const double vel = w[p][1] * r;
      const double pr = ph.gm1 * (w[p][2] - 0.5 * w[p][1] * vel);
      f[p][0] = w[p][1] * pb.iJ;
      f[p][1] = (w[p][1] * vel + pr) * pb.iJ;
      f[p][2] = ((w[p][2] + pr) * vel) * pb.iJ;
    } else {
      f[p][0] = ph.a * w[p][0] * pb.iJ;
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double a = 0, b = 0, c = 0, d = 0;
#pragma unroll
    for (int q = 0; q < NSP; ++q) {
      a += w[q][k] * ll[q];
      b += w[q][k] * lr[q];
      c += f[q][k] * ll[q];
      d += f[q][k] * lr[q];
    }
    uL[k] = a;
    uR[k] = b;
    fL[k] = c;
    fR[k] = d;
    sUR[k * T + tid] = b;
    sFR[k * T + tid] = d;
  }
  __syncthreads();
  // interface flux on the left face of the own element (face between e-1 and e)
  if (tid >= 1 && valid) {
    const bool nb_ok = (ph.bc == FR_BC1_PERIOD) || (first - 2 + tid) >= 0;
    if (nb_ok) {
      if constexpr (EQ == FR_EQ_EULER) {
        double wl[3], fw[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) wl[k] = sUR[k * T + tid - 1];
        riemann_x3<FLUX>(wl, uL, ph.g, ph.gm1, fw);
        // the reference divides the interface flux by J in the correction
#pragma unroll
        for (int k = 0; k < 3; ++k) sFS[k * T + tid] = fw[k] * pb.iJ;
      } else {
        // regularised local Lax-Friedrichs of eq_advection.jl:128-137 (+1e-6 on the periodic face
        // :163); f is already divided by J
        const double fl = sFR[tid - 1], ul = sUR[tid - 1];
        const int ge = first - 1 + tid;  // global element right of the face
        const double eps = (ge == 0 || ge == n) ? 1e-6 : 1e-8;
        const double au = (fL[0] - fl) / (uL[0] - ul + eps);
        sFS[tid] = 0.5 * (fL[0] + fl) - 0.5 * fabs(au) * (uL[0] - ul);
      }
    }
  }
  __syncthreads();
  if (interior) {
    const int ge = first - 1 + tid;
    const bool zero = (ph.bc == FR_BC1_DIRICHLET) && (ge == 0 || ge == n - 1);
    double *__restrict__ out = sa.out + pb.off;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double cl = sFS[k * T + tid] - fL[k];
      const double cr = sFS[k * T + tid + 1] - fR[k];
#pragma unroll
      for (int p = 0; p < NSP; ++p) {
        double d = 0.0;
#pragma unroll
        for (int q = 0; q < NSP; ++q) d += f[q][k] * dl[p + NSP * q];
        double du = -(d + cl * dhl[p] + cr * dhr[p]);
        if (zero) du = 0.0;
        const long long idx = ge + (long long)n * (p + NSP * k);
        if (sa.residual) {
          out[idx] = du;
        } else {
          double v = sa.a1 * w[p][k];
          if (sa.a0 != 0.0) v = sa.a0 * sa.u0[pb.off + idx] + v;
          out[idx] = v + sa.b * du;
          if (sa.acc_out) sa.acc_out[pb.off + idx] = sa.acc_in[pb.off + idx] + sa.b2 * du;
        }
      }
    }
  }
  __syncthreads();  // the shared buffers are reused by the next chunk of a persistent loop
}

// NMAX: the largest nsp the instance handles.  Problem sets of degree <= 4 use the NMAX = 5
// instance, whose register budget is that of those degrees (119 VGPRs for Euler); sets with
// higher degrees (up to 8) the NMAX = NSP_MAX one (179 VGPRs, fewer waves per SIMD).
template <int EQ, int FLUX, int NMAX>
__device__ __forceinline__ void stage_chunk(const Prob1D &pb, int first, const double *__restrict__ ops,
                                            const Phys1D &ph, const Stage1D &sa, double *sh) {
  const double *o = ops + OPS_STRIDE * (pb.nsp - 2);
  switch (pb.nsp) {
    case 2: chunk1d<2, EQ, FLUX>(pb, first, o, ph, sa, sh); break;
    case 3: chunk1d<3, EQ, FLUX>(pb, first, o, ph, sa, sh); break;
    case 4: chunk1d<4, EQ, FLUX>(pb, first, o, ph, sa, sh); break;
    case 5: chunk1d<5, EQ, FLUX>(pb, first, o, ph, sa, sh); break;
    default:
      if constexpr (NMAX > 5) {
        switch (pb.nsp) {
          case 6: chunk1d<6, EQ, FLUX>(pb, first, o, ph, sa, sh); break;
          case 7: chunk1d<7, EQ, FLUX>(pb, first, o, ph, sa, sh); break;
          case 8: chunk1d<8, EQ, FLUX>(pb, first, o, ph, sa, sh); break;
          default: chunk1d<9, EQ, FLUX>(pb, first, o, ph, sa, sh); break;
        }
      }
      break;
  }
}

template <int EQ, int FLUX, int NMAX>
__global__ __launch_bounds__(256) void fr1d_kernel(const Prob1D *__restrict__ probs, const Work1D *__restrict__ work,
                                                   int nwork, const double *__restrict__ ops, Phys1D ph, Stage1D sa) {
  constexpr int NV = (EQ == FR_EQ_EULER) ? 3 : 1;
  __shared__ double sh[3 * NV * 256];
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const Work1D wk = work[wi];
    stage_chunk<EQ, FLUX, NMAX>(probs[wk.prob], wk.first, ops, ph, sa, sh);
  }
}

// ------------------------------------------------------------------ 1D dissipation
// Per-cell dissipation of 1D Euler states, one thread per cell over the same work lists:
//  mode 1: modal sensor + filter of example/euler_highlevel.jl:37-52 (rho modes u_hat = iV rho,
//          su = u_hat[end]^2 / (sum u_hat^2 + 1e-6), shock_detector(log10 su, deg) with
//          S0 = -3 log10(deg), kappa = 4; on a shock all variables filtered with
//          lambda = dt exp(0.875 deg) 0.15, u_hat[n] /= 1 + lambda (n+1)^2 n^2);
//  mode 2: positivity limiter of src/dissipation.jl:61-123 (density squeeze toward the cell
//          mean, then the pressure squeeze with the closed-form tj_equation root).
// count[0] += flagged (mode 1) / limited (mode 2) cells, count[1] += cells whose density
// assert failed (left untouched).
template <int NSP>
__device__ __forceinline__ void diss1d_cell(double *__restrict__ u, long long n, long long e, const double *ops,
                                            int mode, double g, double dt, bool &flag, bool &bad) {
  const double *ll = ops, *lr = ops + NSP_MAX, *w = ops + OFF_W, *V = ops + OFF_V, *iV = ops + OFF_IV;
  double x[NSP][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int p = 0; p < NSP; ++p) x[p][k] = u[e + n * (p + NSP * k)];
  flag = false;
  bad = false;
  if (mode == 1) {
    double den = 0.0, last = 0.0;
#pragma unroll
    for (int a = 0; a < NSP; ++a) {
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < NSP; ++b) s += iV[a + NSP * b] * x[b][0];
      den += s * s;
      if (a == NSP - 1) last = s * s;
    }
    const int deg = NSP - 1;
    const double S0 = -3.0 * log10((double)deg), kappa = 4.0;
    const double Se = log10(last / (den + 1e-6));
    double sigma;
    if (Se < S0 - kappa)
      sigma = 1.0;
    else if (Se < S0 + kappa)
      sigma = 0.5 * (1.0 - sin(0.5 * M_PI * (Se - S0) / kappa));
    else
      sigma = 0.0;
    if (!(sigma < 0.99)) return;
    flag = true;
    const double lambda = dt * exp(0.875 * deg) * 0.15;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double mh[NSP];
#pragma unroll
      for (int a = 0; a < NSP; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NSP; ++b) s += iV[a + NSP * b] * x[b][k];
        mh[a] = a == 0 ? s : s / (1.0 + lambda * (double)((a + 1) * (a + 1) * a * a));
      }
#pragma unroll
      for (int a = 0; a < NSP; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NSP; ++b) s += V[a + NSP * b] * mh[b];
        u[e + n * (a + NSP * k)] = s;
      }
    }
    return;
  }
  // positivity limiter
  double um[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < NSP; ++p) s += x[p][k] * w[p];
    um[k] = s;
  }
  const double um4[4] = {um[0], um[1], 0.0, um[2]};
  const double p_mean = 0.5 * um[0] / prim_lambda(um4, g);
  double bv[2][4];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int p = 0; p < NSP; ++p) {
      a += x[p][k] * ll[p];
      b += x[p][k] * lr[p];
    }
    bv[0][k < 2 ? k : 3] = a;
    bv[1][k < 2 ? k : 3] = b;
  }
  bv[0][2] = bv[1][2] = 0.0;
  const double eps = fmin(1e-13, fmin(um[0], p_mean));
  double rmin = fmin(bv[0][0], bv[1][0]);
#pragma unroll
  for (int p = 0; p < NSP; ++p) rmin = fmin(rmin, x[p][0]);
  const double t1 = fmin((um[0] - eps) / (um[0] - rmin + 1e-8), 1.0);
  if (!(t1 > 0.0 && t1 <= 1.0)) {
    bad = true;
    return;
  }
  bool have = false;
  double tmin = 1.0e300;
#pragma unroll
  for (int f = 0; f < 2; ++f)
    if (prim_lambda_below(bv[f], g, eps)) {
      have = true;
      tmin = fmin(tmin, tj_root(bv[f], um4, g, eps));
    }
#pragma unroll
  for (int p = 0; p < NSP; ++p) {
    x[p][0] = t1 * (x[p][0] - um[0]) + um[0];
    const double w4[4] = {x[p][0], x[p][1], 0.0, x[p][2]};
    if (prim_lambda_below(w4, g, eps)) {
      have = true;
      tmin = fmin(tmin, tj_root(w4, um4, g, eps));
    }
  }
  const double t2 = have ? fmin(tmin, 1.0) : 1.0;
  flag = t1 < 1.0 || have;
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int p = 0; p < NSP; ++p) {
      const double v = have ? t2 * (x[p][k] - um[k]) + um[k] : x[p][k];
      u[e + n * (p + NSP * k)] = v;
    }
}

// the dissipation pass over the interior elements of one work item
template <int NMAX>
__device__ __forceinline__ void diss_chunk(const Prob1D &pb, const Work1D &wk, const double *__restrict__ ops,
                                           double *u, int mode, double g, double dt, int *count) {
  {
    const int tid = threadIdx.x, T = blockDim.x;
    const int e = wk.first - 1 + tid;
    const bool interior = tid >= 1 && tid <= T - 2 && e < pb.ncell;
    bool flag = false, bad = false;
    if (interior) {
      const double *o = ops + OPS_STRIDE * (pb.nsp - 2);
      double *um = u + pb.off;
      switch (pb.nsp) {
        case 2: diss1d_cell<2>(um, pb.ncell, e, o, mode, g, dt, flag, bad); break;
        case 3: diss1d_cell<3>(um, pb.ncell, e, o, mode, g, dt, flag, bad); break;
        case 4: diss1d_cell<4>(um, pb.ncell, e, o, mode, g, dt, flag, bad); break;
        case 5: diss1d_cell<5>(um, pb.ncell, e, o, mode, g, dt, flag, bad); break;
        default:
          if constexpr (NMAX > 5) {
            switch (pb.nsp) {
              case 6: diss1d_cell<6>(um, pb.ncell, e, o, mode, g, dt, flag, bad); break;
              case 7: diss1d_cell<7>(um, pb.ncell, e, o, mode, g, dt, flag, bad); break;
              case 8: diss1d_cell<8>(um, pb.ncell, e, o, mode, g, dt, flag, bad); break;
              default: diss1d_cell<9>(um, pb.ncell, e, o, mode, g, dt, flag, bad); break;
            }
          }
          break;
      }
    }
    const unsigned long long bf = __ballot(flag), bb = __ballot(bad);
    if ((tid & 63) == 0) {
      if (bf) atomicAdd(count, __popcll(bf));
      if (bb) atomicAdd(count + 1, __popcll(bb));
    }
  }
}

template <int NMAX>
__global__ __launch_bounds__(256) void fr1d_diss_kernel(const Prob1D *__restrict__ probs, const Work1D *__restrict__ work,
                                                        int nwork, const double *__restrict__ ops, double *u, int mode,
                                                        double g, double dt, int *count) {
  for (int wi = blockIdx.x; wi < nwork; wi += gridDim.x) {
    const Work1D wk = work[wi];
    diss_chunk<NMAX>(probs[wk.prob], wk, ops, u, mode, g, dt, count);
  }
}

// ------------------------------------------------------------------ per-member stepping
// Small problem sets (the reference's example runs: one problem of 100 cells is one work item)
// are bound by the latency of a launch per stage and per dissipation pass.  Here a workgroup takes
// one member and runs a whole schedule of operations on it: each stage over the member's work
// items in turn (stage_chunk, the batched kernel's code), each dissipation pass (diss_chunk) and
// each copy, with a workgroup barrier after every operation, so the next one sees the whole
// member's result.  The schedule is staged through LDS 32 operations at a time (one coalesced
// read instead of a dependent scalar load per operation).  The bits are the batched launches'.
struct Op1D {
  int kind;     // kOpStage, kOpDiss, kOpCopy
  int dmode;    // kOpDiss: 1 sensor + filter, 2 positivity limiter
  double ddt;   // kOpDiss: the filter time step
  Stage1D s;    // kOpStage
  double *dst;  // kOpDiss: the state; kOpCopy: the destination
  const double *src;  // kOpCopy
};
constexpr int kOpStage = 0, kOpDiss = 1, kOpCopy = 2;
constexpr int kMemberOpsPerLaunch = 3072;

template <int EQ, int FLUX, int NMAX>
__global__ __launch_bounds__(256) void fr1d_member_kernel(const Prob1D *__restrict__ probs,
                                                          const int *__restrict__ mcount,
                                                          const double *__restrict__ ops, Phys1D ph,
                                                          const Op1D *__restrict__ sched, int nops, int *count) {
  constexpr int NV = (EQ == FR_EQ_EULER) ? 3 : 1;
  constexpr int kWin = 32;                               // operations staged in LDS at a time
  constexpr int kOpWords = (int)(sizeof(Op1D) / sizeof(unsigned long long));
  static_assert(sizeof(Op1D) % sizeof(unsigned long long) == 0, "Op1D is copied as 8-byte words");
  __shared__ double sh[3 * NV * 256];
  __shared__ unsigned long long win[kWin * kOpWords];
  const Prob1D pb = probs[blockIdx.x];
  const int nitem = mcount[blockIdx.x];
  // a member's work items start at cells 0, T - 2, 2 (T - 2), .. (the lists of Batch1D::build)
  const int per = (int)blockDim.x - 2;
  const long long mlen = (long long)pb.ncell * pb.nsp * NV;
  const unsigned long long *sw = reinterpret_cast<const unsigned long long *>(sched);
  for (int q0 = 0; q0 < nops; q0 += kWin) {
    const int nq = min(kWin, nops - q0);
    for (int i = threadIdx.x; i < nq * kOpWords; i += blockDim.x) win[i] = sw[(long long)q0 * kOpWords + i];
    __syncthreads();
    for (int q = 0; q < nq; ++q) {
      const Op1D op = reinterpret_cast<const Op1D *>(win)[q];
      if (op.kind == kOpStage) {
        for (int k = 0; k < nitem; ++k) stage_chunk<EQ, FLUX, NMAX>(pb, k * per, ops, ph, op.s, sh);
      } else if (op.kind == kOpDiss) {
        for (int k = 0; k < nitem; ++k)
          diss_chunk<NMAX>(pb, Work1D{(int)blockIdx.x, k * per}, ops, op.dst, op.dmode, ph.g, op.ddt, count);
      } else {
        for (long long v = threadIdx.x; v < mlen; v += blockDim.x) op.dst[pb.off + v] = op.src[pb.off + v];
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ batched problem set
struct Batch1D {
  int eq = FR_EQ_EULER, flux = FR_FLUX_HLL, bc = FR_BC1_DIRICHLET, correction = FR_CORR_RADAU;
  double gamma = 1.4, a = 1.0;
  std::vector<Prob1D> probs;
  int max_nsp = 2;  // largest nsp of the problem set (selects the kernel instance)
  size_t len = 0;
  hipStream_t stream = nullptr;      // the stream every call runs on
  hipStream_t own_stream = nullptr;  // the context's own non-blocking stream
  Prob1D *d_probs = nullptr;
  double *d_ops = nullptr;
  double *d_u = nullptr, *d_s1 = nullptr, *d_s2 = nullptr, *d_acc = nullptr;
  double *d_k = nullptr;  // Tsit5: the six stage derivatives, len doubles each
  double *d_in = nullptr, *d_out = nullptr;
  int *d_count = nullptr;
  // work lists per thread count (64, 128, 256): concatenated list and per-member sub-lists
  int threads = 256, grid = 0;  // grid 0 = one workgroup per work item
  Work1D *d_work[3] = {nullptr, nullptr, nullptr};
  int nwork[3] = {0, 0, 0};
  std::vector<int> member_first[3], member_count[3];
  int *d_mcount[3] = {nullptr, nullptr, nullptr};  // member_count on the device
  // per-member stepping (fr1d_member_kernel): the path setting, the schedule being recorded by
  // enqueue_steps (null: launch instead), its host copy (kept until the next call has synchronised)
  // and its device copy
  int path = FR1D_PATH_AUTO;
  std::vector<Op1D> *rec = nullptr;
  std::vector<Op1D> sched_host;
  Op1D *d_sched = nullptr;
  size_t sched_cap = 0;
  // the last step loop captured by fr_ens_run_graph, replayed while its parameters match
  hipGraphExec_t gexec = nullptr;
  int g_opts = -1, g_nsteps = 0, g_threads = 0, g_grid = -1;
  double g_dt = 0.0;

  int build(int flux_, int eq_) {
    flux = flux_;
    eq = eq_;
    len = 0;
    for (auto &p : probs) {
      max_nsp = std::max(max_nsp, p.nsp);
      p.off = (long long)len;
      len += (size_t)p.ncell * p.nsp * (eq == FR_EQ_EULER ? 3 : 1);
    }
    std::vector<double> tab(OPS_STRIDE * (NSP_MAX - 1), 0.0);
    for (int nsp = 2; nsp <= NSP_MAX; ++nsp) {
      HostOps o;
      if (!build_ops(nsp - 1, correction, o)) return fail(FR_E_INVALID, "fr1d: operator construction failed");
      double *t = tab.data() + OPS_STRIDE * (nsp - 2);
      for (int q = 0; q < nsp; ++q) {
        t[q] = o.ll[q];
        t[NSP_MAX + q] = o.lr[q];
        t[2 * NSP_MAX + q] = o.dhl[q];
        t[3 * NSP_MAX + q] = o.dhr[q];
      }
      for (int q = 0; q < nsp * nsp; ++q) t[4 * NSP_MAX + q] = o.dl[q];
      for (int q = 0; q < nsp; ++q) t[OFF_W + q] = 0.5 * o.w[q];
      for (int q = 0; q < nsp * nsp; ++q) {
        t[OFF_V + q] = o.V1[q];
        t[OFF_IV + q] = o.iV1[q];
      }
    }
    FR_HIP(hipMalloc(&d_ops, tab.size() * sizeof(double)));
    FR_HIP(hipMemcpy(d_ops, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    FR_HIP(hipMalloc(&d_probs, probs.size() * sizeof(Prob1D)));
    FR_HIP(hipMemcpy(d_probs, probs.data(), probs.size() * sizeof(Prob1D), hipMemcpyHostToDevice));
    const int Ts[3] = {64, 128, 256};
    for (int v = 0; v < 3; ++v) {
      const int per = Ts[v] - 2;
      std::vector<Work1D> wl;
      member_first[v].assign(probs.size(), 0);
      member_count[v].assign(probs.size(), 0);
      for (size_t b = 0; b < probs.size(); ++b) {
        member_first[v][b] = (int)wl.size();
        for (int f = 0; f < probs[b].ncell; f += per) wl.push_back(Work1D{(int)b, f});
        member_count[v][b] = (int)wl.size() - member_first[v][b];
      }
      nwork[v] = (int)wl.size();
      FR_HIP(hipMalloc(&d_work[v], wl.size() * sizeof(Work1D)));
      FR_HIP(hipMemcpy(d_work[v], wl.data(), wl.size() * sizeof(Work1D), hipMemcpyHostToDevice));
      FR_HIP(hipMalloc(&d_mcount[v], probs.size() * sizeof(int)));
      FR_HIP(hipMemcpy(d_mcount[v], member_count[v].data(), probs.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    FR_HIP(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
    stream = own_stream;
    FR_HIP(hipMalloc(&d_u, len * sizeof(double)));
    FR_HIP(hipMemset(d_u, 0, len * sizeof(double)));
    return FR_OK;
  }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    double *bufs[] = {d_ops, d_u, d_s1, d_s2, d_acc, d_k, d_in, d_out};
    for (double *b : bufs)
      if (b) (void)hipFree(b);
    if (d_probs) (void)hipFree(d_probs);
    if (d_count) (void)hipFree(d_count);
    for (auto *w : d_work)
      if (w) (void)hipFree(w);
    for (int v = 0; v < 3; ++v)
      if (d_mcount[v]) (void)hipFree(d_mcount[v]);
    if (d_sched) (void)hipFree(d_sched);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }

  static int tindex(int T) { return T == 64 ? 0 : (T == 128 ? 1 : 2); }

  template <int NMAX>
  void launch_eq(const Stage1D &sa, int T, int g, const Work1D *wl, int nw, const Phys1D &ph, hipStream_t st) const {
    if (eq == FR_EQ_ADVECTION)
      hipLaunchKernelGGL((fr1d_kernel<FR_EQ_ADVECTION, FR_FLUX_HLL, NMAX>), dim3(g), dim3(T), 0, st, d_probs, wl, nw,
                         d_ops, ph, sa);
    else if (flux == FR_FLUX_HLL)
      hipLaunchKernelGGL((fr1d_kernel<FR_EQ_EULER, FR_FLUX_HLL, NMAX>), dim3(g), dim3(T), 0, st, d_probs, wl, nw, d_ops,
                         ph, sa);
    else if (flux == FR_FLUX_RUSANOV)
      hipLaunchKernelGGL((fr1d_kernel<FR_EQ_EULER, FR_FLUX_RUSANOV, NMAX>), dim3(g), dim3(T), 0, st, d_probs, wl, nw,
                         d_ops, ph, sa);
    else
      hipLaunchKernelGGL((fr1d_kernel<FR_EQ_EULER, FR_FLUX_ROE, NMAX>), dim3(g), dim3(T), 0, st, d_probs, wl, nw, d_ops,
                         ph, sa);
  }

  // one launch over work items [w0, w0 + nw) of the list for T threads
  int launch(const Stage1D &sa, int T, int grid_, int w0, int nw, hipStream_t st) const {
    const int v = tindex(T);
    Phys1D ph{eq, bc, gamma, gamma - 1.0, a};
    const Work1D *wl = d_work[v] + w0;
    const int g = grid_ > 0 ? std::min(grid_, nw) : nw;
    if (nw <= 0) return FR_OK;
    if (max_nsp <= 5)
      launch_eq<5>(sa, T, g, wl, nw, ph, st);
    else
      launch_eq<NSP_MAX>(sa, T, g, wl, nw, ph, st);
    FR_LAUNCHED("fr1d_kernel");
    return FR_OK;
  }

  // batched (mode 0): whole work list in one launch; looped (mode 1): one launch per member
  int stage(const Stage1D &sa, int mode, hipStream_t st) const {
    const int v = tindex(threads);
    if (mode == 0) return launch(sa, threads, grid, 0, nwork[v], st);
    for (size_t b = 0; b < probs.size(); ++b) {
      int rc = launch(sa, threads, 0, member_first[v][b], member_count[v][b], st);
      if (rc) return rc;
    }
    return FR_OK;
  }

  int ensure_stage_buffers(bool acc, bool tsit = false) {
    if (!d_s1) FR_HIP(hipMalloc(&d_s1, len * sizeof(double)));
    if (!d_s2) FR_HIP(hipMalloc(&d_s2, len * sizeof(double)));
    if (acc && !d_acc) FR_HIP(hipMalloc(&d_acc, len * sizeof(double)));
    if (tsit && !d_k) FR_HIP(hipMalloc(&d_k, 6 * len * sizeof(double)));
    return FR_OK;
  }

  // out = u + dt * sum_{j < nk} row[j] k_{j+1} over the whole state
  int tsit5_comb(double *out, const double *u, const double *row, int nk, double dt, hipStream_t st) const {
    FR_HIP(fr::tsit5_comb(out, u, d_k, len, row, nk, dt, len, st));
    return FR_OK;
  }

  // Per-cell dissipation pass over every member (mode 0: one launch; 1: one launch per member):
  // dmode 1 sensor + filter with the filter time step dt, 2 positivity limiter.  count[0] += cells
  // flagged / limited, count[1] += cells whose mean density is not positive (left unchanged).
  int diss_launch(double *u, int dmode, double dt, int mode, hipStream_t st) const {
    const int v = tindex(threads);
    auto one = [&](int w0, int nw, int grid_) {
      if (nw <= 0) return;
      const int g = grid_ > 0 ? std::min(grid_, nw) : nw;
      if (max_nsp <= 5)
        hipLaunchKernelGGL(fr1d_diss_kernel<5>, dim3(g), dim3(threads), 0, st, d_probs, d_work[v] + w0, nw, d_ops, u,
                           dmode, gamma, dt, d_count);
      else
        hipLaunchKernelGGL(fr1d_diss_kernel<NSP_MAX>, dim3(g), dim3(threads), 0, st, d_probs, d_work[v] + w0, nw, d_ops,
                           u, dmode, gamma, dt, d_count);
    };
    if (mode == 0) {
      one(0, nwork[v], grid);
    } else {
      for (size_t b = 0; b < probs.size(); ++b) one(member_first[v][b], member_count[v][b], 0);
    }
    FR_LAUNCHED("fr1d_diss_kernel");
    return FR_OK;
  }

  // opts (SSP-RK3 and Midpoint): FR_ENS_FILTER runs the sensor + filter on the state before every
  // step (example/euler_highlevel.jl:37-54), FR_ENS_LIMIT the positivity limiter on the input of
  // every stage (the state before the step and the intermediate stage results)
  int enqueue_steps(int scheme, int mode, double dt, int nsteps, hipStream_t st, int opts = 0) {
    double *U = d_u, *S1 = d_s1, *S2 = d_s2, *A = d_acc;
    auto stg = [&](const double *u0, const double *us, double *out, double a0, double a1, double b,
                   const double *ai, double *ao, double b2) {
      Stage1D s{us, out, u0, ai, ao, a0, a1, b, b2, 0};
      if (rec) {
        rec->push_back(Op1D{kOpStage, 0, 0.0, s, nullptr, nullptr});
        return (int)FR_OK;
      }
      return stage(s, mode, st);
    };
    auto dis = [&](double *u, int dmode, double ddt) {
      if (rec) {
        rec->push_back(Op1D{kOpDiss, dmode, ddt, Stage1D{}, u, nullptr});
        return (int)FR_OK;
      }
      return diss_launch(u, dmode, ddt, mode, st);
    };
    const bool filt = (opts & FR_ENS_FILTER) != 0, lim = (opts & FR_ENS_LIMIT) != 0;
    for (int n = 0; n < nsteps; ++n) {
      int rc = FR_OK;
      switch (scheme) {
        case FR_RK_SSPRK3: {
          // consecutive steps alternate the state between U and S1 (every stage writes every
          // cell), so the last stage never writes the u^n it reads; the call's last step writes U
          double *const X = (n & 1) ? S1 : U, *const Y = (n & 1) ? U : S1;
          double *const out = n == nsteps - 1 ? U : Y;
          if (filt && (rc = dis(X, 1, dt))) return rc;
          if (lim && (rc = dis(X, 2, 0.0))) return rc;
          if ((rc = stg(X, X, Y, 0.0, 1.0, dt, nullptr, nullptr, 0.0))) return rc;
          if (lim && (rc = dis(Y, 2, 0.0))) return rc;
          if ((rc = stg(X, Y, S2, 0.75, 0.25, 0.25 * dt, nullptr, nullptr, 0.0))) return rc;
          if (lim && (rc = dis(S2, 2, 0.0))) return rc;
          if ((rc = stg(X, S2, out, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt, nullptr, nullptr, 0.0))) return rc;
          break;
        }
        case FR_RK_RK4:
          if ((rc = stg(U, U, S2, 1.0, 0.0, 0.5 * dt, U, A, dt / 6.0))) return rc;
          if ((rc = stg(U, S2, S1, 1.0, 0.0, 0.5 * dt, A, A, dt / 3.0))) return rc;
          if ((rc = stg(U, S1, S2, 1.0, 0.0, dt, A, A, dt / 3.0))) return rc;
          if ((rc = stg(A, S2, U, 1.0, 0.0, dt / 6.0, nullptr, nullptr, 0.0))) return rc;
          break;
        case FR_RK_MIDPOINT: {
          // consecutive steps alternate the state between U and S2; the call's last step writes U
          double *const X = (n & 1) ? S2 : U;
          double *const out = n == nsteps - 1 ? U : ((n & 1) ? U : S2);
          if (filt && (rc = dis(X, 1, dt))) return rc;
          if (lim && (rc = dis(X, 2, 0.0))) return rc;
          if ((rc = stg(X, X, S1, 0.0, 1.0, 0.5 * dt, nullptr, nullptr, 0.0))) return rc;
          if (lim && (rc = dis(S1, 2, 0.0))) return rc;
          if ((rc = stg(X, S1, out, 1.0, 0.0, dt, nullptr, nullptr, 0.0))) return rc;
          break;
        }
        case FR_RK_EULER: {
          // consecutive steps alternate the state between U and S1; a call that ends in S1
          // copies it back once
          double *const X = (n & 1) ? S1 : U, *const Y = (n & 1) ? U : S1;
          if ((rc = stg(X, X, Y, 0.0, 1.0, dt, nullptr, nullptr, 0.0))) return rc;
          if (n == nsteps - 1 && Y != U) {
            if (rec)
              rec->push_back(Op1D{kOpCopy, 0, 0.0, Stage1D{}, U, Y});
            else
              FR_HIP(hipMemcpyAsync(U, Y, len * sizeof(double), hipMemcpyDeviceToDevice, st));
          }
          break;
        }
        case FR_RK_TSIT5: {
          if (rec) return fail(FR_E_UNSUPPORTED, "fr1d: Tsit5 is stepped by the batched launches only");
          // k_s = L(Y_s) (residual launch), Y_1 = U, Y_{s+1} = U + dt sum_j a_{s+1,j} k_j, and
          // U <- U + dt sum_j b_j k_j (the last combination, in place)
          for (int sg = 0; sg < 6; ++sg) {
            Stage1D r{};
            r.u = sg == 0 ? U : S1;
            r.out = d_k + (size_t)sg * len;
            r.residual = 1;
            if ((rc = stage(r, mode, st))) return rc;
            if ((rc = tsit5_comb(sg == 5 ? U : S1, U, kTsit5[sg], sg + 1, dt, st))) return rc;
          }
          break;
        }
        default:
          return fail(FR_E_INVALID, "fr1d_step: unknown scheme");
      }
    }
    return FR_OK;
  }

  // the per-member path: FR1D_PATH_MEMBER, or FR1D_PATH_AUTO when no member has more than
  // kMemberAutoItems work items (a workgroup runs a member's items one after the other, so the path
  // pays off where a stage of the batched launch is a launch's latency over a few items).  Tsit5
  // and the per-member launch mode always take the batched launches.
  static constexpr int kMemberAutoItems = 2;
  bool use_member(int scheme, int mode) const {
    if (scheme == FR_RK_TSIT5 || mode != 0 || path == FR1D_PATH_BATCHED) return false;
    if (path == FR1D_PATH_MEMBER) return true;
    const auto &mc = member_count[tindex(threads)];
    return !mc.empty() && *std::max_element(mc.begin(), mc.end()) <= kMemberAutoItems;
  }

  template <int NMAX>
  void launch_member_eq(const Phys1D &ph, const Op1D *ops_, int n, hipStream_t st) const {
    const int v = tindex(threads);
    const dim3 g((unsigned)probs.size()), b((unsigned)threads);
    if (eq == FR_EQ_ADVECTION)
      hipLaunchKernelGGL((fr1d_member_kernel<FR_EQ_ADVECTION, FR_FLUX_HLL, NMAX>), g, b, 0, st, d_probs, d_mcount[v], d_ops, ph,
                         ops_, n, d_count);
    else if (flux == FR_FLUX_HLL)
      hipLaunchKernelGGL((fr1d_member_kernel<FR_EQ_EULER, FR_FLUX_HLL, NMAX>), g, b, 0, st, d_probs, d_mcount[v], d_ops, ph,
                         ops_, n, d_count);
    else if (flux == FR_FLUX_RUSANOV)
      hipLaunchKernelGGL((fr1d_member_kernel<FR_EQ_EULER, FR_FLUX_RUSANOV, NMAX>), g, b, 0, st, d_probs, d_mcount[v], d_ops, ph,
                         ops_, n, d_count);
    else
      hipLaunchKernelGGL((fr1d_member_kernel<FR_EQ_EULER, FR_FLUX_ROE, NMAX>), g, b, 0, st, d_probs, d_mcount[v], d_ops, ph,
                         ops_, n, d_count);
  }

  // nsteps steps as one schedule run by fr1d_member_kernel (a workgroup per member), in launches of
  // at most kMemberOpsPerLaunch operations
  int member_steps(int scheme, double dt, int nsteps, hipStream_t st, int opts) {
    // the previous call's upload may still read sched_host
    FR_HIP(hipStreamSynchronize(st));
    sched_host.clear();
    rec = &sched_host;
    int rc = enqueue_steps(scheme, 0, dt, nsteps, st, opts);
    rec = nullptr;
    if (rc) return rc;
    if (sched_host.empty()) return FR_OK;
    if (sched_host.size() > sched_cap) {
      if (d_sched) FR_HIP(hipFree(d_sched));
      d_sched = nullptr;
      sched_cap = 0;
      FR_HIP(hipMalloc(&d_sched, sched_host.size() * sizeof(Op1D)));
      sched_cap = sched_host.size();
    }
    FR_HIP(hipMemcpyAsync(d_sched, sched_host.data(), sched_host.size() * sizeof(Op1D), hipMemcpyHostToDevice, st));
    Phys1D ph{eq, bc, gamma, gamma - 1.0, a};
    const int nops = (int)sched_host.size();
    for (int q0 = 0; q0 < nops; q0 += kMemberOpsPerLaunch) {
      const int n = std::min(kMemberOpsPerLaunch, nops - q0);
      if (max_nsp <= 5)
        launch_member_eq<5>(ph, d_sched + q0, n, st);
      else
        launch_member_eq<NSP_MAX>(ph, d_sched + q0, n, st);
      FR_LAUNCHED("fr1d_member_kernel");
    }
    return FR_OK;
  }

  // the steps on the path use_member selects
  int run_steps(int scheme, int mode, double dt, int nsteps, hipStream_t st, int opts = 0) {
    if (use_member(scheme, mode)) return member_steps(scheme, dt, nsteps, st, opts);
    return enqueue_steps(scheme, mode, dt, nsteps, st, opts);
  }

  // mode 1 sensor + filter, 2 positivity; n_out (optional) receives the two counters
  int dissipation(int mode, double dt, int *n_out) {
    if (!d_count) FR_HIP(hipMalloc(&d_count, 2 * sizeof(int)));
    FR_HIP(hipMemsetAsync(d_count, 0, 2 * sizeof(int), stream));
    int rc = diss_launch(d_u, mode, dt, 0, stream);
    if (rc) return rc;
    int h[2] = {0, 0};
    FR_HIP(hipMemcpyAsync(h, d_count, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
    FR_HIP(hipStreamSynchronize(stream));
    if (n_out) {
      n_out[0] = h[0];
      n_out[1] = h[1];
    }
    if (mode == 2 && h[1] > 0)
      return fail(FR_E_LIMITER, "fr1d_positivity: " + std::to_string(h[1]) + " cell(s) with non-positive mean density");
    return FR_OK;
  }

  int residual(const double *d_u_, double *d_du_, hipStream_t st) const {
    Stage1D s{};
    s.u = d_u_;
    s.out = d_du_;
    s.residual = 1;
    return stage(s, 0, st);
  }
};

}  // namespace fr

using namespace fr;

struct fr1d_ctx : Batch1D {};
struct fr_ens_ctx : Batch1D {};

extern "C" {

int fr1d_create(fr1d_ctx **out, const fr1d_desc *d) {
  FR_REQUIRE(out && d, FR_E_INVALID, "fr1d_create: null argument");
  *out = nullptr;
  FR_REQUIRE(d->equation == FR_EQ_EULER || d->equation == FR_EQ_ADVECTION, FR_E_INVALID, "fr1d_create: unknown equation");
  FR_REQUIRE(d->ncell >= 2, FR_E_INVALID, "fr1d_create: ncell must be >= 2");
  FR_REQUIRE(d->deg >= 1 && d->deg <= NSP_MAX - 1, FR_E_UNSUPPORTED, "fr1d_create: degree must be in 1..8");
  FR_REQUIRE(d->dx > 0.0 && std::isfinite(d->dx), FR_E_INVALID, "fr1d_create: dx must be > 0");
  FR_REQUIRE(d->bc == FR_BC1_DIRICHLET || d->bc == FR_BC1_PERIOD, FR_E_INVALID, "fr1d_create: unknown bc");
  FR_REQUIRE(d->correction >= FR_CORR_RADAU && d->correction <= FR_CORR_HUYNH, FR_E_INVALID, "fr1d_create: unknown correction");
  FR_REQUIRE(d->flux >= FR_FLUX_HLL && d->flux <= FR_FLUX_ROE, FR_E_INVALID, "fr1d_create: unknown flux");
  FR_REQUIRE(d->equation == FR_EQ_ADVECTION || d->gamma > 1.0, FR_E_INVALID, "fr1d_create: gamma must be > 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FR_E_HIP, "fr1d_create: no HIP device visible");
  fr1d_ctx *c = new (std::nothrow) fr1d_ctx();
  FR_REQUIRE(c, FR_E_NOMEM, "fr1d_create: out of host memory");
  c->bc = d->bc;
  c->gamma = d->gamma;
  c->a = d->a;
  c->correction = d->correction;
  c->probs.push_back(Prob1D{0, d->ncell, d->deg + 1, 2.0 / d->dx});
  int rc = c->build(d->flux, d->equation);
  if (rc) {
    c->release();
    delete c;
    return rc;
  }
  *out = c;
  return FR_OK;
}

int fr1d_destroy(fr1d_ctx *c) {
  if (!c) return FR_OK;
  c->release();
  delete c;
  return FR_OK;
}

size_t fr1d_state_len(const fr1d_ctx *c) { return c ? c->len : 0; }

int fr1d_set_stream(fr1d_ctx *c, void *s) {
  FR_REQUIRE(c, FR_E_INVALID, "fr1d_set_stream: null context");
  c->stream = s == FR_OWN_STREAM ? c->own_stream : (hipStream_t)s;
  return FR_OK;
}

int fr1d_rhs_dev(fr1d_ctx *c, const double *d_u, double *d_du, double t) {
  (void)t;
  FR_REQUIRE(c && d_u && d_du && d_u != d_du, FR_E_INVALID, "fr1d_rhs_dev: bad arguments");
  return c->residual(d_u, d_du, c->stream);
}

int fr1d_rhs(fr1d_ctx *c, const double *u, double *du, double t) {
  (void)t;
  FR_REQUIRE(c && u && du, FR_E_INVALID, "fr1d_rhs: null argument");
  const size_t bytes = c->len * sizeof(double);
  if (!c->d_in) FR_HIP(hipMalloc(&c->d_in, bytes));
  if (!c->d_out) FR_HIP(hipMalloc(&c->d_out, bytes));
  FR_HIP(hipMemcpyAsync(c->d_in, u, bytes, hipMemcpyHostToDevice, c->stream));
  int rc = c->residual(c->d_in, c->d_out, c->stream);
  if (rc) return rc;
  FR_HIP(hipMemcpyAsync(du, c->d_out, bytes, hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr1d_set_state(fr1d_ctx *c, const double *u) {
  FR_REQUIRE(c && u, FR_E_INVALID, "fr1d_set_state: null argument");
  FR_HIP(hipMemcpyAsync(c->d_u, u, c->len * sizeof(double), hipMemcpyHostToDevice, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr1d_get_state(fr1d_ctx *c, double *u) {
  FR_REQUIRE(c && u, FR_E_INVALID, "fr1d_get_state: null argument");
  FR_HIP(hipMemcpyAsync(u, c->d_u, c->len * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr1d_step(fr1d_ctx *c, int scheme, double dt, int nsteps) {
  FR_REQUIRE(c && nsteps >= 0 && std::isfinite(dt), FR_E_INVALID, "fr1d_step: bad arguments");
  FR_REQUIRE(scheme >= FR_RK_SSPRK3 && scheme <= FR_RK_TSIT5, FR_E_INVALID, "fr1d_step: unknown scheme");
  int rc = c->ensure_stage_buffers(scheme == FR_RK_RK4, scheme == FR_RK_TSIT5);
  if (rc) return rc;
  return c->run_steps(scheme, 0, dt, nsteps, c->stream);
}

int fr1d_set_path(fr1d_ctx *c, int path) {
  FR_REQUIRE(c, FR_E_INVALID, "fr1d_set_path: null context");
  FR_REQUIRE(path >= FR1D_PATH_AUTO && path <= FR1D_PATH_MEMBER, FR_E_INVALID,
             "fr1d_set_path: path must be FR1D_PATH_AUTO, _BATCHED or _MEMBER");
  c->path = path;
  return FR_OK;
}

int fr1d_get_path(const fr1d_ctx *c, int scheme, int *path) {
  FR_REQUIRE(c && path, FR_E_INVALID, "fr1d_get_path: null argument");
  *path = c->use_member(scheme, 0) ? FR1D_PATH_MEMBER : FR1D_PATH_BATCHED;
  return FR_OK;
}

// ------------------------------------------------------------------ ensembles
int fr_ens_create(fr_ens_ctx **out, int nprob, const int *ncell, const int *deg, const double *dx, double gamma,
                  int flux, int bc) {
  FR_REQUIRE(out && ncell && deg && dx && nprob >= 1, FR_E_INVALID, "fr_ens_create: bad arguments");
  *out = nullptr;
  FR_REQUIRE(gamma > 1.0, FR_E_INVALID, "fr_ens_create: gamma must be > 1");
  FR_REQUIRE(flux >= FR_FLUX_HLL && flux <= FR_FLUX_ROE, FR_E_INVALID, "fr_ens_create: unknown flux");
  FR_REQUIRE(bc == FR_BC1_DIRICHLET || bc == FR_BC1_PERIOD, FR_E_INVALID, "fr_ens_create: unknown bc");
  for (int b = 0; b < nprob; ++b) {
    FR_REQUIRE(ncell[b] >= 2, FR_E_INVALID, "fr_ens_create: every member needs ncell >= 2");
    FR_REQUIRE(deg[b] >= 1 && deg[b] <= NSP_MAX - 1, FR_E_UNSUPPORTED, "fr_ens_create: member degree must be in 1..8");
    FR_REQUIRE(dx[b] > 0.0, FR_E_INVALID, "fr_ens_create: member dx must be > 0");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FR_E_HIP, "fr_ens_create: no HIP device visible");
  fr_ens_ctx *c = new (std::nothrow) fr_ens_ctx();
  FR_REQUIRE(c, FR_E_NOMEM, "fr_ens_create: out of host memory");
  c->bc = bc;
  c->gamma = gamma;
  for (int b = 0; b < nprob; ++b) c->probs.push_back(Prob1D{0, ncell[b], deg[b] + 1, 2.0 / dx[b]});
  int rc = c->build(flux, FR_EQ_EULER);
  if (rc) {
    c->release();
    delete c;
    return rc;
  }
  *out = c;
  return FR_OK;
}

int fr_ens_destroy(fr_ens_ctx *c) {
  if (!c) return FR_OK;
  c->release();
  delete c;
  return FR_OK;
}

size_t fr_ens_state_len(const fr_ens_ctx *c) { return c ? c->len : 0; }

size_t fr_ens_offset(const fr_ens_ctx *c, int member) {
  if (!c || member < 0 || member >= (int)c->probs.size()) return (size_t)-1;
  return (size_t)c->probs[member].off;
}

int fr_ens_set_state(fr_ens_ctx *c, const double *u) {
  FR_REQUIRE(c && u, FR_E_INVALID, "fr_ens_set_state: null argument");
  FR_HIP(hipMemcpyAsync(c->d_u, u, c->len * sizeof(double), hipMemcpyHostToDevice, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr_ens_get_state(fr_ens_ctx *c, double *u) {
  FR_REQUIRE(c && u, FR_E_INVALID, "fr_ens_get_state: null argument");
  FR_HIP(hipMemcpyAsync(u, c->d_u, c->len * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr_ens_step(fr_ens_ctx *c, int mode, double dt, int nsteps) {
  FR_REQUIRE(c && (mode == 0 || mode == 1) && nsteps >= 0 && std::isfinite(dt), FR_E_INVALID,
             "fr_ens_step: bad arguments");
  int rc = c->ensure_stage_buffers(false);
  if (rc) return rc;
  rc = c->run_steps(FR_RK_SSPRK3, mode, dt, nsteps, c->stream);
  if (rc) return rc;
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr_ens_set_path(fr_ens_ctx *c, int path) {
  FR_REQUIRE(c, FR_E_INVALID, "fr_ens_set_path: null context");
  FR_REQUIRE(path >= FR1D_PATH_AUTO && path <= FR1D_PATH_MEMBER, FR_E_INVALID,
             "fr_ens_set_path: path must be FR1D_PATH_AUTO, _BATCHED or _MEMBER");
  c->path = path;
  return FR_OK;
}

// counters of the loop entry points: zeroed before the loop, read after it
static int ens_read_counts(fr_ens_ctx *c, const char *who, int *n_flagged, int *n_limited) {
  int h[2] = {0, 0};
  FR_HIP(hipMemcpyAsync(h, c->d_count, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  if (n_flagged) *n_flagged = h[0];
  if (n_limited) *n_limited = h[1];
  if (h[1] > 0)
    return fail(FR_E_LIMITER, std::string(who) + ": " + std::to_string(h[1]) +
                                  " limiter failure(s) (cell mean density not positive) over the loop");
  return FR_OK;
}

int fr_ens_run(fr_ens_ctx *c, int mode, int opts, double dt, int nsteps, int *n_flagged, int *n_limited) {
  FR_REQUIRE(c && (mode == 0 || mode == 1) && nsteps >= 0 && std::isfinite(dt) && dt >= 0.0, FR_E_INVALID,
             "fr_ens_run: bad arguments");
  FR_REQUIRE((opts & ~(FR_ENS_FILTER | FR_ENS_LIMIT | FR_ENS_MIDPOINT)) == 0, FR_E_INVALID,
             "fr_ens_run: unknown option bits");
  int rc = c->ensure_stage_buffers(false);
  if (rc) return rc;
  if (!c->d_count) FR_HIP(hipMalloc(&c->d_count, 2 * sizeof(int)));
  FR_HIP(hipMemsetAsync(c->d_count, 0, 2 * sizeof(int), c->stream));
  const int scheme = (opts & FR_ENS_MIDPOINT) ? FR_RK_MIDPOINT : FR_RK_SSPRK3;
  if ((rc = c->run_steps(scheme, mode, dt, nsteps, c->stream, opts))) return rc;
  return ens_read_counts(c, "fr_ens_run", n_flagged, n_limited);
}

int fr_ens_run_graph(fr_ens_ctx *c, int opts, double dt, int nsteps, int nrep, int *n_flagged, int *n_limited) {
  FR_REQUIRE(c && nsteps >= 1 && nrep >= 0 && std::isfinite(dt) && dt >= 0.0, FR_E_INVALID,
             "fr_ens_run_graph: bad arguments");
  FR_REQUIRE((opts & ~(FR_ENS_FILTER | FR_ENS_LIMIT | FR_ENS_MIDPOINT)) == 0, FR_E_INVALID,
             "fr_ens_run_graph: unknown option bits");
  int rc = c->ensure_stage_buffers(false);
  if (rc) return rc;
  if (!c->d_count) FR_HIP(hipMalloc(&c->d_count, 2 * sizeof(int)));
  // The graph holds the context's own buffers and work lists, which never move once allocated, so a
  // capture stays valid while the options, dt, the step count and the work mapping are unchanged.
  if (!(c->gexec && c->g_opts == opts && c->g_dt == dt && c->g_nsteps == nsteps && c->g_threads == c->threads &&
        c->g_grid == c->grid)) {
    if (c->gexec) {
      FR_HIP(hipStreamSynchronize(c->stream));
      (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
    hipGraph_t graph = nullptr;
    FR_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = c->enqueue_steps((opts & FR_ENS_MIDPOINT) ? FR_RK_MIDPOINT : FR_RK_SSPRK3, 0, dt, nsteps, c->stream, opts);
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
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
    c->g_opts = opts;
    c->g_dt = dt;
    c->g_nsteps = nsteps;
    c->g_threads = c->threads;
    c->g_grid = c->grid;
  }
  // the counters are cleared outside the graph (no memset node is recorded into it)
  FR_HIP(hipMemsetAsync(c->d_count, 0, 2 * sizeof(int), c->stream));
  for (int r = 0; r < nrep; ++r) FR_HIP(hipGraphLaunch(c->gexec, c->stream));
  return ens_read_counts(c, "fr_ens_run_graph", n_flagged, n_limited);
}

int fr_ens_step_graph(fr_ens_ctx *c, double dt, int nsteps, int nrep) {
  FR_REQUIRE(c && nsteps >= 1 && nrep >= 0 && std::isfinite(dt), FR_E_INVALID, "fr_ens_step_graph: bad arguments");
  return fr_ens_run_graph(c, 0, dt, nsteps, nrep, nullptr, nullptr);
}

// Times candidate (threads per workgroup, grid) mappings with HIP events over a few SSP-RK3
// steps on a scratch copy of the state, keeps the fastest, and restores the state.
int fr_ens_autotune(fr_ens_ctx *c, double dt, int *chosen) {
  FR_REQUIRE(c && std::isfinite(dt), FR_E_INVALID, "fr_ens_autotune: bad arguments");
  int rc = c->ensure_stage_buffers(false);
  if (rc) return rc;
  int ndev = 0, dev = 0, ncu = 256;
  (void)ndev;
  FR_HIP(hipGetDevice(&dev));
  FR_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  double *save = nullptr;
  FR_HIP(hipMalloc(&save, c->len * sizeof(double)));
  FR_HIP(hipMemcpyAsync(save, c->d_u, c->len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  const int Ts[3] = {64, 128, 256};
  const int mult[5] = {0, 1, 2, 4, 8};  // 0: one workgroup per work item
  hipEvent_t e0, e1;
  FR_HIP(hipEventCreate(&e0));
  FR_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int best_i = 0, idx = 0;
  const int keepT = c->threads, keepG = c->grid;
  for (int ti = 0; ti < 3; ++ti)
    for (int mi = 0; mi < 5; ++mi, ++idx) {
      c->threads = Ts[ti];
      c->grid = mult[mi] * ncu;
      rc = c->enqueue_steps(FR_RK_SSPRK3, 0, dt, 1, c->stream);  // warm-up
      if (rc) break;
      FR_HIP(hipEventRecord(e0, c->stream));
      rc = c->enqueue_steps(FR_RK_SSPRK3, 0, dt, 3, c->stream);
      if (rc) break;
      FR_HIP(hipEventRecord(e1, c->stream));
      FR_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      FR_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_i = idx;
      }
      FR_HIP(hipMemcpyAsync(c->d_u, save, c->len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  FR_HIP(hipStreamSynchronize(c->stream));
  (void)hipFree(save);
  if (rc) {
    c->threads = keepT;
    c->grid = keepG;
    return rc;
  }
  c->threads = Ts[best_i / 5];
  c->grid = mult[best_i % 5] * ncu;
  if (chosen) *chosen = best_i;
  return FR_OK;
}

int fr1d_sensor_filter(fr1d_ctx *c, double dt, int *n_flagged) {
  FR_REQUIRE(c && std::isfinite(dt) && dt >= 0.0, FR_E_INVALID, "fr1d_sensor_filter: bad arguments");
  FR_REQUIRE(c->eq == FR_EQ_EULER, FR_E_UNSUPPORTED, "fr1d_sensor_filter: Euler states only");
  int h[2];
  int rc = c->dissipation(1, dt, h);
  if (n_flagged) *n_flagged = h[0];
  return rc;
}

int fr1d_positivity(fr1d_ctx *c, int *n_limited) {
  FR_REQUIRE(c, FR_E_INVALID, "fr1d_positivity: null context");
  FR_REQUIRE(c->eq == FR_EQ_EULER, FR_E_UNSUPPORTED, "fr1d_positivity: Euler states only");
  int h[2] = {0, 0};
  int rc = c->dissipation(2, 0.0, h);
  if (n_limited) *n_limited = h[0];
  return rc;
}

int fr_ens_sensor_filter(fr_ens_ctx *c, double dt, int *n_flagged) {
  FR_REQUIRE(c && std::isfinite(dt) && dt >= 0.0, FR_E_INVALID, "fr_ens_sensor_filter: bad arguments");
  int h[2];
  int rc = c->dissipation(1, dt, h);
  if (n_flagged) *n_flagged = h[0];
  return rc;
}

int fr_ens_positivity(fr_ens_ctx *c, int *n_limited) {
  FR_REQUIRE(c, FR_E_INVALID, "fr_ens_positivity: null context");
  int h[2] = {0, 0};
  int rc = c->dissipation(2, 0.0, h);
  if (n_limited) *n_limited = h[0];
  return rc;
}

}  // extern "C"
