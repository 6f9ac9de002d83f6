// fr_bgk.hip -- 2D discrete-velocity BGK flux reconstruction on periodic uniform quads, and
// batched ensembles of such problems that differ in the relaxation time, for gfx950.
//
// Tensorises example/bgk_wave.jl:69-129: per solution point the conserved moments over the
// velocity grid give the Maxwellian (KitBase moments_conserve / conserve_prim / maxwellian in 2D);
// every discrete velocity is advected by FR with the upwind (Heaviside-blended) interface flux and
// relaxes to the Maxwellian, du = -(div f + corrections) + (M - f) / tau.
//
// State f(i, j, k, l, a, b, member), column-major: the cell index is fastest, so a wavefront of
// consecutive cells reads contiguous memory for a fixed (point, velocity).  Two kernels per stage:
// moments (one thread per solution point, loop over the velocities, FP64 sums; it also writes the
// separable Maxwellian factors of the point) and transport + relaxation + RK stage (one workgroup
// per (member, velocity, 16 x 16 cell tile), one thread per cell; the upwind traces of the
// neighbours come from LDS or from the halo threads).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "fr_common.h"

namespace fr {

constexpr int BT = 16;  // tile edge in cells (BT x BT threads)

template <int NSP>
struct BgkOps {
  double ll[NSP], lr[NSP], dl[NSP * NSP], dhl[NSP], dhr[NSP];
};

struct BgkGeo {
  int nx, ny, nu, nv, nmem;
  long long npts;   // nx * ny * nsp^2 (one velocity plane)
  long long mstride;  // doubles per member
  double iJx, iJy, gamma;
};

struct BgkStage {
  const double *u, *u0;
  double *out;
  double a0, a1, b;
  int residual;
};

// moments -> prim (rho, U, V, lambda) per solution point, prim[member][c][pt], and the separable
// Maxwellian M(a, b) = Ea[a] * Eb[b] of every point: Ea[member][a][pt] = exp(-lambda (u_a - U)^2),
// Eb[member][b][pt] = rho lambda / pi exp(-lambda (v_b - V)^2), so that the transport kernel spends
// one multiplication per value on it.  A wavefront takes 64 consecutive points (the lanes); nq
// wavefronts of the workgroup share a point group and split its velocity rows b between them (nq
// > 1 when the points alone give the chip too few wavefronts).  The partial sums are combined in
// LDS in a fixed order, and each wavefront then writes its share of the factors.
constexpr int MQ = 4;  // wavefronts per workgroup of the moments kernel
__global__ __launch_bounds__(64 * MQ) void bgk_moments_kernel(BgkGeo g, const double *__restrict__ f,
                                                              const double *__restrict__ uv,
                                                              const double *__restrict__ vv,
                                                              const double *__restrict__ wu,
                                                              const double *__restrict__ wv, double *prim,
                                                              double *__restrict__ tab, int nq) {
  __shared__ double sm[MQ][4][64];
  __shared__ double sp[MQ][4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, q = w % nq, grp = w / nq;
  const long long t = (blockIdx.x * (long long)(MQ / nq) + grp) * 64 + lane;
  const bool act = t < g.npts * g.nmem;
  const long long mem = act ? t / g.npts : 0, p = act ? t % g.npts : 0;
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  if (act) {
    const double *fm = f + mem * g.mstride + p;
    double r4[4][4] = {};  // four independent chains per moment
    const int b0 = q * g.nv / nq, b1 = (q + 1) * g.nv / nq;
    for (int b = b0; b < b1; ++b) {
      const double vb = vv[b], wb = wv[b];
      const double *fb = fm + g.npts * (long long)g.nu * b;
      int a = 0;
      for (; a + 4 <= g.nu; a += 4) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double fv = fb[g.npts * (long long)(a + r)];
          const double ua = uv[a + r], wt = wu[a + r] * wb;
          r4[r][0] += wt * fv;
          r4[r][1] += wt * ua * fv;
          r4[r][2] += wt * vb * fv;
          r4[r][3] += wt * 0.5 * (ua * ua + vb * vb) * fv;
        }
      }
      for (; a < g.nu; ++a) {
        const double fv = fb[g.npts * (long long)a];
        const double ua = uv[a], wt = wu[a] * wb;
        r4[0][0] += wt * fv;
        r4[0][1] += wt * ua * fv;
        r4[0][2] += wt * vb * fv;
        r4[0][3] += wt * 0.5 * (ua * ua + vb * vb) * fv;
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) m[c] = (r4[0][c] + r4[1][c]) + (r4[2][c] + r4[3][c]);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) sm[w][c][lane] = m[c];
  __syncthreads();
  if (q == 0) {
    double mm[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double v = sm[w][c][lane];
      for (int r = 1; r < nq; ++r) v += sm[w + r][c][lane];
      mm[c] = v;
    }
    const double rho = mm[0], U = mm[1] / mm[0], V = mm[2] / mm[0];
    const double lam = 0.5 * mm[0] / (g.gamma - 1.0) / (mm[3] - 0.5 * (mm[1] * mm[1] + mm[2] * mm[2]) / mm[0]);
    sp[grp][0][lane] = rho;
    sp[grp][1][lane] = U;
    sp[grp][2][lane] = V;
    sp[grp][3][lane] = lam;
    if (act) {
      double *pr = prim + mem * 4 * g.npts + p;
      pr[0] = rho;
      pr[g.npts] = U;
      pr[2 * g.npts] = V;
      pr[3 * g.npts] = lam;
    }
  }
  __syncthreads();
  if (!act) return;
  const double rho = sp[grp][0][lane], U = sp[grp][1][lane], V = sp[grp][2][lane], lam = sp[grp][3][lane];
  double *ta = tab + mem * (long long)(g.nu + g.nv) * g.npts + p;
  for (int a = q * g.nu / nq; a < (q + 1) * g.nu / nq; ++a) {
    const double du = uv[a] - U;
    ta[g.npts * (long long)a] = exp(-lam * (du * du));
  }
  const double c0 = rho * (lam / M_PI);
  for (int b = q * g.nv / nq; b < (q + 1) * g.nv / nq; ++b) {
    const double dv = vv[b] - V;
    ta[g.npts * (long long)(g.nu + b)] = c0 * exp(-lam * (dv * dv));
  }
}

template <int NSP>
__global__ __launch_bounds__(BT *BT) void bgk_rhs_kernel(BgkGeo g, BgkOps<NSP> ops, const double *__restrict__ uv,
                                                         const double *__restrict__ vv,
                                                         const double *__restrict__ tau,
                                                         const double *__restrict__ tab, BgkStage sa) {
  __shared__ double sx[BT + 1][BT][NSP];  // x-face upwind trace, face fx between cells fx-1 and fx
  __shared__ double sy[BT + 1][BT][NSP];  // y-face upwind trace, face fy between rows fy-1 and fy
  const int tiles_x = (g.nx + BT - 1) / BT, tiles_y = (g.ny + BT - 1) / BT;
  long long bid = blockIdx.x;
  const int tile = (int)(bid % (tiles_x * tiles_y));
  bid /= (tiles_x * tiles_y);
  const int vel = (int)(bid % ((long long)g.nu * g.nv));
  const int mem = (int)(bid / ((long long)g.nu * g.nv));
  const int a = vel % g.nu, b = vel / g.nu;
  const double ua = uv[a], vb = vv[b];
  const bool posx = ua > 0.0, posy = vb > 0.0;  // heaviside of the velocity components
  const int tx = threadIdx.x % BT, ty = threadIdx.x / BT;
  const int i0 = (tile % tiles_x) * BT, j0 = (tile / tiles_x) * BT;
  const int ncx = min(BT, g.nx - i0), ncy = min(BT, g.ny - j0);
  const int i = i0 + tx, j = j0 + ty;
  const bool act = tx < ncx && ty < ncy;
  const long long cells = (long long)g.nx * g.ny;
  const long long vbase = mem * g.mstride + g.npts * (long long)vel;
  const double *__restrict__ fv = sa.u + vbase;
  auto idx = [&](int ci, int cj, int k, int l) -> long long {
    return (long long)ci + (long long)g.nx * cj + cells * (k + NSP * l);
  };
  double w[NSP][NSP];  // w[k][l]
  if (act) {
#pragma unroll
    for (int l = 0; l < NSP; ++l)
#pragma unroll
      for (int k = 0; k < NSP; ++k) w[k][l] = fv[idx(i, j, k, l)];
  }
  // own flux traces (f_flux = velocity * f / J)
  double fL[NSP], fR[NSP], gB[NSP], gT[NSP];
  if (act) {
#pragma unroll
    for (int l = 0; l < NSP; ++l) {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; ++q) {
        const double fl = ua * w[q][l] * g.iJx;
        s0 += fl * ops.ll[q];
        s1 += fl * ops.lr[q];
      }
      fL[l] = s0;
      fR[l] = s1;
      // downwind face of the cell carries its trace to the neighbour
      if (posx)
        sx[tx + 1][ty][l] = s1;
      else
        sx[tx][ty][l] = s0;
    }
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; ++q) {
        const double gl = vb * w[k][q] * g.iJy;
        s0 += gl * ops.ll[q];
        s1 += gl * ops.lr[q];
      }
      gB[k] = s0;
      gT[k] = s1;
      if (posy)
        sy[ty + 1][tx][k] = s1;
      else
        sy[ty][tx][k] = s0;
    }
  }
  // halo: the upwind neighbour cells of the tile edges (periodic wrap)
  const int t = threadIdx.x;
  if (t < 2 * BT) {
    const bool xside = t < BT;
    const int pos = t % BT;
    if (xside && pos < ncy) {
      const int hi = posx ? (i0 - 1 + g.nx) % g.nx : (i0 + ncx) % g.nx;
      const int hj = j0 + pos;
      const int fx = posx ? 0 : ncx;
#pragma unroll
      for (int l = 0; l < NSP; ++l) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < NSP; ++q) s += ua * fv[idx(hi, hj, q, l)] * g.iJx * (posx ? ops.lr[q] : ops.ll[q]);
        sx[fx][pos][l] = s;
      }
    } else if (!xside && pos < ncx) {
      const int hj = posy ? (j0 - 1 + g.ny) % g.ny : (j0 + ncy) % g.ny;
      const int hi = i0 + pos;
      const int fy = posy ? 0 : ncy;
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < NSP; ++q) s += vb * fv[idx(hi, hj, k, q)] * g.iJy * (posy ? ops.lr[q] : ops.ll[q]);
        sy[fy][pos][k] = s;
      }
    }
  }
  __syncthreads();
  if (!act) return;
  const double itau = 1.0 / tau[mem];
  const double *__restrict__ Ea = tab + (long long)mem * (g.nu + g.nv) * g.npts + g.npts * (long long)a;
  const double *__restrict__ Eb = tab + (long long)mem * (g.nu + g.nv) * g.npts + g.npts * (long long)(g.nu + b);
  double *__restrict__ out = sa.out + vbase;
#pragma unroll
  for (int l = 0; l < NSP; ++l)
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      // x: upwind interface fluxes on the left / right faces of row l
      const double FsL = posx ? sx[tx][ty][l] : fL[l];
      const double FsR = posx ? fR[l] : sx[tx + 1][ty][l];
      const double GsB = posy ? sy[ty][tx][k] : gB[k];
      const double GsT = posy ? gT[k] : sy[ty + 1][tx][k];
      double dfx = 0.0, dfy = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; ++q) {
        dfx += ua * w[q][l] * g.iJx * ops.dl[k + NSP * q];
        dfy += vb * w[k][q] * g.iJy * ops.dl[l + NSP * q];
      }
      const long long pt = (long long)i + (long long)g.nx * j + cells * (k + NSP * l);
      const double M = Eb[pt] * Ea[pt];
      const double fc = w[k][l];
      const double r = -(dfx + (FsL - fL[l]) * ops.dhl[k] + (FsR - fR[l]) * ops.dhr[k] + dfy +
                         (GsB - gB[k]) * ops.dhl[l] + (GsT - gT[k]) * ops.dhr[l]) +
                       (M - fc) * itau;
      const long long o = idx(i, j, k, l);
      if (sa.residual) {
        out[o] = r;
      } else {
        double v = sa.a1 * fc;
        if (sa.a0 != 0.0) v = sa.a0 * sa.u0[vbase + o] + v;
        out[o] = v + sa.b * r;
      }
    }
}

// ------------------------------------------------------------------ context
struct BgkCtx {
  int nx = 0, ny = 0, nsp = 0, nu = 0, nv = 0, nmem = 1;
  double dx = 1, dy = 1, gamma = 2.0;
  HostOps ops;
  size_t len = 0;
  hipStream_t stream = nullptr;
  double *d_vel = nullptr;  // uv | vv | wu | wv
  double *d_tau = nullptr;
  double *d_prim = nullptr;
  double *d_tab = nullptr;  // Maxwellian factors Ea | Eb per member
  double *d_u = nullptr, *d_s1 = nullptr, *d_s2 = nullptr, *d_in = nullptr, *d_out = nullptr;

  BgkGeo geo() const {
    BgkGeo g;
    g.nx = nx;
    g.ny = ny;
    g.nu = nu;
    g.nv = nv;
    g.nmem = nmem;
    g.npts = (long long)nx * ny * nsp * nsp;
    g.mstride = g.npts * nu * nv;
    g.iJx = 2.0 / dx;
    g.iJy = 2.0 / dy;
    g.gamma = gamma;
    return g;
  }

  template <int NSP>
  int launch_rhs(const BgkStage &sa, hipStream_t st) const {
    BgkOps<NSP> o;
    for (int q = 0; q < NSP; ++q) {
      o.ll[q] = ops.ll[q];
      o.lr[q] = ops.lr[q];
      o.dhl[q] = ops.dhl[q];
      o.dhr[q] = ops.dhr[q];
    }
    for (int q = 0; q < NSP * NSP; ++q) o.dl[q] = ops.dl[q];
    const BgkGeo g = geo();
    const long long tiles = (long long)((nx + BT - 1) / BT) * ((ny + BT - 1) / BT);
    const long long nb = tiles * nu * nv * nmem;
    hipLaunchKernelGGL(bgk_rhs_kernel<NSP>, dim3((unsigned)nb), dim3(BT * BT), 0, st, g, o, d_vel, d_vel + nu,
                       d_tau, d_tab, sa);
    FR_LAUNCHED("bgk_rhs_kernel");
    return FR_OK;
  }

  // moments kernel launch: nq wavefronts per group of 64 points, enough for about eight
  // wavefronts per SIMD; MQ / nq point groups per workgroup
  int launch_moments(const double *d_f, hipStream_t st) const {
    const BgkGeo g = geo();
    const long long n = g.npts * nmem, groups = (n + 63) / 64;
    int nq = 1;
    while (nq < MQ && groups * nq < 8LL * 1024) nq *= 2;
    const long long nb = (groups + (MQ / nq) - 1) / (MQ / nq);
    hipLaunchKernelGGL(bgk_moments_kernel, dim3((unsigned)nb), dim3(64 * MQ), 0, st, g, d_f, d_vel, d_vel + nu,
                       d_vel + nu + nv, d_vel + 2 * nu + nv, d_prim, d_tab, nq);
    FR_LAUNCHED("bgk_moments_kernel");
    return FR_OK;
  }

  int stage(const BgkStage &sa, hipStream_t st) const {
    int rc0 = launch_moments(sa.u, st);
    if (rc0) return rc0;
    switch (nsp) {
      case 2: return launch_rhs<2>(sa, st);
      case 3: return launch_rhs<3>(sa, st);
      case 4: return launch_rhs<4>(sa, st);
      default: return launch_rhs<5>(sa, st);
    }
  }

  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    double *bufs[] = {d_vel, d_tau, d_prim, d_tab, d_u, d_s1, d_s2, d_in, d_out};
    for (double *b : bufs)
      if (b) (void)hipFree(b);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace fr

using namespace fr;

struct fr_bgk_ctx : BgkCtx {};

extern "C" {

int fr_bgk_create(fr_bgk_ctx **out, const fr_bgk_desc *d) {
  FR_REQUIRE(out && d, FR_E_INVALID, "fr_bgk_create: null argument");
  *out = nullptr;
  FR_REQUIRE(d->nx >= 1 && d->ny >= 1 && d->nu >= 1 && d->nv >= 1 && d->nmem >= 1, FR_E_INVALID,
             "fr_bgk_create: sizes must be >= 1");
  FR_REQUIRE(d->deg >= 1 && d->deg <= 4, FR_E_UNSUPPORTED, "fr_bgk_create: degree must be in 1..4");
  FR_REQUIRE(d->dx > 0 && d->dy > 0 && d->gamma > 1.0, FR_E_INVALID, "fr_bgk_create: dx, dy > 0, gamma > 1");
  FR_REQUIRE(d->u && d->v && d->wu && d->wv && d->tau, FR_E_INVALID, "fr_bgk_create: null velocity grid or tau");
  for (int m = 0; m < d->nmem; ++m) FR_REQUIRE(d->tau[m] > 0.0, FR_E_INVALID, "fr_bgk_create: tau must be > 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FR_E_HIP, "fr_bgk_create: no HIP device visible");
  fr_bgk_ctx *c = new (std::nothrow) fr_bgk_ctx();
  FR_REQUIRE(c, FR_E_NOMEM, "fr_bgk_create: out of host memory");
  c->nx = d->nx;
  c->ny = d->ny;
  c->nsp = d->deg + 1;
  c->nu = d->nu;
  c->nv = d->nv;
  c->nmem = d->nmem;
  c->dx = d->dx;
  c->dy = d->dy;
  c->gamma = d->gamma;
  if (!build_ops(d->deg, FR_CORR_RADAU, c->ops)) {
    delete c;
    return fail(FR_E_INVALID, "fr_bgk_create: operator construction failed");
  }
  const BgkGeo g = c->geo();
  c->len = (size_t)g.mstride * d->nmem;
  std::vector<double> vel;
  vel.insert(vel.end(), d->u, d->u + d->nu);
  vel.insert(vel.end(), d->v, d->v + d->nv);
  vel.insert(vel.end(), d->wu, d->wu + d->nu);
  vel.insert(vel.end(), d->wv, d->wv + d->nv);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->d_vel, vel.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(c->d_vel, vel.data(), vel.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->d_tau, d->nmem * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(c->d_tau, d->tau, d->nmem * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->d_prim, (size_t)g.npts * 4 * d->nmem * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->d_tab, (size_t)g.npts * (d->nu + d->nv) * d->nmem * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->d_u, c->len * sizeof(double));
  if (e != hipSuccess) {
    int rc = hip_fail(e, "fr_bgk_create", __FILE__, __LINE__);
    c->release();
    delete c;
    return rc;
  }
  *out = c;
  return FR_OK;
}

int fr_bgk_destroy(fr_bgk_ctx *c) {
  if (!c) return FR_OK;
  c->release();
  delete c;
  return FR_OK;
}

size_t fr_bgk_state_len(const fr_bgk_ctx *c) { return c ? c->len : 0; }

int fr_bgk_rhs(fr_bgk_ctx *c, const double *f, double *df) {
  FR_REQUIRE(c && f && df, FR_E_INVALID, "fr_bgk_rhs: null argument");
  const size_t bytes = c->len * sizeof(double);
  if (!c->d_in) FR_HIP(hipMalloc(&c->d_in, bytes));
  if (!c->d_out) FR_HIP(hipMalloc(&c->d_out, bytes));
  FR_HIP(hipMemcpyAsync(c->d_in, f, bytes, hipMemcpyHostToDevice, c->stream));
  BgkStage sa{c->d_in, nullptr, c->d_out, 0.0, 0.0, 0.0, 1};
  int rc = c->stage(sa, c->stream);
  if (rc) return rc;
  FR_HIP(hipMemcpyAsync(df, c->d_out, bytes, hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr_bgk_set_state(fr_bgk_ctx *c, const double *f) {
  FR_REQUIRE(c && f, FR_E_INVALID, "fr_bgk_set_state: null argument");
  FR_HIP(hipMemcpyAsync(c->d_u, f, c->len * sizeof(double), hipMemcpyHostToDevice, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr_bgk_set_state_dev(fr_bgk_ctx *c, const double *d_f) {
  FR_REQUIRE(c && d_f, FR_E_INVALID, "fr_bgk_set_state_dev: null argument");
  FR_HIP(hipMemcpyAsync(c->d_u, d_f, c->len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

int fr_bgk_get_state(fr_bgk_ctx *c, double *f) {
  FR_REQUIRE(c && f, FR_E_INVALID, "fr_bgk_get_state: null argument");
  FR_HIP(hipMemcpyAsync(f, c->d_u, c->len * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

// nsteps Shu-Osher SSP-RK3 steps of all members; each stage = moments + transport/relaxation
int fr_bgk_step(fr_bgk_ctx *c, double dt, int nsteps) {
  FR_REQUIRE(c && nsteps >= 0 && std::isfinite(dt), FR_E_INVALID, "fr_bgk_step: bad arguments");
  if (!c->d_s1) FR_HIP(hipMalloc(&c->d_s1, c->len * sizeof(double)));
  if (!c->d_s2) FR_HIP(hipMalloc(&c->d_s2, c->len * sizeof(double)));
  // Consecutive steps alternate the state between d_u and d_s1 (every stage writes every value),
  // so the last stage never writes the u^n it reads; the call's last step writes d_u.
  for (int n = 0; n < nsteps; ++n) {
    int rc;
    double *const X = (n & 1) ? c->d_s1 : c->d_u, *const Y = (n & 1) ? c->d_u : c->d_s1;
    double *const out = n == nsteps - 1 ? c->d_u : Y;
    if ((rc = c->stage(BgkStage{X, X, Y, 0.0, 1.0, dt, 0}, c->stream))) return rc;
    if ((rc = c->stage(BgkStage{Y, X, c->d_s2, 0.75, 0.25, 0.25 * dt, 0}, c->stream))) return rc;
    if ((rc = c->stage(BgkStage{c->d_s2, X, out, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt, 0}, c->stream)))
      return rc;
  }
  FR_HIP(hipStreamSynchronize(c->stream));
  return FR_OK;
}

// conserved moments (rho, rho U, rho V, E) of the device state at every solution point, host
// output w[member][c][pt]
int fr_bgk_moments(fr_bgk_ctx *c, double *w) {
  FR_REQUIRE(c && w, FR_E_INVALID, "fr_bgk_moments: null argument");
  const BgkGeo g = c->geo();
  const long long n = g.npts * c->nmem;
  int rc = c->launch_moments(c->d_u, c->stream);
  if (rc) return rc;
  std::vector<double> prim((size_t)n * 4);
  FR_HIP(hipMemcpyAsync(prim.data(), c->d_prim, prim.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  // prim -> conserved (KitBase prim_conserve)
  for (long long mem = 0; mem < c->nmem; ++mem)
    for (long long p = 0; p < g.npts; ++p) {
      const double *pr = prim.data() + mem * 4 * g.npts + p;
      const double rho = pr[0], U = pr[g.npts], V = pr[2 * g.npts], lam = pr[3 * g.npts];
      double *o = w + mem * 4 * g.npts + p;
      o[0] = rho;
      o[g.npts] = rho * U;
      o[2 * g.npts] = rho * V;
      o[3 * g.npts] = 0.5 * rho / lam / (c->gamma - 1.0) + 0.5 * rho * (U * U + V * V);
    }
  return FR_OK;
}

// Gauss-Hermite rule for the weight exp(-x^2): nodes ascending, weights (host, native).
// The orthonormal Hermite functions p_j follow p_{j+1} = x sqrt(2/(j+1)) p_j - sqrt(j/(j+1)) p_{j-1},
// p_0 = pi^(-1/4), with p_n' = sqrt(2n) p_{n-1}.  All n zeros lie in |x| < sqrt(2n+1); a sign scan
// of p_n on a grid finer than the closest pair of zeros brackets each one, bisection narrows the
// bracket and Newton polishes it; the weight is 2 / p_n'(x)^2.
static void hermite_pn(int n, double x, double &pn, double &pn1) {
  double p1 = 0.7511255444649425, p0 = 0.0;  // pi^(-1/4), p_{-1}
  for (int j = 0; j < n; ++j) {
    const double p2 = x * std::sqrt(2.0 / (j + 1)) * p1 - std::sqrt((double)j / (j + 1)) * p0;
    p0 = p1;
    p1 = p2;
  }
  pn = p1;
  pn1 = p0;
}

int fr_gauss_hermite(int n, double *x, double *w) {
  FR_REQUIRE(n >= 1 && n <= 200 && x && w, FR_E_INVALID, "fr_gauss_hermite: bad arguments");
  const double R = std::sqrt(2.0 * n + 1.0);
  std::vector<double> lo, hi;
  for (int cells = 64 * n; (int)lo.size() != n; cells *= 2) {
    if (cells > (1 << 24)) return fail(FR_E_INVALID, "fr_gauss_hermite: zeros not separated");
    lo.clear();
    hi.clear();
    double xa = -R, pa, q;
    hermite_pn(n, xa, pa, q);
    for (int c = 1; c <= cells; ++c) {
      const double xb = -R + 2.0 * R * c / cells;
      double pb;
      hermite_pn(n, xb, pb, q);
      if ((pa < 0.0) != (pb < 0.0) || pb == 0.0) {
        lo.push_back(xa);
        hi.push_back(xb);
        if (pb == 0.0) {  // step past an exact zero so it is bracketed once
          xa = xb + 1e-300;
          hermite_pn(n, xa, pa, q);
          continue;
        }
      }
      xa = xb;
      pa = pb;
    }
  }
  for (int i = 0; i < n; ++i) {
    double a = lo[i], b = hi[i], pa, q;
    hermite_pn(n, a, pa, q);
    for (int it = 0; it < 200 && b - a > 1e-13 * std::max(1.0, std::fabs(a)); ++it) {
      const double m = 0.5 * (a + b);
      double pm;
      hermite_pn(n, m, pm, q);
      if ((pm < 0.0) == (pa < 0.0)) {
        a = m;
        pa = pm;
      } else {
        b = m;
      }
    }
    double z = 0.5 * (a + b), d = 0.0;
    for (int it = 0; it < 5; ++it) {
      double pn, pn1;
      hermite_pn(n, z, pn, pn1);
      d = std::sqrt(2.0 * n) * pn1;
      z -= pn / d;
    }
    double pn, pn1;
    hermite_pn(n, z, pn, pn1);
    d = std::sqrt(2.0 * n) * pn1;
    x[i] = z;
    w[i] = 2.0 / (d * d);
  }
  // symmetric rule: pair the nodes exactly
  for (int i = 0; i < n / 2; ++i) {
    const double z = 0.5 * (x[n - 1 - i] - x[i]), ww = 0.5 * (w[i] + w[n - 1 - i]);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = ww;
  }
  if (n & 1) x[n / 2] = 0.0;
  return FR_OK;
}

}  // extern "C"
