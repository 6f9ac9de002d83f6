// fr2d_api.hip -- C ABI of the 2D Euler solver: context, drop-in residual, device-resident
// integrator, dissipation calls (declared in include/fr.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "fr2d_internal.h"
#include "fr_tsit5.h"

namespace fr {

// ------------------------------------------------------------------ scratch of the allocation-free step
// fr2d_step_dev_noalloc takes its two stage buffers from this array: kNoAllocStateMax doubles
// each, one copy per GPU, held from module load until the process ends.  Host code never passes
// the array's name to a kernel (in host code it is a stand-in, not a device address): the device
// address comes from hipGetSymbolAddress, called once per device at context creation with that
// device current, and is kept in noalloc_ptr.
constexpr size_t kNoAllocStateMax = 6291456;  // 48 MiB of doubles per stage buffer
static __device__ __align__(256) double scratch[2 * kNoAllocStateMax];
constexpr int kNoAllocDevices = 64;
static std::mutex noalloc_mu;
static double *noalloc_ptr[kNoAllocDevices] = {};

// resolves the scratch address of the current device once (fr2d_create)
static void noalloc_setup(int dev) {
  if (dev < 0 || dev >= kNoAllocDevices) return;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (noalloc_ptr[dev]) return;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) noalloc_ptr[dev] = static_cast<double *>(p);
  else (void)hipGetLastError();  // reported by fr2d_step_dev_noalloc, which cannot run without it
}


// layout of Fr2dCtx::d_dops: V (np^2) | iV (np^2) | wp/4 (np) | ll (nsp) | lr (nsp)
DissOps fr2d_diss_ops(const Fr2dCtx &c) {
  const int np = c.nsp * c.nsp;
  DissOps o;
  o.V = c.d_dops;
  o.iV = c.d_dops + np * np;
  o.wts = c.d_dops + 2 * np * np;
  o.ll = o.wts + np;
  o.lr = o.ll + c.nsp;
  return o;
}

static hipStream_t stream_of(const Fr2dCtx &c) { return fr2d_stream(c); }

// one padded device array (Fr2dCtx: pitch, ldpi): raw is the allocation, the array starts
// kPadFront doubles into it; the padding is zeroed once and never read.  The zero fill completes
// before the array is handed out, so a caller that writes the array through fr2d_state_dev on
// another stream is not overtaken by it.
static int alloc_padded(const Fr2dCtx &c, double *&raw, double *&arr) {
  const size_t n = c.lenp + Fr2dCtx::kPadFront + 1;
  FR_HIP(hipMalloc(&raw, n * sizeof(double)));
  hipStream_t st = fr2d_stream(c);
  FR_HIP(hipMemsetAsync(raw, 0, n * sizeof(double), st));
  FR_HIP(hipStreamSynchronize(st));
  arr = raw + Fr2dCtx::kPadFront;
  return FR_OK;
}

static int ensure_state(Fr2dCtx &c) {
  if (!c.d_u) return alloc_padded(c, c.raw_u, c.d_u);
  return FR_OK;
}

static int ensure_stage_buffers(Fr2dCtx &c, bool acc, bool tsit = false) {
  int rc = FR_OK;
  if (!c.d_s1 && (rc = alloc_padded(c, c.raw_s1, c.d_s1))) return rc;
  if (!c.d_s2 && (rc = alloc_padded(c, c.raw_s2, c.d_s2))) return rc;
  if (acc && !c.d_acc && (rc = alloc_padded(c, c.raw_acc, c.d_acc))) return rc;
  if (tsit && !c.d_k) {
    // six arrays with the state's alignment: slots of a multiple of 256 bytes, each array
    // kPadFront doubles into its slot as the state is into its allocation
    const size_t slot = (c.lenp + Fr2dCtx::kPadFront + 1 + 31) / 32 * 32;
    FR_HIP(hipMalloc(&c.raw_k, 6 * slot * sizeof(double)));
    hipStream_t st = fr2d_stream(c);
    FR_HIP(hipMemsetAsync(c.raw_k, 0, 6 * slot * sizeof(double), st));
    FR_HIP(hipStreamSynchronize(st));
    c.d_k = c.raw_k + Fr2dCtx::kPadFront;
    c.kstride = slot;
  }
  return FR_OK;
}

// copies between a reference-layout array (rows of nx+2) and a padded one (rows of pitch)
static int copy_to_padded(const Fr2dCtx &c, double *dst, const double *src, hipMemcpyKind kind, hipStream_t st) {
  const size_t w = (size_t)(c.nx + 2) * sizeof(double);
  FR_HIP(hipMemcpy2DAsync(dst, (size_t)c.pitch * sizeof(double), src, w, w, (size_t)(c.ny + 2) * c.nsp * c.nsp * 4,
                          kind, st));
  return FR_OK;
}
static int copy_from_padded(const Fr2dCtx &c, double *dst, const double *src, hipMemcpyKind kind, hipStream_t st) {
  const size_t w = (size_t)(c.nx + 2) * sizeof(double);
  FR_HIP(hipMemcpy2DAsync(dst, w, src, (size_t)c.pitch * sizeof(double), w, (size_t)(c.ny + 2) * c.nsp * c.nsp * 4,
                          kind, st));
  return FR_OK;
}

// Stage launches of one explicit step on the device state (see fr.h: fr2d_step).  The first stage
// of every step copies the state's ghost layer into both stage buffers (from its halo loads), and
// no stage writes a ghost cell of its result, so FR_BC_GHOST keeps the ghost layer fixed during a
// step as the reference does (du = 0 there).
static int enqueue_steps(const Fr2dCtx &c, int scheme, double dt, int nsteps, hipStream_t st) {
  double *U = c.d_u, *S1 = c.d_s1, *S2 = c.d_s2;
  for (int n = 0; n < nsteps; ++n) {
    int rc = FR_OK;
    switch (scheme) {
      case FR_RK_SSPRK3: {
        // Shu-Osher: u1 = u + dt L(u); u2 = 3/4 u + 1/4 u1 + dt/4 L(u1); u = 1/3 u + 2/3 u2 + 2dt/3 L(u2).
        // Consecutive steps of one call alternate the state between U and S1: the last stage
        // writes u^{n+1} over u1 instead of over the u^n it reads (at 256^2, p = 3 an in-place
        // last stage measured 22.3 us against 20.8 us).  Odd steps of the call end in U, and the
        // call's last step writes U, so the state is in U between calls.
        double *const A = (n & 1) ? S1 : U;  // u^n
        double *const B = (n & 1) ? U : S1;  // u1
        double *const out = n == nsteps - 1 ? U : B;
        if ((rc = fr2d_launch_stage(c, A, A, B, 0.0, 1.0, dt, nullptr, nullptr, 0.0, st, B, S2))) return rc;
        if ((rc = fr2d_launch_stage(c, A, B, S2, 0.75, 0.25, 0.25 * dt, nullptr, nullptr, 0.0, st))) return rc;
        if ((rc = fr2d_launch_stage(c, A, S2, out, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt, nullptr, nullptr, 0.0, st)))
          return rc;
        break;
      }
      case FR_RK_RK4: {
        double *A = c.d_acc;
        if ((rc = fr2d_launch_stage(c, U, U, S2, 1.0, 0.0, 0.5 * dt, U, A, dt / 6.0, st, S1, S2))) return rc;
        if ((rc = fr2d_launch_stage(c, U, S2, S1, 1.0, 0.0, 0.5 * dt, A, A, dt / 3.0, st))) return rc;
        if ((rc = fr2d_launch_stage(c, U, S1, S2, 1.0, 0.0, dt, A, A, dt / 3.0, st))) return rc;
        if ((rc = fr2d_launch_stage(c, A, S2, U, 1.0, 0.0, dt / 6.0, nullptr, nullptr, 0.0, st))) return rc;
        break;
      }
      case FR_RK_MIDPOINT: {
        // as SSP-RK3: consecutive steps alternate the state between U and S2, the call's last
        // step writes U (the first stage gives both S1 and the step's output its ghost layer)
        double *const X = (n & 1) ? S2 : U;
        double *const out = n == nsteps - 1 ? U : ((n & 1) ? U : S2);
        if ((rc = fr2d_launch_stage(c, X, X, S1, 0.0, 1.0, 0.5 * dt, nullptr, nullptr, 0.0, st, S1,
                                    out == X ? S2 : out)))
          return rc;
        if ((rc = fr2d_launch_stage(c, X, S1, out, 1.0, 0.0, dt, nullptr, nullptr, 0.0, st))) return rc;
        break;
      }
      case FR_RK_EULER: {
        // consecutive steps alternate the state between U and S1 (a stage cannot write the array
        // whose halo it reads); a call that ends in S1 copies it back once
        double *const X = (n & 1) ? S1 : U, *const Y = (n & 1) ? U : S1;
        if ((rc = fr2d_launch_stage(c, X, X, Y, 0.0, 1.0, dt, nullptr, nullptr, 0.0, st, Y, S2))) return rc;
        if (n == nsteps - 1 && Y != U)
          FR_HIP(hipMemcpyAsync(U, Y, c.lenp * sizeof(double), hipMemcpyDeviceToDevice, st));
        break;
      }
      case FR_RK_TSIT5: {
        // k_s = L(Y_s) into the padded k arrays (a stage launch with a0 = a1 = 0, b = 1, which
        // writes no ghost cell, so the zeroed ghost layer of every k stays zero), Y_1 = U,
        // Y_{s+1} = U + dt sum_j a_{s+1,j} k_j and U <- U + dt sum_j b_j k_j over the whole padded
        // arrays: the ghost layer of every Y and of U stays that of u^n, as the reference holds it
        for (int sg = 0; sg < 6; ++sg) {
          const double *Y = sg == 0 ? U : S1;
          double *K = c.d_k + (size_t)sg * c.kstride;
          if ((rc = fr2d_launch_stage(c, nullptr, Y, K, 0.0, 0.0, 1.0, nullptr, nullptr, 0.0, st))) return rc;
          FR_HIP(tsit5_comb(sg == 5 ? U : S1, U, c.d_k, c.kstride, kTsit5[sg], sg + 1, dt, c.lenp, st));
        }
        break;
      }
      default:
        return fail(FR_E_INVALID, "fr2d_step: unknown scheme");
    }
  }
  return FR_OK;
}

static int prepare_step(Fr2dCtx &c, int scheme, hipStream_t st) {
  FR_REQUIRE(c.d_u, FR_E_INVALID, "fr2d_step: no state (call fr2d_set_state first)");
  // the stage buffers take the state's ghost layer inside the first stage of every step
  (void)st;
  int rc = ensure_stage_buffers(c, scheme == FR_RK_RK4, scheme == FR_RK_TSIT5);
  // the first stepping call of a context picks the staggered-start delay on the device, on the
  // context's own stream (so a caller stream under capture records nothing of it); it reads the
  // state and writes only the first stage buffer, which the steps overwrite
  if (!rc && c.stagger < 0) {
    if ((rc = fr2d_tune_stagger(c, c.own_stream))) return rc;
    FR_HIP(hipStreamSynchronize(c.own_stream));
  }
  return rc;
}

static int finish_step(const Fr2dCtx &c, hipStream_t st) {
  if (c.bc == FR_BC_PERIODIC) return fr2d_launch_fill_ghost(c, c.d_u, FR_GHOST_PERIODIC, st);
  return FR_OK;
}

}  // namespace fr

using namespace fr;


extern "C" {

int fr2d_create(fr2d_ctx **out, const fr2d_desc *d) {
  FR_REQUIRE(out && d, FR_E_INVALID, "fr2d_create: null argument");
  *out = nullptr;
  FR_REQUIRE(d->nx >= 1 && d->ny >= 1, FR_E_INVALID, "fr2d_create: nx, ny must be >= 1");
  FR_REQUIRE(d->deg >= 1 && d->deg <= 5, FR_E_UNSUPPORTED, "fr2d_create: degree must be in 1..5");
  FR_REQUIRE(d->flux >= FR_FLUX_HLL && d->flux <= FR_FLUX_ROE, FR_E_INVALID, "fr2d_create: unknown flux");
  FR_REQUIRE(d->bc == FR_BC_GHOST || d->bc == FR_BC_PERIODIC, FR_E_INVALID, "fr2d_create: unknown bc");
  FR_REQUIRE(d->gamma > 1.0 && std::isfinite(d->gamma), FR_E_INVALID, "fr2d_create: gamma must be > 1");
  FR_REQUIRE(d->iJ11 > 0.0 && d->iJ22 > 0.0, FR_E_INVALID, "fr2d_create: iJ11, iJ22 must be > 0");
  // the kernels address one (k, l, m) plane with a 32-bit byte offset per element
  FR_REQUIRE((double)((d->nx + 2 + 15) / 16 * 16) * (d->ny + 2) * 8.0 < 4294967296.0, FR_E_INVALID,
             "fr2d_create: grid too large (a padded plane, (nx+2 rounded up to 16)(ny+2), must be below 2^29)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(FR_E_HIP, "fr2d_create: no HIP device visible");
  fr2d_ctx *c = new (std::nothrow) fr2d_ctx();
  FR_REQUIRE(c, FR_E_NOMEM, "fr2d_create: out of host memory");
  c->nx = d->nx;
  c->ny = d->ny;
  c->deg = d->deg;
  c->nsp = d->deg + 1;
  c->flux = d->flux;
  c->bc = d->bc;
  c->gamma = d->gamma;
  c->iJ11 = d->iJ11;
  c->iJ22 = d->iJ22;
  c->len = (size_t)(d->nx + 2) * (d->ny + 2) * c->nsp * c->nsp * 4;
  c->pitch = (d->nx + 2 + 15) / 16 * 16;
  c->ldpi = (long long)c->pitch * (d->ny + 2);
  c->lenp = (size_t)c->ldpi * c->nsp * c->nsp * 4;
  if (!build_ops(d->deg, FR_CORR_RADAU, c->ops)) {
    delete c;
    return fail(FR_E_INVALID, "fr2d_create: operator construction failed");
  }
  const int n = c->nsp, np = n * n;
  // Caller-provided residual operators (FRPSpace2D fields) must be the Gauss-Legendre / Radau
  // ones that FRPSpace2D builds (src/struct.jl:193) and that the kernels carry as compile-time
  // constants; a caller's independently computed copy differs from them by round-off only.
  {
    auto same = [&](const double *mine, const double *theirs, int cnt) {
      if (!theirs) return true;
      double scale = 0.0;
      for (int q = 0; q < cnt; ++q) scale = std::max(scale, std::fabs(mine[q]));
      for (int q = 0; q < cnt; ++q)
        if (!(std::fabs(mine[q] - theirs[q]) <= 1e-12 * scale)) return false;
      return true;
    };
    if (!same(c->ops.ll, d->ll, n) || !same(c->ops.lr, d->lr, n) || !same(c->ops.dl, d->dl, n * n) ||
        !same(c->ops.dhl, d->dhl, n) || !same(c->ops.dhr, d->dhr, n)) {
      delete c;
      return fail(FR_E_UNSUPPORTED,
                  "fr2d_create: ll/lr/dl/dhl/dhr differ from the Gauss-Legendre / Radau operators of this degree");
    }
  }
  if (d->V) std::memcpy(c->ops.V, d->V, sizeof(double) * np * np);
  if (d->iV) std::memcpy(c->ops.iV, d->iV, sizeof(double) * np * np);
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) noalloc_setup(c->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->d_count, 4 * sizeof(int));
  const size_t nd = 2 * (size_t)np * np + np + 2 * n;
  if (e == hipSuccess) e = hipMalloc(&c->d_dops, nd * sizeof(double));
  if (e == hipSuccess) {
    double host[2 * kMaxNp2D * kMaxNp2D + kMaxNp2D + 12];
    std::memcpy(host, c->ops.V, sizeof(double) * np * np);
    std::memcpy(host + np * np, c->ops.iV, sizeof(double) * np * np);
    for (int q = 0; q < np; ++q) host[2 * np * np + q] = c->ops.wp[q] / 4.0;
    std::memcpy(host + 2 * np * np + np, c->ops.ll, sizeof(double) * n);
    std::memcpy(host + 2 * np * np + np + n, c->ops.lr, sizeof(double) * n);
    e = hipMemcpy(c->d_dops, host, nd * sizeof(double), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    int rc = hip_fail(e, "fr2d_create", __FILE__, __LINE__);
    fr2d_destroy(c);
    return rc;
  }
  *out = c;
  return FR_OK;
}

int fr2d_destroy(fr2d_ctx *c) {
  if (!c) return FR_OK;
  hipStream_t st = stream_of(*c);
  if (st) (void)hipStreamSynchronize(st);
  double *bufs[] = {c->raw_u,  c->raw_s1,   c->raw_s2,   c->raw_acc,  c->raw_k,
                    c->d_in,   c->d_out,    c->d_dops,   c->d_mj_ref, c->raw_mj_pad};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  if (c->d_count) (void)hipFree(c->d_count);
  // d_mlp refers to the caller's network (fr2d_set_sensor_mlp), which the context does not own
  if (c->d_feat) (void)hipFree(c->d_feat);
  if (c->d_strength) (void)hipFree(c->d_strength);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return FR_OK;
}

int fr2d_set_stream(fr2d_ctx *c, void *s) {
  FR_REQUIRE(c, FR_E_INVALID, "fr2d_set_stream: null context");
  if (s == FR_OWN_STREAM) {
    c->caller_stream = false;
    c->stream = nullptr;
  } else {
    c->caller_stream = true;
    c->stream = (hipStream_t)s;
  }
  return FR_OK;
}

void *fr2d_get_stream(fr2d_ctx *c) { return c ? (void *)stream_of(*c) : nullptr; }

size_t fr2d_state_len(const fr2d_ctx *c) { return c ? c->len : 0; }

int fr2d_euler_rhs_dev(fr2d_ctx *c, const double *d_u, double *d_du, double t) {
  (void)t;  // autonomous system: the reference dudt! ignores t
  FR_REQUIRE(c && d_u && d_du, FR_E_INVALID, "fr2d_euler_rhs_dev: null argument");
  FR_REQUIRE(d_u != d_du, FR_E_INVALID, "fr2d_euler_rhs_dev: u and du must not alias");
  return fr2d_launch_residual(*c, d_u, d_du, stream_of(*c));
}

int fr2d_euler_rhs(fr2d_ctx *c, const double *u, double *du, double t) {
  (void)t;
  FR_REQUIRE(c && u && du, FR_E_INVALID, "fr2d_euler_rhs: null argument");
  hipStream_t st = stream_of(*c);
  const size_t bytes = c->len * sizeof(double);
  if (!c->d_in) FR_HIP(hipMalloc(&c->d_in, bytes));
  if (!c->d_out) FR_HIP(hipMalloc(&c->d_out, bytes));
  FR_HIP(hipMemcpyAsync(c->d_in, u, bytes, hipMemcpyHostToDevice, st));
  int rc = fr2d_launch_residual(*c, c->d_in, c->d_out, st);
  if (rc) return rc;
  FR_HIP(hipMemcpyAsync(du, c->d_out, bytes, hipMemcpyDeviceToHost, st));
  FR_HIP(hipStreamSynchronize(st));
  return FR_OK;
}

// Host states cross the bus as one contiguous copy of the reference layout through the staging
// array d_in; the pitch changes on the device.  (A pitched copy between pageable host memory and
// the padded array moves one row of nx+2 doubles at a time: 117 ms for a 256 x 256, p = 3 state.)
static int ensure_staging(Fr2dCtx &c) {
  if (!c.d_in) FR_HIP(hipMalloc(&c.d_in, c.len * sizeof(double)));
  return FR_OK;
}

int fr2d_set_state(fr2d_ctx *c, const double *u) {
  FR_REQUIRE(c && u, FR_E_INVALID, "fr2d_set_state: null argument");
  int rc = ensure_state(*c);
  if (!rc) rc = ensure_staging(*c);
  if (rc) return rc;
  hipStream_t st = stream_of(*c);
  FR_HIP(hipMemcpyAsync(c->d_in, u, c->len * sizeof(double), hipMemcpyHostToDevice, st));
  if ((rc = copy_to_padded(*c, c->d_u, c->d_in, hipMemcpyDeviceToDevice, st))) return rc;
  FR_HIP(hipStreamSynchronize(st));
  return FR_OK;
}

int fr2d_get_state(fr2d_ctx *c, double *u) {
  FR_REQUIRE(c && u, FR_E_INVALID, "fr2d_get_state: null argument");
  FR_REQUIRE(c->d_u, FR_E_INVALID, "fr2d_get_state: no state");
  int rc = ensure_staging(*c);
  if (rc) return rc;
  hipStream_t st = stream_of(*c);
  if ((rc = copy_from_padded(*c, c->d_in, c->d_u, hipMemcpyDeviceToDevice, st))) return rc;
  FR_HIP(hipMemcpyAsync(u, c->d_in, c->len * sizeof(double), hipMemcpyDeviceToHost, st));
  FR_HIP(hipStreamSynchronize(st));
  return FR_OK;
}

int fr2d_set_state_dev(fr2d_ctx *c, const double *d_u) {
  FR_REQUIRE(c && d_u, FR_E_INVALID, "fr2d_set_state_dev: null argument");
  int rc = ensure_state(*c);
  if (rc) return rc;
  return copy_to_padded(*c, c->d_u, d_u, hipMemcpyDeviceToDevice, stream_of(*c));
}

int fr2d_get_state_dev(fr2d_ctx *c, double *d_u) {
  FR_REQUIRE(c && d_u, FR_E_INVALID, "fr2d_get_state_dev: null argument");
  FR_REQUIRE(c->d_u, FR_E_INVALID, "fr2d_get_state_dev: no state");
  return copy_from_padded(*c, d_u, c->d_u, hipMemcpyDeviceToDevice, stream_of(*c));
}

double *fr2d_state_dev(fr2d_ctx *c) {
  if (!c) return nullptr;
  if (ensure_state(*c)) return nullptr;
  return c->d_u;
}

int fr2d_state_pitch(const fr2d_ctx *c, int *pitch, long long *plane) {
  FR_REQUIRE(c && pitch && plane, FR_E_INVALID, "fr2d_state_pitch: null argument");
  *pitch = c->pitch;
  *plane = c->ldpi;
  return FR_OK;
}

int fr2d_step(fr2d_ctx *c, int scheme, double dt, int nsteps) {
  FR_REQUIRE(c, FR_E_INVALID, "fr2d_step: null context");
  FR_REQUIRE(nsteps >= 0 && std::isfinite(dt), FR_E_INVALID, "fr2d_step: bad dt / nsteps");
  FR_REQUIRE(scheme >= FR_RK_SSPRK3 && scheme <= FR_RK_TSIT5, FR_E_INVALID, "fr2d_step: unknown scheme");
  hipStream_t st = stream_of(*c);
  int rc = prepare_step(*c, scheme, st);
  if (rc) return rc;
  if ((rc = enqueue_steps(*c, scheme, dt, nsteps, st))) return rc;
  return finish_step(*c, st);
}

int fr2d_step_graph(fr2d_ctx *c, int scheme, double dt, int nsteps, int nrep) {
  FR_REQUIRE(c, FR_E_INVALID, "fr2d_step_graph: null context");
  FR_REQUIRE(nsteps >= 1 && nrep >= 0 && std::isfinite(dt), FR_E_INVALID, "fr2d_step_graph: bad arguments");
  FR_REQUIRE(scheme >= FR_RK_SSPRK3 && scheme <= FR_RK_TSIT5, FR_E_INVALID, "fr2d_step_graph: unknown scheme");
  hipStream_t st = stream_of(*c);
  int rc = prepare_step(*c, scheme, st);
  if (rc) return rc;
  // The graph holds the context's own buffers (state and stage arrays never move once allocated),
  // so a capture stays valid while scheme, dt and nsteps are unchanged.
  if (!(c->gexec && c->g_scheme == scheme && c->g_dt == dt && c->g_nsteps == nsteps)) {
    if (c->gexec) {
      (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
    // capture on the private stream so a caller stream in relaxed use is never put in capture
    // mode; nothing runs during a capture, and the replays below are ordered on st
    hipStream_t cap = c->own_stream;
    hipGraph_t graph = nullptr;
    FR_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    rc = enqueue_steps(*c, scheme, dt, nsteps, cap);
    hipError_t e = hipStreamEndCapture(cap, &graph);
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
  return finish_step(*c, st);
}

int fr2d_get_stagger(const fr2d_ctx *c, int *stagger) {
  FR_REQUIRE(c && stagger, FR_E_INVALID, "fr2d_get_stagger: null argument");
  *stagger = c->stagger;
  return FR_OK;
}

int fr2d_set_stagger(fr2d_ctx *c, int stagger) {
  FR_REQUIRE(c, FR_E_INVALID, "fr2d_set_stagger: null context");
  FR_REQUIRE(stagger >= -1 && stagger <= 16, FR_E_INVALID, "fr2d_set_stagger: stagger must be in -1..16");
  // a captured step graph holds the delay of its capture
  if (c->gexec && stagger != c->stagger) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
  c->stagger = stagger;
  return FR_OK;
}

int fr2d_step_dev_noalloc(fr2d_ctx *c, double *d_u, int scheme, double dt, int nsteps) {
  FR_REQUIRE(c && d_u, FR_E_INVALID, "fr2d_step_dev_noalloc: null argument");
  FR_REQUIRE(nsteps >= 0 && std::isfinite(dt), FR_E_INVALID, "fr2d_step_dev_noalloc: bad dt / nsteps");
  FR_REQUIRE(scheme == FR_RK_SSPRK3 || scheme == FR_RK_MIDPOINT || scheme == FR_RK_EULER, FR_E_UNSUPPORTED,
             "fr2d_step_dev_noalloc: SSP-RK3, Midpoint and Euler only (RK4 needs a third array)");
  FR_REQUIRE(c->len <= kNoAllocStateMax, FR_E_UNSUPPORTED,
             "fr2d_step_dev_noalloc: the state has " + std::to_string(c->len) + " doubles, above the " +
                 std::to_string(kNoAllocStateMax) + " of each scratch stage buffer");
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (c->device >= 0 && c->device < kNoAllocDevices) base = noalloc_ptr[c->device];
  }
  FR_REQUIRE(base, FR_E_HIP, "fr2d_step_dev_noalloc: the scratch array of this device has no address");
  // the context's launch parameters with the caller's array in the reference layout (rows of
  // nx+2) as the state and the scratch as the stage buffers, which need no initial values: every
  // interior value is written before it is read, the first stage of every step copies the state's
  // ghost layer into both buffers, and periodic grids never read a ghost layer
  Fr2dCtx r = *c;
  r.pitch = c->nx + 2;
  r.ldpi = (long long)(c->nx + 2) * (c->ny + 2);
  r.lenp = c->len;
  r.d_u = d_u;
  r.d_s1 = base;
  r.d_s2 = base + kNoAllocStateMax;
  r.d_acc = nullptr;
  r.gexec = nullptr;
  hipStream_t st = stream_of(*c);
  int rc = enqueue_steps(r, scheme, dt, nsteps, st);
  if (rc) return rc;
  return finish_step(r, st);
}

int fr2d_fill_ghost(fr2d_ctx *c, int rule) {
  FR_REQUIRE(c && c->d_u, FR_E_INVALID, "fr2d_fill_ghost: no state");
  FR_REQUIRE(rule >= FR_GHOST_PERIODIC && rule <= FR_GHOST_WAVE_Y, FR_E_INVALID, "fr2d_fill_ghost: unknown rule");
  return fr2d_launch_fill_ghost(*c, c->d_u, rule, stream_of(*c));
}

int fr2d_sensor_filter(fr2d_ctx *c, int mode, double lambda, double S0, double kappa, int *n_flagged) {
  FR_REQUIRE(c && c->d_u, FR_E_INVALID, "fr2d_sensor_filter: no state");
  FR_REQUIRE(mode >= FR_SENSOR_OFF && mode <= FR_SENSOR_LEARNED, FR_E_INVALID, "fr2d_sensor_filter: unknown mode");
  FR_REQUIRE(lambda >= 0.0 && kappa > 0.0, FR_E_INVALID, "fr2d_sensor_filter: lambda >= 0 and kappa > 0 required");
  hipStream_t st = stream_of(*c);
  if (mode == FR_SENSOR_OFF) {
    if (n_flagged) *n_flagged = 0;
    return FR_OK;
  }
  if (mode == FR_SENSOR_LEARNED && !c->mlp_loaded) mode = FR_SENSOR_MODAL;  // plain sensor is the fallback
  FR_HIP(hipMemsetAsync(c->d_count, 0, sizeof(int), st));
  int rc = FR_OK;
  if (mode == FR_SENSOR_LEARNED) {
    if ((rc = fr2d_launch_mlp_sensor(*c, c->d_u, st))) return rc;
  }
  if ((rc = fr2d_launch_sensor_filter(*c, c->d_u, mode, lambda, S0, kappa, c->d_count, nullptr, st))) return rc;
  if (n_flagged) {
    FR_HIP(hipMemcpyAsync(n_flagged, c->d_count, sizeof(int), hipMemcpyDeviceToHost, st));
    FR_HIP(hipStreamSynchronize(st));
  }
  return FR_OK;
}

int fr2d_filter_refresh_limit(fr2d_ctx *c, int mode, double lambda, double S0, double kappa, int rule,
                              int *n_filtered) {
  FR_REQUIRE(c && c->d_u, FR_E_INVALID, "fr2d_filter_refresh_limit: no state");
  FR_REQUIRE(mode >= FR_SENSOR_OFF && mode <= FR_SENSOR_LEARNED, FR_E_INVALID, "fr2d_filter_refresh_limit: unknown mode");
  FR_REQUIRE(lambda >= 0.0 && kappa > 0.0, FR_E_INVALID, "fr2d_filter_refresh_limit: lambda >= 0 and kappa > 0 required");
  FR_REQUIRE(rule >= FR_GHOST_PERIODIC && rule <= FR_GHOST_WAVE_Y, FR_E_INVALID, "fr2d_filter_refresh_limit: unknown rule");
  hipStream_t st = stream_of(*c);
  if (mode == FR_SENSOR_LEARNED && !c->mlp_loaded) mode = FR_SENSOR_MODAL;  // plain sensor is the fallback
  FR_HIP(hipMemsetAsync(c->d_count, 0, 2 * sizeof(int), st));
  int rc = FR_OK;
  if (mode == FR_SENSOR_LEARNED && (rc = fr2d_launch_mlp_sensor(*c, c->d_u, st))) return rc;
  if ((rc = fr2d_launch_filter_refresh_limit(*c, c->d_u, mode, lambda, S0, kappa, rule, c->d_count, st))) return rc;
  int h[2] = {0, 0};
  FR_HIP(hipMemcpyAsync(h, c->d_count, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  FR_HIP(hipStreamSynchronize(st));
  if (n_filtered) *n_filtered = h[0];
  if (h[1] > 0)
    return fail(FR_E_LIMITER, "fr2d_filter_refresh_limit: " + std::to_string(h[1]) +
                                  " element(s) with non-positive mean density (left as filtered)");
  return FR_OK;
}

int fr2d_sensor_eval(fr2d_ctx *c, int mode, double S0, double kappa, float *d_out) {
  FR_REQUIRE(c && c->d_u && d_out, FR_E_INVALID, "fr2d_sensor_eval: null argument or no state");
  FR_REQUIRE(mode == FR_SENSOR_MODAL || mode == FR_SENSOR_LEARNED, FR_E_INVALID, "fr2d_sensor_eval: unknown mode");
  FR_REQUIRE(mode == FR_SENSOR_MODAL || c->mlp_loaded, FR_E_INVALID, "fr2d_sensor_eval: no MLP loaded");
  hipStream_t st = stream_of(*c);
  int rc = FR_OK;
  if (mode == FR_SENSOR_LEARNED && (rc = fr2d_launch_mlp_sensor(*c, c->d_u, st))) return rc;
  return fr2d_launch_sensor_filter(*c, c->d_u, mode, 1.0, S0, kappa, nullptr, d_out, st);
}

int fr2d_run_host_loop(fr2d_ctx *c, int scheme, double dt, int nsteps, int mode, double lambda, double S0,
                       double kappa, int rule, int *n_filtered) {
  FR_REQUIRE(c && c->d_u, FR_E_INVALID, "fr2d_run_host_loop: no state");
  FR_REQUIRE(nsteps >= 0 && std::isfinite(dt), FR_E_INVALID, "fr2d_run_host_loop: bad dt / nsteps");
  FR_REQUIRE(scheme >= FR_RK_SSPRK3 && scheme <= FR_RK_TSIT5, FR_E_INVALID, "fr2d_run_host_loop: unknown scheme");
  FR_REQUIRE(mode >= FR_SENSOR_OFF && mode <= FR_SENSOR_LEARNED, FR_E_INVALID, "fr2d_run_host_loop: unknown mode");
  FR_REQUIRE(lambda >= 0.0 && kappa > 0.0, FR_E_INVALID, "fr2d_run_host_loop: lambda >= 0 and kappa > 0 required");
  FR_REQUIRE(rule >= FR_GHOST_PERIODIC && rule <= FR_GHOST_WAVE_Y, FR_E_INVALID, "fr2d_run_host_loop: unknown rule");
  hipStream_t st = stream_of(*c);
  if (mode == FR_SENSOR_LEARNED && !c->mlp_loaded) mode = FR_SENSOR_MODAL;  // plain sensor is the fallback
  // d_count[0]: filtered cells, d_count[1]: limiter failures, both summed over the loop
  FR_HIP(hipMemsetAsync(c->d_count, 0, 2 * sizeof(int), st));
  int rc = fr2d_launch_positivity(*c, c->d_u, c->d_count + 1, st);
  for (int n = 0; n < nsteps && rc == FR_OK; ++n) {
    if ((rc = prepare_step(*c, scheme, st))) break;
    if ((rc = enqueue_steps(*c, scheme, dt, 1, st))) break;
    if ((rc = finish_step(*c, st))) break;
    if (mode == FR_SENSOR_LEARNED && (rc = fr2d_launch_mlp_sensor(*c, c->d_u, st))) break;
    if (n + 1 < nsteps) {
      rc = fr2d_launch_filter_refresh_limit(*c, c->d_u, mode, lambda, S0, kappa, rule, c->d_count, st);
    } else {
      if (mode != FR_SENSOR_OFF &&
          (rc = fr2d_launch_sensor_filter(*c, c->d_u, mode, lambda, S0, kappa, c->d_count, nullptr, st)))
        break;
      rc = fr2d_launch_fill_ghost(*c, c->d_u, rule, st);
    }
  }
  if (rc) return rc;
  int h[2] = {0, 0};
  FR_HIP(hipMemcpyAsync(h, c->d_count, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  FR_HIP(hipStreamSynchronize(st));
  if (n_filtered) *n_filtered = h[0];
  if (h[1] > 0)
    return fail(FR_E_LIMITER, "fr2d_run_host_loop: " + std::to_string(h[1]) +
                                  " limiter failure(s) (element mean density not positive) over the loop");
  return FR_OK;
}

int fr2d_positivity(fr2d_ctx *c) {
  FR_REQUIRE(c && c->d_u, FR_E_INVALID, "fr2d_positivity: no state");
  hipStream_t st = stream_of(*c);
  FR_HIP(hipMemsetAsync(c->d_count, 0, sizeof(int), st));
  int rc = fr2d_launch_positivity(*c, c->d_u, c->d_count, st);
  if (rc) return rc;
  int bad = 0;
  FR_HIP(hipMemcpyAsync(&bad, c->d_count, sizeof(int), hipMemcpyDeviceToHost, st));
  FR_HIP(hipStreamSynchronize(st));
  if (bad) return fail(FR_E_LIMITER, "fr2d_positivity: " + std::to_string(bad) + " element(s) failed 0 < t1 <= 1");
  return FR_OK;
}

int fr2d_set_geometry(fr2d_ctx *c, const double *iJ, const double *n1, const double *n2) {
  FR_REQUIRE(c && iJ, FR_E_INVALID, "fr2d_set_geometry: null argument");
  // iJ column-major like Julia's 2x2 Matrix: iJ[0] = r_x, iJ[1] = s_x, iJ[2] = r_y, iJ[3] = s_y
  const double a = iJ[0], b = iJ[2], cc = iJ[1], d = iJ[3];
  for (int q = 0; q < 4; ++q) FR_REQUIRE(std::isfinite(iJ[q]), FR_E_INVALID, "fr2d_set_geometry: iJ not finite");
  FR_REQUIRE(a * d - b * cc > 0.0, FR_E_INVALID, "fr2d_set_geometry: iJ must have a positive determinant");
  double m1[2] = {a, b}, m2[2] = {cc, d};  // grad r, grad s: the default face normals
  if (n1) m1[0] = n1[0], m1[1] = n1[1];
  if (n2) m2[0] = n2[0], m2[1] = n2[1];
  const double l1 = std::hypot(m1[0], m1[1]), l2 = std::hypot(m2[0], m2[1]);
  FR_REQUIRE(std::isfinite(l1) && std::isfinite(l2) && l1 > 0.0 && l2 > 0.0, FR_E_INVALID,
             "fr2d_set_geometry: face normals must be finite and non-zero");
  c->geom = 1;
  c->iJ11 = a;
  c->iJ12 = b;
  c->iJ21 = cc;
  c->iJ22 = d;
  c->n1[0] = m1[0] / l1;
  c->n1[1] = m1[1] / l1;
  c->n2[0] = m2[0] / l2;
  c->n2[1] = m2[1] / l2;
  // a captured step sequence holds the old geometry in its kernel arguments
  if (c->gexec) {
    FR_HIP(hipStreamSynchronize(stream_of(*c)));
    FR_HIP(hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
  }
  return FR_OK;
}

int fr2d_set_metrics(fr2d_ctx *c, const double *iJ) {
  FR_REQUIRE(c && iJ, FR_E_INVALID, "fr2d_set_metrics: null argument");
  const size_t nx2 = (size_t)c->nx + 2, cells = nx2 * (size_t)(c->ny + 2), np = (size_t)c->nsp * c->nsp;
  // entry (i, j, k, l, e) at i + (nx+2) (j + (ny+2) (k + nsp (l + nsp e))), e = a + 2 b of the 2x2
  // matrix; the interior cells must hold finite matrices of positive determinant
  for (size_t p = 0; p < np; ++p)
    for (int j = 1; j <= c->ny; ++j)
      for (int i = 1; i <= c->nx; ++i) {
        const size_t at = (size_t)i + nx2 * (size_t)j + cells * p;
        const double a = iJ[at], cc = iJ[at + cells * np], b = iJ[at + 2 * cells * np], d = iJ[at + 3 * cells * np];
        FR_REQUIRE(std::isfinite(a) && std::isfinite(b) && std::isfinite(cc) && std::isfinite(d), FR_E_INVALID,
                   "fr2d_set_metrics: iJ not finite");
        FR_REQUIRE(a * d - b * cc > 0.0, FR_E_INVALID, "fr2d_set_metrics: iJ must have a positive determinant");
      }
  hipStream_t st = stream_of(*c);
  if (!c->d_mj_ref) FR_HIP(hipMalloc(&c->d_mj_ref, c->len * sizeof(double)));
  if (!c->d_mj_pad) {
    int rc = alloc_padded(*c, c->raw_mj_pad, c->d_mj_pad);
    if (rc) return rc;
  }
  // a captured step sequence holds the old geometry in its kernel arguments
  if (c->gexec) {
    FR_HIP(hipStreamSynchronize(st));
    FR_HIP(hipGraphExecDestroy(c->gexec));
    c->gexec = nullptr;
  }
  FR_HIP(hipMemcpyAsync(c->d_mj_ref, iJ, c->len * sizeof(double), hipMemcpyHostToDevice, st));
  int rc = copy_to_padded(*c, c->d_mj_pad, c->d_mj_ref, hipMemcpyDeviceToDevice, st);
  if (rc) return rc;
  FR_HIP(hipStreamSynchronize(st));
  c->geom = 2;
  return FR_OK;
}

int fr2d_primitives_dev(fr2d_ctx *c, double *d_sol) {
  FR_REQUIRE(c && d_sol, FR_E_INVALID, "fr2d_primitives_dev: null argument");
  FR_REQUIRE(c->d_u, FR_E_INVALID, "fr2d_primitives_dev: no state");
  return fr2d_launch_primitives(*c, c->d_u, d_sol, stream_of(*c));
}

int fr2d_primitives(fr2d_ctx *c, double *sol) {
  FR_REQUIRE(c && sol, FR_E_INVALID, "fr2d_primitives: null argument");
  FR_REQUIRE(c->d_u, FR_E_INVALID, "fr2d_primitives: no state");
  hipStream_t st = stream_of(*c);
  // the staging buffer of the host-pointer residual holds the ghost-inclusive state, which is
  // larger than the interior point field
  if (!c->d_out) FR_HIP(hipMalloc(&c->d_out, c->len * sizeof(double)));
  const size_t n = (size_t)c->nx * c->ny * c->nsp * c->nsp * 4;
  int rc = fr2d_launch_primitives(*c, c->d_u, c->d_out, st);
  if (rc) return rc;
  FR_HIP(hipMemcpyAsync(sol, c->d_out, n * sizeof(double), hipMemcpyDeviceToHost, st));
  FR_HIP(hipStreamSynchronize(st));
  return FR_OK;
}

int fr2d_check_state(fr2d_ctx *c) {
  FR_REQUIRE(c && c->d_u, FR_E_INVALID, "fr2d_check_state: no state");
  hipStream_t st = stream_of(*c);
  FR_HIP(hipMemsetAsync(c->d_count, 0, 2 * sizeof(int), st));
  int rc = fr2d_launch_check(*c, c->d_u, c->d_count, st);
  if (rc) return rc;
  int cnt[2] = {0, 0};
  FR_HIP(hipMemcpyAsync(cnt, c->d_count, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  FR_HIP(hipStreamSynchronize(st));
  if (cnt[0]) return fail(FR_E_NONFINITE, "fr2d_check_state: " + std::to_string(cnt[0]) + " non-finite value(s)");
  if (cnt[1])
    return fail(FR_E_NEGATIVE_STATE,
                "fr2d_check_state: " + std::to_string(cnt[1]) + " point(s) with rho <= 0 or p <= 0");
  return FR_OK;
}

}  // extern "C"
