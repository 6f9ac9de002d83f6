// fr_mlp.hip -- learned troubled-cell sensor for the 2D solver on gfx950 MFMA.
//
// Network: 16 normalised density modal coefficients -> 32 -> 32 -> 1 (ReLU hidden layers; the
// output o is the filter-strength multiplier, max(o, 0) at inference).  Weights and activations
// enter the matrix cores as BF16, products accumulate in FP32 (v_mfma_f32_16x16x16_bf16).
//
// Every layer is evaluated transposed, Z^T = W X^T, on tiles of 16 cells per wavefront: the MFMA
// result layout (lane = cell, four consecutive output rows per lane) is exactly the B-operand
// layout of the next layer, so the forward pass needs no data movement between layers.  The
// weight gradients reduce over cells, so their operands are transposed through LDS.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "fr2d_internal.h"

namespace fr {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 16, NH = 32;
// parameter vector (FP32): W1[32][16] b1[32] W2[32][32] b2[32] W3[32] b3
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH, NPARAM = OFF_B3 + 1;

__host__ __device__ __forceinline__ unsigned short f2bf(float f) {
  // round to nearest even (finite inputs)
  unsigned int u = __builtin_bit_cast(unsigned int, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) {
  return __builtin_bit_cast(float, (unsigned int)h << 16);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 r;
  r[0] = (short)f2bf(a);
  r[1] = (short)f2bf(b);
  r[2] = (short)f2bf(c);
  r[3] = (short)f2bf(d);
  return r;
}

__device__ __forceinline__ bf16x4 load4(const unsigned short *p) {
  bf16x4 r;
  r[0] = (short)p[0];
  r[1] = (short)p[1];
  r[2] = (short)p[2];
  r[3] = (short)p[3];
  return r;
}

__device__ __forceinline__ floatx4 mfma16(bf16x4 a, bf16x4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// device parameter block: FP32 master copy, BF16 operand copies, Adam moments, gradients
struct MlpDev {
  float *p;               // NPARAM master parameters
  float *g;               // NPARAM gradient accumulators
  float *m, *v;           // Adam moments
  unsigned short *w1b;    // [32][16]
  unsigned short *w2b;    // [32][32]
  unsigned short *w2tb;   // [32][32], w2tb[i][j] = W2[j][i]
  float *loss;            // accumulated loss of the last batch
};

// ------------------------------------------------------------------ forward of one 16-cell tile
struct TileFwd {
  bf16x4 h1b[2];    // H1^T, M tiles 0/1, MFMA output layout, BF16
  floatx4 z1[2];    // pre-activations of layer 1
  floatx4 z2[2];    // pre-activations of layer 2
  floatx4 h2[2];
  float o;          // output of the cell held by this lane (lane % 16)
};

__device__ __forceinline__ void tile_forward(const MlpDev &w, bf16x4 xb, TileFwd &t) {
  const int lane = threadIdx.x & 63, c = lane & 15, kb = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    const bf16x4 wa = load4(w.w1b + (mt * 16 + c) * NIN + 4 * kb);
    floatx4 z = mfma16(wa, xb, floatx4{0.f, 0.f, 0.f, 0.f});
    float h[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      z[j] += w.p[OFF_B1 + mt * 16 + 4 * kb + j];
      h[j] = fmaxf(z[j], 0.f);
    }
    t.z1[mt] = z;
    t.h1b[mt] = pack4(h[0], h[1], h[2], h[3]);
  }
  float part = 0.f;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    floatx4 acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) acc = mfma16(load4(w.w2b + (mt * 16 + c) * NH + kc * 16 + 4 * kb), t.h1b[kc], acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[j] += w.p[OFF_B2 + mt * 16 + 4 * kb + j];
      const float h = fmaxf(acc[j], 0.f);
      t.h2[mt][j] = h;
      part += w.p[OFF_W3 + mt * 16 + 4 * kb + j] * h;
    }
    t.z2[mt] = acc;
  }
  part += __shfl_xor(part, 16);
  part += __shfl_xor(part, 32);
  t.o = part + w.p[OFF_B3];
}

__device__ __forceinline__ bf16x4 load_x(const unsigned short *X, long long row, long long n, int kb) {
  if (row < n) return load4(X + row * NIN + 4 * kb);
  return bf16x4{0, 0, 0, 0};
}

// strength s = max(o, 0) of every row of X (inference)
__global__ __launch_bounds__(256) void mlp_forward_kernel(MlpDev w, const unsigned short *__restrict__ X, float *out,
                                                          long long n) {
  const int lane = threadIdx.x & 63, c = lane & 15, kb = lane >> 4;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const long long ntiles = (n + 15) / 16;
  for (long long tl = wave; tl < ntiles; tl += nwaves) {
    const long long row = tl * 16 + c;
    TileFwd t;
    tile_forward(w, load_x(X, row, n, kb), t);
    if (kb == 0 && row < n) out[row] = fmaxf(t.o, 0.f);
  }
}

// ------------------------------------------------------------------ training: forward + backward
// MSE loss L = (1/B) sum (o - y)^2 over the batch rows [0, n).  Gradients accumulate into w.g.
__global__ __launch_bounds__(256) void mlp_train_kernel(MlpDev w, const unsigned short *__restrict__ X,
                                                        const float *__restrict__ Y, long long n, float inv_b) {
  __shared__ unsigned short sT[4][3][NH * 16];  // per wave: transposes of delta2, h1, delta1 ([row][cell])
  const int lane = threadIdx.x & 63, c = lane & 15, kb = lane >> 4, wv = threadIdx.x >> 6;
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
  const long long ntiles = (n + 15) / 16;
  floatx4 gW1[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
  floatx4 gW2[2][2] = {{{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}}, {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}}};
  float gb1[2][4] = {}, gb2[2][4] = {}, gW3[2][4] = {}, gb3 = 0.f, loss = 0.f;
  unsigned short *tD2 = sT[wv][0], *tH1 = sT[wv][1], *tD1 = sT[wv][2];
  for (long long tl = wave; tl < ntiles; tl += nwaves) {
    const long long row = tl * 16 + c;
    const bool valid = row < n;
    TileFwd t;
    tile_forward(w, load_x(X, row, n, kb), t);
    const float err = valid ? (t.o - Y[row]) : 0.f;
    if (kb == 0) loss += err * err * inv_b;
    const float dout = 2.f * err * inv_b;
    gb3 += (kb == 0) ? dout : 0.f;
    // delta2 = W3 dout (z2 > 0), layout of z2
    floatx4 d2[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int h = mt * 16 + 4 * kb + j;
        gW3[mt][j] += dout * t.h2[mt][j];
        const float d = (t.z2[mt][j] > 0.f) ? w.p[OFF_W3 + h] * dout : 0.f;
        d2[mt][j] = d;
        gb2[mt][j] += d;
      }
    bf16x4 d2b[2] = {pack4(d2[0][0], d2[0][1], d2[0][2], d2[0][3]), pack4(d2[1][0], d2[1][1], d2[1][2], d2[1][3])};
    // delta1 = (W2^T delta2) (z1 > 0)
    bf16x4 d1b[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      floatx4 acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < 2; ++kc) acc = mfma16(load4(w.w2tb + (mt * 16 + c) * NH + kc * 16 + 4 * kb), d2b[kc], acc);
      float d[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        d[j] = (t.z1[mt][j] > 0.f) ? acc[j] : 0.f;
        gb1[mt][j] += d[j];
      }
      d1b[mt] = pack4(d[0], d[1], d[2], d[3]);
    }
    // stage delta2, h1 and delta1 as [row][cell] BF16 tiles for the cell-reduction GEMMs
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = mt * 16 + 4 * kb + j;
        tD2[r * 16 + c] = (unsigned short)d2b[mt][j];
        tH1[r * 16 + c] = (unsigned short)t.h1b[mt][j];
        tD1[r * 16 + c] = (unsigned short)d1b[mt][j];
      }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // dW2[a][b] += sum_cells delta2[a][cell] h1[b][cell];  A = delta2 (rows a, K = cells),
    // B[cell][b] = h1[b][cell]
    bf16x4 bh1[2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      bf16x4 r;
#pragma unroll
      for (int j = 0; j < 4; ++j) r[j] = (short)tH1[(nt * 16 + c) * 16 + 4 * kb + j];
      bh1[nt] = r;
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      bf16x4 a2, a1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a2[j] = (short)tD2[(mt * 16 + c) * 16 + 4 * kb + j];
        a1[j] = (short)tD1[(mt * 16 + c) * 16 + 4 * kb + j];
      }
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) gW2[mt][nt] = mfma16(a2, bh1[nt], gW2[mt][nt]);
      // dW1[a][f] += sum_cells delta1[a][cell] x[cell][f];  B[cell][f] = x[cell][f]
      bf16x4 bx;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long long rr = tl * 16 + 4 * kb + j;
        bx[j] = (rr < n) ? (short)X[rr * NIN + c] : (short)0;
      }
      gW1[mt] = mfma16(a1, bx, gW1[mt]);
    }
    __builtin_amdgcn_wave_barrier();
  }
  // flush this lane's partial sums (output layout: row 4*kb + j of the tile, column c)
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = mt * 16 + 4 * kb + j;
      atomicAdd(&w.g[OFF_W1 + r * NIN + c], gW1[mt][j]);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) atomicAdd(&w.g[OFF_W2 + r * NH + nt * 16 + c], gW2[mt][nt][j]);
      // per-hidden-unit sums over the 16 cells of this lane group
      float b1 = gb1[mt][j], b2 = gb2[mt][j], w3 = gW3[mt][j];
#pragma unroll
      for (int s = 1; s < 16; s <<= 1) {
        b1 += __shfl_xor(b1, s);
        b2 += __shfl_xor(b2, s);
        w3 += __shfl_xor(w3, s);
      }
      if (c == 0) {
        atomicAdd(&w.g[OFF_B1 + r], b1);
        atomicAdd(&w.g[OFF_B2 + r], b2);
        atomicAdd(&w.g[OFF_W3 + r], w3);
      }
    }
  for (int s = 1; s < 64; s <<= 1) {
    gb3 += __shfl_xor(gb3, s);
    loss += __shfl_xor(loss, s);
  }
  if (lane == 0) {
    atomicAdd(&w.g[OFF_B3], gb3);
    atomicAdd(w.loss, loss);
  }
}

// Adam step on the master parameters; refreshes the BF16 operand copies and clears gradients.
__global__ void mlp_adam_kernel(MlpDev w, float lr, float b1, float b2, float eps, float c1, float c2, int apply) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  if (apply) {
    const float g = w.g[i];
    const float m = b1 * w.m[i] + (1.f - b1) * g;
    const float v = b2 * w.v[i] + (1.f - b2) * g * g;
    w.m[i] = m;
    w.v[i] = v;
    w.p[i] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
  }
  w.g[i] = 0.f;
  if (i < NH * NIN) w.w1b[i] = f2bf(w.p[OFF_W1 + i]);
  if (i < NH * NH) {
    const float p = w.p[OFF_W2 + i];
    w.w2b[i] = f2bf(p);
    w.w2tb[(i % NH) * NH + i / NH] = f2bf(p);
  }
  if (i == 0) *w.loss = 0.f;
}

// features of every cell: x = u_hat / |u_hat|, u_hat = iV * rho nodal (nsp = 4), BF16
__global__ __launch_bounds__(256) void mlp_features_kernel(const double *u, int ncell, int nxg, int pitch, long long ldp,
                                                           const double *__restrict__ iV, unsigned short *X) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;  // cell, reference order
  if (e >= ncell) return;
  const long long ep = e % nxg + (long long)pitch * (e / nxg);             // its place in a plane
  double nod[16], s2 = 0.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) nod[q] = u[ep + ldp * q];
  double mod[16];
#pragma unroll
  for (int a = 0; a < 16; ++a) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < 16; ++b) s += iV[a + 16 * b] * nod[b];
    mod[a] = s;
    s2 += s * s;
  }
  const double inv = 1.0 / sqrt(s2 + 1e-300);
#pragma unroll
  for (int a = 0; a < 16; ++a) X[e * 16 + a] = f2bf((float)(mod[a] * inv));
}

// MFMA lane-layout check: D = A B for A[i][k] = (i + 2k) mod 7 - 3, B[k][j] = (3k + j) mod 5 - 2
// (small integers, exact in BF16 and FP32); err[0] = max |D_mfma - D_loop|.
__global__ void mfma_selftest_kernel(float *err) {
  const int lane = threadIdx.x & 63, c = lane & 15, kb = lane >> 4;
  bf16x4 a, b;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = 4 * kb + j;
    a[j] = (short)f2bf((float)((c + 2 * k) % 7 - 3));
    b[j] = (short)f2bf((float)((3 * k + c) % 5 - 2));
  }
  floatx4 d = mfma16(a, b, floatx4{0.f, 0.f, 0.f, 0.f});
  float e = 0.f;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = 4 * kb + j;  // row of D held in register j
    float ref = 0.f;
    for (int k = 0; k < 16; ++k) ref += (float)((i + 2 * k) % 7 - 3) * (float)((3 * k + c) % 5 - 2);
    e = fmaxf(e, fabsf(d[j] - ref));
  }
  for (int s = 1; s < 64; s <<= 1) e = fmaxf(e, __shfl_xor(e, s));
  if (lane == 0) err[0] = e;
}

// ------------------------------------------------------------------ host context
struct MlpCtx {
  MlpDev d{};
  void *block = nullptr;
  hipStream_t stream = nullptr;
  int step = 0;
  float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;

  int init() {
    const size_t fl = 4 * (size_t)NPARAM + 1;
    const size_t bf = NH * NIN + 2 * NH * NH;
    const size_t bytes = fl * sizeof(float) + bf * sizeof(unsigned short) + 256;
    FR_HIP(hipMalloc(&block, bytes));
    FR_HIP(hipMemset(block, 0, bytes));
    float *f = (float *)block;
    d.p = f;
    d.g = f + NPARAM;
    d.m = f + 2 * NPARAM;
    d.v = f + 3 * NPARAM;
    d.loss = f + 4 * NPARAM;
    unsigned short *h = (unsigned short *)(f + fl);
    d.w1b = h;
    d.w2b = h + NH * NIN;
    d.w2tb = h + NH * NIN + NH * NH;
    FR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    return FR_OK;
  }

  int refresh(bool apply, float lr, hipStream_t st) {
    ++step;
    const float c1 = 1.f / (1.f - std::pow(beta1, (float)step));
    const float c2 = 1.f / (1.f - std::pow(beta2, (float)step));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, d, lr, beta1, beta2, eps, c1, c2,
                       apply ? 1 : 0);
    FR_LAUNCHED("mlp_adam_kernel");
    return FR_OK;
  }
};

// training: few workgroups, each wave loops over several tiles before its gradient flush
static int train_grid(long long n) {
  const long long tiles = (n + 15) / 16;
  long long wg = (tiles + 7) / 8;
  if (wg > 64) wg = 64;
  if (wg < 1) wg = 1;
  return (int)wg;
}

static int grid_for_tiles(long long n) {
  const long long tiles = (n + 15) / 16;
  long long wg = (tiles + 3) / 4;
  if (wg > 1024) wg = 1024;
  if (wg < 1) wg = 1;
  return (int)wg;
}

int fr2d_launch_mlp_sensor(const Fr2dCtx &c, const double *d_u, hipStream_t st) {
  if (c.nsp != 4) return fail(FR_E_UNSUPPORTED, "learned sensor: trained for p = 3 (16 modes)");
  const long long ldp = (long long)(c.nx + 2) * (c.ny + 2);
  const DissOps o = fr2d_diss_ops(c);
  hipLaunchKernelGGL(mlp_features_kernel, dim3((unsigned)((ldp + 255) / 256)), dim3(256), 0, st, d_u, (int)ldp,
                     c.nx + 2, c.pitch, c.ldpi,
                     o.iV, (unsigned short *)c.d_feat);
  FR_LAUNCHED("mlp_features_kernel");
  const MlpCtx *m = (const MlpCtx *)c.d_mlp;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid_for_tiles(ldp)), dim3(256), 0, st, m->d,
                     (const unsigned short *)c.d_feat, (float *)c.d_strength, ldp);
  FR_LAUNCHED("mlp_forward_kernel");
  return FR_OK;
}

}  // namespace fr

using namespace fr;

struct fr_mlp : MlpCtx {};

extern "C" {

int fr_mlp_nparam(void) { return NPARAM; }

int fr_mlp_create(fr_mlp **out) {
  FR_REQUIRE(out, FR_E_INVALID, "fr_mlp_create: null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FR_E_HIP, "fr_mlp_create: no HIP device visible");
  fr_mlp *m = new (std::nothrow) fr_mlp();
  FR_REQUIRE(m, FR_E_NOMEM, "fr_mlp_create: out of host memory");
  int rc = m->init();
  if (rc) {
    if (m->block) (void)hipFree(m->block);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return rc;
  }
  *out = m;
  return FR_OK;
}

int fr_mlp_destroy(fr_mlp *m) {
  if (!m) return FR_OK;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->block) (void)hipFree(m->block);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return FR_OK;
}

int fr_mlp_set_params(fr_mlp *m, const float *params) {
  FR_REQUIRE(m && params, FR_E_INVALID, "fr_mlp_set_params: null argument");
  FR_HIP(hipMemcpyAsync(m->d.p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
  FR_HIP(hipMemsetAsync(m->d.m, 0, 2 * NPARAM * sizeof(float), m->stream));
  m->step = 0;
  int rc = m->refresh(false, 0.f, m->stream);
  if (rc) return rc;
  m->step = 0;
  FR_HIP(hipStreamSynchronize(m->stream));
  return FR_OK;
}

int fr_mlp_get_params(fr_mlp *m, float *params) {
  FR_REQUIRE(m && params, FR_E_INVALID, "fr_mlp_get_params: null argument");
  FR_HIP(hipMemcpyAsync(params, m->d.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  FR_HIP(hipStreamSynchronize(m->stream));
  return FR_OK;
}

int fr_mlp_forward(fr_mlp *m, const uint16_t *d_x, float *d_out, long long n) {
  FR_REQUIRE(m && d_x && d_out && n >= 0, FR_E_INVALID, "fr_mlp_forward: bad arguments");
  if (n == 0) return FR_OK;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid_for_tiles(n)), dim3(256), 0, m->stream, m->d,
                     (const unsigned short *)d_x, d_out, n);
  FR_LAUNCHED("mlp_forward_kernel");
  FR_HIP(hipStreamSynchronize(m->stream));
  return FR_OK;
}

// gradients of the batch loss (host copy, NPARAM floats) without updating the parameters
int fr_mlp_grad(fr_mlp *m, const uint16_t *d_x, const float *d_y, long long n, float *grads, float *loss) {
  FR_REQUIRE(m && d_x && d_y && n >= 1 && grads, FR_E_INVALID, "fr_mlp_grad: bad arguments");
  FR_HIP(hipMemsetAsync(m->d.g, 0, NPARAM * sizeof(float), m->stream));
  FR_HIP(hipMemsetAsync(m->d.loss, 0, sizeof(float), m->stream));
  hipLaunchKernelGGL(mlp_train_kernel, dim3(train_grid(n)), dim3(256), 0, m->stream, m->d,
                     (const unsigned short *)d_x, d_y, n, 1.0f / (float)n);
  FR_LAUNCHED("mlp_train_kernel");
  FR_HIP(hipMemcpyAsync(grads, m->d.g, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  float l = 0.f;
  FR_HIP(hipMemcpyAsync(&l, m->d.loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  FR_HIP(hipStreamSynchronize(m->stream));
  if (loss) *loss = l;
  return FR_OK;
}

// nsteps Adam steps over consecutive batches of `batch` rows of (X, Y) (wrapping around n rows);
// returns the mean batch loss of the last epoch-equivalent in *loss
int fr_mlp_train(fr_mlp *m, const uint16_t *d_x, const float *d_y, long long n, int batch, int nsteps, float lr,
                 float *loss) {
  FR_REQUIRE(m && d_x && d_y && n >= 1 && batch >= 16 && nsteps >= 0 && lr > 0.f, FR_E_INVALID,
             "fr_mlp_train: bad arguments");
  FR_HIP(hipMemsetAsync(m->d.g, 0, NPARAM * sizeof(float), m->stream));
  FR_HIP(hipMemsetAsync(m->d.loss, 0, sizeof(float), m->stream));
  long long pos = 0;
  double lsum = 0.0;
  int lcount = 0;
  const int report = nsteps < 32 ? nsteps : 32;
  for (int s = 0; s < nsteps; ++s) {
    long long nb = batch;
    if (pos + nb > n) pos = 0;
    if (nb > n) nb = n;
    hipLaunchKernelGGL(mlp_train_kernel, dim3(train_grid(nb)), dim3(256), 0, m->stream, m->d,
                       (const unsigned short *)(d_x + pos * NIN), d_y + pos, nb, 1.0f / (float)nb);
    FR_LAUNCHED("mlp_train_kernel");
    if (s >= nsteps - report) {
      float l = 0.f;
      FR_HIP(hipMemcpyAsync(&l, m->d.loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
      FR_HIP(hipStreamSynchronize(m->stream));
      lsum += l;
      ++lcount;
    }
    int rc = m->refresh(true, lr, m->stream);
    if (rc) return rc;
    pos += nb;
  }
  FR_HIP(hipStreamSynchronize(m->stream));
  if (loss) *loss = lcount ? (float)(lsum / lcount) : 0.f;
  return FR_OK;
}

int fr_mfma_selftest(float *max_err) {
  FR_REQUIRE(max_err, FR_E_INVALID, "fr_mfma_selftest: null argument");
  float *d = nullptr;
  FR_HIP(hipMalloc(&d, sizeof(float)));
  hipLaunchKernelGGL(mfma_selftest_kernel, dim3(1), dim3(64), 0, 0, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(max_err, d, sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  FR_HIP(e);
  return FR_OK;
}

int fr2d_set_sensor_mlp(fr2d_ctx *c, fr_mlp *m) {
  FR_REQUIRE(c, FR_E_INVALID, "fr2d_set_sensor_mlp: null context");
  Fr2dCtx *x = c;
  if (!m) {
    x->mlp_loaded = false;
    x->d_mlp = nullptr;
    return FR_OK;
  }
  FR_REQUIRE(x->nsp == 4, FR_E_UNSUPPORTED, "fr2d_set_sensor_mlp: the learned sensor is defined for p = 3");
  const long long ldp = (long long)(x->nx + 2) * (x->ny + 2);
  if (!x->d_feat) FR_HIP(hipMalloc(&x->d_feat, ldp * NIN * sizeof(unsigned short)));
  if (!x->d_strength) FR_HIP(hipMalloc(&x->d_strength, ldp * sizeof(float)));
  FR_HIP(hipStreamSynchronize(m->stream));
  x->d_mlp = (void *)(MlpCtx *)m;
  x->mlp_loaded = true;
  return FR_OK;
}

// features (BF16, 16 per cell, every cell of the layout including ghosts) of the device state
int fr2d_sensor_features(fr2d_ctx *c, uint16_t *d_x) {
  FR_REQUIRE(c && d_x, FR_E_INVALID, "fr2d_sensor_features: null argument");
  Fr2dCtx *x = c;
  FR_REQUIRE(x->d_u, FR_E_INVALID, "fr2d_sensor_features: no state");
  FR_REQUIRE(x->nsp == 4, FR_E_UNSUPPORTED, "fr2d_sensor_features: defined for p = 3");
  const long long ldp = (long long)(x->nx + 2) * (x->ny + 2);
  const DissOps o = fr2d_diss_ops(*x);
  hipStream_t st = fr2d_stream(*x);
  hipLaunchKernelGGL(mlp_features_kernel, dim3((unsigned)((ldp + 255) / 256)), dim3(256), 0, st, x->d_u, (int)ldp,
                     x->nx + 2, x->pitch, x->ldpi,
                     o.iV, (unsigned short *)d_x);
  FR_LAUNCHED("mlp_features_kernel");
  return FR_OK;
}

}  // extern "C"
