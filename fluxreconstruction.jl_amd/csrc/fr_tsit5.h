// fr_tsit5.h -- the fixed-step Tsit5 tableau and its stage-combination kernel, shared by the 1D
// and 2D steppers.
#pragma once

#include <cstddef>

#include <hip/hip_runtime.h>

namespace fr {

// Tsitouras' 5(4) pair (Ch. Tsitouras, Comput. Math. Appl. 62 (2011) 770-775), the Tsit5 of
// OrdinaryDiffEq that example/advection_highlevel.jl:26 and example/euler1d_convergence.jl:133
// step with a fixed dt (adaptive = false).  Row s (0..4) holds a_{s+2, 1..s+1}, the weights of
// k_1..k_{s+1} in the input of stage s+2; row 5 holds b_1..b_6 of the new state (b_7 = 0: the
// seventh evaluation is the next step's first, so a fixed-step run needs six per step).
constexpr double kTsit5[6][6] = {
    {0.161, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401, -0.028269050394068383, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081, 2.324710524099774},
};

struct Coef6 {
  double c[6];
};

// out = u + dt * (c_1 k_1 + ... + c_nk k_nk), k_j at k + (j-1)*stride: one thread per value
// (coalesced, HBM-bound).  out may be u (the last combination updates the state in place: every
// thread reads its own u[i] before writing out[i]).
static __global__ __launch_bounds__(256) void tsit5_comb_kernel(double *out, const double *u,
                                                                const double *__restrict__ k, size_t stride, int nk,
                                                                Coef6 c, double dt, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double s = c.c[0] * k[i];
    for (int j = 1; j < nk; ++j) s += c.c[j] * k[j * stride + i];
    out[i] = u[i] + dt * s;
  }
}

// enqueues out = u + dt * sum_{j < nk} row[j] k_{j+1} over n values
inline hipError_t tsit5_comb(double *out, const double *u, const double *k, size_t stride, const double *row, int nk,
                             double dt, size_t n, hipStream_t st) {
  Coef6 c{};
  for (int j = 0; j < nk; ++j) c.c[j] = row[j];
  const size_t blocks = (n + 255) / 256 < 8192 ? (n + 255) / 256 : 8192;
  hipLaunchKernelGGL(tsit5_comb_kernel, dim3((unsigned)blocks), dim3(256), 0, st, out, u, k, stride, nk, c, dt, n);
  return hipGetLastError();
}

}  // namespace fr
