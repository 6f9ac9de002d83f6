// fr_ops_dev.hip -- the library-level contractions of FluxReconstruction.jl on device arrays:
// interp_face! (src/interpolate.jl:1-10) and poly_derivative! (src/derivative.jl:1-10) over every
// row of a batch of matrices, and the nodal exponential filter of filter_exp
// (src/Polynomial/poly_filter.jl:11-47), built on the host and applied on the device with the same
// row contraction.
//
// Layout: f(row, q, b), column-major (nrow, n, nbatch), the layout of a Julia Array{Float64,3}
// (a 1D state u(i, p, k) is f with rows = cells, q = solution point, b = variable; a 2D state
// u(i, j, k, l, m) contracted along k is rows = (i, j), q = k, b = (l, m)).  One thread per
// (row, batch), or two consecutive rows with 16-byte accesses where the rows and the arrays allow:
// the n loads of a thread are n coalesced row vectors of the wavefront, so the kernel streams f
// once and writes the outputs once (HBM-bound, 8 (n + nout) bytes per row).
#include <cmath>
#include <cstdint>
#include <vector>

#include "fr_common.h"

namespace fr {

constexpr int kRowMaxN = 16;

// the operator of one call, passed by value as kernel arguments (at most 2 KB)
struct RowOp {
  double m[kRowMaxN * kRowMaxN];  // m[i + nout q]: output i, input q
};

template <int N, int NOUT>
__global__ __launch_bounds__(256) void row_contract_kernel(const double *__restrict__ f, long long nrow,
                                                           long long total, const RowOp op,
                                                           double *__restrict__ out) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long b = t / nrow, row = t - b * nrow;
    const double *fr = f + row + nrow * (N * b);
    double x[N];
#pragma unroll
    for (int q = 0; q < N; ++q) x[q] = fr[nrow * q];
    double *o = out + row + nrow * (NOUT * b);
#pragma unroll
    for (int i = 0; i < NOUT; ++i) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < N; ++q) s += x[q] * op.m[i + NOUT * q];
      o[nrow * i] = s;
    }
  }
}

// the same contraction for two consecutive rows per thread with 16-byte loads and stores (nrow
// even, both arrays 16-byte aligned): half the memory instructions per byte, the same arithmetic
// per row
template <int N, int NOUT>
__global__ __launch_bounds__(256) void row_contract2_kernel(const double *__restrict__ f, long long nrow,
                                                            long long total2, const RowOp op,
                                                            double *__restrict__ out) {
  const long long half = nrow >> 1;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total2;
       t += (long long)gridDim.x * blockDim.x) {
    const long long b = t / half, row = 2 * (t - b * half);
    const double *fr = f + row + nrow * (N * b);
    double2 x[N];
#pragma unroll
    for (int q = 0; q < N; ++q) x[q] = *reinterpret_cast<const double2 *>(fr + nrow * q);
    double *o = out + row + nrow * (NOUT * b);
#pragma unroll
    for (int i = 0; i < NOUT; ++i) {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int q = 0; q < N; ++q) {
        s0 += x[q].x * op.m[i + NOUT * q];
        s1 += x[q].y * op.m[i + NOUT * q];
      }
      *reinterpret_cast<double2 *>(o + nrow * i) = make_double2(s0, s1);
    }
  }
}

template <int N>
static void launch_row(const double *f, long long nrow, long long total, int nout, const RowOp &op, double *out,
                       hipStream_t st) {
  const bool pairs = (nrow & 1) == 0 && ((reinterpret_cast<uintptr_t>(f) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
  if (pairs) {
    const long long total2 = total >> 1;
    const int blocks = (int)std::min<long long>((total2 + 255) / 256, 65536);
    if (nout == 2)
      hipLaunchKernelGGL((row_contract2_kernel<N, 2>), dim3(blocks), dim3(256), 0, st, f, nrow, total2, op, out);
    else
      hipLaunchKernelGGL((row_contract2_kernel<N, N>), dim3(blocks), dim3(256), 0, st, f, nrow, total2, op, out);
    return;
  }
  const long long want = (total + 255) / 256;
  const int blocks = (int)std::min<long long>(want, 65536);
  if (nout == 2)
    hipLaunchKernelGGL((row_contract_kernel<N, 2>), dim3(blocks), dim3(256), 0, st, f, nrow, total, op, out);
  else
    hipLaunchKernelGGL((row_contract_kernel<N, N>), dim3(blocks), dim3(256), 0, st, f, nrow, total, op, out);
}

// out(row, i, b) = sum_q f(row, q, b) m(i, q) for i < nout (nout = 2 or n)
static int row_contract(const char *who, const double *d_f, long long nrow, int n, int nbatch, int nout,
                        const RowOp &op, double *d_out, void *stream) {
  FR_REQUIRE(d_f && d_out, FR_E_INVALID, std::string(who) + ": null device pointer");
  FR_REQUIRE(d_f != d_out, FR_E_INVALID, std::string(who) + ": input and output must not alias");
  FR_REQUIRE(nrow >= 0 && nbatch >= 0, FR_E_INVALID, std::string(who) + ": negative size");
  FR_REQUIRE(n >= 1 && n <= kRowMaxN, FR_E_UNSUPPORTED, std::string(who) + ": n must be in 1..16");
  const long long total = nrow * (long long)nbatch;
  if (total == 0) return FR_OK;
  hipStream_t st = (hipStream_t)stream;
  switch (n) {
#define FR_ROW_CASE(K) \
  case K: launch_row<K>(d_f, nrow, total, nout, op, d_out, st); break;
    FR_ROW_CASE(1) FR_ROW_CASE(2) FR_ROW_CASE(3) FR_ROW_CASE(4) FR_ROW_CASE(5) FR_ROW_CASE(6) FR_ROW_CASE(7)
    FR_ROW_CASE(8) FR_ROW_CASE(9) FR_ROW_CASE(10) FR_ROW_CASE(11) FR_ROW_CASE(12) FR_ROW_CASE(13)
    FR_ROW_CASE(14) FR_ROW_CASE(15) FR_ROW_CASE(16)
#undef FR_ROW_CASE
    default: break;
  }
  FR_LAUNCHED("row_contract_kernel");
  return FR_OK;
}

}  // namespace fr

using namespace fr;

extern "C" {

int fr_interp_face_dev(const double *d_f, long long nrow, int n, int nbatch, const double *ll, const double *lr,
                       double *d_ff, void *stream) {
  FR_REQUIRE(ll && lr, FR_E_INVALID, "fr_interp_face_dev: null ll / lr");
  FR_REQUIRE(n >= 1 && n <= kRowMaxN, FR_E_UNSUPPORTED, "fr_interp_face_dev: n must be in 1..16");
  RowOp op{};
  for (int q = 0; q < n; ++q) {
    op.m[0 + 2 * q] = ll[q];
    op.m[1 + 2 * q] = lr[q];
  }
  return row_contract("fr_interp_face_dev", d_f, nrow, n, nbatch, 2, op, d_ff, stream);
}

int fr_poly_derivative_dev(const double *d_f, long long nrow, int n, int nbatch, const double *pdm, double *d_df,
                           void *stream) {
  FR_REQUIRE(pdm, FR_E_INVALID, "fr_poly_derivative_dev: null pdm");
  FR_REQUIRE(n >= 1 && n <= kRowMaxN, FR_E_UNSUPPORTED, "fr_poly_derivative_dev: n must be in 1..16");
  RowOp op{};
  for (int q = 0; q < n * n; ++q) op.m[q] = pdm[q];  // pdm(i, q) at i + n q
  return row_contract("fr_poly_derivative_dev", d_f, nrow, n, nbatch, n, op, d_df, stream);
}

int fr_filter_exp(int N, int s, int Nc, const double *V, int nv, const double *invV, double *F) {
  FR_REQUIRE(V && F, FR_E_INVALID, "fr_filter_exp: null V / F");
  FR_REQUIRE(N >= 1 && s >= 1 && Nc >= 0 && Nc < N, FR_E_INVALID, "fr_filter_exp: need N >= 1, s >= 1, 0 <= Nc < N");
  FR_REQUIRE(nv >= 1 && nv <= 1024, FR_E_INVALID, "fr_filter_exp: nv out of range");
  // KitBase filter_exp1d for a line (nv = N + 1), filter_exp2d of poly_filter.jl:32-47 for a
  // triangle (nv = (N + 1)(N + 2) / 2, modes i outer, j inner, total degree i + j): the
  // exponential filter of Hesthaven & Warburton (Nodal DG Methods, sec. 5.3),
  // sigma = exp(-alpha ((d - Nc) / (N - Nc))^s) for mode degree d >= Nc, alpha = -log(eps)
  const double alpha = -std::log(2.220446049250313e-16);
  std::vector<double> sigma(nv, 1.0);
  if (nv == N + 1) {
    for (int i = Nc; i <= N; ++i) sigma[i] = std::exp(-alpha * std::pow((double)(i - Nc) / (N - Nc), s));
  } else if (nv == (N + 1) * (N + 2) / 2) {
    int sk = 0;
    for (int i = 0; i <= N; ++i)
      for (int j = 0; j <= N - i; ++j, ++sk)
        if (i + j >= Nc) sigma[sk] = std::exp(-alpha * std::pow((double)(i + j - Nc) / (N - Nc), s));
  } else {
    return fail(FR_E_UNSUPPORTED,
                "fr_filter_exp: V must have N + 1 (line) or (N + 1)(N + 2) / 2 (triangle) rows; the reference "
                "defines no filter for other bases");
  }
  std::vector<double> iv;
  if (!invV) {
    iv.resize((size_t)nv * nv);
    int rc = fr_inv(nv, V, iv.data());
    if (rc) return rc;
    invV = iv.data();
  }
  // F = V diag(sigma) invV, column-major
  for (int j = 0; j < nv; ++j)
    for (int i = 0; i < nv; ++i) {
      double a = 0.0;
      for (int k = 0; k < nv; ++k) a += V[i + (size_t)nv * k] * sigma[k] * invV[k + (size_t)nv * j];
      F[i + (size_t)nv * j] = a;
    }
  return FR_OK;
}

}  // extern "C"
