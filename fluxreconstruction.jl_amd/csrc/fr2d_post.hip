// fr2d_post.hip -- post-processing of the 2D state: the primitive variables of every solution
// point, assembled on the global grid of solution points (example/shock-vortex.jl:333-352).
#include <algorithm>

#include "fr2d_internal.h"

namespace fr {

// One thread per (element, solution point), element fastest, so the reads of each (k, l, m)
// plane are coalesced.  Output, column-major (nx*nsp, ny*nsp, 4) like the reference's `sol`:
// sol[(i-1)*nsp + k, (j-1)*nsp + l, :] = (rho, U, V, 1/lambda) with KitBase's lambda = rho/(2p),
// i.e. the last entry is 2p/rho.  Plain IEEE division: this is output, not the hot path.
template <int NSP>
__global__ void fr2d_primitives_kernel(const double *__restrict__ u, double *__restrict__ sol, int nx, int ny,
                                       int pitch, long long ldp, double gm1) {
  const long long ne = (long long)nx * ny;
  const long long n = ne * NSP * NSP;
  const long long NX = (long long)nx * NSP, NXY = NX * ny * NSP;
  const long long mstride = ldp * NSP * NSP;  // plane distance between two variables
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long e = t % ne;
    const int pt = (int)(t / ne);  // k + NSP*l
    const int i = (int)(e % nx) + 1, j = (int)(e / nx) + 1;
    const int k = pt % NSP, l = pt / NSP;
    const long long b = i + (long long)pitch * j + ldp * pt;
    const double rho = u[b], mx = u[b + mstride], my = u[b + 2 * mstride], E = u[b + 3 * mstride];
    const double U = mx / rho, V = my / rho;
    const double p = gm1 * (E - 0.5 * (mx * U + my * V));
    const long long o = ((long long)(i - 1) * NSP + k) + NX * ((long long)(j - 1) * NSP + l);
    sol[o] = rho;
    sol[o + NXY] = U;
    sol[o + 2 * NXY] = V;
    sol[o + 3 * NXY] = 2.0 * p / rho;
  }
}

template <int NSP>
static int launch_prim(const Fr2dCtx &c, const double *d_u, double *d_sol, hipStream_t st) {
  const long long n = (long long)c.nx * c.ny * NSP * NSP;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 8LL * std::max(c.num_cu, 1) * 4);
  hipLaunchKernelGGL(fr2d_primitives_kernel<NSP>, dim3(blocks), dim3(256), 0, st, d_u, d_sol, c.nx, c.ny, c.pitch,
                     c.ldpi, c.gamma - 1.0);
  FR_LAUNCHED("fr2d_primitives_kernel");
  return FR_OK;
}

int fr2d_launch_primitives(const Fr2dCtx &c, const double *d_u, double *d_sol, hipStream_t st) {
  switch (c.nsp) {
    case 2: return launch_prim<2>(c, d_u, d_sol, st);
    case 3: return launch_prim<3>(c, d_u, d_sol, st);
    case 4: return launch_prim<4>(c, d_u, d_sol, st);
    case 5: return launch_prim<5>(c, d_u, d_sol, st);
    case 6: return launch_prim<6>(c, d_u, d_sol, st);
    default: return fail(FR_E_UNSUPPORTED, "fr2d_primitives: degree out of range");
  }
}

}  // namespace fr
