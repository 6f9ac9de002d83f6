// fr2d_euler.hip -- 2D Euler flux-reconstruction residual on uniform quads for gfx950.
//
// One fused kernel evaluates the residual of example/shock-vortex.jl:26-118 (discontinuous flux,
// ll/lr interpolation to flux points, interface Riemann flux, dl derivative, dhl/dhr correction)
// and, in stage mode, the explicit RK update out = a0*u0 + a1*u + b*L(u) in its epilogue.
//
// Mapping: a workgroup owns a tile of TX x TY = 16 x 4 elements and has NSP wavefronts.  Lane =
// element of the tile (16 consecutive elements along x per row, so every global access of a
// wavefront covers four 128-byte segments of the element-fastest SoA layout), wave index = one
// line of solution points.  Each thread first holds row `line` of its element (x sweep: flux,
// traces, derivative), the tile is transposed through LDS, and the same thread then holds column
// `line` (y sweep).  Face traces are exchanged through LDS; the traces of the 2*(TX+TY) halo
// elements around the tile are computed by extra line tasks, so every element is read from HBM
// once and the halo re-reads hit L2 (tiles are ordered XCD-aware).
#include <algorithm>
#include <cstring>
#include <new>

#include "fr_common.h"
#include "fr2d_internal.h"

namespace fr {

// tile shape (elements along x and y); TX * TY = 64 so that one lane holds one element
#ifndef FR2D_TX
#define FR2D_TX 16
#endif
constexpr int TX = FR2D_TX, TY = 64 / FR2D_TX, TE = TX * TY;
static_assert(TE == 64 && (TX & (TX - 1)) == 0, "tile must hold 64 elements, TX a power of two");
static_assert(2 * (TX + TY) <= 64, "halo line tasks must fit in one pass of the workgroup");

template <int NSP>
struct Ops2D {
  double ll[NSP], lr[NSP], dl[NSP * NSP], dhl[NSP], dhr[NSP];
};

struct Geo2D {
  int nx, ny, ntx, nty;
  long long ldp;  // elements per (k,l,m) plane: (nx+2)(ny+2)
  double iJ11, iJ22, g, gm1;
  int periodic;
};

enum { EPI_RESIDUAL = 0, EPI_STAGE = 1 };

struct StageArgs {
  const double *u;   // state whose residual is taken
  double *out;       // du (residual) or the stage result
  const double *u0;  // a0 operand (unused when a0 == 0)
  const double *acc_in;
  double *acc_out;   // optional second output acc_out = acc_in + b2 * L(u)
  double a0, a1, b, b2;
};

template <int NSP, int FLUX, int EPI>
__global__ __launch_bounds__(NSP * 64) void fr2d_kernel(const Geo2D geo, const Ops2D<NSP> ops,
                                                        const StageArgs sa) {
  constexpr int NL = NSP * 4;  // values of one line (NSP points x 4 variables)
  __shared__ double sU[NSP * NSP * 4 * TE];
  __shared__ double sX[4 * NSP * TY * (TX + 1)];
  __shared__ double sY[4 * NSP * (TY + 1) * TX];

  const int tid = threadIdx.x;
  const int line = tid >> 6, lane = tid & 63;
  const int ex = lane % TX, ey = lane / TX;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int bx = tile % geo.ntx, by = tile / geo.ntx;
  const int i0 = 1 + bx * TX, j0 = 1 + by * TY;
  const int ncx = min(TX, geo.nx - i0 + 1), ncy = min(TY, geo.ny - j0 + 1);
  const bool act = ex < ncx && ey < ncy;
  const long long nxg = geo.nx + 2;
  const long long ldp = geo.ldp;
  const long long e = (long long)(i0 + ex) + nxg * (j0 + ey);
  const double *__restrict__ u = sa.u;
  const double g = geo.g, gm1 = geo.gm1, iJ11 = geo.iJ11, iJ22 = geo.iJ22;

  auto SX = [&](int m, int l, int y, int fx) -> double & {
    return sX[((m * NSP + l) * TY + y) * (TX + 1) + fx];
  };
  auto SY = [&](int m, int c, int fy, int x) -> double & {
    return sY[((m * NSP + c) * (TY + 1) + fy) * TX + x];
  };
  auto SU = [&](int k, int l, int m) -> double & { return sU[((k * NSP + l) * 4 + m) * TE + lane]; };

  // ---------------------------------------------------------------- P0: own row + halo lines
  double ur[NSP][4];
  if (act) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) ur[k][m] = u[e + ldp * (k + NSP * (line + NSP * m))];
  } else {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) ur[k][m] = (m == 0 || m == 3) ? 1.0 : 0.0;
  }

  // stage epilogue operand a0*u0 + a1*u of the own row, formed while the row is in registers
  double pre[NSP][4];
  if constexpr (EPI == EPI_STAGE) {
    if (act) {
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int k = 0; k < NSP; ++k) {
          const long long idx = e + ldp * (k + NSP * (line + NSP * m));
          double v = sa.a1 * ur[k][m];
          if (sa.a0 != 0.0) v = sa.a0 * sa.u0[idx] + v;
          pre[k][m] = v;
        }
    }
  }

  // halo task: kind 0 bottom (column c, uT), 1 top (column c, uB), 2 left (row l, uR), 3 right (row l, uL)
  int hkind = -1, hpos = 0, hline = 0;
  {
    int t = tid;
    if (t < TX * NSP) {
      hkind = 0;
      hpos = t % TX;
      hline = t / TX;
    } else if ((t -= TX * NSP) < TX * NSP) {
      hkind = 1;
      hpos = t % TX;
      hline = t / TX;
    } else if ((t -= TX * NSP) < TY * NSP) {
      hkind = 2;
      hpos = t % TY;
      hline = t / TY;
    } else if ((t -= TY * NSP) < TY * NSP) {
      hkind = 3;
      hpos = t % TY;
      hline = t / TY;
    }
    if (hkind == 0 || hkind == 1) {
      if (hpos >= ncx) hkind = -1;
    } else if (hkind >= 2) {
      if (hpos >= ncy) hkind = -1;
    }
  }
  double htr[4] = {0.0, 0.0, 0.0, 0.0};
  if (hkind >= 0) {
    int hi, hj;
    if (hkind == 0) {
      hi = i0 + hpos;
      hj = j0 - 1;
    } else if (hkind == 1) {
      hi = i0 + hpos;
      hj = j0 + ncy;
    } else if (hkind == 2) {
      hi = i0 - 1;
      hj = j0 + hpos;
    } else {
      hi = i0 + ncx;
      hj = j0 + hpos;
    }
    int si = hi, sj = hj;
    if (geo.periodic) {
      si = (hi == 0) ? geo.nx : (hi == geo.nx + 1 ? 1 : hi);
      sj = (hj == 0) ? geo.ny : (hj == geo.ny + 1 ? 1 : hj);
    }
    const long long he = (long long)si + nxg * sj;
    const bool column = hkind <= 1;
    const double *w = (hkind == 0 || hkind == 2) ? ops.lr : ops.ll;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; ++q) {
        const long long off = column ? (hline + NSP * (q + NSP * m)) : (q + NSP * (hline + NSP * m));
        s += u[he + ldp * off] * w[q];
      }
      htr[m] = s;
    }
    if (hkind == 0) {
#pragma unroll
      for (int m = 0; m < 4; ++m) SY(m, hline, 0, hpos) = htr[m];
    } else if (hkind == 2) {
#pragma unroll
      for (int m = 0; m < 4; ++m) SX(m, hline, hpos, 0) = htr[m];
    }
    if constexpr (EPI == EPI_RESIDUAL) {
      // du vanishes on the ghost layer (example/shock-vortex.jl:27)
      const bool ghost = hi == 0 || hi == geo.nx + 1 || hj == 0 || hj == geo.ny + 1;
      if (ghost) {
        const long long ge = (long long)hi + nxg * hj;
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int q = 0; q < NSP; ++q) {
            const long long off = column ? (hline + NSP * (q + NSP * m)) : (q + NSP * (hline + NSP * m));
            sa.out[ge + ldp * off] = 0.0;
          }
      }
    }
  }
  if constexpr (EPI == EPI_RESIDUAL) {
    if (blockIdx.x == 0) {  // the four ghost corners
      constexpr int per = NSP * NSP * 4;
      for (int v = tid; v < 4 * per; v += NSP * 64) {
        const int c = v / per, o = v % per;
        const long long ci = (c & 1) ? geo.nx + 1 : 0, cj = (c & 2) ? geo.ny + 1 : 0;
        sa.out[ci + nxg * cj + ldp * o] = 0.0;
      }
    }
  }

  // row: LDS transpose copy, x traces, contravariant x flux, its traces and dl derivative
  double uLx[4], dux[NSP][4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      a += ur[k][m] * ops.ll[k];
      b += ur[k][m] * ops.lr[k];
    }
    uLx[m] = a;
    if (act) SX(m, line, ey, ex + 1) = b;
  }
  if (act) {
#pragma unroll
    for (int k = 0; k < NSP; ++k)
#pragma unroll
      for (int m = 0; m < 4; ++m) SU(k, line, m) = ur[k][m];
  }
  {
    double f1[NSP][4];
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      XState s;
      xstate(ur[k], gm1, s);
#pragma unroll
      for (int m = 0; m < 4; ++m) f1[k][m] = iJ11 * s.F[m];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; ++q) {
        a += f1[q][m] * ops.ll[q];
        b += f1[q][m] * ops.lr[q];
      }
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        double d = 0.0;
#pragma unroll
        for (int q = 0; q < NSP; ++q) d += f1[q][m] * ops.dl[k + NSP * q];
        // derivative minus the discontinuous-flux traces of the correction terms
        dux[k][m] = d - a * ops.dhl[k] - b * ops.dhr[k];
      }
    }
  }
  __syncthreads();  // B1: sU rows, x traces and left/bottom halo traces are visible

  // ---------------------------------------------------------------- P1: column sweep + x Riemann
  double uBy[4], duy[NSP][4];
  {
    double uc[NSP][4];
    if (act) {
#pragma unroll
      for (int q = 0; q < NSP; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m) uc[q][m] = SU(line, q, m);
    } else {
#pragma unroll
      for (int q = 0; q < NSP; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m) uc[q][m] = (m == 0 || m == 3) ? 1.0 : 0.0;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; ++q) {
        a += uc[q][m] * ops.ll[q];
        b += uc[q][m] * ops.lr[q];
      }
      uBy[m] = a;
      if (act) SY(m, line, ey + 1, ex) = b;
    }
    double f2[NSP][4];
#pragma unroll
    for (int q = 0; q < NSP; ++q) {
      const double r = fr_rcp(uc[q][0]);
      const double uu = uc[q][1] * r, vv = uc[q][2] * r;
      const double p = gm1 * (uc[q][3] - 0.5 * (uc[q][1] * uu + uc[q][2] * vv));
      f2[q][0] = iJ22 * uc[q][2];
      f2[q][1] = iJ22 * (uc[q][1] * vv);
      f2[q][2] = iJ22 * (uc[q][2] * vv + p);
      f2[q][3] = iJ22 * ((uc[q][3] + p) * vv);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; ++q) {
        a += f2[q][m] * ops.ll[q];
        b += f2[q][m] * ops.lr[q];
      }
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        double d = 0.0;
#pragma unroll
        for (int q = 0; q < NSP; ++q) d += f2[q][m] * ops.dl[k + NSP * q];
        duy[k][m] = d - a * ops.dhl[k] - b * ops.dhr[k];
      }
    }
  }
  // x-face on the left of the own element: left state = right trace of the neighbour
  if (act) {
    double wl[4], fw[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) wl[m] = SX(m, line, ey, ex);
    riemann_x4<FLUX>(wl, uLx, g, gm1, fw);
#pragma unroll
    for (int m = 0; m < 4; ++m) SX(m, line, ey, ex) = fw[m];
  }
  if (hkind == 3) {  // right boundary face of the tile row, solved by the right halo task
    double wl[4], fw[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) wl[m] = SX(m, hline, hpos, ncx);
    riemann_x4<FLUX>(wl, htr, g, gm1, fw);
#pragma unroll
    for (int m = 0; m < 4; ++m) SX(m, hline, hpos, ncx) = fw[m];
  }
  __syncthreads();  // B2: y traces visible

  // ---------------------------------------------------------------- P2: y Riemann
  if (act) {
    double wb[4], fw[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) wb[m] = SY(m, line, ey, ex);
    riemann_y4<FLUX>(wb, uBy, g, gm1, fw);
#pragma unroll
    for (int m = 0; m < 4; ++m) SY(m, line, ey, ex) = fw[m];
  }
  if (hkind == 1) {  // top boundary face of the tile column
    double wb[4], fw[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) wb[m] = SY(m, hline, ncy, hpos);
    riemann_y4<FLUX>(wb, htr, g, gm1, fw);
#pragma unroll
    for (int m = 0; m < 4; ++m) SY(m, hline, ncy, hpos) = fw[m];
  }
  __syncthreads();  // B3: interface fluxes visible; sU is free

  // ---------------------------------------------------------------- P3: corrections
  if (act) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double cb = SY(m, line, ey, ex) * iJ22;
      const double ct = SY(m, line, ey + 1, ex) * iJ22;
#pragma unroll
      for (int k = 0; k < NSP; ++k) SU(line, k, m) = duy[k][m] + cb * ops.dhl[k] + ct * ops.dhr[k];
      const double cl = SX(m, line, ey, ex) * iJ11;
      const double cr = SX(m, line, ey, ex + 1) * iJ11;
#pragma unroll
      for (int k = 0; k < NSP; ++k) dux[k][m] += cl * ops.dhl[k] + cr * ops.dhr[k];
    }
  }
  __syncthreads();  // B4: y contributions of every point visible

  // ---------------------------------------------------------------- P4: assemble + epilogue
  if (act) {
    double *__restrict__ out = sa.out;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        const long long idx = e + ldp * (k + NSP * (line + NSP * m));
        const double du = -(dux[k][m] + SU(k, line, m));
        if constexpr (EPI == EPI_RESIDUAL) {
          out[idx] = du;
        } else {
          out[idx] = pre[k][m] + sa.b * du;
          if (sa.acc_out) sa.acc_out[idx] = sa.acc_in[idx] + sa.b2 * du;
        }
      }
  }
  (void)NL;
}

// ------------------------------------------------------------------ ghost-layer helpers
// rule FR_GHOST_PERIODIC / FR_GHOST_OUTFLOW over every ghost cell and every (k,l,m) plane.
__global__ void fr2d_fill_ghost_kernel(double *u, int nx, int ny, long long ldp, int planes, int rule) {
  const int nring = 2 * (nx + 2) + 2 * ny;
  const long long total = (long long)nring * planes;
  const long long nxg = nx + 2;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int g = (int)(t % nring);
    const long long pl = t / nring;
    int i, j;
    if (g < nx + 2) {
      i = g;
      j = 0;
    } else if (g < 2 * (nx + 2)) {
      i = g - (nx + 2);
      j = ny + 1;
    } else if (g < 2 * (nx + 2) + ny) {
      i = 0;
      j = g - 2 * (nx + 2) + 1;
    } else {
      i = nx + 1;
      j = g - 2 * (nx + 2) - ny + 1;
    }
    int si = i, sj = j;
    if (rule == FR_GHOST_PERIODIC) {
      si = (i == 0) ? nx : (i == nx + 1 ? 1 : i);
      sj = (j == 0) ? ny : (j == ny + 1 ? 1 : j);
    } else {
      si = (i == 0) ? 1 : (i == nx + 1 ? nx : i);
      sj = (j == 0) ? 1 : (j == ny + 1 ? ny : j);
    }
    u[(long long)i + nxg * j + ldp * pl] = u[(long long)si + nxg * sj + ldp * pl];
  }
}

// ------------------------------------------------------------------ launch dispatch
template <int NSP, int FLUX, int EPI>
static int launch_nsp(const Fr2dCtx &c, const StageArgs &sa, hipStream_t st) {
  Geo2D geo;
  geo.nx = c.nx;
  geo.ny = c.ny;
  geo.ntx = (c.nx + TX - 1) / TX;
  geo.nty = (c.ny + TY - 1) / TY;
  geo.ldp = (long long)(c.nx + 2) * (c.ny + 2);
  geo.iJ11 = c.iJ11;
  geo.iJ22 = c.iJ22;
  geo.g = c.gamma;
  geo.gm1 = c.gamma - 1.0;
  geo.periodic = c.bc == FR_BC_PERIODIC;
  Ops2D<NSP> ops;
  for (int k = 0; k < NSP; ++k) {
    ops.ll[k] = c.ops.ll[k];
    ops.lr[k] = c.ops.lr[k];
    ops.dhl[k] = c.ops.dhl[k];
    ops.dhr[k] = c.ops.dhr[k];
  }
  for (int q = 0; q < NSP * NSP; ++q) ops.dl[q] = c.ops.dl[q];
  const int nb = geo.ntx * geo.nty;
  hipLaunchKernelGGL((fr2d_kernel<NSP, FLUX, EPI>), dim3(nb), dim3(NSP * 64), 0, st, geo, ops, sa);
  FR_LAUNCHED("fr2d_kernel");
  return FR_OK;
}

template <int NSP>
static int launch_flux(const Fr2dCtx &c, const StageArgs &sa, bool residual, hipStream_t st) {
  switch (c.flux) {
    case FR_FLUX_HLL:
      return residual ? launch_nsp<NSP, FR_FLUX_HLL, EPI_RESIDUAL>(c, sa, st)
                      : launch_nsp<NSP, FR_FLUX_HLL, EPI_STAGE>(c, sa, st);
    case FR_FLUX_RUSANOV:
      return residual ? launch_nsp<NSP, FR_FLUX_RUSANOV, EPI_RESIDUAL>(c, sa, st)
                      : launch_nsp<NSP, FR_FLUX_RUSANOV, EPI_STAGE>(c, sa, st);
    default:
      return residual ? launch_nsp<NSP, FR_FLUX_ROE, EPI_RESIDUAL>(c, sa, st)
                      : launch_nsp<NSP, FR_FLUX_ROE, EPI_STAGE>(c, sa, st);
  }
}

static int launch(const Fr2dCtx &c, const StageArgs &sa, bool residual, hipStream_t st) {
  switch (c.nsp) {
    case 2: return launch_flux<2>(c, sa, residual, st);
    case 3: return launch_flux<3>(c, sa, residual, st);
    case 4: return launch_flux<4>(c, sa, residual, st);
    case 5: return launch_flux<5>(c, sa, residual, st);
    default: return fail(FR_E_UNSUPPORTED, "fr2d: degree out of range");
  }
}

int fr2d_launch_residual(const Fr2dCtx &c, const double *d_u, double *d_du, hipStream_t st) {
  StageArgs sa{};
  sa.u = d_u;
  sa.out = d_du;
  return launch(c, sa, true, st);
}

int fr2d_launch_stage(const Fr2dCtx &c, const double *d_u0, const double *d_us, double *d_out, double a0,
                      double a1, double b, const double *acc_in, double *acc_out, double b2, hipStream_t st) {
  StageArgs sa{};
  sa.u = d_us;
  sa.out = d_out;
  sa.u0 = d_u0;
  sa.a0 = a0;
  sa.a1 = a1;
  sa.b = b;
  sa.acc_in = acc_in;
  sa.acc_out = acc_out;
  sa.b2 = b2;
  return launch(c, sa, false, st);
}

int fr2d_launch_fill_ghost(const Fr2dCtx &c, double *d_u, int rule, hipStream_t st) {
  const long long ldp = (long long)(c.nx + 2) * (c.ny + 2);
  const int planes = c.nsp * c.nsp * 4;
  const long long total = (long long)(2 * (c.nx + 2) + 2 * c.ny) * planes;
  const int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(fr2d_fill_ghost_kernel, dim3(blocks), dim3(256), 0, st, d_u, c.nx, c.ny, ldp, planes, rule);
  FR_LAUNCHED("fr2d_fill_ghost_kernel");
  return FR_OK;
}

}  // namespace fr
