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
#include "fr_ops_table.h"

namespace fr {

// tile shape (elements along x and y); TX * TY = 64 so that one lane holds one element
#ifndef FR2D_TX
#define FR2D_TX 16
#endif
constexpr int TX = FR2D_TX, TY = 64 / FR2D_TX, TE = TX * TY;
static_assert(TE == 64 && (TX & (TX - 1)) == 0, "tile must hold 64 elements, TX a power of two");
// 1: each workgroup computes two tiles and loads the second one's rows during the first
#ifndef FR2D_PAIR
#define FR2D_PAIR 0
#endif
// 1: the tile's right x-faces and top y-faces share one Riemann pass after the y traces
#ifndef FR2D_MERGE_HALO
#define FR2D_MERGE_HALO 1
#endif
// tile order: 0 = raster over (bx, by); W > 0 = vertical strips of W tiles
#ifndef FR2D_STRIP
#define FR2D_STRIP 0
#endif
// global accesses: 0 = flat global addressing, 1 = one scalar buffer resource per plane,
// 2 = one buffer resource per array with the plane offset as scalar offset (arrays < 4 GiB)
#ifndef FR2D_BUFFER
#define FR2D_BUFFER 0
#endif
static_assert(2 * (TX + TY) <= 64, "the halo tasks of one line must fit in one wavefront");


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
__global__ __launch_bounds__(NSP * 64) void fr2d_kernel(const Geo2D geo, const StageArgs sa) {
  // operators are compile-time constants (Gauss-Legendre points, Radau correction), so they
  // occupy no scalar registers for the length of the kernel
  constexpr StdOps2D<NSP> ops{};
  __shared__ double sU[NSP * NSP * 4 * TE];
  __shared__ double sX[4 * NSP * TY * (TX + 1)];
  __shared__ double sY[4 * NSP * (TY + 1) * TX];

  const int tid = threadIdx.x;
  // the line index is uniform over a wavefront, so every plane offset below is a scalar
  const int line = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int ex = lane % TX, ey = lane / TX;
  const int ntiles = geo.ntx * geo.nty;
  const unsigned nxg = geo.nx + 2;
  long long ldp = geo.ldp;
  const double *u = sa.u, *u0p = sa.u0, *accin = sa.acc_in;
  double *outp = sa.out, *accout = sa.acc_out;
  const double g = geo.g, gm1 = geo.gm1, iJ11 = geo.iJ11, iJ22 = geo.iJ22;
  const bool with_u0 = EPI == EPI_STAGE && sa.a0 != 0.0;

  auto SX = [&](int m, int l, int y, int fx) -> double & {
    return sX[((m * NSP + l) * TY + y) * (TX + 1) + fx];
  };
  auto SY = [&](int m, int c, int fy, int x) -> double & {
    return sY[((m * NSP + c) * (TY + 1) + fy) * TX + x];
  };
  auto SU = [&](int k, int l, int m) -> double & { return sU[((k * NSP + l) * 4 + m) * TE + lane]; };
  // value of plane (k, l, m) at the element with byte offset bo inside the plane
#if FR2D_BUFFER == 2
  // one buffer resource per array (arrays below 4 GiB), the plane offset in the scalar offset
  const unsigned ldp8 = (unsigned)(ldp * 8);
  auto LD = [&](const double *base, int k, int l, int m, unsigned bo) -> double {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(
        plane_rsrc(base, 0xFFFFFFFFu), bo, (int)(ldp8 * (unsigned)(k + NSP * (l + NSP * m))), 0));
  };
  auto ST = [&](double *base, int k, int l, int m, unsigned bo, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(fr_u2, v), plane_rsrc(base, 0xFFFFFFFFu), bo,
                                          (int)(ldp8 * (unsigned)(k + NSP * (l + NSP * m))), 0);
  };
#elif FR2D_BUFFER
  const unsigned pbytes = (unsigned)(ldp * 8);
  auto LD = [&](const double *base, int k, int l, int m, unsigned bo) -> double {
    return bload(plane_rsrc(base + ldp * (k + NSP * (l + NSP * m)), pbytes), bo);
  };
  auto ST = [&](double *base, int k, int l, int m, unsigned bo, double v) {
    bstore(plane_rsrc(base + ldp * (k + NSP * (l + NSP * m)), pbytes), bo, v);
  };
#else
  auto LD = [&](const double *base, int k, int l, int m, unsigned bo) -> double {
    return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base + ldp * (k + NSP * (l + NSP * m))) + bo);
  };
  auto ST = [&](double *base, int k, int l, int m, unsigned bo, double v) {
    *reinterpret_cast<double *>(reinterpret_cast<char *>(base + ldp * (k + NSP * (l + NSP * m))) + bo) = v;
  };
#endif
  // geometry of one tile for this thread
  struct TileGeo {
    int i0, j0, ncx, ncy;
    bool act;
    unsigned ebo;  // byte offset of the own element inside a plane ((nx+2)(ny+2) < 2^29 on the host)
  };
  auto tile_geo = [&](int vt) {
    TileGeo t;
    const int tile = xcd_tile(vt, ntiles);
    int bx, by;
#if FR2D_STRIP > 0
    // strip-major order: vertical strips FR2D_STRIP tiles wide, rows inside a strip, so that the
    // tile below a tile was processed FR2D_STRIP tiles earlier (its lines are still in L2)
    {
      const int full = geo.ntx / FR2D_STRIP, per = FR2D_STRIP * geo.nty, s = tile / per;
      if (s < full) {
        const int r = tile - s * per;
        by = r / FR2D_STRIP;
        bx = s * FR2D_STRIP + r % FR2D_STRIP;
      } else {
        const int wl = geo.ntx - full * FR2D_STRIP, r = tile - full * per;
        by = r / wl;
        bx = full * FR2D_STRIP + r % wl;
      }
    }
#else
    bx = tile % geo.ntx;
    by = tile / geo.ntx;
#endif
    t.i0 = 1 + bx * TX;
    t.j0 = 1 + by * TY;
    t.ncx = min(TX, geo.nx - t.i0 + 1);
    t.ncy = min(TY, geo.ny - t.j0 + 1);
    t.act = ex < t.ncx && ey < t.ncy;
    t.ebo = 8u * ((unsigned)(t.i0 + ex) + nxg * (unsigned)(t.j0 + ey));
    return t;
  };
  // own row `line` of the element (and the u0 operand of the stage epilogue)
  auto load_row = [&](const TileGeo &t, double (&r)[NSP][4], double (&r0)[NSP][4]) {
    if (t.act) {
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int k = 0; k < NSP; ++k) r[k][m] = LD(u, k, line, m, t.ebo);
      if (with_u0) {
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int k = 0; k < NSP; ++k) r0[k][m] = LD(u0p, k, line, m, t.ebo);
      }
    }
  };

  if constexpr (EPI == EPI_RESIDUAL) {
    if (blockIdx.x == 0) {  // du vanishes on the four ghost corners (example/shock-vortex.jl:27)
      constexpr int per = NSP * NSP * 4;
      for (int v = tid; v < 4 * per; v += NSP * 64) {
        const int c = v / per, o = v % per;
        const long long ci = (c & 1) ? geo.nx + 1 : 0, cj = (c & 2) ? geo.ny + 1 : 0;
        sa.out[ci + (long long)nxg * cj + ldp * o] = 0.0;
      }
    }
  }

  // One tile of the workgroup; with `more`, the own rows of the next tile NT are loaded into
  // nur / nu0r as soon as the current rows are consumed, so they arrive during this tile's work.
  auto run_tile = [&](const TileGeo &T, double (&ur)[NSP][4], double (&u0r)[NSP][4], const bool more,
                      const TileGeo &NT, double (&nur)[NSP][4], double (&nu0r)[NSP][4]) {
    const int i0 = T.i0, j0 = T.j0, ncx = T.ncx, ncy = T.ncy;
    const bool act = T.act;
    const unsigned ebo = T.ebo;
    if (!act) {
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int k = 0; k < NSP; ++k) ur[k][m] = (m == 0 || m == 3) ? 1.0 : 0.0;
    }

    // -------------------------------------------------------------- P0: own row + halo lines
    // Halo tasks of this wavefront's line: lanes [0, TX) bottom (column `line` of the element
    // below the tile, its top trace), [TX, 2TX) top (bottom trace of the element above),
    // [2TX, 2TX+TY) left (row `line`, right trace of the left neighbour), [2TX+TY, 2TX+2TY) right.
    int hkind = -1, hpos = 0;
    if (lane < TX) {
      hkind = 0;
      hpos = lane;
    } else if (lane < 2 * TX) {
      hkind = 1;
      hpos = lane - TX;
    } else if (lane < 2 * TX + TY) {
      hkind = 2;
      hpos = lane - 2 * TX;
    } else if (lane < 2 * TX + 2 * TY) {
      hkind = 3;
      hpos = lane - 2 * TX - TY;
    }
    if (hkind >= 0 && hpos >= (hkind <= 1 ? ncx : ncy)) hkind = -1;
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
      const unsigned hbo = 8u * ((unsigned)si + nxg * (unsigned)sj);
      const bool lo = hkind == 0 || hkind == 2;  // weights lr (trace on the far side) or ll
      double w[NSP];
#pragma unroll
      for (int q = 0; q < NSP; ++q) w[q] = lo ? ops.lr[q] : ops.ll[q];
      if (hkind <= 1) {  // column `line`: points (k = line, l = q)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < NSP; ++q) s += LD(u, line, q, m, hbo) * w[q];
          htr[m] = s;
        }
      } else {  // row `line`: points (k = q, l = line)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < NSP; ++q) s += LD(u, q, line, m, hbo) * w[q];
          htr[m] = s;
        }
      }
      if (hkind == 0) {
#pragma unroll
        for (int m = 0; m < 4; ++m) SY(m, line, 0, hpos) = htr[m];
      } else if (hkind == 2) {
#pragma unroll
        for (int m = 0; m < 4; ++m) SX(m, line, hpos, 0) = htr[m];
      }
      if constexpr (EPI == EPI_RESIDUAL) {
        // du vanishes on the ghost layer (example/shock-vortex.jl:27)
        const bool ghost = hi == 0 || hi == geo.nx + 1 || hj == 0 || hj == geo.ny + 1;
        if (ghost) {
          const unsigned gbo = 8u * ((unsigned)hi + nxg * (unsigned)hj);
          if (hkind <= 1) {
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
              for (int q = 0; q < NSP; ++q) ST(outp, line, q, m, gbo, 0.0);
          } else {
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
              for (int q = 0; q < NSP; ++q) ST(outp, q, line, m, gbo, 0.0);
          }
        }
      }
    }

    // stage epilogue operand a0*u0 + a1*u of the own row
    double pre[NSP][4];
    if constexpr (EPI == EPI_STAGE) {
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int k = 0; k < NSP; ++k) {
          double v = sa.a1 * ur[k][m];
          if (with_u0) v = sa.a0 * u0r[k][m] + v;
          pre[k][m] = v;
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
        XState xs;
        xstate(ur[k], gm1, xs);
#pragma unroll
        for (int m = 0; m < 4; ++m) f1[k][m] = iJ11 * xs.F[m];
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

    // the own rows are consumed: start loading the next tile's
    if (more) load_row(NT, nur, nu0r);
    __syncthreads();  // B1: sU rows, x traces and left/bottom halo traces are visible

    // -------------------------------------------------------------- P1: column sweep + x Riemann
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
#if !FR2D_MERGE_HALO
    if (hkind == 3) {  // right boundary face of the tile row, solved by the right halo task
      double wl[4], fw[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) wl[m] = SX(m, line, hpos, ncx);
      riemann_x4<FLUX>(wl, htr, g, gm1, fw);
#pragma unroll
      for (int m = 0; m < 4; ++m) SX(m, line, hpos, ncx) = fw[m];
    }
#endif
    __syncthreads();  // B2: y traces visible

    // -------------------------------------------------------------- P2: y Riemann
    if (act) {
      double wb[4], fw[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) wb[m] = SY(m, line, ey, ex);
      riemann_y4<FLUX>(wb, uBy, g, gm1, fw);
#pragma unroll
      for (int m = 0; m < 4; ++m) SY(m, line, ey, ex) = fw[m];
    }
#if FR2D_MERGE_HALO
    // One divergent pass for the tile's right x-faces (hkind 3, solved by the right halo task)
    // and top y-faces (hkind 1): a y-face is solved in its normal frame n = (0, 1), i.e. on
    // (rho, m_y, -m_x, E) with the flux rotated back as (f0, -f2, f1, f3).
    if (hkind == 1 || hkind == 3) {
      const bool xf = hkind == 3;
      double *slot[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) slot[m] = xf ? &SX(m, line, hpos, ncx) : &SY(m, line, ncy, hpos);
      double w[4], a[4], b[4], fw[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) w[m] = *slot[m];
      a[0] = w[0];
      a[1] = xf ? w[1] : w[2];
      a[2] = xf ? w[2] : -w[1];
      a[3] = w[3];
      b[0] = htr[0];
      b[1] = xf ? htr[1] : htr[2];
      b[2] = xf ? htr[2] : -htr[1];
      b[3] = htr[3];
      riemann_x4<FLUX>(a, b, g, gm1, fw);
      *slot[0] = fw[0];
      *slot[1] = xf ? fw[1] : -fw[2];
      *slot[2] = xf ? fw[2] : fw[1];
      *slot[3] = fw[3];
    }
#else
    if (hkind == 1) {  // top boundary face of the tile column
      double wb[4], fw[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) wb[m] = SY(m, line, ncy, hpos);
      riemann_y4<FLUX>(wb, htr, g, gm1, fw);
#pragma unroll
      for (int m = 0; m < 4; ++m) SY(m, line, ncy, hpos) = fw[m];
    }
#endif
    __syncthreads();  // B3: interface fluxes visible; sU is free

    // -------------------------------------------------------------- P3: corrections
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

    // -------------------------------------------------------------- P4: assemble + epilogue
    if (act) {
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int k = 0; k < NSP; ++k) {
          const double du = -(dux[k][m] + SU(k, line, m));
          if constexpr (EPI == EPI_RESIDUAL) {
            ST(outp, k, line, m, ebo, du);
          } else {
            ST(outp, k, line, m, ebo, pre[k][m] + sa.b * du);
            if (accout) ST(accout, k, line, m, ebo, LD(accin, k, line, m, ebo) + sa.b2 * du);
          }
        }
    }
  };

  const int ntg = gridDim.x;
  const int vt0 = blockIdx.x;
  if (vt0 >= ntiles) return;  // whole workgroup: grid rounded up to a multiple of 8
  double ua[NSP][4], ua0[NSP][4], ub[NSP][4], ub0[NSP][4];
  const TileGeo T0 = tile_geo(vt0);
  load_row(T0, ua, ua0);
#if FR2D_PAIR
  // two tiles per workgroup (vt0 and vt0 + gridDim.x, the same XCD range when gridDim.x is a
  // multiple of 8): the second tile's rows are in flight while the first is computed
  const int vt1 = vt0 + ntg;
  const bool two = vt1 < ntiles;
  const TileGeo T1 = tile_geo(two ? vt1 : vt0);
  run_tile(T0, ua, ua0, two, T1, ub, ub0);
  if (two) {
    // the second tile recomputes its plane addresses instead of keeping the first tile's live
    asm volatile("" : "+s"(ldp), "+s"(u), "+s"(u0p), "+s"(outp), "+s"(accin), "+s"(accout));
    __syncthreads();  // sU is read at the end of the first tile before the second writes it
    run_tile(T1, ub, ub0, false, T1, ua, ua0);
  }
#else
  (void)ntg;
  run_tile(T0, ua, ua0, false, T0, ub, ub0);
#endif
}

// ------------------------------------------------------------------ ghost-layer helpers
// rule FR_GHOST_PERIODIC / _OUTFLOW / _INFLOW_LEFT over every ghost cell and every (k,l,m) plane.
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
    } else if (rule == FR_GHOST_OUTFLOW) {
      si = (i == 0) ? 1 : (i == nx + 1 ? nx : i);
      sj = (j == 0) ? 1 : (j == ny + 1 ? ny : j);
    } else {
      // bottom/top rows first (every i), then the right column (every j); the left column
      // 1..ny is left alone, so each ghost reads a cell that this pass does not write
      if (i == 0 && j >= 1 && j <= ny) continue;
      si = (i == nx + 1) ? nx : i;
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
  // one tile per workgroup (two with FR2D_PAIR); a multiple of 8 workgroups keeps both tiles of a
  // pair in one XCD's range of the tile order
  const long long nb = (long long)geo.ntx * geo.nty;
  const long long tpw = FR2D_PAIR ? 2 : 1;
  long long grid = (nb + tpw - 1) / tpw;
  grid = std::min(nb, (grid + 7) / 8 * 8);
  hipLaunchKernelGGL((fr2d_kernel<NSP, FLUX, EPI>), dim3((unsigned)grid), dim3(NSP * 64), 0, st, geo, sa);
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
