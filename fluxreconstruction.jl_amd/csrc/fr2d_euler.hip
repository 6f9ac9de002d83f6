// fr2d_euler.hip -- 2D Euler flux-reconstruction residual on uniform quads for gfx950.
//
// One fused kernel evaluates the residual of example/shock-vortex.jl:26-118 (discontinuous flux,
// ll/lr interpolation to flux points, interface Riemann flux, dl derivative, dhl/dhr correction)
// and, in stage mode, the explicit RK update out = a0*u0 + a1*u + b*L(u) in its epilogue.
//
// Mapping: a workgroup owns a tile of TX x TY = 16 x 4 elements and has NSP wavefronts.  Lane =
// element of the tile (16 consecutive elements along x per row, so every global access of a
// wavefront covers four 128-byte segments of the element-fastest SoA layout), wave index = one
// line of solution points.  Each thread first holds row `line` of its element and finishes the
// whole x direction before the first barrier (flux, traces, derivative, the x-face Riemann
// problems inside the row and the correction terms); the tile is transposed through LDS and the
// same thread then holds column `line` for the y direction.  All face data of a line stays inside
// its wavefront (the neighbours of an element are lanes -1/+1 along x and -TX/+TX along y; the
// traces of the halo elements around the tile come from halo tasks on lanes of the same
// wavefront), so faces need only wavefront-level LDS ordering: x faces in a per-line slice,
// y faces in the line's own column block of the transpose buffer once that column is read.
// At nsp = 4 a workgroup takes 40 KB of LDS and 128 VGPRs (Rusanov), so four workgroups share
// a CU and a 256 x 256 grid (1024 tiles) is resident at once.  Every element is read from HBM
// once; halo re-reads hit L2 (tiles are ordered XCD-aware).
#include <algorithm>
#include <cstring>
#include <new>

#include "fr_common.h"
#include "fr2d_internal.h"
#include "fr_ops_table.h"

namespace fr {

// tile shape (elements along x and y); TX * TY = 64 so that one lane holds one element
constexpr int TX = 16, TY = 4, TE = TX * TY;
static_assert(TE == 64, "one lane per element of the tile");
static_assert(2 * (TX + TY) <= 64, "the halo tasks of one line must fit in one wavefront");
// global addressing of the state arrays: flat 64-bit pointers, or (arrays below 4 GiB) one buffer
// resource per array with the plane offset in the scalar offset, which spends no vector
// instruction on addresses
enum { ADDR_FLAT = 0, ADDR_BUFFER = 1 };

// cell geometry: rectangles (iJ diagonal, faces normal to the axes), uniform parallelograms
// (full 2x2 iJ, faces solved in the frames of the unit normals n1, n2), or point-varying metrics
// (a 2x2 iJ per solution point, read from an array shaped like a state whose four planes per
// point hold r_x, s_x, r_y, s_y; faces as for rectangles, the face flux of a point scaled by its
// own r_x (x faces) or s_y (y faces), as example/shock-vortex.jl:102-115 writes it)
enum { GEO_RECT = 0, GEO_AFFINE = 1, GEO_POINT = 2 };

struct Geo2D {
  int nx, ny, ntx, nty;
  int strip;      // tile-order strip width in tiles (strip_tile)
  int stag_cu;    // staggered start: workgroups per slot (the CU count), 0 = off
  int stag;       // staggered start: units of s_sleep(8) per slot
  int pitch;      // elements per row of a plane: nx+2 (caller arrays) or the padded pitch
  long long ldp;  // elements per (k,l,m) plane: pitch (ny+2)
  double iJ11, iJ22, g, gm1;
  int periodic;
  double iJ12, iJ21;          // off-diagonal inverse-Jacobian entries (GEO_AFFINE)
  double n1c, n1s, n2c, n2s;  // unit normals of the x faces and y faces (GEO_AFFINE)
  double q1, q2;              // contravariant scales of the face fluxes, (iJ n1)_1 and (iJ n2)_2
  const double *mj;           // per-point metric in the layout of the state arrays (GEO_POINT)
};

// epilogues: residual du; RK stage out = a0*u0 + a1*u + b*du (b != 0); the general stage form
// with du kept apart, optionally with the RK4 accumulator acc_out = acc_in + b2*du
enum { EPI_RESIDUAL = 0, EPI_STAGE = 1, EPI_ACC = 2 };

struct StageArgs {
  const double *u;   // state whose residual is taken
  double *out;       // du (residual) or the stage result
  const double *u0;  // a0 operand (unused when a0 == 0)
  const double *acc_in;
  double *acc_out;   // optional second output acc_out = acc_in + b2 * L(u)
  double a0, a1, b, b2;
  // optional copies of u's ghost layer (non-periodic grids): the stage buffers that later stages
  // read take the ghost cells from the halo loads of the first stage
  double *gc0, *gc1;
};

// Tile order: column strips of `sw` tiles, row by row inside a strip (then cut into one
// contiguous range per XCD by xcd_tile).  The tile below a workgroup's tile was then issued sw
// tiles earlier on the same XCD, so its halo lines are still in that XCD's L2 on wide grids;
// sw = ntx keeps plain row-major order.
__device__ __forceinline__ void strip_tile(int t, int ntx, int nty, int sw, int &bx, int &by) {
  const int full = ntx / sw;
  const int s = t / (sw * nty);
  const int ws = s < full ? sw : ntx - full * sw;  // > 0 for every t < ntx * nty
  const int r = t - s * sw * nty;
  by = r / ws;
  bx = s * sw + r % ws;
}

// Four workgroups per CU at nsp = 4: HLL and Rusanov fit 128 VGPRs (HLL with at most a few
// spilled values); Roe and the RK4 accumulator form would spill heavily and keep the default.
// The point-metric instances keep the default as well: their metric values stay live from the
// x sweep to the corrections.
template <int NSP, int FLUX, int EPI, int GEO>
constexpr int fr2d_min_blocks() {
  return (NSP == 4 && FLUX != FR_FLUX_ROE && EPI != EPI_ACC && GEO != GEO_POINT) ? 4 : 1;
}

template <int NSP, int FLUX, int EPI, int ADDR, int GEO>
__global__ __launch_bounds__(NSP * 64, (fr2d_min_blocks<NSP, FLUX, EPI, GEO>())) void fr2d_kernel(const Geo2D geo,
                                                                                              const StageArgs sa) {
  // operators are compile-time constants (Gauss-Legendre points, Radau correction), so they
  // occupy no scalar registers for the length of the kernel
  constexpr StdOps2D<NSP> ops{};
  // LDS: the tile transpose sU (rows -> columns, then the y contributions columns -> rows) and
  // the x faces of every line (sX).  The y faces of a line live in that line's own column block
  // of sU between the column read and the write-back of the y contributions.
  __shared__ double sU[NSP * NSP * 4 * TE];
  __shared__ double sX[4 * NSP * TY * TX];

  const int tid = threadIdx.x;
  // the line index is uniform over a wavefront, so every plane offset below is a scalar
  const int line = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int ex = lane % TX, ey = lane / TX;
  const unsigned nxg = geo.pitch;
  const long long ldp = geo.ldp;
  const double *u = sa.u;
  const double g = geo.g, gm1 = geo.gm1;
  // scale of the x and y contributions in the result: -iJ (residual, RK4 form) or -b*iJ (stage,
  // b != 0 on the host); the stage carries a1*u inside the x part as (a1 / sxs)*u.  Parallelogram
  // cells carry the metric inside the contravariant fluxes and the face scales q1, q2 instead;
  // point metrics carry it inside the fluxes and scale each point's face terms.
  constexpr bool AFF = GEO == GEO_AFFINE;
  constexpr bool PT = GEO == GEO_POINT;
  const double bsc = EPI == EPI_STAGE ? sa.b : 1.0;
  const double sxs = (AFF || PT) ? -bsc : -bsc * geo.iJ11;
  const double sys = (AFF || PT) ? -bsc : -bsc * geo.iJ22;
  const double ua = EPI == EPI_STAGE ? sa.a1 / sxs : 0.0;

  auto SU = [&](int k, int l, int m) -> double & { return sU[((k * NSP + l) * 4 + m) * TE + lane]; };
  // x-face slot s of row y holds face s + 1 (between elements s and s + 1 of the row): first the
  // right trace of element s, then the face flux
  auto SX = [&](int m, int y, int s) -> double & { return sX[((m * NSP + line) * TY + y) * TX + s]; };
  double *const sYl = sU + line * (NSP * 4 * TE);
  static_assert(4 * (TY + 1) * TX <= 4 * 2 * TE, "the y faces of a line fit in its column block");
  // y-face slot fy of column x: face between elements fy - 1 and fy of the column
  auto SY = [&](int m, int fy, int x) -> double & { return sYl[(m * (TY + 1) + fy) * TX + x]; };
  // value of plane (k, l, m) at the element with byte offset bo inside the plane
  const unsigned ldp8 = (unsigned)(ldp * 8);
  auto LD = [&](const double *base, int k, int l, int m, unsigned bo) -> double {
    const int plane = k + NSP * (l + NSP * m);
    if constexpr (ADDR == ADDR_BUFFER)
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(plane_rsrc(base, 0xFFFFFFFFu), bo,
                                                                             (int)(ldp8 * (unsigned)plane), 0));
    else
      return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base + ldp * plane) + bo);
  };
  // stage results go to the padded arrays (aligned row segments) and are read by the next launch
  // only: stored through L2 (kStoreAgent); residuals go to caller arrays in the reference layout
  constexpr int kStAux = EPI == EPI_RESIDUAL ? 0 : kStoreAgent;
  auto ST = [&](double *base, int k, int l, int m, unsigned bo, double v) {
    const int plane = k + NSP * (l + NSP * m);
    if constexpr (ADDR == ADDR_BUFFER)
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(fr_u2, v), plane_rsrc(base, 0xFFFFFFFFu), bo,
                                            (int)(ldp8 * (unsigned)plane), kStAux);
    else
      *reinterpret_cast<double *>(reinterpret_cast<char *>(base + ldp * plane) + bo) = v;
  };

  // the same with a per-lane plane pl (of the first variable) and a uniform variable index m
  auto LDL = [&](const double *base, int pl, int m, unsigned bo) -> double {
    if constexpr (ADDR == ADDR_BUFFER)
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(plane_rsrc(base, 0xFFFFFFFFu),
                                                                             bo + ldp8 * (unsigned)pl,
                                                                             (int)(ldp8 * (unsigned)(NSP * NSP * m)), 0));
    else
      return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base + ldp * (pl + NSP * NSP * m)) + bo);
  };
  auto STL = [&](double *base, int pl, int m, unsigned bo, double v) {
    if constexpr (ADDR == ADDR_BUFFER)
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(fr_u2, v), plane_rsrc(base, 0xFFFFFFFFu),
                                            bo + ldp8 * (unsigned)pl, (int)(ldp8 * (unsigned)(NSP * NSP * m)), kStAux);
    else
      *reinterpret_cast<double *>(reinterpret_cast<char *>(base + ldp * (pl + NSP * NSP * m)) + bo) = v;
  };

  if (EPI != EPI_RESIDUAL && geo.stag_cu) {
    // staggered start (single-round grids): the workgroup's slot on its CU times the delay
    const int n = (int)(blockIdx.x / (unsigned)geo.stag_cu) * geo.stag;
    for (int q = 0; q < n; ++q) __builtin_amdgcn_s_sleep(8);
  }
  int bx, by;
  strip_tile(xcd_tile(blockIdx.x, gridDim.x), geo.ntx, geo.nty, geo.strip, bx, by);
  const int i0 = 1 + bx * TX, j0 = 1 + by * TY;
  const int ncx = min(TX, geo.nx - i0 + 1), ncy = min(TY, geo.ny - j0 + 1);
  const bool act = ex < ncx && ey < ncy;
  // byte offset of the own element inside a plane (pitch (ny+2) < 2^29 on the host)
  const unsigned ebo = 8u * ((unsigned)(i0 + ex) + nxg * (unsigned)(j0 + ey));

  if constexpr (EPI == EPI_RESIDUAL) {
    if (blockIdx.x == 0) {  // du vanishes on the four ghost corners (example/shock-vortex.jl:27)
      constexpr int per = NSP * NSP * 4;
      for (int v = tid; v < 4 * per; v += NSP * 64) {
        const int c = v / per, o = v % per;
        const long long ci = (c & 1) ? geo.nx + 1 : 0, cj = (c & 2) ? geo.ny + 1 : 0;
        sa.out[ci + (long long)nxg * cj + ldp * o] = 0.0;
      }
    }
  } else {
    if (blockIdx.x == 0 && sa.gc0) {  // the four ghost corners, which no halo task loads
      constexpr int per = NSP * NSP * 4;
      for (int v = tid; v < 4 * per; v += NSP * 64) {
        const int c = v / per, o = v % per;
        const long long ci = (c & 1) ? geo.nx + 1 : 0, cj = (c & 2) ? geo.ny + 1 : 0;
        const long long at = ci + (long long)nxg * cj + ldp * o;
        sa.gc0[at] = u[at];
        sa.gc1[at] = u[at];
      }
    }
  }

  // ---------------------------------------------------------------- P0: own row + halo lines
  double ur[NSP][4];
  if (act) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) ur[k][m] = LD(u, k, line, m, ebo);
  } else {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) ur[k][m] = (m == 0 || m == 3) ? 1.0 : 0.0;
  }

  // Halo tasks of this wavefront's line.  The left task (right trace of the left neighbour, row
  // `line`) runs on the lane of the tile's first element of that row, which solves that face
  // itself; the right task (left trace of the right neighbour) on the last lane of the row; the
  // bottom tasks (top trace of the element below, column `line`) and top tasks (bottom trace of
  // the element above) on the lanes between.
  int hkind = -1, hpos = 0;  // 0 bottom, 1 top, 2 left, 3 right
  if (ex == 0) {
    hkind = 2;
    hpos = ey;
  } else if (ex == TX - 1) {
    hkind = 3;
    hpos = ey;
  } else {
    const int t = ey * (TX - 2) + ex - 1;
    if (t < TX) {
      hkind = 0;
      hpos = t;
    } else if (t < 2 * TX) {
      hkind = 1;
      hpos = t - TX;
    }
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
    // point q of the task's line: (k = line, l = q) for column tasks, (k = q, l = line) for row
    // tasks, i.e. plane pb + ps*q + NSP*NSP*m; one set of loads serves both kinds
    const bool col = hkind <= 1;
    const int pb = col ? line : NSP * line, ps = col ? NSP : 1;
    double w[NSP];
#pragma unroll
    for (int q = 0; q < NSP; ++q) w[q] = lo ? ops.lr[q] : ops.ll[q];
    // the ghost element's line is copied to gc0, gc1 when asked (then hi, hj = si, sj)
    const bool ghost = hi == 0 || hi == geo.nx + 1 || hj == 0 || hj == geo.ny + 1;
    const bool gcopy = EPI != EPI_RESIDUAL && ghost && sa.gc0 != nullptr;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; ++q) {
        const double v = LDL(u, pb + ps * q, m, hbo);
        s += v * w[q];
        if (gcopy) {
          STL(sa.gc0, pb + ps * q, m, hbo, v);
          STL(sa.gc1, pb + ps * q, m, hbo, v);
        }
      }
      htr[m] = s;
    }
    if constexpr (EPI == EPI_RESIDUAL) {
      // du vanishes on the ghost layer (example/shock-vortex.jl:27)
      if (ghost) {
        const unsigned gbo = 8u * ((unsigned)hi + nxg * (unsigned)hj);
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int q = 0; q < NSP; ++q) STL(sa.out, pb + ps * q, m, gbo, 0.0);
      }
    }
  }

  // x sweep of the row: traces (right traces into the line's x-face slots), LDS transpose copy,
  // x flux and its derivative with the flux-trace parts of the corrections folded in (dc); the
  // stage seeds it with (a1 / sxs)*u, so that sxs times it carries a1*u and the row dies here
  double uLx[4], dxr[NSP][4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      a += ur[k][m] * ops.ll[k];
      b += ur[k][m] * ops.lr[k];
    }
    uLx[m] = a;
    if (act) SX(m, ey, ex) = b;
  }
  if (act) {
#pragma unroll
    for (int k = 0; k < NSP; ++k)
#pragma unroll
      for (int m = 0; m < 4; ++m) SU(k, line, m) = ur[k][m];
  }
  // point metrics: r_x of the row's points, which also scales their x-face terms
  double jx[NSP];
  {
    double f1[NSP][4];
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
      XState xs;
      xstate(ur[k], gm1, xs);
      if constexpr (AFF || PT) {  // contravariant flux iJ11 F + iJ12 G
        const double v = ur[k][2] * xs.rinv;
        const double G[4] = {ur[k][2], ur[k][1] * v, ur[k][2] * v + xs.p, (ur[k][3] + xs.p) * v};
        double a = geo.iJ11, b = geo.iJ12;
        if constexpr (PT) {
          a = act ? LD(geo.mj, k, line, 0, ebo) : 0.0;
          b = act ? LD(geo.mj, k, line, 2, ebo) : 0.0;
          jx[k] = a;
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) f1[k][m] = a * xs.F[m] + b * G[m];
      } else {
#pragma unroll
        for (int m = 0; m < 4; ++m) f1[k][m] = xs.F[m];
      }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        double d = 0.0;
#pragma unroll
        for (int q = 0; q < NSP; ++q) d += f1[q][m] * ops.dc[k + NSP * q];
        dxr[k][m] = EPI == EPI_STAGE ? ua * ur[k][m] + d : d;
      }
  }
  // RK4 form: stage operand a1*u of the own row
  double pre[NSP][4];
  if constexpr (EPI == EPI_ACC) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) pre[k][m] = sa.a1 * ur[k][m];
  }
  wave_lds_sync();  // the right traces of this line are visible to its lanes

  // own left x-face: the left neighbour's right trace (the left halo on the row's first lane);
  // the flux replaces the trace in its slot, where the left neighbour reads it
  double cl[4] = {0.0, 0.0, 0.0, 0.0};
  if (act) {
    double wl[4];
    const int s = ex > 0 ? ex - 1 : 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) wl[m] = ex > 0 ? SX(m, ey, s) : htr[m];
    if constexpr (AFF) {
      riemann_n4<FLUX>(wl, uLx, geo.n1c, geo.n1s, g, gm1, cl);
#pragma unroll
      for (int m = 0; m < 4; ++m) cl[m] *= geo.q1;
    } else {
      riemann_x4<FLUX>(wl, uLx, g, gm1, cl);
    }
    if (ex > 0) {
#pragma unroll
      for (int m = 0; m < 4; ++m) SX(m, ey, s) = cl[m];
    }
  }
  wave_lds_sync();
  // x corrections (the tile's last element of a row takes its right face after the halo pass);
  // the x part of the result: -iJ11*x (residual, RK4 form) or a1*u - b*iJ11*x (stage)
  double res[NSP][4];
  {
    const bool has_r = ex < ncx - 1;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double cr = has_r ? SX(m, ey, ex) : 0.0;
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        if constexpr (PT)
          res[k][m] = sxs * (dxr[k][m] + jx[k] * (cl[m] * ops.dhl[k] + cr * ops.dhr[k]));
        else
          res[k][m] = sxs * (dxr[k][m] + cl[m] * ops.dhl[k] + cr * ops.dhr[k]);
        fr_pin(res[k][m]);
      }
    }
  }
  __syncthreads();  // B1: the rows of every line are in sU

  // ---------------------------------------------------------------- P1: column sweep
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
  wave_lds_sync();  // the column block is read: it now carries this line's y faces
  double uBy[4], dyr[NSP][4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int q = 0; q < NSP; ++q) {
      a += uc[q][m] * ops.ll[q];
      b += uc[q][m] * ops.lr[q];
    }
    uBy[m] = a;
    if (act) SY(m, ey + 1, ex) = b;  // top trace, at the y-slot of the face it borders
  }
  if (hkind == 0) {
#pragma unroll
    for (int m = 0; m < 4; ++m) SY(m, 0, hpos) = htr[m];
  }
  // the y part of the result, scaled like the x part; point metrics: s_y of the column's points
  double jy[NSP];
  {
    double f2[NSP][4];
#pragma unroll
    for (int q = 0; q < NSP; ++q) {
      const double r = fr_rcp(uc[q][0]);
      const double vv = uc[q][2] * r;
      const double p = gm1 * (uc[q][3] - 0.5 * (uc[q][1] * (uc[q][1] * r) + uc[q][2] * vv));
      f2[q][0] = uc[q][2];
      f2[q][1] = uc[q][1] * vv;
      f2[q][2] = uc[q][2] * vv + p;
      f2[q][3] = (uc[q][3] + p) * vv;
      if constexpr (AFF || PT) {  // contravariant flux iJ21 F + iJ22 G
        const double uu = uc[q][1] * r;
        const double F[4] = {uc[q][1], uc[q][1] * uu + p, uc[q][2] * uu, (uc[q][3] + p) * uu};
        double a = geo.iJ21, b = geo.iJ22;
        if constexpr (PT) {  // point (k, l) = (line, q) of the own element
          a = act ? LD(geo.mj, line, q, 1, ebo) : 0.0;
          b = act ? LD(geo.mj, line, q, 3, ebo) : 0.0;
          jy[q] = b;
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) f2[q][m] = a * F[m] + b * f2[q][m];
      }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        double d = 0.0;
#pragma unroll
        for (int q = 0; q < NSP; ++q) d += f2[q][m] * ops.dc[k + NSP * q];
        dyr[k][m] = sys * d;
        fr_pin(dyr[k][m]);
      }
  }
  wave_lds_sync();  // this line's y traces are visible

  // ---------------------------------------------------------------- P2: y faces, tile edges
  // own bottom y-face: the slot holds the lower neighbour's top trace (or the bottom halo) and is
  // replaced by the flux
  if (act) {
    double wb[4], fw[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) wb[m] = SY(m, ey, ex);
    if constexpr (AFF) {
      riemann_n4<FLUX>(wb, uBy, geo.n2c, geo.n2s, g, gm1, fw);
#pragma unroll
      for (int m = 0; m < 4; ++m) fw[m] *= geo.q2;
    } else {
      riemann_y4<FLUX>(wb, uBy, g, gm1, fw);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) SY(m, ey, ex) = fw[m];
  }
  // One divergent pass for the tile's right x-faces (right tasks) and top y-faces (top tasks): a
  // y-face is solved in its normal frame n = (0, 1), i.e. on (rho, m_y, -m_x, E) with the flux
  // rotated back as (f0, -f2, f1, f3).
  if (hkind == 1 || hkind == 3) {
    const bool xf = hkind == 3;
    double *slot[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) slot[m] = xf ? &SX(m, hpos, ncx - 1) : &SY(m, ncy, hpos);
    double w[4], fw[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) w[m] = *slot[m];
    if constexpr (AFF) {
      const double c = xf ? geo.n1c : geo.n2c, sn = xf ? geo.n1s : geo.n2s, q = xf ? geo.q1 : geo.q2;
      riemann_n4<FLUX>(w, htr, c, sn, g, gm1, fw);
#pragma unroll
      for (int m = 0; m < 4; ++m) *slot[m] = q * fw[m];
    } else {
      double a[4], b[4];
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
  }
  wave_lds_sync();

  // ---------------------------------------------------------------- P3: corrections
  if (act) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double cb = sys * SY(m, ey, ex), ct = sys * SY(m, ey + 1, ex);
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        if constexpr (PT)
          dyr[k][m] += jy[k] * (cb * ops.dhl[k] + ct * ops.dhr[k]);
        else
          dyr[k][m] += cb * ops.dhl[k] + ct * ops.dhr[k];
      }
    }
    if (ex == ncx - 1) {  // right face of the row's last element, solved by the right task
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const double cr = sxs * SX(m, ey, ex);
#pragma unroll
        for (int k = 0; k < NSP; ++k) res[k][m] += (PT ? cr * jx[k] : cr) * ops.dhr[k];
      }
    }
  }
  wave_lds_sync();  // the y faces are read: the column block takes the y contributions
  // each thread writes the entries of its own column that it read in P1, so no other wavefront
  // is touched before the barrier
  if (act) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) SU(line, k, m) = dyr[k][m];
  }
  // the u0 row is read here, when the live set is smallest (4 workgroups per CU at nsp = 4), and
  // its loads are in flight during the wait at the barrier
  const bool with_u0 = EPI != EPI_RESIDUAL && sa.a0 != 0.0;
  double u0r[NSP][4];
  if (EPI != EPI_RESIDUAL && act && with_u0) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) u0r[k][m] = LD(sa.u0, k, line, m, ebo);
  }
  __syncthreads();  // B2: y contributions of every point visible

  // ---------------------------------------------------------------- P4: assemble + epilogue
  if (act) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < NSP; ++k) {
        const double v = SU(k, line, m) + res[k][m];
        if constexpr (EPI == EPI_ACC) {  // v = du; out = a0*u0 + a1*u + b*du, acc += b2*du
          double o = pre[k][m] + sa.b * v;
          if (with_u0) o = sa.a0 * u0r[k][m] + o;
          ST(sa.out, k, line, m, ebo, o);
          if (sa.acc_out) ST(sa.acc_out, k, line, m, ebo, LD(sa.acc_in, k, line, m, ebo) + sa.b2 * v);
        } else if constexpr (EPI == EPI_STAGE) {
          ST(sa.out, k, line, m, ebo, with_u0 ? sa.a0 * u0r[k][m] + v : v);
        } else {
          ST(sa.out, k, line, m, ebo, v);
        }
      }
  }
}

// ------------------------------------------------------------------ ghost-layer helpers
// rule FR_GHOST_PERIODIC / _OUTFLOW / _INFLOW_LEFT / _WAVE / _WAVE_Y over every ghost cell and every
// (k,l,m) plane (planes = 4 nsp^2, variable m = plane / nsp^2).
__global__ void fr2d_fill_ghost_kernel(double *u, int nx, int ny, int pitch, long long ldp, int planes, int rule) {
  const int nring = 2 * (nx + 2) + 2 * ny;
  const long long total = (long long)nring * planes;
  const long long nxg = pitch;
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
    bool neg = false;
    if (rule == FR_GHOST_PERIODIC || rule == FR_GHOST_WAVE || rule == FR_GHOST_WAVE_Y) {
      // euler2d_wave.jl:127-134 fills the x ghosts first and then the whole y ghost rows from the
      // filled columns (:153-160 the other way round), so a corner takes the diagonally opposite
      // interior cell; the wave rules negate the y-momentum of the y ghost rows (along x) or the
      // x-momentum of the x ghost columns (along y)
      si = (i == 0) ? nx : (i == nx + 1 ? 1 : i);
      sj = (j == 0) ? ny : (j == ny + 1 ? 1 : j);
      const long long var = pl / (planes / 4);
      neg = (rule == FR_GHOST_WAVE && (j == 0 || j == ny + 1) && var == 2) ||
            (rule == FR_GHOST_WAVE_Y && (i == 0 || i == nx + 1) && var == 1);
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
    const double v = u[(long long)si + nxg * sj + ldp * pl];
    u[(long long)i + nxg * j + ldp * pl] = neg ? -v : v;
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
  // strips of 16 tiles on grids of 32 tiles (512 elements) across and wider, measured 1-9 %
  // faster there at nsp = 4 (padded rows); narrower grids keep row-major order
  geo.strip = geo.ntx >= 32 ? 16 : geo.ntx;
  // the drop-in residual reads and writes caller arrays in the reference layout; the stages work
  // on the context's padded arrays
  const bool caller = EPI == EPI_RESIDUAL;
  geo.pitch = caller ? c.nx + 2 : c.pitch;
  geo.ldp = caller ? (long long)(c.nx + 2) * (c.ny + 2) : c.ldpi;
  geo.iJ11 = c.iJ11;
  geo.iJ22 = c.iJ22;
  geo.g = c.gamma;
  geo.gm1 = c.gamma - 1.0;
  geo.periodic = c.bc == FR_BC_PERIODIC;
  geo.iJ12 = c.iJ12;
  geo.iJ21 = c.iJ21;
  geo.n1c = c.n1[0];
  geo.n1s = c.n1[1];
  geo.n2c = c.n2[0];
  geo.n2s = c.n2[1];
  geo.q1 = c.iJ11 * c.n1[0] + c.iJ12 * c.n1[1];
  geo.q2 = c.iJ21 * c.n2[0] + c.iJ22 * c.n2[1];
  // the metric copy in the layout of this launch's arrays (the padded copy equals the
  // reference-layout one when the padded pitch is nx + 2)
  geo.mj = geo.pitch == c.nx + 2 ? c.d_mj_ref : c.d_mj_pad;
  if (c.geom == 2 && !geo.mj) return fail(FR_E_INVALID, "fr2d: point metrics missing on the device");
  const unsigned nb = (unsigned)(geo.ntx * geo.nty);  // one tile per workgroup
  const bool stag = !caller && c.stagger > 0 && fr2d_single_round(c);
  geo.stag_cu = stag ? c.num_cu : 0;
  geo.stag = stag ? c.stagger : 0;
  const bool buf = (caller ? c.len : c.lenp) * sizeof(double) < (size_t)0xFFFFFFFFu;
  if (c.geom == 0 && buf)
    hipLaunchKernelGGL((fr2d_kernel<NSP, FLUX, EPI, ADDR_BUFFER, GEO_RECT>), dim3(nb), dim3(NSP * 64), 0, st, geo, sa);
  else if (c.geom == 0)
    hipLaunchKernelGGL((fr2d_kernel<NSP, FLUX, EPI, ADDR_FLAT, GEO_RECT>), dim3(nb), dim3(NSP * 64), 0, st, geo, sa);
  else if (c.geom == 1 && buf)
    hipLaunchKernelGGL((fr2d_kernel<NSP, FLUX, EPI, ADDR_BUFFER, GEO_AFFINE>), dim3(nb), dim3(NSP * 64), 0, st, geo, sa);
  else if (c.geom == 1)
    hipLaunchKernelGGL((fr2d_kernel<NSP, FLUX, EPI, ADDR_FLAT, GEO_AFFINE>), dim3(nb), dim3(NSP * 64), 0, st, geo, sa);
  else if (buf)
    hipLaunchKernelGGL((fr2d_kernel<NSP, FLUX, EPI, ADDR_BUFFER, GEO_POINT>), dim3(nb), dim3(NSP * 64), 0, st, geo, sa);
  else
    hipLaunchKernelGGL((fr2d_kernel<NSP, FLUX, EPI, ADDR_FLAT, GEO_POINT>), dim3(nb), dim3(NSP * 64), 0, st, geo, sa);
  FR_LAUNCHED("fr2d_kernel");
  return FR_OK;
}

template <int NSP>
static int launch_flux(const Fr2dCtx &c, const StageArgs &sa, bool residual, hipStream_t st) {
  switch (c.flux) {
    case FR_FLUX_HLL:
      return residual ? launch_nsp<NSP, FR_FLUX_HLL, EPI_RESIDUAL>(c, sa, st)
             : (sa.acc_out || sa.b == 0.0) ? launch_nsp<NSP, FR_FLUX_HLL, EPI_ACC>(c, sa, st)
                          : launch_nsp<NSP, FR_FLUX_HLL, EPI_STAGE>(c, sa, st);
    case FR_FLUX_RUSANOV:
      return residual ? launch_nsp<NSP, FR_FLUX_RUSANOV, EPI_RESIDUAL>(c, sa, st)
             : (sa.acc_out || sa.b == 0.0) ? launch_nsp<NSP, FR_FLUX_RUSANOV, EPI_ACC>(c, sa, st)
                          : launch_nsp<NSP, FR_FLUX_RUSANOV, EPI_STAGE>(c, sa, st);
    default:
      return residual ? launch_nsp<NSP, FR_FLUX_ROE, EPI_RESIDUAL>(c, sa, st)
             : (sa.acc_out || sa.b == 0.0) ? launch_nsp<NSP, FR_FLUX_ROE, EPI_ACC>(c, sa, st)
                          : launch_nsp<NSP, FR_FLUX_ROE, EPI_STAGE>(c, sa, st);
  }
}

static int launch(const Fr2dCtx &c, const StageArgs &sa, bool residual, hipStream_t st) {
  switch (c.nsp) {
    case 2: return launch_flux<2>(c, sa, residual, st);
    case 3: return launch_flux<3>(c, sa, residual, st);
    case 4: return launch_flux<4>(c, sa, residual, st);
    case 5: return launch_flux<5>(c, sa, residual, st);
    case 6: return launch_flux<6>(c, sa, residual, st);
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
                      double a1, double b, const double *acc_in, double *acc_out, double b2, hipStream_t st,
                      double *ghost_copy0, double *ghost_copy1) {
  StageArgs sa{};
  if (c.bc != FR_BC_PERIODIC) {  // periodic grids never read the ghost layer
    sa.gc0 = ghost_copy0;
    sa.gc1 = ghost_copy1 ? ghost_copy1 : ghost_copy0;
  }
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

bool fr2d_single_round(const Fr2dCtx &c) {
  // four workgroups per CU for the nsp = 4 stage instances (Rusanov, HLL, cells without point
  // metrics): more tiles than one workgroup per CU and no more than all four slots
  const long long nb = (long long)((c.nx + TX - 1) / TX) * ((c.ny + TY - 1) / TY);
  return c.nsp == 4 && c.flux != FR_FLUX_ROE && c.geom != 2 && c.num_cu > 0 && nb > c.num_cu &&
         nb <= 4LL * c.num_cu;
}

// Times the stage launches of SSP-RK3 steps (state and second stage buffer -> first stage buffer,
// both buffers scratch here) for each stagger candidate, interleaved over four passes after a
// warm-up, and keeps the fastest.  The state is only read; st must not be capturing (the timing
// synchronises on an event).
int fr2d_tune_stagger(Fr2dCtx &c, hipStream_t st) {
  if (!fr2d_single_round(c)) {
    c.stagger = 0;
    return FR_OK;
  }
  static const int cand[] = {0, 3, 4, 5, 6, 7};  // units of s_sleep(8): 0 to 3584 clocks per slot
  constexpr int NC = sizeof(cand) / sizeof(cand[0]), LAUNCH = 18, PASSES = 4;
  float best[NC];
  for (int k = 0; k < NC; ++k) best[k] = 1e30f;
  hipEvent_t e0, e1;
  FR_HIP(hipEventCreate(&e0));
  FR_HIP(hipEventCreate(&e1));
  const int keep = c.stagger;
  int rc = FR_OK;
  // the launch mix of an SSP-RK3 step: one stage reading the state (16 B/DOF), two reading the
  // state and a second array (24 B/DOF); results go to the first stage buffer
  auto run = [&](int n) {
    for (int q = 0; q < n && rc == FR_OK; ++q)
      rc = (q % 3 == 0) ? fr2d_launch_stage(c, c.d_u, c.d_u, c.d_s1, 0.0, 1.0, 1e-9, nullptr, nullptr, 0.0, st)
                        : fr2d_launch_stage(c, c.d_s2, c.d_u, c.d_s1, 0.75, 0.25, 1e-9, nullptr, nullptr, 0.0, st);
  };
  c.stagger = 0;
  run(1500);  // the clocks settle under sustained load (about 27 ms at 256^2)
  for (int pass = 0; pass < PASSES && rc == FR_OK; ++pass)
    for (int k = 0; k < NC && rc == FR_OK; ++k) {
      c.stagger = cand[k];
      hipError_t e = hipEventRecord(e0, st);
      run(LAUNCH);
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = hip_fail(e, "fr2d_tune_stagger", __FILE__, __LINE__);
        break;
      }
      best[k] = std::min(best[k], ms);
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    c.stagger = keep;
    return rc;
  }
  int bi = 0;
  for (int k = 1; k < NC; ++k)
    if (best[k] < best[bi]) bi = k;
  c.stagger = cand[bi];
  return FR_OK;
}

int fr2d_launch_fill_ghost(const Fr2dCtx &c, double *d_u, int rule, hipStream_t st) {
  const int planes = c.nsp * c.nsp * 4;
  const long long total = (long long)(2 * (c.nx + 2) + 2 * c.ny) * planes;
  const int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(fr2d_fill_ghost_kernel, dim3(blocks), dim3(256), 0, st, d_u, c.nx, c.ny, c.pitch, c.ldpi, planes,
                     rule);
  FR_LAUNCHED("fr2d_fill_ghost_kernel");
  return FR_OK;
}

}  // namespace fr
