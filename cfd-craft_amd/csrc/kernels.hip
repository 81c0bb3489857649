// CDNA4 (gfx950) kernels of the pressure solve: the masked 7-point operator,
// fused PCG vector updates with deterministic wave64 reductions, segmented
// gauge means, IC(0) hyperplane sweeps, the multigrid V-cycles and the
// projection steps either side of the solve.
//
// Built with -ffp-contract=off: the FP64 operator and IC(0) sweeps repeat the
// reference's per-cell arithmetic (pressure.rs:223-241, 273-285) operation for
// operation, so their results are bitwise equal to the serial CPU code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"

namespace cfdgpu {

thread_local LaunchProbe* t_probe = nullptr;

constexpr int BS = 256;
constexpr int RB = kReduceBlock;  // threads of a reduction block
constexpr int RP = kReduceGrid;   // blocks of a reduction grid
static_assert(kReduceGrid % 8 == 0, "the reduction grid deals whole rounds of blocks to the 8 XCDs");
constexpr uint8_t FLUID = 1u << 6;
constexpr uint8_t CLOSED = 1u << 7;

// ---------------------------------------------------------------- reductions
// Device-wide sums without fences: a reduction kernel runs on a fixed grid of
// RP blocks of RB threads and writes one partial per block; the kernel that
// needs the total re-reduces the RP partials in a fixed order (the kernel
// boundary makes them visible). Results are deterministic for a given grid.
// (An agent-scope fence per block would write back and invalidate the L2 of
// every XCD once per wave.)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// Sum over a block of RB threads, returned to every thread (lane 0 of each wave,
// then lane 0 of the first wave: fixed association order).
__device__ double block_sum(double v) {
  __shared__ double s[RB / 64];
  __shared__ double total;
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x < 64) {
    double t = threadIdx.x < RB / 64 ? s[threadIdx.x] : 0.0;
    t = wave_sum(t);
    if (threadIdx.x == 0) total = t;
  }
  __syncthreads();
  const double r = total;
  __syncthreads();
  return r;
}
// Total of n <= RB partials, in every thread.
__device__ double reduce_part(const double* part, int n, int stride = 1) {
  double v = (int)threadIdx.x < n ? part[threadIdx.x * stride] : 0.0;
  return block_sum(v);
}
// The block's view of the solve status, read once so that every thread agrees.
__device__ bool block_running(const PcgState* st) {
  __shared__ int status;
  if (threadIdx.x == 0) status = st->status;
  __syncthreads();
  const bool run = status == ST_RUNNING;
  __syncthreads();
  return run;
}

// ---------------------------------------------------------------- geometry helpers
// Cell coordinates from a linear index; grids hold at most 1e8 cells, so 32-bit
// unsigned division is exact (and far cheaper than 64-bit division).
__device__ __forceinline__ void coords(const Geo& g, long long c, int q[3]) {
  const unsigned u = (unsigned)c;
  const unsigned nx = (unsigned)g.n[0], ny = (unsigned)g.n[1];
  const unsigned r = u / nx;
  q[0] = (int)(u - r * nx);
  const unsigned k = r / ny;
  q[1] = (int)(r - k * ny);
  q[2] = (int)k;
}
__device__ __forceinline__ long long stride_of(const Geo& g, int a) { return a == 0 ? 1 : (a == 1 ? g.sy : g.sz); }
__device__ __forceinline__ long long lo_nb(const Geo& g, long long c, int a, int t) {
  long long s = stride_of(g, a);
  return t > 0 ? c - s : c + (long long)(g.n[a] - 1) * s;
}
__device__ __forceinline__ long long hi_nb(const Geo& g, long long c, int a, int t) {
  long long s = stride_of(g, a);
  return t < g.n[a] - 1 ? c + s : c - (long long)(g.n[a] - 1) * s;
}

// (A x)_c in the reference's face order: per axis the lower face, then the upper
// face, except on the last cell of a periodic axis whose wrapping upper face
// comes first in the face list. Interior: d = w (x_lo - x_hi), y_lo += d,
// y_hi -= d; outlet: y += w x (pressure.rs:273-285).
// Neighbour index of cell c on axis a: side 0 lower, 1 upper (wrapping on a
// periodic axis); c itself where there is no neighbour, so that a load from the
// returned index is always valid and can be issued before the coupling is known.
__device__ __forceinline__ long long nb_or_self(const Geo& g, long long c, int a, int t, int side) {
  const long long s = stride_of(g, a);
  const int n = g.n[a];
  if (side == 0) return t > 0 ? c - s : (g.per[a] ? c + (long long)(n - 1) * s : c);
  return t < n - 1 ? c + s : (g.per[a] ? c - (long long)(n - 1) * s : c);
}

// (A x)_c from the cell's value xc and its six neighbour values nv (axis-major,
// lower side first; entries of uncoupled sides are not read).
__device__ __forceinline__ double apply_vals(const Geo& g, uint8_t code, const int q[3], double xc, const double nv[6]) {
  double y = 0.0;
  if (!(code & 0x3f)) return y;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const bool lo = code & (1u << (2 * a));
    const bool hi = code & (2u << (2 * a));
    if (!lo && !hi) continue;
    const int t = q[a], n = g.n[a];
    const bool P = g.per[a];
    const bool wrap_first = P && t == n - 1;
    for (int k = 0; k < 2; ++k) {
      const bool do_hi = (k == 0) == wrap_first;
      if (do_hi) {
        if (!hi) continue;
        if (t == n - 1 && !P) {
          y += g.wo[a] * xc;
        } else {
          double d = g.wi[a] * (xc - nv[2 * a + 1]);
          y += d;
        }
      } else {
        if (!lo) continue;
        if (t == 0 && !P) {
          y += g.wo[a] * xc;
        } else {
          double d = g.wi[a] * (nv[2 * a] - xc);
          y -= d;
        }
      }
    }
  }
  return y;
}

template <class XV>
__device__ __forceinline__ double apply_cell(const Geo& g, uint8_t code, long long c, const int q[3], XV xv) {
  // every value first (independent loads), then the reference's operation order
  const double xc = xv(c);
  double nv[6];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    nv[2 * a] = xv(nb_or_self(g, c, a, q[a], 0));
    nv[2 * a + 1] = xv(nb_or_self(g, c, a, q[a], 1));
  }
  return apply_vals(g, code, q, xc, nv);
}

__device__ __forceinline__ double mean_of(uint8_t code, long long c, const double* mean, const int* cidx) {
  if (!(code & CLOSED)) return 0.0;
  return mean[cidx ? cidx[c] : 0];
}

// ---------------------------------------------------------------- operator
__global__ void __launch_bounds__(BS) k_apply(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ x,
                                              double* __restrict__ y) {
  long long c = (long long)blockIdx.x * BS + threadIdx.x;
  if (c >= g.N) return;
  int q[3];
  coords(g, c, q);
  y[c] = apply_cell(g, codes[c], c, q, [&](long long i) { return x[i]; });
}

// The share of `items` work units (in order) of the block's XCD: the blocks of
// a launch are dealt to the 8 XCDs in turn (block b runs on XCD b % 8), so XCD x
// takes the x-th contiguous eighth of the items and its blocks take them in turn
// (first, first + step, ...). Neighbouring rows and layers of a stencil are then
// read by the same XCD, from its own L2. The grid must be a multiple of 8 blocks.
// floor(n / d) from m = ceil(2^40 / d), exact for n d < 2^40 (n < 2^32): the
// error n (m - 2^40 / d) / 2^40 < n / 2^40 stays below 1 / d, less than the
// distance from n / d up to the next integer.
__device__ __forceinline__ unsigned div_magic(unsigned n, unsigned long long m) {
  return (unsigned)(((unsigned long long)n * m) >> 40);
}

struct XcdShare {
  unsigned begin, end, first, step;
};
// (bid, nblk: the block's index and the block count of its grid; a batched
// launch gives each grid its own range of blocks, starting at a multiple of 8)
__device__ __forceinline__ XcdShare xcd_share_at(unsigned bid, unsigned nblk, unsigned items, unsigned per_block,
                                                 unsigned rank_in_block) {
  const unsigned x = bid & 7, blocks = nblk >> 3;
  const unsigned chunk = (items + 7) >> 3;
  XcdShare s;
  s.begin = min(items, x * chunk);
  s.end = min(items, s.begin + chunk);
  s.first = s.begin + (bid >> 3) * per_block + rank_in_block;
  s.step = blocks * per_block;
  return s;
}
__device__ __forceinline__ XcdShare xcd_share(unsigned items, unsigned per_block, unsigned rank_in_block) {
  return xcd_share_at(blockIdx.x, gridDim.x, items, per_block, rank_in_block);
}

// Every cell of the grid, once, over a fixed reduction grid: each wave takes
// segments of 64 cells of one x row in turn, from its XCD's contiguous share;
// a segment's row is wave-uniform (scalar arithmetic), so no cell needs a
// division.
template <class F>
__device__ __forceinline__ void for_cells_at(const Geo& g, unsigned bid, unsigned nblk, F f) {
  const unsigned nx = (unsigned)g.n[0], ny = (unsigned)g.n[1];
  const unsigned sx = (nx + 63) >> 6, segs = sx * ny * (unsigned)g.n[2];
  const unsigned lane = threadIdx.x & 63;
  const XcdShare sh = xcd_share_at(bid, nblk, segs, blockDim.x >> 6, threadIdx.x >> 6);
  for (unsigned sg = __builtin_amdgcn_readfirstlane(sh.first); sg < sh.end; sg += sh.step) {
    const unsigned row = sg / sx;
    const unsigned i = (sg - row * sx) * 64 + lane;
    if (i >= nx) continue;
    const unsigned k = row / ny;
    const int q[3] = {(int)i, (int)(row - k * ny), (int)k};
    f((long long)i + (long long)nx * row, q);
  }
}
template <class F>
__device__ __forceinline__ void for_cells(const Geo& g, F f) {
  for_cells_at(g, blockIdx.x, gridDim.x, f);
}

// The cells of the operator kernel: p_new = z_c (FIRST) or z_c + beta p_old,
// Ap = A p_new, and the thread's part of p_new . Ap. FIRST is a template
// argument so that the seven p values of a cell are loaded without a branch
// between the loads (they are then all in flight at once).
// RES: z in the colour-separated layout of the fine red-black level (cells with
// x + y + z even first, then the others, each in natural order: cell c at c / 2
// of its colour's half, nx even), zhalf entries per colour. Every neighbour of
// a cell has the other colour (periodic axes have even lengths there), except
// the cell itself standing in at a closed boundary.
// TP: the type p is stored in (float in the mixed-precision solve: each p value
// is rounded once, and A p and p . Ap are formed from the rounded values, so x
// += alpha p and r -= alpha A p stay consistent).
template <bool HAS_CLOSED, bool FIRST, bool RES = false, class TZ = double, class TP = double>
__device__ __forceinline__ double pcg_op_cells(const Geo& g, const uint8_t* __restrict__ codes,
                                               const TZ* __restrict__ z, const TP* __restrict__ pold,
                                               TP* __restrict__ pnew, double* __restrict__ ap,
                                               const double* __restrict__ mean, const int* __restrict__ cidx,
                                               double beta, long long zhalf, unsigned bid, unsigned nblk) {
  auto pv = [&](long long i, int colour) {
    double zc = RES ? (double)z[(colour ? zhalf : 0) + (i >> 1)] : (double)z[i];
    // closed components: z centered (without any, the code load is skipped)
    if (HAS_CLOSED && (codes[i] & CLOSED)) zc -= mean[cidx ? cidx[i] : 0];
    if (FIRST) return (double)(TP)zc;
    return (double)(TP)(zc + beta * (double)pold[i]);
  };
  double acc = 0.0;
  for_cells_at(g, bid, nblk, [&](long long c, const int q[3]) {
    const int cc = (q[0] + q[1] + q[2]) & 1;
    auto pvc = [&](long long i) { return pv(i, i == c ? cc : cc ^ 1); };
    uint8_t code = codes[c];
    double pc = pvc(c);
    double y = apply_cell(g, code, c, q, pvc);
    pnew[c] = (TP)pc;
    ap[c] = y;
    acc += pc * y;
  });
  return acc;
}

// p_new = z_c + beta p (p_new = z_c on the first iteration) with beta = rz_next / rz,
// Ap = A p_new, and per-block partials of p_new . Ap (pressure.rs:391-393,426-432).
// (registers sized for 8 waves per SIMD: two blocks per CU)
template <bool HAS_CLOSED, bool RES, class TZ, class TP = double>
__device__ __forceinline__ void pcg_op_body(const Geo& g, const uint8_t* __restrict__ codes, const TZ* __restrict__ z,
                                            const TP* __restrict__ pold, TP* __restrict__ pnew,
                                            double* __restrict__ ap, const double* __restrict__ mean,
                                            const int* __restrict__ cidx, PcgState* st, double* part, long long zhalf,
                                            unsigned bid, unsigned nblk) {
  if (!block_running(st)) return;
  const bool first = st->first;
  double beta = 0.0;
  if (!first) beta = reduce_part(part + P_RZ, RP) / st->rz;
  const double acc =
      first ? pcg_op_cells<HAS_CLOSED, true, RES, TZ, TP>(g, codes, z, pold, pnew, ap, mean, cidx, beta, zhalf, bid,
                                                          nblk)
            : pcg_op_cells<HAS_CLOSED, false, RES, TZ, TP>(g, codes, z, pold, pnew, ap, mean, cidx, beta, zhalf, bid,
                                                           nblk);
  double s = block_sum(acc);
  if (threadIdx.x == 0) {
    part[P_PAP + bid] = s;
    if (bid == 0 && !first) st->beta = beta;
  }
}
template <bool HAS_CLOSED, bool RES, class TZ = double, class TP = double>
__global__ void __launch_bounds__(RB, 8)
    k_pcg_op(Geo g, const uint8_t* __restrict__ codes, const TZ* __restrict__ z, const TP* __restrict__ pold,
             TP* __restrict__ pnew, double* __restrict__ ap, const double* __restrict__ mean,
             const int* __restrict__ cidx, PcgState* st, double* part, long long zhalf) {
  pcg_op_body<HAS_CLOSED, RES, TZ, TP>(g, codes, z, pold, pnew, ap, mean, cidx, st, part, zhalf, blockIdx.x,
                                       gridDim.x);
}

// The same kernel for large grids, marching through z: work items are tiles of
// MX x MY columns (x, y) over MZ layers, dealt to the blocks of the fixed grid
// from each XCD's contiguous share (tiles of one layer range side by side). A
// thread keeps its column's p at layers k - 1, k and k + 1 in registers and
// writes layer k into an LDS tile with a one-cell rim (the x and y neighbours
// of the edge cells, formed again by the first threads), so each p value of the
// grid is formed once per item and every neighbour comes from registers or LDS.
// Every p value and every (A p)_c repeats pcg_op_cells' arithmetic bit for bit.
constexpr int MX = 64, MY = 16, MZ = 16;
static_assert(MX * MY == RB, "one thread per column of a tile");
struct MarchShape {
  int tx, ty, tz;  // tiles in x, y and layer ranges in z
  int mz;          // layers of a range (at most MZ)
};
// The shortest layer range tried (CFD_MARCH_MZ, 1 .. 16; 4: a 192x192x96 grid
// marches in ranges of 4 layers, which took the served mix of 64^3 .. 192x192x96
// grids from 1137 to 1167 solves/s against the cell-by-cell kernel; 2 gained
// nothing more, and the 128x128x64 urban block solves at the same rate either way).
inline int march_min_mz() {
  static const int v = [] {
    const char* e = std::getenv("CFD_MARCH_MZ");
    const int m = e ? std::atoi(e) : 4;
    return m >= 1 && m <= MZ ? m : MZ;
  }();
  return v;
}
// The fine restriction with two bins per thread where nx is a multiple of 4
// (CFD_RB_RESTRICT_PAIRS=0: one bin per thread).
inline bool rb_restrict_pairs() {
  static const bool v = [] {
    const char* e = std::getenv("CFD_RB_RESTRICT_PAIRS");
    return !(e && std::atoi(e) == 0);
  }();
  return v;
}
// Ranges of MZ layers, halved while the grid gives fewer items than the fixed
// grid has blocks (down to march_min_mz()).
inline MarchShape march_shape(const Geo& g) {
  const int tx = (g.n[0] + MX - 1) / MX, ty = (g.n[1] + MY - 1) / MY;
  int mz = MZ;
  while (mz / 2 >= march_min_mz() && (long long)tx * ty * ((g.n[2] + mz - 1) / mz) < RP) mz /= 2;
  return MarchShape{tx, ty, (g.n[2] + mz - 1) / mz, mz};
}

template <bool HAS_CLOSED, bool FIRST, bool RES, class TZ, class TP>
__device__ __forceinline__ double pcg_op_march(const Geo& g, MarchShape ms, const uint8_t* __restrict__ codes,
                                               const TZ* __restrict__ z, const TP* __restrict__ pold,
                                               TP* __restrict__ pnew, double* __restrict__ ap,
                                               const double* __restrict__ mean, const int* __restrict__ cidx,
                                               double beta, double (*tile)[MY + 2][MX + 2], long long zhalf,
                                               unsigned bid, unsigned nblk) {
  // colour: that of cell i (used with RES, z in the colour-separated layout)
  auto pv = [&](long long i, int colour) {
    double zc = RES ? (double)z[(colour ? zhalf : 0) + (i >> 1)] : (double)z[i];
    if (HAS_CLOSED && (codes[i] & CLOSED)) zc -= mean[cidx ? cidx[i] : 0];
    if (FIRST) return (double)(TP)zc;
    return (double)(TP)(zc + beta * (double)pold[i]);
  };
  const int lx = (int)threadIdx.x % MX, ly = (int)threadIdx.x / MX;
  const unsigned items = (unsigned)(ms.tx * ms.ty * ms.tz);
  const XcdShare sh = xcd_share_at(bid, nblk, items, 1, 0);
  double acc = 0.0;
  for (unsigned item = sh.first; item < sh.end; item += sh.step) {
    const int txi = (int)(item % (unsigned)ms.tx);
    const int tyi = (int)((item / (unsigned)ms.tx) % (unsigned)ms.ty);
    const int tzi = (int)(item / (unsigned)(ms.tx * ms.ty));
    const int ox = txi * MX, oy = tyi * MY;
    const int wx = min(MX, g.n[0] - ox), wy = min(MY, g.n[1] - oy);  // live columns of the tile
    const int i = ox + lx, j = oy + ly;
    const int k0 = tzi * ms.mz, k1 = min(g.n[2], k0 + ms.mz);
    const bool live = lx < wx && ly < wy;
    const long long col = (long long)i + (long long)j * g.sy;
    // The rim: the x and y neighbours (nb_or_self: wrapping, or the edge cell
    // itself where the axis ends) of the tile's edge cells, one slot each,
    // formed by thread r < 2 (wx + wy): the bottom and top rows, then the left
    // and right columns.
    const int r = (int)threadIdx.x;
    const bool rim = r < 2 * (wx + wy);
    int sx = 0, sy = 0, rpar = 0;
    long long rcol = 0;
    if (rim) {
      int ex, ey, axis, side;
      if (r < wx) {
        ex = r, ey = 0, axis = 1, side = 0, sx = r + 1, sy = 0;
      } else if (r < 2 * wx) {
        ex = r - wx, ey = wy - 1, axis = 1, side = 1, sx = r - wx + 1, sy = wy + 1;
      } else if (r < 2 * wx + wy) {
        ex = 0, ey = r - 2 * wx, axis = 0, side = 0, sx = 0, sy = r - 2 * wx + 1;
      } else {
        ex = wx - 1, ey = r - 2 * wx - wy, axis = 0, side = 1, sx = wx + 1, sy = r - 2 * wx - wy + 1;
      }
      const int eq[2] = {ox + ex, oy + ey};
      const int n = g.n[axis], t0 = eq[axis];
      const int tn = side == 0 ? (t0 > 0 ? t0 - 1 : (g.per[axis] ? n - 1 : t0))
                               : (t0 < n - 1 ? t0 + 1 : (g.per[axis] ? 0 : t0));
      const int nxy[2] = {axis == 0 ? tn : eq[0], axis == 1 ? tn : eq[1]};
      rcol = (long long)nxy[0] + (long long)nxy[1] * g.sy;
      rpar = (nxy[0] + nxy[1]) & 1;
    }
    const int cpar = (i + j) & 1;
    // software pipeline: the values of the next PD layers (the column's p, its
    // code and the rim slot's p) are in flight while the current layer is
    // worked on (one layer ahead: with eight waves per SIMD the other waves hide
    // the latency, and deeper pipelines spill registers: 185 against 191 and
    // 215 us a launch for 1, 2 and 3 layers on the 384x384x160 canyon). p of
    // layer L = k0 - 1 .. k1 is the column's z neighbour at the ends
    // (nb_or_self), the column's own value in between.
    constexpr int PD = 1;
    auto p_at = [&](int L) {
      if (L < k0) {
        const long long c0 = col + (long long)k0 * g.sz, n = nb_or_self(g, c0, 2, k0, 0);
        return pv(n, (cpar + k0 + (n == c0 ? 0 : 1)) & 1);
      }
      if (L >= k1) {
        const long long c1 = col + (long long)(k1 - 1) * g.sz, n = nb_or_self(g, c1, 2, k1 - 1, 1);
        return pv(n, (cpar + k1 - 1 + (n == c1 ? 0 : 1)) & 1);
      }
      return pv(col + (long long)L * g.sz, (cpar + L) & 1);
    };
    double pm = 0.0, pc = 0.0, rv = 0.0;
    double pq[PD] = {}, rq[PD] = {};
    uint8_t code = 0, cq[PD] = {};
    if (live) {
      pm = p_at(k0 - 1);
      pc = p_at(k0);
      code = codes[col + (long long)k0 * g.sz];
#pragma unroll
      for (int d = 0; d < PD; ++d) {
        if (k0 + 1 + d <= k1) pq[d] = p_at(k0 + 1 + d);
        if (k0 + 1 + d < k1) cq[d] = codes[col + (long long)(k0 + 1 + d) * g.sz];
      }
    }
    if (rim) {
      rv = pv(rcol + (long long)k0 * g.sz, (rpar + k0) & 1);
#pragma unroll
      for (int d = 0; d < PD; ++d)
        if (k0 + 1 + d < k1) rq[d] = pv(rcol + (long long)(k0 + 1 + d) * g.sz, (rpar + k0 + 1 + d) & 1);
    }
    for (int k = k0; k < k1; ++k) {
      const long long c = col + (long long)k * g.sz;
      const int q[3] = {i, j, k};
      // layer k + 1 + PD enters the pipeline
      const int L = k + 1 + PD;
      double pin = 0.0, rin = 0.0;
      uint8_t cin = 0;
      if (live && L <= k1) pin = p_at(L);
      if (live && L < k1) cin = codes[col + (long long)L * g.sz];
      if (rim && L < k1) rin = pv(rcol + (long long)L * g.sz, (rpar + L) & 1);
      const double pn = pq[0];
      double(*t)[MX + 2] = tile[k & 1];
      if (live) t[ly + 1][lx + 1] = pc;
      if (rim) t[sy][sx] = rv;
      __syncthreads();
      if (live) {
        double nv[6];
        nv[0] = t[ly + 1][lx];
        nv[1] = t[ly + 1][lx + 2];
        nv[2] = t[ly][lx + 1];
        nv[3] = t[ly + 2][lx + 1];
        nv[4] = pm;
        nv[5] = pn;
        const double y = apply_vals(g, code, q, pc, nv);
        pnew[c] = (TP)pc;
        ap[c] = y;
        acc += pc * y;
      }
      pm = pc;
      pc = pn;
      rv = rq[0];
      code = cq[0];
#pragma unroll
      for (int d = 0; d + 1 < PD; ++d) {
        pq[d] = pq[d + 1];
        rq[d] = rq[d + 1];
        cq[d] = cq[d + 1];
      }
      pq[PD - 1] = pin;
      rq[PD - 1] = rin;
      cq[PD - 1] = cin;
    }
    __syncthreads();
  }
  return acc;
}

template <bool HAS_CLOSED, bool RES, class TZ, class TP = double>
__device__ __forceinline__ void pcg_op_march_body(const Geo& g, MarchShape ms, const uint8_t* __restrict__ codes,
                                                  const TZ* __restrict__ z, const TP* __restrict__ pold,
                                                  TP* __restrict__ pnew, double* __restrict__ ap,
                                                  const double* __restrict__ mean, const int* __restrict__ cidx,
                                                  PcgState* st, double* part, long long zhalf, unsigned bid,
                                                  unsigned nblk) {
  __shared__ double tile[2][MY + 2][MX + 2];
  if (!block_running(st)) return;
  const bool first = st->first;
  double beta = 0.0;
  if (!first) beta = reduce_part(part + P_RZ, RP) / st->rz;
  const double acc =
      first ? pcg_op_march<HAS_CLOSED, true, RES, TZ, TP>(g, ms, codes, z, pold, pnew, ap, mean, cidx, beta, tile,
                                                           zhalf, bid, nblk)
            : pcg_op_march<HAS_CLOSED, false, RES, TZ, TP>(g, ms, codes, z, pold, pnew, ap, mean, cidx, beta, tile,
                                                            zhalf, bid, nblk);
  double s = block_sum(acc);
  if (threadIdx.x == 0) {
    part[P_PAP + bid] = s;
    if (bid == 0 && !first) st->beta = beta;
  }
}
template <bool HAS_CLOSED, bool RES, class TZ = double, class TP = double>
__global__ void __launch_bounds__(RB, 8)
    k_pcg_op_march(Geo g, MarchShape ms, const uint8_t* __restrict__ codes, const TZ* __restrict__ z,
                   const TP* __restrict__ pold, TP* __restrict__ pnew, double* __restrict__ ap,
                   const double* __restrict__ mean, const int* __restrict__ cidx, PcgState* st, double* part,
                   long long zhalf) {
  pcg_op_march_body<HAS_CLOSED, RES, TZ, TP>(g, ms, codes, z, pold, pnew, ap, mean, cidx, st, part, zhalf, blockIdx.x,
                                             gridDim.x);
}

// alpha = rz / (p . Ap) with CFD-E032 (pressure.rs:393-397); x += alpha p;
// r -= alpha Ap; partials of |r|^2 (pressure.rs:398-402).
template <class TP>
__device__ __forceinline__ void pcg_axpy_body(long long N, const TP* __restrict__ p,
                                              const double* __restrict__ ap, double* __restrict__ x,
                                              double* __restrict__ r, PcgState* st, double* part, unsigned bid,
                                              unsigned nblk) {
  if (!block_running(st)) return;
  const double rz = reduce_part(part + P_RZ, RP);
  const double denom = reduce_part(part + P_PAP, RP);
  if (denom <= 0.0 || !isfinite(denom)) {
    if (bid == 0 && threadIdx.x == 0) {
      st->denom = denom;
      st->status = ST_E032;
    }
    return;
  }
  const double alpha = rz / denom;
  double acc = 0.0;
  for (long long i = (long long)bid * RB + threadIdx.x; i < N; i += (long long)nblk * RB) {
    x[i] += alpha * (double)p[i];
    double ri = r[i] - alpha * ap[i];
    r[i] = ri;
    acc += ri * ri;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) {
    part[P_RR + bid] = s;
    if (bid == 0) {
      st->rz = rz;
      st->alpha = alpha;
      st->denom = denom;
      st->first = 0;
    }
  }
}
__global__ void __launch_bounds__(RB) k_pcg_axpy(long long N, const double* __restrict__ p, const double* __restrict__ ap,
                                                 double* __restrict__ x, double* __restrict__ r, PcgState* st,
                                                 double* part) {
  pcg_axpy_body<double>(N, p, ap, x, r, st, part, blockIdx.x, gridDim.x);
}

// The same with r in the colour-separated layout of the fine red-black level
// (cell c at c / 2 of the half of its colour, (x + y + z) & 1; nx even): the
// same cells per thread in the same order, so x, r and the partials of |r|^2
// are those of k_pcg_axpy bit for bit. mnx, mny: ceil(2^40 / nx), ceil(2^40 / ny)
// (N nx < 2^40).
template <class TP>
__device__ __forceinline__ void pcg_axpy_rb_body(long long N, int nx, int ny, unsigned long long mnx,
                                                 unsigned long long mny, const TP* __restrict__ p,
                                                 const double* __restrict__ ap, double* __restrict__ x,
                                                 double* __restrict__ r, PcgState* st, double* part, unsigned bid,
                                                 unsigned nblk) {
  if (!block_running(st)) return;
  const double rz = reduce_part(part + P_RZ, RP);
  const double denom = reduce_part(part + P_PAP, RP);
  if (denom <= 0.0 || !isfinite(denom)) {
    if (bid == 0 && threadIdx.x == 0) {
      st->denom = denom;
      st->status = ST_E032;
    }
    return;
  }
  const double alpha = rz / denom;
  const long long half = N >> 1;
  double acc = 0.0;
  for (long long i = (long long)bid * RB + threadIdx.x; i < N; i += (long long)nblk * RB) {
    const unsigned c = (unsigned)i;
    const unsigned row = div_magic(c, mnx), k = div_magic(row, mny);
    const unsigned colour = (c - row * (unsigned)nx + row - k * (unsigned)ny + k) & 1u;
    double* rc = r + (colour ? half : 0) + (c >> 1);
    x[i] += alpha * (double)p[i];
    double ri = *rc - alpha * ap[i];
    *rc = ri;
    acc += ri * ri;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) {
    part[P_RR + bid] = s;
    if (bid == 0) {
      st->rz = rz;
      st->alpha = alpha;
      st->denom = denom;
      st->first = 0;
    }
  }
}
template <class TP = double>
__global__ void __launch_bounds__(RB) k_pcg_axpy_rb(long long N, int nx, int ny, unsigned long long mnx,
                                                    unsigned long long mny, const TP* __restrict__ p,
                                                    const double* __restrict__ ap, double* __restrict__ x,
                                                    double* __restrict__ r, PcgState* st, double* part) {
  pcg_axpy_rb_body<TP>(N, nx, ny, mnx, mny, p, ap, x, r, st, part, blockIdx.x, gridDim.x);
}

// When the recurrent residual |r|/|b_c| reached tol: partials of |A x - b_c|^2,
// the independent recomputed residual (pressure.rs:403-414).
__device__ __forceinline__ void pcg_true_body(const Geo& g, const uint8_t* __restrict__ codes,
                                              const double* __restrict__ x, const double* __restrict__ bc, PcgState* st,
                                              double* part, unsigned bid, unsigned nblk) {
  if (!block_running(st)) return;
  const double rel = sqrt(reduce_part(part + P_RR, RP)) / st->norm;
  if (!(rel <= st->tol)) return;
  double acc = 0.0;
  for_cells_at(g, bid, nblk, [&](long long c, const int q[3]) {
    double y = apply_cell(g, codes[c], c, q, [&](long long i) { return x[i]; });
    double d = y - bc[c];
    acc += d * d;
  });
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[P_TR + bid] = s;
}
__global__ void __launch_bounds__(RB, 8) k_pcg_true(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ x,
                                                 const double* __restrict__ bc, PcgState* st, double* part) {
  pcg_true_body(g, codes, x, bc, st, part, blockIdx.x, gridDim.x);
}

// One block: the iteration's decisions. Acceptance needs the recurrent and the
// recomputed residual (<= 10 tol + 1e-12, pressure.rs:415); the limit raises
// CFD-E033 (pressure.rs:435).
__device__ __forceinline__ void pcg_fin_body(PcgState* st, const double* part) {
  if (!block_running(st)) return;
  const double rel = sqrt(reduce_part(part + P_RR, RP)) / st->norm;
  double tr = 0.0;
  const bool check = rel <= st->tol;
  if (check) tr = sqrt(reduce_part(part + P_TR, RP)) / st->norm;
  if (threadIdx.x == 0) {
    st->rel = rel;
    st->it += 1;
    if (check) {
      st->true_rel = tr;
      st->true_checks += 1;
    }
    if (check && tr <= st->accept)
      st->status = ST_CONVERGED;
    else if (st->it >= st->max_it)
      st->status = ST_E033;
  }
}
__global__ void __launch_bounds__(RB) k_pcg_fin(PcgState* st, const double* part) { pcg_fin_body(st, part); }

// Partials of r . z_c (pressure.rs:428); the next iteration's kernels reduce them.
__global__ void __launch_bounds__(RB) k_pcg_rz(long long N, const uint8_t* __restrict__ codes, const double* __restrict__ r,
                                               const double* __restrict__ z, const double* __restrict__ mean,
                                               const int* __restrict__ cidx, PcgState* st, double* part) {
  if (!block_running(st)) return;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < N; i += (long long)gridDim.x * RB) {
    uint8_t ci = codes[i];
    double zc = z[i];
    if (ci & CLOSED) zc -= mean[cidx ? cidx[i] : 0];
    acc += r[i] * zc;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[P_RZ + blockIdx.x] = s;
}

// ---------------------------------------------------------------- gauges
// Per closed component: sum of src (and of max(|src|, scale) for the CFD-E031
// compatibility test, pressure.rs:351-362); then mean = sum / members.
// One closed component: a grid-stride pass over the cells flagged closed.
__global__ void __launch_bounds__(RB) k_csum1(long long N, const uint8_t* __restrict__ codes, const double* __restrict__ src,
                                              const double* __restrict__ scale, int e031, const PcgState* st,
                                              double* part) {
  if (st && !block_running(st)) return;
  double acc = 0.0, acc2 = 0.0;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < N; i += (long long)gridDim.x * RB) {
    if (!(codes[i] & CLOSED)) continue;
    double v = src[i];
    acc += v;
    if (e031) acc2 += fmax(fabs(v), scale ? scale[i] : 0.0);
  }
  double s = block_sum(acc);
  double s2 = e031 ? block_sum(acc2) : 0.0;
  if (threadIdx.x == 0) {
    part[P_CS + 2 * blockIdx.x] = s;
    part[P_CS + 2 * blockIdx.x + 1] = s2;
  }
}

// Several closed components: one block per chunk of a component's member list.
__global__ void __launch_bounds__(RB) k_csumN(const int* __restrict__ members, const int* __restrict__ chunk_begin,
                                              const double* __restrict__ src, const double* __restrict__ scale, int e031,
                                              const PcgState* st, double* part) {
  if (st && !block_running(st)) return;
  const int ch = blockIdx.x;
  const int b0 = chunk_begin[ch], b1 = chunk_begin[ch + 1];
  double acc = 0.0, acc2 = 0.0;
  for (int k = b0 + threadIdx.x; k < b1; k += RB) {
    int i = members[k];
    double v = src[i];
    acc += v;
    if (e031) acc2 += fmax(fabs(v), scale ? scale[i] : 0.0);
  }
  double s = block_sum(acc);
  double s2 = e031 ? block_sum(acc2) : 0.0;
  if (threadIdx.x == 0) {
    part[P_CS + 2 * ch] = s;
    part[P_CS + 2 * ch + 1] = s2;
  }
}

// One block: component means from the partials, and the CFD-E031 test.
__global__ void __launch_bounds__(RB) k_csum_fin(int n_comp, const int* __restrict__ comp_first_chunk,
                                                 const double* __restrict__ inv_count, double inv_count1, int e031,
                                                 double* mean, PcgState* st, const double* part) {
  if (st && !block_running(st)) return;
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  if (n_comp == 1) {
    const double sum = reduce_part(part + P_CS, RP, 2);
    const double sum2 = e031 ? reduce_part(part + P_CS + 1, RP, 2) : 0.0;
    if (threadIdx.x == 0) {
      mean[0] = sum * inv_count1;
      if (e031 && fabs(sum) > 1e-9 * sum2) bad = 1;
    }
  } else {
    for (int comp = threadIdx.x; comp < n_comp; comp += RB) {
      double v = 0.0, v2 = 0.0;
      for (int k = comp_first_chunk[comp]; k < comp_first_chunk[comp + 1]; ++k) {
        v += part[P_CS + 2 * k];
        v2 += part[P_CS + 2 * k + 1];
      }
      mean[comp] = v * inv_count[comp];
      if (e031 && fabs(v) > 1e-9 * v2) atomicOr(&bad, 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && st && bad) st->status = ST_E031;
}

// r = b - mean (closed cells), b_c = r, x = 0, partials of |r|^2 (pressure.rs:365-372).
__global__ void __launch_bounds__(RB) k_pcg_init(long long N, const uint8_t* __restrict__ codes, const double* __restrict__ b,
                                                 const double* __restrict__ mean, const int* __restrict__ cidx,
                                                 double* __restrict__ r, double* __restrict__ bc, double* __restrict__ x,
                                                 PcgState* st, double* part) {
  if (!block_running(st)) return;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < N; i += (long long)gridDim.x * RB) {
    double v = b[i];
    uint8_t ci = codes[i];
    if (ci & CLOSED) v -= mean[cidx ? cidx[i] : 0];
    r[i] = v;
    bc[i] = v;
    x[i] = 0.0;
    acc += v * v;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[P_RR + blockIdx.x] = s;
}

// One block: |b_c|; a zero right-hand side ends with zero iterations (pressure.rs:372-375).
__global__ void __launch_bounds__(RB) k_init_fin(PcgState* st, const double* part) {
  if (!block_running(st)) return;
  const double norm = sqrt(reduce_part(part + P_RR, RP));
  if (threadIdx.x == 0) {
    st->norm = norm;
    if (norm < 1e-24) st->status = ST_ZERO;
  }
}

// Warm start: x = x0 and r = b_c - A x0 (b_c and |b_c| formed by k_pcg_init).
__global__ void __launch_bounds__(BS) k_warm_residual(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ x0,
                                                      const double* __restrict__ bc, double* __restrict__ x,
                                                      double* __restrict__ r, const PcgState* st) {
  if (st->status != ST_RUNNING) return;
  long long c = (long long)blockIdx.x * BS + threadIdx.x;
  if (c >= g.N) return;
  int q[3];
  coords(g, c, q);
  double y = apply_cell(g, codes[c], c, q, [&](long long i) { return x0[i]; });
  x[c] = x0[c];
  r[c] = bc[c] - y;
}

// dst = src - mean on closed cells (center_nullspaces, pressure.rs:316-325).
__global__ void __launch_bounds__(BS) k_center(long long N, const uint8_t* __restrict__ codes, const double* __restrict__ src,
                                               const double* __restrict__ mean, const int* __restrict__ cidx,
                                               double* __restrict__ dst) {
  long long i = (long long)blockIdx.x * BS + threadIdx.x;
  if (i >= N) return;
  double v = src[i];
  if (codes[i] & CLOSED) v -= mean[cidx ? cidx[i] : 0];
  dst[i] = v;
}

__global__ void __launch_bounds__(BS) k_fill(long long N, double* __restrict__ dst, double v) {
  long long i = (long long)blockIdx.x * BS + threadIdx.x;
  if (i < N) dst[i] = v;
}

__global__ void k_state_init(PcgState* st, double tol, long long max_it) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->rz = 0.0;
  st->denom = 0.0;
  st->alpha = 0.0;
  st->beta = 0.0;
  st->norm = 0.0;
  st->rel = 0.0;
  st->true_rel = 0.0;
  st->tol = tol;
  st->accept = tol * 10.0 + 1e-12;
  st->it = 0;
  st->max_it = max_it;
  st->true_checks = 0;
  st->status = ST_RUNNING;
  st->check = 0;
  st->first = 1;
  st->ticket = 0;
}

// ---------------------------------------------------------------- IC(0)
// The IC(0) neighbours of a cell, in the order the reference's factor visits
// them. Lower neighbours ascend in index (z wrap, z lo, y wrap, y lo, x wrap,
// x lo); upper neighbours descend (z wrap, z hi, y wrap, y hi, x wrap, x hi). A
// periodic axis of two cells joins the pair through two faces, merged into one
// entry of weight 2w (the reference adds the second face to the first entry).
// Slots that do not apply keep index c and weight 0, so every load is valid.
struct Ic0Nbrs {
  long long idx[6];
  double a[6];  // factor numerator: -w or (-w - w); 0 for an unused slot
};
__device__ __forceinline__ Ic0Nbrs ic0_nbrs(const Geo& g, uint8_t code, long long c, const int q[3], bool lower) {
  Ic0Nbrs N;
  int k = 0;
#pragma unroll
  for (int a = 2; a >= 0; --a) {
    const int n = g.n[a], t = q[a];
    const long long s = stride_of(g, a);
    const double w = g.wi[a];
    long long i0 = c, i1 = c;
    double a0 = 0.0, a1 = 0.0;
    if (lower) {
      if (!g.per[a] || n == 1) {
        if (t > 0 && (code & (1u << (2 * a)))) i1 = c - s, a1 = -w;
      } else if (n == 2) {
        if (t == 1 && (code & (1u << (2 * a)))) i1 = c - s, a1 = -w - w;
      } else {
        if (t == n - 1 && (code & (2u << (2 * a)))) i0 = c - (long long)(n - 1) * s, a0 = -w;
        if (t > 0 && (code & (1u << (2 * a)))) i1 = c - s, a1 = -w;
      }
    } else {
      if (!g.per[a] || n == 1) {
        if (t < n - 1 && (code & (2u << (2 * a)))) i1 = c + s, a1 = -w;
      } else if (n == 2) {
        if (t == 0 && (code & (2u << (2 * a)))) i1 = c + s, a1 = -w - w;
      } else {
        if (t == 0 && (code & (1u << (2 * a)))) i0 = c + (long long)(n - 1) * s, a0 = -w;
        if (t < n - 1 && (code & (2u << (2 * a)))) i1 = c + s, a1 = -w;
      }
    }
    N.idx[k] = i0;
    N.a[k] = a0;
    N.idx[k + 1] = i1;
    N.a[k + 1] = a1;
    k += 2;
  }
  return N;
}

// Forward sweep of the hyperplane i + j + k = L, gather form of Factor::apply
// (pressure.rs:225-232): y = (((r - l y_zlo) - l y_ylo) - l y_xlo) / D with the
// lower neighbours in ascending index order and l = a / D[col], the reference's
// factor entries (no fill without a 3-cell periodic axis). All loads are issued
// before the arithmetic.
__global__ void __launch_bounds__(BS) k_ic0_fwd(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ D,
                                                const double* __restrict__ r, double* __restrict__ y, int L, int j0,
                                                int k0, int nj, int nk, const PcgState* st,
                                                const double* __restrict__ fill) {
  const int status = st ? st->status : ST_RUNNING;
  int t = blockIdx.x * BS + threadIdx.x;
  if (t >= nj * nk) return;
  int j = j0 + t % nj, k = k0 + t / nj;
  int i = L - j - k;
  if (i < 0 || i >= g.n[0]) return;
  long long c = (long long)i + j * g.sy + k * g.sz;
  const uint8_t code = codes[c];
  const int q[3] = {i, j, k};
  const Ic0Nbrs N = ic0_nbrs(g, code, c, q, true);
  double dn[6], yn[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    dn[m] = D[N.idx[m]];
    yn[m] = y[N.idx[m]];
  }
  double value = r[c];
  const double dc = D[c];
  if (status != ST_RUNNING) return;
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    if (N.a[m] == 0.0) continue;
    double l = fill ? fill[12 * c + m] : N.a[m] / dn[m];
    value -= l * yn[m];
  }
  y[c] = value / dc;
}

// Backward sweep (pressure.rs:233-240) in gather form: the reference's scatter
// z[col] -= l * x_i over rows i in descending order becomes, for cell c, the
// subtraction of its upper neighbours' terms in descending index order, then
// the division by D[c]; l = a / D[c].
__global__ void __launch_bounds__(BS) k_ic0_bwd(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ D,
                                                double* __restrict__ z, int L, int j0, int k0, int nj, int nk,
                                                const PcgState* st, const double* __restrict__ fill) {
  const int status = st ? st->status : ST_RUNNING;
  int t = blockIdx.x * BS + threadIdx.x;
  if (t >= nj * nk) return;
  int j = j0 + t % nj, k = k0 + t / nj;
  int i = L - j - k;
  if (i < 0 || i >= g.n[0]) return;
  long long c = (long long)i + j * g.sy + k * g.sz;
  const uint8_t code = codes[c];
  const int q[3] = {i, j, k};
  const Ic0Nbrs N = ic0_nbrs(g, code, c, q, false);
  double zn[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) zn[m] = z[N.idx[m]];
  const double dc = D[c];
  double value = z[c];
  if (status != ST_RUNNING) return;
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    if (N.a[m] == 0.0) continue;
    double l = fill ? fill[12 * c + 6 + m] : N.a[m] / dc;
    value -= l * zn[m];
  }
  z[c] = value / dc;
}


// Tiled wavefront of the same sweeps: 8x8x8 tiles; the tiles on one diagonal
// I + J + K = P are independent (a face neighbour lies in the same tile or on
// the neighbouring diagonal, wrap neighbours of a periodic axis on a lower
// diagonal for the forward sweep and a higher one for the backward sweep), so
// one launch per diagonal processes them, and inside a tile one wavefront walks
// its 22 local hyperplanes with LDS between the steps. The tile's factor
// entries, right-hand side and the neighbours outside the tile are loaded into
// LDS first, so a step costs one LDS round trip. Each cell repeats
// k_ic0_fwd / k_ic0_bwd's operations in the same order (bitwise equal).
// Tiles of TX x TY x TZ cells (TX TY TZ = 512 threads, one per cell): 8 x 8 x 8,
// or 16 x 2 x 16 (and its rotations) on grids with an axis of at most 2 cells.
// T is the arithmetic type of the sweep: double repeats the reference's
// operations bit for bit; float is the mixed-precision application (the factor
// pivots D stay FP64: each entry a / D is formed in FP64 and rounded once).
// own: the value the cell starts from (r forward, the forward result backward);
// out: where the sweep writes, also read for the neighbours solved by earlier
// tile diagonals (the forward sweep's lower and the backward sweep's upper
// neighbours).
template <bool FWD, int TX, int TY, int TZ, class T, class TOWN, class TOUT>
__global__ void __launch_bounds__(TX * TY * TZ)
    k_ic0_tile(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ D, const TOWN* own, TOUT* out,
               int P, int J0, int K0, int nJ, int nK, int ntx, const PcgState* st, const double* __restrict__ fill) {
  constexpr int NT = TX * TY * TZ;
  __shared__ T ys[NT];
  if (st && st->status != ST_RUNNING) return;
  const int J = J0 + (int)blockIdx.x % nJ, K = K0 + (int)blockIdx.x / nJ;
  const int I = P - J - K;
  if (I < 0 || I >= ntx) return;
  const int ox = I * TX, oy = J * TY, oz = K * TZ;
  // one thread per tile cell: its factor entries, right-hand side and outside
  // neighbours are loaded by all threads at once, before the steps
  const int li = threadIdx.x;
  const int a = li % TX, b = (li / TX) % TY, cz = li / (TX * TY);
  const int i = ox + a, j = oy + b, k = oz + cz;
  const bool live = i < g.n[0] && j < g.n[1] && k < g.n[2];
  T co[6], ext[6], rv = 0, dc = 1;
  int lix[6];
  long long c = 0;
#pragma unroll
  for (int m = 0; m < 6; ++m) lix[m] = -2;
  if (live) {
    c = (long long)i + j * g.sy + k * g.sz;
    const uint8_t code = codes[c];
    const int q[3] = {i, j, k};
    const Ic0Nbrs N = ic0_nbrs(g, code, c, q, FWD);
    const double d64 = D[c];
    dc = (T)d64;
    rv = (T)own[c];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      if (N.a[m] == 0.0) continue;
      const unsigned nb = (unsigned)N.idx[m];
      co[m] = (T)(fill ? fill[12 * c + (FWD ? 0 : 6) + m] : FWD ? N.a[m] / D[nb] : N.a[m] / d64);
      const unsigned ni = nb % (unsigned)g.n[0], nj = (nb / (unsigned)g.n[0]) % (unsigned)g.n[1];
      const unsigned nk = nb / (unsigned)g.sz;
      if ((int)ni >= ox && (int)ni < ox + TX && (int)nj >= oy && (int)nj < oy + TY && (int)nk >= oz && (int)nk < oz + TZ) {
        lix[m] = (int)(ni - ox) + TX * ((int)(nj - oy) + TY * (int)(nk - oz));
      } else {
        lix[m] = -1;
        ext[m] = (T)out[nb];
      }
    }
  }
  const int plane = a + b + cz;
  // the tile's last local hyperplane holding a cell of the grid (partial tiles
  // at the grid's far faces walk fewer steps)
  const int S = min(TX, g.n[0] - ox) + min(TY, g.n[1] - oy) + min(TZ, g.n[2] - oz) - 3;
  // the steps exchange values through LDS only; each cell stores its value once
  // after the last step (no global store outstanding at the steps' barriers)
  T v = 0;
  for (int step = 0; step <= S; ++step) {
    const int sp = FWD ? step : S - step;
    if (live && plane == sp) {
      T value = rv;
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        if (lix[m] == -2) continue;
        const T ym = lix[m] >= 0 ? ys[lix[m]] : ext[m];
        value -= co[m] * ym;
      }
      v = value / dc;
      ys[li] = v;
    }
    __syncthreads();
  }
  if (live) out[c] = (TOUT)v;
}

// ---------------------------------------------------------------- fine level in the natural layout
// Colour of a fine cell: its parity, and with `seams` (a bit per periodic axis
// of odd length) two more classes per seam for the cells on the last slab of
// that axis, whose wrap neighbour has the same parity.
__device__ __forceinline__ int fine_colour(const Geo& g, int i, int j, int k, int seams) {
  int c = (i + j + k) & 1;
  if (seams) {
    // one bit per seam axis, packed in axis order
    const int q[3] = {i, j, k};
    int s = 0, bit = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if ((seams >> a) & 1) {
        if (q[a] == g.n[a] - 1) s |= 1 << bit;
        ++bit;
      }
    c += 2 * s;
  }
  return c;
}

// Red-black Gauss-Seidel half-sweep on the fine level from the coupling codes.
// `first`: x starts at zero (neighbours contribute nothing). `xc`: fuse the
// prolongation x += P e of the coarse correction into the neighbour reads (the
// cells of this colour are overwritten, so only neighbours need it). `out64`:
// also store the result in FP64.
template <class T, class TB>
__global__ void __launch_bounds__(BS)
    k_fine_gs(Geo g, const uint8_t* __restrict__ codes, const TB* __restrict__ b, T* __restrict__ x, int color, int first,
              const T* __restrict__ xc, Geo gc, double* __restrict__ out64, int all_cells, const PcgState* st,
              T cscale = T(1), const int* __restrict__ fparent = nullptr, int seams = 0) {
  long long t = (long long)blockIdx.x * BS + threadIdx.x;
  int i, j, k;
  if (all_cells) {
    if (t >= g.N) return;
    int q[3];
    coords(g, t, q);
    i = q[0];
    j = q[1];
    k = q[2];
  } else {
    const unsigned hx = (unsigned)(g.n[0] + 1) >> 1;
    const unsigned tu = (unsigned)t;
    const unsigned row = tu / hx;
    if (row >= (unsigned)g.n[1] * (unsigned)g.n[2]) return;
    const int i2 = (int)(tu - row * hx);
    const unsigned kk = row / (unsigned)g.n[1];
    j = (int)(row - kk * (unsigned)g.n[1]);
    k = (int)kk;
    i = 2 * i2 + ((j + k + color) & 1);
    if (i >= g.n[0]) return;
  }
  const long long c = (long long)i + j * g.sy + k * g.sz;
  const int status = st ? st->status : ST_RUNNING;
  if (fine_colour(g, i, j, k, seams) != color) {
    if (out64 && status == ST_RUNNING) out64[c] = (double)x[c];
    return;
  }
  // all loads first (one memory round trip), then the update
  const uint8_t code = codes[c];
  const T bv = (T)b[c];
  const int q[3] = {i, j, k};
  T xn[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      T v = T(0);
      if (!first) {
        const long long nb = nb_or_self(g, c, a, q[a], side);
        v = x[nb];
        if (xc) {
          if (fparent) {
            const int pn = fparent[nb];
            if (pn >= 0) v += cscale * xc[pn];
          } else {
            int qn[3] = {i, j, k};
            qn[a] = side == 0 ? (q[a] > 0 ? q[a] - 1 : g.n[a] - 1) : (q[a] < g.n[a] - 1 ? q[a] + 1 : 0);
            v += cscale * xc[(qn[0] >> 1) + (qn[1] >> 1) * gc.sy + (qn[2] >> 1) * gc.sz];
          }
        }
      }
      xn[2 * a + side] = v;
    }
  if (status != ST_RUNNING) return;
  if (!(code & FLUID)) {
    x[c] = T(0);
    if (out64) out64[c] = 0.0;
    return;
  }
  T diag = T(0);
  T value = bv;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (!(code & (1u << (2 * a + side)))) continue;
      const bool boundary = side == 0 ? q[a] == 0 : q[a] == g.n[a] - 1;
      if (boundary && !g.per[a]) {
        diag += (T)g.wo[a];
        continue;
      }
      const T w = (T)g.wi[a];
      diag += w;
      if (first) continue;
      value += w * xn[2 * a + side];
    }
  }
  T v = diag > T(0) ? value / diag : T(0);
  x[c] = v;
  if (out64) out64[c] = (double)v;
}

// Fine-level residual (b - A x)_c of one cell from the coupling codes.
template <class T, class TB>
__device__ __forceinline__ T fine_residual(const Geo& g, const uint8_t* __restrict__ codes, const TB* __restrict__ b,
                                           const T* __restrict__ x, const int q[3]) {
  const long long c = (long long)q[0] + q[1] * g.sy + q[2] * g.sz;
  const uint8_t code = codes[c];
  const T bv = (T)b[c];
  const T xv = x[c];
  T xn[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) xn[2 * a + side] = x[nb_or_self(g, c, a, q[a], side)];
  if (!(code & FLUID)) return T(0);
  T ax = T(0), off = T(0);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (!(code & (1u << (2 * a + side)))) continue;
      const bool boundary = side == 0 ? q[a] == 0 : q[a] == g.n[a] - 1;
      if (boundary && !g.per[a]) {
        ax += (T)g.wo[a];
        continue;
      }
      ax += (T)g.wi[a];
      off += (T)g.wi[a] * xn[2 * a + side];
    }
  return bv - (ax * xv - off);
}

// Coefficients and neighbour values of a coarse cell, loaded together: the
// coupling to the + neighbour (wp[a][c]) and to the - neighbour (wp[a][lo]) per
// axis, zero where the axis ends.
// Tile index for workgroup b of nb: each XCD (blocks b = x mod 8) gets a
// contiguous range of tiles, so neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// ---------------------------------------------------------------- multigrid (aggregation)
// Coarse levels are CSR matrices over aggregates (connected groups of 2x2x2
// bins, multigrid/masked.rs:127-156). Rows are colour-sorted: one colour class
// is a contiguous range and an independent set, so a class is relaxed in
// parallel with the Gauss-Seidel formula x_i = (b_i - sum a_ij x_j) / a_ii. The
// status word is read first and tested only before the stores, so its latency
// overlaps the other loads.

// a_ij x_j over row i's off-diagonal entries, in entry order; the first eight
// entries are loaded together (predicated), so a row costs one load round trip.
constexpr int kRowPre = 8;
// With xc, each neighbour value carries the next level's correction
// x_j + cs * xc[parent_j] (the prolongation riding on the first post-sweep).
template <class T>
__device__ __forceinline__ T agg_offdiag(const AggDevLevel<T>& L, int e0, int e1, T v, bool subtract,
                                         const T* __restrict__ xc = nullptr, T cs = T(0)) {
  int cols[kRowPre];
  T vals[kRowPre], xv[kRowPre];
#pragma unroll
  for (int k = 0; k < kRowPre; ++k) {
    cols[k] = -1;
    vals[k] = T(0);
    if (e0 + k < e1) {
      cols[k] = L.col[e0 + k];
      vals[k] = L.val[e0 + k];
    }
  }
#pragma unroll
  for (int k = 0; k < kRowPre; ++k) {
    xv[k] = T(0);
    if (cols[k] >= 0) {
      xv[k] = L.x[cols[k]];
      if (xc) xv[k] += cs * xc[L.parent[cols[k]]];
    }
  }
#pragma unroll
  for (int k = 0; k < kRowPre; ++k)
    if (cols[k] >= 0) {
      if (subtract)
        v -= vals[k] * xv[k];
      else
        v += vals[k] * xv[k];
    }
  for (int e = e0 + kRowPre; e < e1; ++e) {
    const int j = L.col[e];
    T xj = L.x[j];
    if (xc) xj += cs * xc[L.parent[j]];
    if (subtract)
      v -= L.val[e] * xj;
    else
      v += L.val[e] * xj;
  }
  return v;
}

// The same sum over the padded rows: no row-pointer load, K is uniform.
template <class T>
__device__ __forceinline__ T ell_offdiag(const AggDevLevel<T>& L, int i, T v, bool subtract,
                                         const T* __restrict__ xc = nullptr, T cs = T(0)) {
  int cols[kEllWidth];
  T vals[kEllWidth], xv[kEllWidth];
#pragma unroll
  for (int k = 0; k < kEllWidth; ++k) {
    cols[k] = i;
    vals[k] = T(0);
    if (k < L.K) {
      const unsigned w = L.ecol[(long long)k * L.n + i];
      cols[k] = (int)(w & 0xFFFFFFu);
      vals[k] = L.etab[w >> 24];
    }
  }
#pragma unroll
  for (int k = 0; k < kEllWidth; ++k) {
    xv[k] = T(0);
    // padding (the row itself, value 0) needs no gather
    if (k < L.K && cols[k] != i) {
      xv[k] = L.x[cols[k]];
      if (xc) xv[k] += cs * xc[L.parent[cols[k]]];
    }
  }
#pragma unroll
  for (int k = 0; k < kEllWidth; ++k)
    if (k < L.K) {
      if (subtract)
        v -= vals[k] * xv[k];
      else
        v += vals[k] * xv[k];
    }
  return v;
}

// b_i - (a_ii x_i + sum a_ij x_j) (Matrix::product, multigrid.rs:263-271)
template <class T>
__device__ __forceinline__ T agg_row_residual(const AggDevLevel<T>& L, int i) {
  const T d = L.diag[i], xi = L.x[i], bi = L.b[i];
  if (L.ecol) return bi - ell_offdiag<T>(L, i, d * xi, false);
  const int e0 = L.ptr[i], e1 = L.ptr[i + 1];
  return bi - agg_offdiag<T>(L, e0, e1, d * xi, false);
}

// (b_i - sum a_ij x_j) / a_ii, 0 for a node without couplings (masked.rs:158-172)
template <class T>
__device__ __forceinline__ T agg_row_relax(const AggDevLevel<T>& L, int i, const T* __restrict__ xc = nullptr,
                                           T cs = T(0)) {
  const T d = L.diag[i], bi = L.b[i];
  T v;
  if (L.ecol) {
    v = ell_offdiag<T>(L, i, bi, true, xc, cs);
  } else {
    const int e0 = L.ptr[i], e1 = L.ptr[i + 1];
    v = agg_offdiag<T>(L, e0, e1, bi, true, xc, cs);
  }
  return d != T(0) ? v / d : T(0);
}

// One colour class of a half-sweep. With xc (two-colour levels only) the
// prolongation of the next level's correction is fused into the neighbour
// reads: this class is overwritten and the other one is recomputed from scratch
// by the next half-sweep, so no node needs the correction stored.
template <class T>
__device__ __forceinline__ void agg_gs_body(const AggDevLevel<T>& L, int colour, const PcgState* st,
                                            const T* __restrict__ xc, T cs, unsigned bid) {
  const int status = st ? st->status : ST_RUNNING;
  const int i = L.cptr[colour] + (int)(bid * BS + threadIdx.x);
  if (i >= L.cptr[colour + 1]) return;
  const T v = agg_row_relax<T>(L, i, xc, cs);
  if (status != ST_RUNNING) return;
  L.x[i] = v;
}
template <class T>
__global__ void __launch_bounds__(BS) k_agg_gs(AggDevLevel<T> L, int colour, const PcgState* st,
                                               const T* __restrict__ xc = nullptr, T cs = T(0)) {
  agg_gs_body<T>(L, colour, st, xc, cs, blockIdx.x);
}

// The first pre-smoothing half-sweep of a coarse level starts from a zero
// iterate, so its colour-0 nodes take b / a_ii without reading neighbours; the
// restriction stores that value at once (x of the other colours stays 0).
template <class T>
__device__ __forceinline__ T first_relax(int node, int red_end, T b, const T* __restrict__ diag) {
  if (node >= red_end) return T(0);
  const T d = diag[node];
  return d != T(0) ? b / d : T(0);
}

// Restriction into level C: wavefront s of a workgroup forms the residuals of
// child slot s (children s, s + 8, ...) of 64 consecutive aggregates, so the
// lanes of a load read the same slot of neighbouring aggregates; the eight
// partial sums meet in LDS and the first wavefront adds them in slot order
// (ascending children when there are at most eight). Children from zero_from
// on (the colour class relaxed last) have a residual that is zero to rounding
// and are taken as zero, except rows without couplings (their residual is b).
// The coarse iterate starts at zero.
constexpr int kRestrictThreads = 512;
template <class T>
__device__ __forceinline__ void agg_restrict_body(const AggDevLevel<T>& F, const AggDevLevel<T>& C, const PcgState* st,
                                                  int zero_from, unsigned bid) {
  __shared__ T part[8][64];
  const int status = st ? st->status : ST_RUNNING;
  const int lane = (int)(threadIdx.x & 63), sub = (int)(threadIdx.x >> 6);
  const int J = (int)bid * 64 + lane;
  T r = T(0);
  if (J < C.n) {
    const int k0 = C.child_ptr[J], k1 = C.child_ptr[J + 1];
    for (int k = k0 + sub; k < k1; k += 8) {
      const int c = C.child[k];
      if (c < zero_from || F.diag[c] == T(0)) r += agg_row_residual<T>(F, c);
    }
  }
  part[sub][lane] = r;
  __syncthreads();
  if (sub != 0 || J >= C.n || status != ST_RUNNING) return;
  T sum = T(0);
#pragma unroll
  for (int s2 = 0; s2 < 8; ++s2) sum += part[s2][lane];
  C.b[J] = sum;
  C.x[J] = first_relax<T>(J, C.cptr[1], sum, C.diag);
}
template <class T>
__global__ void __launch_bounds__(kRestrictThreads) k_agg_restrict(AggDevLevel<T> F, AggDevLevel<T> C,
                                                                   const PcgState* st, int zero_from) {
  agg_restrict_body<T>(F, C, st, zero_from, blockIdx.x);
}

template <class T>
__device__ __forceinline__ void agg_prolong_body(const AggDevLevel<T>& F, const T* __restrict__ xc, T cs,
                                                 const PcgState* st, unsigned bid) {
  const int status = st ? st->status : ST_RUNNING;
  const int i = (int)(bid * BS + threadIdx.x);
  if (i >= F.n) return;
  const T v = F.x[i] + cs * xc[F.parent[i]];
  if (status != ST_RUNNING) return;
  F.x[i] = v;
}
template <class T>
__global__ void __launch_bounds__(BS) k_agg_prolong(AggDevLevel<T> F, const T* __restrict__ xc, T cs,
                                                    const PcgState* st) {
  agg_prolong_body<T>(F, xc, cs, st, blockIdx.x);
}

// Coarsest level: x = G_k b per component block, one wavefront per block.
template <class TI, class TO>
__device__ __forceinline__ void agg_direct_block(const AggCoarsest& D, int k, int lane, const TI* __restrict__ b,
                                                 TO* __restrict__ x, double* __restrict__ out64, bool run) {
  const int r0 = D.bptr[k], m = D.bptr[k + 1] - r0;
  if (m > kSmallBlock || lane >= m) return;
  const double* G = D.G + D.goff[k] + (long long)lane * m;
  double acc = 0.0;
  for (int j = 0; j < m; ++j) acc += G[j] * (double)b[D.rows[r0 + j]];
  if (!run) return;
  const int row = D.rows[r0 + lane];
  x[row] = (TO)acc;
  if (out64) out64[row] = (double)(TO)acc;
}

// Coarsest solve in one launch: the first big_rows workgroups take one row of a
// big block each (the four wavefronts split the columns), the following ones
// four small blocks each (one wavefront per block).
template <class TI, class TO>
__device__ __forceinline__ void agg_direct_body(const AggCoarsest& D, const TI* __restrict__ b, TO* __restrict__ x,
                                                double* __restrict__ out64, const PcgState* st, unsigned bid) {
  __shared__ double part[BS / 64];
  const int status = st ? st->status : ST_RUNNING;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int w = (int)bid;
  if (w >= D.big_rows) {
    const int k = (w - D.big_rows) * (BS / 64) + wave;
    if (k < D.nsmall) agg_direct_block<TI, TO>(D, D.small_block[k], lane, b, x, out64, status == ST_RUNNING);
    return;
  }
  int q = 0;
  while (D.big_first[q + 1] <= w) ++q;
  const int k = D.big_block[q];
  const int r0 = D.bptr[k], m = D.bptr[k + 1] - r0;
  const int row = w - D.big_first[q];
  const double* G = D.G + D.goff[k] + (long long)row * m;
  double acc = 0.0;
#pragma unroll 4
  for (int j = threadIdx.x; j < m; j += BS) acc += G[j] * (double)b[D.rows[r0 + j]];
  acc = wave_sum(acc);
  if (lane == 0) part[wave] = acc;
  __syncthreads();
  if (threadIdx.x != 0 || status != ST_RUNNING) return;
  double total = 0.0;
#pragma unroll
  for (int v = 0; v < BS / 64; ++v) total += part[v];
  const int r = D.rows[r0 + row];
  x[r] = (TO)total;
  if (out64) out64[r] = (double)(TO)total;
}
template <class TI, class TO>
__global__ void __launch_bounds__(BS) k_agg_direct(AggCoarsest D, const TI* __restrict__ b, TO* __restrict__ x,
                                                   double* __restrict__ out64, const PcgState* st) {
  agg_direct_body<TI, TO>(D, b, x, out64, st, blockIdx.x);
}

// Set-up of a big coarsest block on the device: in-place blocked Gauss-Jordan
// inversion of the SPD k x k matrix A (no pivoting needed), then embedding into
// the m x m block (first row and column zero for a closed component) and the
// symmetric centering C G C. A step takes the next nb <= kGJ pivots P:
//   A_PP <- A_PP^-1 (scalar Gauss-Jordan in LDS; a pivot that is not positive
//                    flags the block as no longer definite),
//   A_Pj <- A_PP^-1 A_Pj, A_iP <- -A_iP A_PP^-1, A_ij <- A_ij - A_iP A_PP^-1 A_Pj
// for i, j outside P: three launches and one rank-nb update per step.
constexpr int kGJ = 64;

constexpr int kGJThreads = 1024;
__global__ void __launch_bounds__(kGJThreads) k_gj_panel(const double* __restrict__ A, int k, int p0, int nb,
                                                         double* __restrict__ pinv, int* __restrict__ bad) {
  static_assert(kGJ == 64, "entries are located by shifts");
  __shared__ double M[kGJ * kGJ];
  __shared__ double rowq[kGJ], colq[kGJ];
  const int t = (int)threadIdx.x;
  for (int e = t; e < kGJ * kGJ; e += kGJThreads) {
    const int i = e >> 6, j = e & 63;
    M[e] = (i < nb && j < nb) ? A[(long long)(p0 + i) * k + p0 + j] : 0.0;
  }
  __syncthreads();
  for (int q = 0; q < nb; ++q) {
    if (t < nb) {
      rowq[t] = M[q * kGJ + t];
      colq[t] = M[t * kGJ + q];
    }
    __syncthreads();
    const double d = rowq[q];
    if (t == 0 && (!(d > 0.0) || !isfinite(d))) *bad = 1;
    const double inv = 1.0 / d;
    for (int e = t; e < kGJ * kGJ; e += kGJThreads) {
      const int i = e >> 6, j = e & 63;
      if (i >= nb || j >= nb) continue;
      double v;
      if (i == q && j == q)
        v = inv;
      else if (i == q)
        v = rowq[j] * inv;
      else if (j == q)
        v = -colq[i] * inv;
      else
        v = M[e] - colq[i] * (rowq[j] * inv);
      M[e] = v;
    }
    __syncthreads();
  }
  for (int e = t; e < kGJ * kGJ; e += kGJThreads) pinv[e] = M[e];
}

// rp[s][j] = (A_PP^-1 A_Pj)[s][j] for j outside P, A_PP^-1[s][j - p0] inside;
// cold[i][s] = A[i][p0 + s] (the column panel before the step).
__global__ void __launch_bounds__(BS) k_gj_rows(const double* __restrict__ A, int k, int p0, int nb,
                                                const double* __restrict__ pinv, double* __restrict__ rp,
                                                double* __restrict__ cold) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  const long long nrow = (long long)nb * k;
  if (t < nrow) {
    const int s = (int)(t / k), j = (int)(t - (long long)s * k);
    double v;
    if (j >= p0 && j < p0 + nb) {
      v = pinv[s * kGJ + (j - p0)];
    } else {
      v = 0.0;
      for (int q = 0; q < nb; ++q) v += pinv[s * kGJ + q] * A[(long long)(p0 + q) * k + j];
    }
    rp[(long long)s * k + j] = v;
  } else if (t < 2 * nrow) {
    const long long e = t - nrow;
    const int i = (int)(e / nb), s = (int)(e - (long long)i * nb);
    cold[(long long)i * kGJ + s] = A[(long long)i * k + p0 + s];
  }
}

// The rank-nb update over 64 x 64 tiles of A (each thread 4 x 4 entries).
__global__ void __launch_bounds__(BS) k_gj_update(double* __restrict__ A, int k, int p0, int nb,
                                                  const double* __restrict__ rp, const double* __restrict__ cold) {
  __shared__ double Cs[64 * kGJ];  // Cs[r * kGJ + s] = cold[i0 + r][s]
  __shared__ double Rs[kGJ * 64];  // Rs[s * 64 + c] = rp[s][j0 + c]
  const int tilesx = (k + 63) / 64;
  const int i0 = (int)(blockIdx.x / tilesx) * 64, j0 = (int)(blockIdx.x % tilesx) * 64;
  for (int e = (int)threadIdx.x; e < 64 * kGJ; e += BS) {
    const int r = e / kGJ, s2 = e - r * kGJ;
    Cs[e] = (i0 + r < k && s2 < nb) ? cold[(long long)(i0 + r) * kGJ + s2] : 0.0;
    const int s3 = e / 64, c = e - s3 * 64;
    Rs[e] = (j0 + c < k && s3 < nb) ? rp[(long long)s3 * k + j0 + c] : 0.0;
  }
  __syncthreads();
  const int tx = (int)(threadIdx.x & 15), ty = (int)(threadIdx.x >> 4);
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int q = 0; q < nb; ++q) {
    double cv[4], rv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) cv[a] = Cs[(ty + 16 * a) * kGJ + q];
#pragma unroll
    for (int b = 0; b < 4; ++b) rv[b] = Rs[q * 64 + tx + 16 * b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += cv[a] * rv[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ty + 16 * a;
    if (i >= k) continue;
    const bool irow = i >= p0 && i < p0 + nb;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int j = j0 + tx + 16 * b;
      if (j >= k) continue;
      const long long at = (long long)i * k + j;
      if (irow)
        A[at] = rp[(long long)(i - p0) * k + j];
      else
        A[at] = ((j >= p0 && j < p0 + nb) ? 0.0 : A[at]) - acc[a][b];
    }
  }
}
__global__ void __launch_bounds__(BS) k_block_embed(double* __restrict__ G, int m, int skip, const double* __restrict__ A) {
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  if (t >= (long long)m * m) return;
  const int i = (int)(t / m), j = (int)(t - (long long)i * m);
  const int k = m - skip;
  G[t] = (i >= skip && j >= skip) ? A[(long long)(i - skip) * k + (j - skip)] : 0.0;
}
__global__ void __launch_bounds__(BS) k_block_colmean(const double* __restrict__ G, int m, double* __restrict__ mu) {
  const int j = (int)(blockIdx.x * BS + threadIdx.x);
  if (j >= m) return;
  double s = 0.0;
  for (int i = 0; i < m; ++i) s += G[(long long)i * m + j];
  mu[j] = s / (double)m;
}
__global__ void __launch_bounds__(BS) k_block_center(double* __restrict__ G, int m, const double* __restrict__ mu) {
  __shared__ double sh;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int j = 0; j < m; ++j) s += mu[j];
    sh = s / (double)m;
  }
  __syncthreads();
  const long long t = (long long)blockIdx.x * BS + threadIdx.x;
  if (t >= (long long)m * m) return;
  const int i = (int)(t / m), j = (int)(t - (long long)i * m);
  G[t] = G[t] - mu[i] - mu[j] + sh;
}

// Fine level -> first coarse level: eight lanes per 2x2x2 bin form the fine
// residuals; the bin's first lane adds them up per aggregate (a bin holds one
// aggregate per connected group of its fluid cells) in ascending cell order.
// zero_colour: the colour class relaxed last (its residuals vanish in exact
// arithmetic and are taken as zero), -1 for none.
template <class T, class TB>
__global__ void __launch_bounds__(BS) k_agg_fine_restrict(Geo g, const uint8_t* __restrict__ codes,
                                                          const TB* __restrict__ b, const T* __restrict__ x,
                                                          const int* __restrict__ fparent, T* __restrict__ bc,
                                                          T* __restrict__ xc, const T* __restrict__ cdiag, int red_end,
                                                          const PcgState* st, int zero_colour) {
  const int status = st ? st->status : ST_RUNNING;
  const unsigned t = blockIdx.x * BS + threadIdx.x;
  const unsigned bx = (unsigned)(g.n[0] + 1) >> 1, by = (unsigned)(g.n[1] + 1) >> 1, bz = (unsigned)(g.n[2] + 1) >> 1;
  const unsigned B = t >> 3;
  const int sub = (int)(t & 7);
  T r = T(0);
  int p = -1;
  if (B < bx * by * bz) {
    const unsigned Bx = B % bx, Byz = B / bx;
    const unsigned By = Byz % by, Bz = Byz / by;
    const int q[3] = {(int)(2 * Bx) + (sub & 1), (int)(2 * By) + ((sub >> 1) & 1), (int)(2 * Bz) + (sub >> 2)};
    if (q[0] < g.n[0] && q[1] < g.n[1] && q[2] < g.n[2]) {
      p = fparent[(long long)q[0] + q[1] * g.sy + q[2] * g.sz];
      if (((q[0] + q[1] + q[2]) & 1) != zero_colour) r = fine_residual<T, TB>(g, codes, b, x, q);
    }
  }
  const int base = (int)(threadIdx.x & 63) & ~7;
  T rs[8];
  int ps[8];
#pragma unroll
  for (int s2 = 0; s2 < 8; ++s2) {
    rs[s2] = __shfl(r, base + s2, 64);
    ps[s2] = __shfl(p, base + s2, 64);
  }
  if (sub != 0 || status != ST_RUNNING) return;
  unsigned done = 0;
#pragma unroll
  for (int s2 = 0; s2 < 8; ++s2) {
    if (ps[s2] < 0 || (done & (1u << s2))) continue;
    T sum = T(0);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u >= s2 && ps[u] == ps[s2]) {
        sum += rs[u];
        done |= 1u << u;
      }
    bc[ps[s2]] = sum;
    xc[ps[s2]] = first_relax<T>(ps[s2], red_end, sum, cdiag);
  }
}

// x += cs * e[parent] on the fine level (used when the fine level has more than
// two colour classes, so the prolongation cannot ride on the first half-sweep).
template <class T>
__global__ void __launch_bounds__(BS) k_agg_fine_prolong(long long N, const int* __restrict__ fparent,
                                                         const T* __restrict__ xc, T cs, T* __restrict__ x,
                                                         const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  const long long c = (long long)blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  const int p = fparent[c];
  if (p < 0) return;
  const T v = x[c] + cs * xc[p];
  if (status != ST_RUNNING) return;
  x[c] = v;
}

template <class T>
__global__ void __launch_bounds__(BS) k_fill_t(long long N, T* __restrict__ x, const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  const long long c = (long long)blockIdx.x * BS + threadIdx.x;
  if (c >= N || status != ST_RUNNING) return;
  x[c] = T(0);
}

// ---------------------------------------------------------------- colour-separated fine level
__device__ __forceinline__ long long rb_index(const Geo& g, int hx, int i, int j, int k) {
  return (long long)(i >> 1) + (long long)hx * (j + (long long)g.n[1] * k);
}

// Diagonal of a fine cell from its code, accumulated in T as k_fine_gs does.
template <class T>
__device__ __forceinline__ T fine_diag(const Geo& g, uint8_t code, const int q[3]) {
  T diag = T(0);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (!(code & (1u << (2 * a + side)))) continue;
      const bool boundary = side == 0 ? q[a] == 0 : q[a] == g.n[a] - 1;
      diag += (boundary && !g.per[a]) ? (T)g.wo[a] : (T)g.wi[a];
    }
  return diag;
}

// The neighbour of q across side `side` of axis a (wrapping on periodic axes;
// q itself where the axis ends, whose value is then never used).
__device__ __forceinline__ void nb_coords(const Geo& g, const int q[3], int a, int side, int out[3]) {
  out[0] = q[0];
  out[1] = q[1];
  out[2] = q[2];
  const int n = g.n[a];
  if (side == 0)
    out[a] = q[a] > 0 ? q[a] - 1 : (g.per[a] ? n - 1 : q[a]);
  else
    out[a] = q[a] < n - 1 ? q[a] + 1 : (g.per[a] ? 0 : q[a]);
}

// First pre-smoothing half-sweep (x = 0): scatters the right-hand side into the
// colour-separated layout and sets the red cells to b / a_ii.
template <class T>
__global__ void __launch_bounds__(BS) k_rb_first(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ r,
                                                 FineRB<T> F, const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  // tiles of 64 x 4 cells (x, y) per z layer, located by block-uniform arithmetic
  const unsigned nx = (unsigned)g.n[0], ny = (unsigned)g.n[1];
  const unsigned gx = (nx + 63) >> 6, gy = (ny + 3) >> 2;
  const unsigned tile = blockIdx.x, txy = tile % (gx * gy);
  const int q[3] = {(int)((txy % gx) * 64 + (threadIdx.x & 63)), (int)((txy / gx) * 4 + (threadIdx.x >> 6)),
                    (int)(tile / (gx * gy))};
  if (q[0] >= (int)nx || q[1] >= (int)ny) return;
  const unsigned c = (unsigned)q[0] + nx * ((unsigned)q[1] + ny * (unsigned)q[2]);
  const uint8_t code = codes[c];
  const T bv = (T)r[c];
  const int colour = (q[0] + q[1] + q[2]) & 1;
  const long long at =
      colour * F.half + (long long)((unsigned)(q[0] >> 1) + (unsigned)F.hx * ((unsigned)q[1] + ny * (unsigned)q[2]));
  if (status != ST_RUNNING) return;
  F.b[at] = bv;
  if (colour == 0) {
    const T d = fine_diag<T>(g, code, q);
    F.x[at] = (code & FLUID) && d > T(0) ? bv / d : T(0);
  }
}

// k_rb_first with one thread per pair of natural cells (2 i2, 2 i2 + 1 of a row:
// one red, one black; nx even): r and the codes are read as pairs and each
// colour array is written by whole wavefronts (tiles of 64 pairs x 4 rows).
template <class T>
__device__ __forceinline__ void rb_first2_body(const Geo& g, const uint8_t* __restrict__ codes,
                                               const double* __restrict__ r, const FineRB<T>& F, const PcgState* st,
                                               unsigned bid) {
  const int status = st ? st->status : ST_RUNNING;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const unsigned gx = (hx + 63) >> 6, gy = (ny + 3) >> 2;
  const unsigned tile = bid, txy = tile % (gx * gy);
  const int i2 = (int)((txy % gx) * 64 + (threadIdx.x & 63));
  const int j = (int)((txy / gx) * 4 + (threadIdx.x >> 6));
  const int k = (int)(tile / (gx * gy));
  if (i2 >= (int)hx || j >= (int)ny) return;
  const unsigned c = 2u * (unsigned)i2 + (unsigned)g.sy * (unsigned)j + (unsigned)g.sz * (unsigned)k;
  const double2 rv = *reinterpret_cast<const double2*>(r + c);
  const uint8_t c0 = codes[c], c1 = codes[c + 1];
  const int par = (j + k) & 1;  // the red cell of the pair is 2 i2 + par
  const uint8_t code = par ? c1 : c0;
  const T bred = (T)(par ? rv.y : rv.x), bblack = (T)(par ? rv.x : rv.y);
  const int q[3] = {2 * i2 + par, j, k};
  const unsigned u = (unsigned)i2 + hx * ((unsigned)j + ny * (unsigned)k);
  const T d = fine_diag<T>(g, code, q);
  const T v = (code & FLUID) && d > T(0) ? bred / d : T(0);
  if (status != ST_RUNNING) return;
  F.b[u] = bred;
  F.b[F.half + u] = bblack;
  F.x[u] = v;
}
template <class T>
__global__ void __launch_bounds__(BS) k_rb_first2(Geo g, const uint8_t* __restrict__ codes,
                                                  const double* __restrict__ r, FineRB<T> F, const PcgState* st) {
  rb_first2_body<T>(g, codes, r, F, st, blockIdx.x);
}

// The first red half-sweep (from x = 0) when the right-hand side is already in
// the colour-separated layout (F.b, where the PCG keeps its residual): x_red =
// b / diagonal, the arithmetic of k_rb_first2; consecutive threads on
// consecutive pairs of red entries (nx a multiple of 4).
template <class T>
__device__ __forceinline__ void rb_first_lin_body(const Geo& g, const FineRB<T>& F, const PcgState* st, unsigned bid,
                                                  unsigned nblk) {
  using T2 = std::conditional_t<std::is_same<T, double>::value, double2, float2>;
  const int status = st ? st->status : ST_RUNNING;
  const unsigned pair = (unsigned)xcd_tile(bid, nblk) * BS + threadIdx.x;
  const unsigned u = 2 * pair;
  if (u >= (unsigned)F.half) return;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const unsigned row = div_magic(u, F.mhx);
  const int i2 = (int)(u - row * hx);
  const unsigned k = div_magic(row, F.mny);
  const int j = (int)(row - k * ny);
  const int par = (j + (int)k) & 1;
  const T2 bv = *reinterpret_cast<const T2*>(F.b + u);
  const uint8_t ca = F.codes[u], cb = F.codes[u + 1];
  const int qa[3] = {2 * i2 + par, j, (int)k}, qb[3] = {2 * i2 + 2 + par, j, (int)k};
  const T da = fine_diag<T>(g, ca, qa), db = fine_diag<T>(g, cb, qb);
  if (status != ST_RUNNING) return;
  T2 out;
  out.x = (ca & FLUID) && da > T(0) ? bv.x / da : T(0);
  out.y = (cb & FLUID) && db > T(0) ? bv.y / db : T(0);
  *reinterpret_cast<T2*>(F.x + u) = out;
}
template <class T>
__global__ void __launch_bounds__(BS) k_rb_first_lin(Geo g, FineRB<T> F, const PcgState* st) {
  rb_first_lin_body<T>(g, F, st, blockIdx.x, gridDim.x);
}

// The same from an FP64 right-hand side kept in the same colour-separated
// layout (rres, the mixed-precision solve's residual): both colours' b rounded
// into F.b and the red x = b / diagonal, k_rb_first2's arithmetic.
template <class T>
__device__ __forceinline__ void rb_first_cvt_body(const Geo& g, const FineRB<T>& F, const double* __restrict__ rres,
                                                  const PcgState* st, unsigned bid, unsigned nblk) {
  using T2 = std::conditional_t<std::is_same<T, double>::value, double2, float2>;
  const int status = st ? st->status : ST_RUNNING;
  const unsigned pair = (unsigned)xcd_tile(bid, nblk) * BS + threadIdx.x;
  const unsigned u = 2 * pair;
  if (u >= (unsigned)F.half) return;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const unsigned row = div_magic(u, F.mhx);
  const int i2 = (int)(u - row * hx);
  const unsigned k = div_magic(row, F.mny);
  const int j = (int)(row - k * ny);
  const int par = (j + (int)k) & 1;
  const double2 rr = *reinterpret_cast<const double2*>(rres + u);
  const double2 rk = *reinterpret_cast<const double2*>(rres + F.half + u);
  const uint8_t ca = F.codes[u], cb = F.codes[u + 1];
  const int qa[3] = {2 * i2 + par, j, (int)k}, qb[3] = {2 * i2 + 2 + par, j, (int)k};
  const T da = fine_diag<T>(g, ca, qa), db = fine_diag<T>(g, cb, qb);
  T2 red, black, out;
  red.x = (T)rr.x;
  red.y = (T)rr.y;
  black.x = (T)rk.x;
  black.y = (T)rk.y;
  out.x = (ca & FLUID) && da > T(0) ? red.x / da : T(0);
  out.y = (cb & FLUID) && db > T(0) ? red.y / db : T(0);
  if (status != ST_RUNNING) return;
  *reinterpret_cast<T2*>(F.b + u) = red;
  *reinterpret_cast<T2*>(F.b + F.half + u) = black;
  *reinterpret_cast<T2*>(F.x + u) = out;
}
template <class T>
__global__ void __launch_bounds__(BS) k_rb_first_cvt(Geo g, FineRB<T> F, const double* __restrict__ rres,
                                                     const PcgState* st) {
  rb_first_cvt_body<T>(g, F, rres, st, blockIdx.x, gridDim.x);
}

// The x/r update of k_pcg_axpy_rb (the same cells per thread in the same order:
// x, r and the partials of |r|^2 bit for bit) with the next V-cycle's first
// half-sweep from a zero iterate done on the way: each red cell's x in the fine
// level F set to r / a_ii (T double: F's b arrays are r itself; T float: both
// colours of r rounded into F's b first), the values of k_rb_first_lin and
// k_rb_first_cvt.
template <class TP, class T>
__device__ __forceinline__ void pcg_axpy_first_body(const Geo& g, const FineRB<T>& F, unsigned long long mnx,
                                                    unsigned long long mny, const TP* __restrict__ p,
                                                    const double* __restrict__ ap, double* __restrict__ x,
                                                    double* __restrict__ r, PcgState* st, double* part, unsigned bid,
                                                    unsigned nblk) {
  if (!block_running(st)) return;
  const double rz = reduce_part(part + P_RZ, RP);
  const double denom = reduce_part(part + P_PAP, RP);
  if (denom <= 0.0 || !isfinite(denom)) {
    if (bid == 0 && threadIdx.x == 0) {
      st->denom = denom;
      st->status = ST_E032;
    }
    return;
  }
  const double alpha = rz / denom;
  const long long N = g.N, half = N >> 1;
  const unsigned nx = (unsigned)g.n[0], ny = (unsigned)g.n[1];
  double acc = 0.0;
  for (long long i = (long long)bid * RB + threadIdx.x; i < N; i += (long long)nblk * RB) {
    const unsigned c = (unsigned)i;
    const unsigned row = div_magic(c, mnx), k = div_magic(row, mny);
    const unsigned ci = c - row * nx, cj = row - k * ny;
    const unsigned colour = (ci + cj + k) & 1u;
    const long long at = (colour ? half : 0) + (c >> 1);
    double* rc = r + at;
    x[i] += alpha * (double)p[i];
    double ri = *rc - alpha * ap[i];
    *rc = ri;
    acc += ri * ri;
    const T bv = (T)ri;
    if constexpr (!std::is_same<T, double>::value) F.b[at] = bv;
    if (colour == 0) {
      const uint8_t code = F.codes[c >> 1];
      const int q[3] = {(int)ci, (int)cj, (int)k};
      const T d = fine_diag<T>(g, code, q);
      F.x[c >> 1] = (code & FLUID) && d > T(0) ? bv / d : T(0);
    }
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) {
    part[P_RR + bid] = s;
    if (bid == 0) {
      st->rz = rz;
      st->alpha = alpha;
      st->denom = denom;
      st->first = 0;
    }
  }
}
template <class TP, class T>
__global__ void __launch_bounds__(RB) k_pcg_axpy_first(Geo g, FineRB<T> F, unsigned long long mnx,
                                                       unsigned long long mny, const TP* __restrict__ p,
                                                       const double* __restrict__ ap, double* __restrict__ x,
                                                       double* __restrict__ r, PcgState* st, double* part) {
  pcg_axpy_first_body<TP, T>(g, F, mnx, mny, p, ap, x, r, st, part, blockIdx.x, gridDim.x);
}

// The Gauss-Seidel value of a cell from its code, right-hand side and the six
// neighbour values xn (axis-major, lower side first; entries of uncoupled or
// boundary faces are not read).
template <class T>
__device__ __forceinline__ T rb_relax(const Geo& g, uint8_t code, T bv, const int q[3], const T xn[6]) {
  if (!(code & FLUID)) return T(0);
  T diag = T(0);
  T value = bv;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (!(code & (1u << (2 * a + side)))) continue;
      const bool boundary = side == 0 ? q[a] == 0 : q[a] == g.n[a] - 1;
      if (boundary && !g.per[a]) {
        diag += (T)g.wo[a];
        continue;
      }
      const T w = (T)g.wi[a];
      diag += w;
      value += w * xn[2 * a + side];
    }
  }
  return diag > T(0) ? value / diag : T(0);
}

// The new value of the colour-`colour` cell at q (half-row index i2; u its offset
// in its colour's array): Gauss-Seidel from the other colour's values, with the
// prolongation x += cs e[parent] riding on the neighbour reads when xc is set.
template <class T>
__device__ __forceinline__ T rb_update(const Geo& g, const FineRB<T>& F, int colour, const int q[3], unsigned u,
                                       const T* __restrict__ xc, const int* __restrict__ fparent, T cs) {
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const T* __restrict__ xo = F.x + (1 - colour) * F.half;
  const long long own = colour * F.half + u;
  const uint8_t code = F.codes[own];
  const T bv = F.b[own];
  T xn[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      int qn[3];
      nb_coords(g, q, a, side, qn);
      T v = xo[(unsigned)(qn[0] >> 1) + hx * ((unsigned)qn[1] + ny * (unsigned)qn[2])];
      if (xc) {
        const int p = fparent[(unsigned)qn[0] + (unsigned)g.sy * (unsigned)qn[1] + (unsigned)g.sz * (unsigned)qn[2]];
        if (p >= 0) v += cs * xc[p];
      }
      xn[2 * a + side] = v;
    }
  return rb_relax<T>(g, code, bv, q, xn);
}

// One half-sweep of colour `colour` (rb_update per cell). out64: also store the
// cell's value, in the natural layout and FP64, there.
template <class T>
__global__ void __launch_bounds__(BS) k_rb_gs(Geo g, FineRB<T> F, int colour, const T* __restrict__ xc,
                                              const int* __restrict__ fparent, T cs, double* __restrict__ out64,
                                              const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  // a workgroup covers 64 cells of each of 4 rows (x and y); tiles are numbered
  // x, then y, then z, and each XCD takes a contiguous range of them, so the rows
  // above and below a workgroup's rows are in the same L2. The tile's position
  // is block-uniform (scalar arithmetic); no per-cell division.
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const unsigned gx = (hx + 63) >> 6, gy = (ny + 3) >> 2;
  const unsigned tile = (unsigned)xcd_tile(blockIdx.x, gridDim.x);
  const unsigned txy = tile % (gx * gy);
  const int i2 = (int)((txy % gx) * 64 + (threadIdx.x & 63));
  const int j = (int)((txy / gx) * 4 + (threadIdx.x >> 6));
  const int k = (int)(tile / (gx * gy));
  if (i2 >= (int)hx || j >= (int)ny) return;
  const int i = 2 * i2 + ((j + k + colour) & 1);
  if (i >= g.n[0]) return;
  const int q[3] = {i, j, k};
  const unsigned u = (unsigned)i2 + hx * ((unsigned)j + ny * (unsigned)k);
  const T v = rb_update<T>(g, F, colour, q, u, xc, fparent, cs);
  if (status != ST_RUNNING) return;
  F.x[colour * F.half + u] = v;
  if (out64) out64[(long long)i + j * g.sy + k * g.sz] = (double)v;
}

// Gauss-Seidel values of the colour-`colour` pair at entries i2 and i2 + 1 of row
// (j, k) (offset u in the colour's array; nx a multiple of 4, so the pair is
// aligned and complete): its right-hand sides, its neighbours in y and z and
// the row segment around it move as 16-byte (FP64) or 8-byte (FP32) pairs;
// each cell's arithmetic is rb_update's.
template <class T>
__device__ __forceinline__ void rb_pair(const Geo& g, const FineRB<T>& F, int colour, int i2, int j, int k, unsigned u,
                                        T& va, T& vb) {
  using T2 = std::conditional_t<std::is_same<T, double>::value, double2, float2>;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const int par = (j + k + colour) & 1;
  const int ia = 2 * i2 + par, ib = ia + 2;
  const T* __restrict__ xo = F.x + (1 - colour) * F.half;
  const long long own = colour * F.half + u;
  const T2 bv = *reinterpret_cast<const T2*>(F.b + own);
  const uint8_t ca = F.codes[own], cbc = F.codes[own + 1];
  // x neighbours: entries i2 - 1 + par .. i2 + 1 + par of the other colour's row
  const unsigned row = hx * ((unsigned)j + ny * (unsigned)k);
  const T2 mid = *reinterpret_cast<const T2*>(xo + row + (unsigned)i2);
  int e = par ? i2 + 2 : i2 - 1;
  if (e < 0) e = g.per[0] ? (int)hx - 1 : 0;
  if (e >= (int)hx) e = g.per[0] ? 0 : (int)hx - 1;
  const T edge = xo[row + (unsigned)e];
  T yz[4][2];
#pragma unroll
  for (int a = 1; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int qa[3] = {ia, j, k};
      int qn[3];
      nb_coords(g, qa, a, side, qn);
      const T2 v = *reinterpret_cast<const T2*>(xo + (unsigned)i2 + hx * ((unsigned)qn[1] + ny * (unsigned)qn[2]));
      yz[2 * (a - 1) + side][0] = v.x;
      yz[2 * (a - 1) + side][1] = v.y;
    }
  T xa[6], xb[6];
  xa[0] = par ? mid.x : edge;
  xa[1] = par ? mid.y : mid.x;
  xb[0] = par ? mid.y : mid.x;
  xb[1] = par ? edge : mid.y;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    xa[2 + f] = yz[f][0];
    xb[2 + f] = yz[f][1];
  }
  const int qa[3] = {ia, j, k}, qb[3] = {ib, j, k};
  va = rb_relax<T>(g, ca, bv.x, qa, xa);
  vb = rb_relax<T>(g, cbc, bv.y, qb, xb);
}

// One half-sweep with two cells of the colour per thread (entries i2 and i2 + 1
// of a colour row, tiles of 128 x 4). Needs nx a multiple of 4. With xc (the
// prolongation riding on the reads) each cell takes rb_update itself; the
// V-cycle launches that pass one cell per thread.
template <class T>
__global__ void __launch_bounds__(BS) k_rb_gs2(Geo g, FineRB<T> F, int colour, const T* __restrict__ xc,
                                               const int* __restrict__ fparent, T cs, double* __restrict__ out64,
                                               const PcgState* st) {
  using T2 = std::conditional_t<std::is_same<T, double>::value, double2, float2>;
  const int status = st ? st->status : ST_RUNNING;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const unsigned gx = (hx + 127) >> 7, gy = (ny + 3) >> 2;
  const unsigned tile = (unsigned)xcd_tile(blockIdx.x, gridDim.x);
  const unsigned txy = tile % (gx * gy);
  const int i2 = (int)((txy % gx) * 128 + 2 * (threadIdx.x & 63));
  const int j = (int)((txy / gx) * 4 + (threadIdx.x >> 6));
  const int k = (int)(tile / (gx * gy));
  if (i2 >= (int)hx || j >= (int)ny) return;
  const int par = (j + k + colour) & 1;
  const int ia = 2 * i2 + par, ib = ia + 2;  // ib < nx: nx is a multiple of 4
  const unsigned u = (unsigned)i2 + hx * ((unsigned)j + ny * (unsigned)k);
  T va, vb;
  if (xc) {
    const int qa[3] = {ia, j, k}, qb[3] = {ib, j, k};
    va = rb_update<T>(g, F, colour, qa, u, xc, fparent, cs);
    vb = rb_update<T>(g, F, colour, qb, u + 1, xc, fparent, cs);
  } else {
    rb_pair<T>(g, F, colour, i2, j, k, u, va, vb);
  }
  if (status != ST_RUNNING) return;
  T2 out;
  out.x = va;
  out.y = vb;
  *reinterpret_cast<T2*>(F.x + colour * F.half + u) = out;
  if (out64) {
    const long long c = (long long)ia + j * g.sy + k * g.sz;
    out64[c] = (double)va;
    out64[c + 2] = (double)vb;
  }
}


// The same half-sweep with consecutive threads on consecutive pairs of the
// colour's array (no idle lanes where a row is not a multiple of the tile
// width, whole-wave linear streams); each XCD takes a contiguous range of blocks.
// Threads of a pair half-sweep workgroup: 512 for the FP64 level of one grid
// (69.4 -> 67.8 us on the 384x384x160 canyon), 256 for the FP32 level (39.2
// against 41.3 us with 512) and for batches of smaller grids.
template <class T>
constexpr int gs2_threads() {
  return std::is_same<T, double>::value ? 512 : BS;
}
template <class T>
__device__ __forceinline__ void rb_gs2_lin_body(const Geo& g, const FineRB<T>& F, int colour,
                                                double* __restrict__ out64, const PcgState* st, unsigned bid,
                                                unsigned nblk) {
  using T2 = std::conditional_t<std::is_same<T, double>::value, double2, float2>;
  const int status = st ? st->status : ST_RUNNING;
  const unsigned pair = (unsigned)xcd_tile(bid, nblk) * blockDim.x + threadIdx.x;
  const unsigned u = 2 * pair;
  if (u >= (unsigned)F.half) return;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const unsigned row = div_magic(u, F.mhx);
  const int i2 = (int)(u - row * hx);
  const unsigned k = div_magic(row, F.mny);
  const int j = (int)(row - k * ny);
  T va, vb;
  rb_pair<T>(g, F, colour, i2, j, (int)k, u, va, vb);
  if (status != ST_RUNNING) return;
  T2 out;
  out.x = va;
  out.y = vb;
  *reinterpret_cast<T2*>(F.x + colour * F.half + u) = out;
  if (out64) {
    const int ia = 2 * i2 + ((j + (int)k + colour) & 1);
    const long long c = (long long)ia + j * g.sy + (long long)k * g.sz;
    out64[c] = (double)va;
    out64[c + 2] = (double)vb;
  }
}
template <class T>
__global__ void __launch_bounds__(512) k_rb_gs2_lin(Geo g, FineRB<T> F, int colour, double* __restrict__ out64,
                                                   const PcgState* st) {
  rb_gs2_lin_body<T>(g, F, colour, out64, st, blockIdx.x, gridDim.x);
}

// The coarse correction of the red cells, x += cs e[parent] (the value that the
// black half-sweep after the prolongation would otherwise form at each of its
// neighbour reads, with the same arithmetic): the red cells are recomputed from
// scratch by the next red half-sweep, so only that black half-sweep reads these
// values.
template <class T>
__device__ __forceinline__ void rb_prolong_body(long long half, T* __restrict__ xr, const int* __restrict__ rpar,
                                                const T* __restrict__ xc, T cs, const PcgState* st, unsigned bid) {
  const int status = st ? st->status : ST_RUNNING;
  const long long u = (long long)bid * BS + threadIdx.x;
  if (u >= half) return;
  const int p = rpar[u];
  if (p < 0 || status != ST_RUNNING) return;
  T v = xr[u];
  v += cs * xc[p];
  xr[u] = v;
}
template <class T>
__global__ void __launch_bounds__(BS) k_rb_prolong(long long half, T* __restrict__ xr, const int* __restrict__ rpar,
                                                   const T* __restrict__ xc, T cs, const PcgState* st) {
  rb_prolong_body<T>(half, xr, rpar, xc, cs, st, blockIdx.x);
}
// The same with two red cells per thread (pair loads of x and the parents;
// half even).
template <class T>
__device__ __forceinline__ void rb_prolong2_body(long long half, T* __restrict__ xr, const int* __restrict__ rpar,
                                                 const T* __restrict__ xc, T cs, const PcgState* st, unsigned bid) {
  using T2 = std::conditional_t<std::is_same<T, double>::value, double2, float2>;
  const int status = st ? st->status : ST_RUNNING;
  const long long u = 2 * ((long long)bid * BS + threadIdx.x);
  if (u >= half || status != ST_RUNNING) return;
  const int2 p = *reinterpret_cast<const int2*>(rpar + u);
  if (p.x < 0 && p.y < 0) return;
  T2 v = *reinterpret_cast<const T2*>(xr + u);
  if (p.x >= 0) v.x += cs * xc[p.x];
  if (p.y >= 0) v.y += cs * xc[p.y];
  *reinterpret_cast<T2*>(xr + u) = v;
}
template <class T>
__global__ void __launch_bounds__(BS) k_rb_prolong2(long long half, T* __restrict__ xr, const int* __restrict__ rpar,
                                                    const T* __restrict__ xc, T cs, const PcgState* st) {
  rb_prolong2_body<T>(half, xr, rpar, xc, cs, st, blockIdx.x);
}

// The last post-smoothing half-sweep (colour 0) together with the PCG's r . z
// (pressure.rs:428) for a grid without closed components, whose z needs no
// centering: each thread relaxes its red cell and takes the black cell beside it
// (final since the previous half-sweep), writes both into z (out64, natural
// layout: each wave fills whole cache lines) and adds both to r . z; the fixed
// reduction grid (RP blocks of RB threads, tiles of 64 x 16 cells taken in turn)
// writes the partials that k_pcg_rz would; each XCD takes a contiguous share of
// the tiles.
template <class T>
__global__ void __launch_bounds__(RB, 8) k_rb_gs_rz(Geo g, FineRB<T> F, double* __restrict__ out64,
                                                 const double* __restrict__ r, const PcgState* st, double* part) {
  if (!block_running(st)) return;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1], nx = (unsigned)g.n[0];
  const unsigned gx = (hx + 63) >> 6, gy = (ny + 15) >> 4, tiles = gx * gy * (unsigned)g.n[2];
  double acc = 0.0;
  const XcdShare sh = xcd_share(tiles, 1, 0);
  for (unsigned tile = sh.first; tile < sh.end; tile += sh.step) {
    const unsigned txy = tile % (gx * gy);
    const int i2 = (int)((txy % gx) * 64 + (threadIdx.x & 63));
    const int j = (int)((txy / gx) * 16 + (threadIdx.x >> 6));
    const int k = (int)(tile / (gx * gy));
    if (i2 >= (int)hx || j >= (int)ny) continue;
    const unsigned u = (unsigned)i2 + hx * ((unsigned)j + ny * (unsigned)k);
    const long long row = (long long)j * g.sy + k * g.sz;
    const int ir = 2 * i2 + ((j + k) & 1), ib = 2 * i2 + ((j + k + 1) & 1);
    if (ir < (int)nx) {
      const int q[3] = {ir, j, k};
      const T v = rb_update<T>(g, F, 0, q, u, (const T*)nullptr, (const int*)nullptr, T(0));
      F.x[u] = v;
      out64[row + ir] = (double)v;
      acc += r[row + ir] * (double)v;
    }
    if (ib < (int)nx) {
      const T vb = F.x[F.half + u];
      out64[row + ib] = (double)vb;
      acc += r[row + ib] * (double)vb;
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[P_RZ + blockIdx.x] = s;
}

// The red pair at entries i2, i2 + 1 of row (j, k) (offset u) relaxed, and the
// four natural cells 2 i2 .. 2 i2 + 3 (red and black alternating) with z written
// as two 16-byte pairs and r . z added to acc in the order red a, black a, red
// b, black b. RES: r is the fine level's right-hand side F.b (both colours; the
// PCG keeps its residual there) and z stays in F.x (nothing written in the
// natural layout).
// TR (RES): the type of the colour-separated residual; TR == T reads it from
// F.b (the FP64 solve), TR double with T float from r (the mixed-precision
// solve keeps its FP64 residual in the FP64 level's b arrays, in the same
// layout).
template <class T, bool RES, class TR = T>
__device__ __forceinline__ void rz_pair(const Geo& g, const FineRB<T>& F, double* __restrict__ out64,
                                        float* __restrict__ out32, const double* __restrict__ r, int i2, int j, int k,
                                        unsigned u, double& acc) {
  using T2 = std::conditional_t<std::is_same<T, double>::value, double2, float2>;
  const long long row = (long long)j * g.sy + k * g.sz;
  const int par = (j + k) & 1;  // red at 2 i2 + par and 2 i2 + 2 + par
  T va, vb;
  rb_pair<T>(g, F, 0, i2, j, k, u, va, vb);
  T2 red;
  red.x = va;
  red.y = vb;
  *reinterpret_cast<T2*>(F.x + u) = red;
  const T2 black = *reinterpret_cast<const T2*>(F.x + F.half + u);
  double2 r01, r23;
  if (RES) {
    using R2 = std::conditional_t<std::is_same<TR, double>::value, double2, float2>;
    const TR* rs;
    if constexpr (std::is_same<TR, T>::value)
      rs = F.b;
    else
      rs = r;
    const R2 rr = *reinterpret_cast<const R2*>(rs + u);
    const R2 rk = *reinterpret_cast<const R2*>(rs + F.half + u);
    r01.x = (double)(par ? rk.x : rr.x);
    r01.y = (double)(par ? rr.x : rk.x);
    r23.x = (double)(par ? rk.y : rr.y);
    r23.y = (double)(par ? rr.y : rk.y);
  } else {
    r01 = *reinterpret_cast<const double2*>(r + row + 2 * i2);
    r23 = *reinterpret_cast<const double2*>(r + row + 2 * i2 + 2);
  }
  double2 z01, z23;
  if (par == 0) {
    z01.x = (double)va;
    z01.y = (double)black.x;
    z23.x = (double)vb;
    z23.y = (double)black.y;
    acc += r01.x * z01.x;
    acc += r01.y * z01.y;
    acc += r23.x * z23.x;
    acc += r23.y * z23.y;
  } else {
    z01.x = (double)black.x;
    z01.y = (double)va;
    z23.x = (double)black.y;
    z23.y = (double)vb;
    acc += r01.y * z01.y;
    acc += r01.x * z01.x;
    acc += r23.y * z23.y;
    acc += r23.x * z23.x;
  }
  if (!RES && out64) {
    *reinterpret_cast<double2*>(out64 + row + 2 * i2) = z01;
    *reinterpret_cast<double2*>(out64 + row + 2 * i2 + 2) = z23;
  } else if (!RES && out32) {
    // z of an FP32 cycle: every value is an FP32 value, stored exactly
    float4 zf;
    zf.x = (float)z01.x;
    zf.y = (float)z01.y;
    zf.z = (float)z23.x;
    zf.w = (float)z23.y;
    *reinterpret_cast<float4*>(out32 + row + 2 * i2) = zf;
  }
}

// k_rb_gs_rz with two red cells per thread (entries i2 and i2 + 1 of a colour
// row, tiles of 128 x 16 taken in turn by the fixed reduction grid; needs nx a
// multiple of 4).
template <class T, bool RES>
__global__ void __launch_bounds__(RB, 8) k_rb_gs_rz2(Geo g, FineRB<T> F, double* __restrict__ out64,
                                                  float* __restrict__ out32, const double* __restrict__ r,
                                                  const PcgState* st, double* part) {
  if (!block_running(st)) return;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const unsigned gx = (hx + 127) >> 7, gy = (ny + 15) >> 4, tiles = gx * gy * (unsigned)g.n[2];
  double acc = 0.0;
  const XcdShare sh = xcd_share(tiles, 1, 0);
  for (unsigned tile = sh.first; tile < sh.end; tile += sh.step) {
    const unsigned txy = tile % (gx * gy);
    const int i2 = (int)((txy % gx) * 128 + 2 * (threadIdx.x & 63));
    const int j = (int)((txy / gx) * 16 + (threadIdx.x >> 6));
    const int k = (int)(tile / (gx * gy));
    if (i2 >= (int)hx || j >= (int)ny) continue;
    const unsigned u = (unsigned)i2 + hx * ((unsigned)j + ny * (unsigned)k);
    rz_pair<T, RES>(g, F, out64, out32, r, i2, j, k, u, acc);
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[P_RZ + blockIdx.x] = s;
}

// The same with consecutive threads on consecutive red pairs: items of RB pairs
// in order, each XCD taking a contiguous share of them (no idle lanes where a
// row is not a multiple of the tile width).
template <class T, bool RES, class TR = T>
__device__ __forceinline__ void rb_gs_rz2_lin_body(const Geo& g, const FineRB<T>& F, double* __restrict__ out64,
                                                   float* __restrict__ out32, const double* __restrict__ r,
                                                   const PcgState* st, double* part, unsigned bid, unsigned nblk) {
  if (!block_running(st)) return;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const unsigned pairs = (unsigned)(F.half / 2);
  double acc = 0.0;
  const XcdShare sh = xcd_share_at(bid, nblk, (pairs + RB - 1) / RB, 1, 0);
  for (unsigned item = sh.first; item < sh.end; item += sh.step) {
    const unsigned pair = item * RB + threadIdx.x;
    if (pair >= pairs) continue;
    const unsigned u = 2 * pair;
    const unsigned row = div_magic(u, F.mhx);
    const int i2 = (int)(u - row * hx);
    const unsigned k = div_magic(row, F.mny);
    rz_pair<T, RES, TR>(g, F, out64, out32, r, i2, (int)(row - k * ny), (int)k, u, acc);
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[P_RZ + bid] = s;
}
template <class T, bool RES, class TR = T>
__global__ void __launch_bounds__(RB, 8) k_rb_gs_rz2_lin(Geo g, FineRB<T> F, double* __restrict__ out64,
                                                      float* __restrict__ out32, const double* __restrict__ r,
                                                      const PcgState* st, double* part) {
  rb_gs_rz2_lin_body<T, RES, TR>(g, F, out64, out32, r, st, part, blockIdx.x, gridDim.x);
}

// Workgroups of a half-sweep: tiles of 64 x 4 cells of a colour (x, y) per z layer.
template <class T>
unsigned rb_blocks(const Geo& g, const FineRB<T>& F) {
  return (unsigned)(((F.hx + 63) / 64) * ((g.n[1] + 3) / 4) * g.n[2]);
}

// One fine half-sweep: two cells per thread when the grid chose so at set-up.
template <class T>
void launch_rb_half(const Geo& g0, const FineRB<T>& F, int colour, const T* xc, const int* fparent, T cs,
                    double* out64, const PcgState* st, hipStream_t s) {
  // the pass carrying the prolongation keeps one cell per thread (its neighbour
  // reads gather the coarse correction cell by cell)
  // per swept cell: its code, b and new x, and the other colour's x (each value
  // read once), and z when written
  const double bytes = (double)F.half * (1 + 3 * sizeof(T) + (out64 ? 8 : 0));
  if (F.two_cells == 2 && g0.n[0] % 4 == 0 && !xc)
    PROBED_LAUNCH(PROBE_FINE_HALF, bytes, (k_rb_gs2_lin<T>), dim3((unsigned)((F.half / 2 + gs2_threads<T>() - 1) / gs2_threads<T>())), dim3(gs2_threads<T>()), 0,
                  s, g0, F, colour, out64, st);
  else if (F.two_cells && g0.n[0] % 4 == 0 && !xc)
    PROBED_LAUNCH(PROBE_FINE_HALF, bytes, (k_rb_gs2<T>), dim3(((F.hx + 127) / 128) * ((g0.n[1] + 3) / 4) * g0.n[2]),
                  dim3(BS), 0, s, g0, F, colour, xc, fparent, cs, out64, st);
  else
    PROBED_LAUNCH(PROBE_FINE_HALF, bytes, (k_rb_gs<T>), dim3(rb_blocks(g0, F)), dim3(BS), 0, s, g0, F, colour, xc,
                  fparent, cs, out64, st);
}

// Fine residual of cell q from the colour-separated layout.
template <class T>
__device__ __forceinline__ T rb_resid(const Geo& g, uint8_t code, T bv, T xv, const int q[3], const T xn[6]);
template <class T>
__device__ __forceinline__ T rb_residual(const Geo& g, const FineRB<T>& F, const int q[3]) {
  const int colour = (q[0] + q[1] + q[2]) & 1;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  auto at = [&](const int v[3]) { return (unsigned)(v[0] >> 1) + hx * ((unsigned)v[1] + ny * (unsigned)v[2]); };
  const long long own = colour * F.half + at(q);
  const T* __restrict__ xo = F.x + (1 - colour) * F.half;
  const uint8_t code = F.codes[own];
  const T bv = F.b[own];
  const T xv = F.x[own];
  T xn[6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      int qn[3];
      nb_coords(g, q, a, side, qn);
      xn[2 * a + side] = xo[at(qn)];
    }
  return rb_resid<T>(g, code, bv, xv, q, xn);
}

// The residual b - A x of a cell from its code, b, x and the six neighbour
// values (axis-major, lower side first; uncoupled sides not used).
template <class T>
__device__ __forceinline__ T rb_resid(const Geo& g, uint8_t code, T bv, T xv, const int q[3], const T xn[6]) {
  if (!(code & FLUID)) return T(0);
  T ax = T(0), off = T(0);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      if (!(code & (1u << (2 * a + side)))) continue;
      const bool boundary = side == 0 ? q[a] == 0 : q[a] == g.n[a] - 1;
      if (boundary && !g.per[a]) {
        ax += (T)g.wo[a];
        continue;
      }
      ax += (T)g.wi[a];
      off += (T)g.wi[a] * xn[2 * a + side];
    }
  return bv - (ax * xv - off);
}

// Restriction from the colour-separated fine level, as k_agg_fine_restrict: a
// workgroup takes one row of 64 2x2x2 bins and wavefront w the cells at
// (dy, dz) = (w & 1, w >> 1) of every bin (the bin's red cell there and the black
// one beside it), so its loads are as coalesced as a half-sweep's. A bin at even
// coordinates holds its red cells at subs 0, 3, 5, 6 and its black ones at 1, 2,
// 4, 7 (sub = dx + 2 dy + 4 dz); the four red residuals and eight parents meet in
// LDS and the first wavefront sums each parent's share in sub order.
template <class T>
__device__ __forceinline__ void rb_restrict_body(const Geo& g, const FineRB<T>& F, const int* __restrict__ fparent,
                                                 T* __restrict__ bc, T* __restrict__ xc, const T* __restrict__ cdiag,
                                                 int red_end, const PcgState* st, unsigned bid) {
  __shared__ T sr[4][64];
  __shared__ int sp[4][64], spb[4][64];
  const int status = st ? st->status : ST_RUNNING;
  const unsigned bx = (unsigned)(g.n[0] + 1) >> 1, by = (unsigned)(g.n[1] + 1) >> 1;
  const unsigned gbx = (bx + 63) >> 6;
  const unsigned tile = bid, rest = tile / gbx;
  const unsigned lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned Bx = (tile % gbx) * 64 + lane, By = rest % by, Bz = rest / by;
  const int dy = (int)(w & 1), dz = (int)(w >> 1);
  // red cell of the sub-row: i + j + k even with j = 2 By + dy, k = 2 Bz + dz
  const int dxr = (dy + dz) & 1;
  const int j = (int)(2 * By) + dy, k = (int)(2 * Bz) + dz;
  T r = T(0);
  int p = -1, pb = -1;
  // a block of one parent needs no per-cell parent load
  const int bp = Bx < bx ? F.bpar[Bx + bx * (By + by * Bz)] : kMixedParent;
  if (Bx < bx && j < g.n[1] && k < g.n[2]) {
    const int ir = (int)(2 * Bx) + dxr, ib = (int)(2 * Bx) + (1 - dxr);
    const unsigned row = (unsigned)g.sy * (unsigned)j + (unsigned)g.sz * (unsigned)k;
    if (ir < g.n[0]) {
      const int q[3] = {ir, j, k};
      p = bp != kMixedParent ? bp : fparent[row + (unsigned)ir];
      r = rb_residual<T>(g, F, q);
    }
    // the black cells were relaxed last: their residuals vanish (taken as zero)
    if (ib < g.n[0]) pb = bp != kMixedParent ? bp : fparent[row + (unsigned)ib];
  }
  sr[w][lane] = r;
  sp[w][lane] = p;
  spb[w][lane] = pb;
  __syncthreads();
  if (w != 0 || Bx >= bx || status != ST_RUNNING) return;
  // sub = dx + 2 dy + 4 dz: red subs 0, 3, 5, 6 are sub-rows 0, 1, 2, 3; black
  // subs 1, 2, 4, 7 likewise
  const T rs[8] = {sr[0][lane], T(0), T(0), sr[1][lane], T(0), sr[2][lane], sr[3][lane], T(0)};
  const int ps[8] = {sp[0][lane], spb[0][lane], spb[1][lane], sp[1][lane], spb[2][lane], sp[2][lane], sp[3][lane],
                     spb[3][lane]};
  unsigned done = 0;
#pragma unroll
  for (int s2 = 0; s2 < 8; ++s2) {
    if (ps[s2] < 0 || (done & (1u << s2))) continue;
    T sum = T(0);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u >= s2 && ps[u] == ps[s2]) {
        sum += rs[u];
        done |= 1u << u;
      }
    bc[ps[s2]] = sum;
    xc[ps[s2]] = first_relax<T>(ps[s2], red_end, sum, cdiag);
  }
}
template <class T>
__global__ void __launch_bounds__(BS) k_rb_restrict(Geo g, FineRB<T> F, const int* __restrict__ fparent,
                                                     T* __restrict__ bc, T* __restrict__ xc,
                                                     const T* __restrict__ cdiag, int red_end, const PcgState* st) {
  rb_restrict_body<T>(g, F, fparent, bc, xc, cdiag, red_end, st, blockIdx.x);
}

// The residuals of the red cells at entries i2, i2 + 1 of colour row (j, k)
// (offset u): rb_pair's loads (16-byte pairs of b, x and the black neighbours),
// rb_residual's arithmetic, so the values are bitwise rb_residual's.
template <class T>
__device__ __forceinline__ void rb_residual_pair(const Geo& g, const FineRB<T>& F, int i2, int j, int k, unsigned u,
                                                 T& ra, T& rb) {
  using T2 = std::conditional_t<std::is_same<T, double>::value, double2, float2>;
  const unsigned hx = (unsigned)F.hx, ny = (unsigned)g.n[1];
  const int par = (j + k) & 1;
  const int ia = 2 * i2 + par, ib = ia + 2;
  const T* __restrict__ xo = F.x + F.half;
  const T2 bv = *reinterpret_cast<const T2*>(F.b + u);
  const T2 xv = *reinterpret_cast<const T2*>(F.x + u);
  const uint8_t ca = F.codes[u], cbc = F.codes[u + 1];
  const unsigned row = hx * ((unsigned)j + ny * (unsigned)k);
  const T2 mid = *reinterpret_cast<const T2*>(xo + row + (unsigned)i2);
  int e = par ? i2 + 2 : i2 - 1;
  if (e < 0) e = g.per[0] ? (int)hx - 1 : 0;
  if (e >= (int)hx) e = g.per[0] ? 0 : (int)hx - 1;
  const T edge = xo[row + (unsigned)e];
  T yz[4][2];
#pragma unroll
  for (int a = 1; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int qa[3] = {ia, j, k};
      int qn[3];
      nb_coords(g, qa, a, side, qn);
      const T2 v = *reinterpret_cast<const T2*>(xo + (unsigned)i2 + hx * ((unsigned)qn[1] + ny * (unsigned)qn[2]));
      yz[2 * (a - 1) + side][0] = v.x;
      yz[2 * (a - 1) + side][1] = v.y;
    }
  T xa[6], xb[6];
  xa[0] = par ? mid.x : edge;
  xa[1] = par ? mid.y : mid.x;
  xb[0] = par ? mid.y : mid.x;
  xb[1] = par ? edge : mid.y;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    xa[2 + f] = yz[f][0];
    xb[2 + f] = yz[f][1];
  }
  const int qa[3] = {ia, j, k}, qb[3] = {ib, j, k};
  ra = rb_resid<T>(g, ca, bv.x, xv.x, qa, xa);
  rb = rb_resid<T>(g, cbc, bv.y, xv.y, qb, xb);
}

// rb_restrict_body with two bins per thread (nx a multiple of 4): a workgroup of
// 128 threads takes one row of 64 2x2x2 bins, each half-wavefront one (dy, dz)
// sub-row of it, a thread the bins 2 h and 2 h + 1 of the row (their red cells
// at consecutive entries of the colour row: pair loads, as the pair
// half-sweeps). The residuals and parents meet in LDS as before and the first
// wavefront sums each bin's parents in sub order: bitwise rb_restrict_body.
constexpr int RS2 = 128;
template <class T>
__device__ __forceinline__ void rb_restrict2_body(const Geo& g, const FineRB<T>& F, const int* __restrict__ fparent,
                                                  T* __restrict__ bc, T* __restrict__ xc, const T* __restrict__ cdiag,
                                                  int red_end, const PcgState* st, unsigned bid) {
  __shared__ T sr[4][64];
  __shared__ int sp[4][64], spb[4][64];
  const int status = st ? st->status : ST_RUNNING;
  const unsigned bx = (unsigned)g.n[0] >> 1, by = (unsigned)(g.n[1] + 1) >> 1;
  const unsigned gbx = (bx + 63) >> 6;
  const unsigned tile = bid, rest = tile / gbx;
  const unsigned sub = threadIdx.x >> 5, h = threadIdx.x & 31;
  const unsigned B0 = (tile % gbx) * 64 + 2 * h, By = rest % by, Bz = rest / by;
  const int dy = (int)(sub & 1), dz = (int)(sub >> 1);
  const int dxr = (dy + dz) & 1;
  const int j = (int)(2 * By) + dy, k = (int)(2 * Bz) + dz;
  T r0 = T(0), r1 = T(0);
  int p0 = -1, p1 = -1, pb0 = -1, pb1 = -1;
  // bx is even (nx a multiple of 4), so both bins of a pair are in the row or neither
  if (B0 < bx && j < g.n[1] && k < g.n[2]) {
    const int2 bp = *reinterpret_cast<const int2*>(F.bpar + B0 + bx * (By + by * Bz));
    const unsigned row = (unsigned)g.sy * (unsigned)j + (unsigned)g.sz * (unsigned)k;
    const int ir = (int)(2 * B0) + dxr, ib = (int)(2 * B0) + (1 - dxr);
    p0 = bp.x != kMixedParent ? bp.x : fparent[row + (unsigned)ir];
    p1 = bp.y != kMixedParent ? bp.y : fparent[row + (unsigned)ir + 2];
    // the black cells were relaxed last: their residuals vanish (taken as zero)
    pb0 = bp.x != kMixedParent ? bp.x : fparent[row + (unsigned)ib];
    pb1 = bp.y != kMixedParent ? bp.y : fparent[row + (unsigned)ib + 2];
    const unsigned u = B0 + (unsigned)F.hx * ((unsigned)j + (unsigned)g.n[1] * (unsigned)k);
    rb_residual_pair<T>(g, F, (int)B0, j, k, u, r0, r1);
  }
  sr[sub][2 * h] = r0;
  sr[sub][2 * h + 1] = r1;
  sp[sub][2 * h] = p0;
  sp[sub][2 * h + 1] = p1;
  spb[sub][2 * h] = pb0;
  spb[sub][2 * h + 1] = pb1;
  __syncthreads();
  const unsigned lane = threadIdx.x, Bx = (tile % gbx) * 64 + lane;
  if (lane >= 64 || Bx >= bx || status != ST_RUNNING) return;
  const T rs[8] = {sr[0][lane], T(0), T(0), sr[1][lane], T(0), sr[2][lane], sr[3][lane], T(0)};
  const int ps[8] = {sp[0][lane], spb[0][lane], spb[1][lane], sp[1][lane], spb[2][lane], sp[2][lane], sp[3][lane],
                     spb[3][lane]};
  unsigned done = 0;
#pragma unroll
  for (int s2 = 0; s2 < 8; ++s2) {
    if (ps[s2] < 0 || (done & (1u << s2))) continue;
    T sum = T(0);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (u >= s2 && ps[u] == ps[s2]) {
        sum += rs[u];
        done |= 1u << u;
      }
    bc[ps[s2]] = sum;
    xc[ps[s2]] = first_relax<T>(ps[s2], red_end, sum, cdiag);
  }
}
template <class T>
__global__ void __launch_bounds__(RS2) k_rb_restrict2(Geo g, FineRB<T> F, const int* __restrict__ fparent,
                                                       T* __restrict__ bc, T* __restrict__ xc,
                                                       const T* __restrict__ cdiag, int red_end, const PcgState* st) {
  rb_restrict2_body<T>(g, F, fparent, bc, xc, cdiag, red_end, st, blockIdx.x);
}

// ---------------------------------------------------------------- multigrid (the reference's cycle)
// The reference's V-cycle (multigrid.rs:667-726, masked.rs:304-342) with its
// lexicographic Gauss-Seidel sweeps executed by dependency level: on the fine
// grid the hyperplanes i + j + k (a row's lower-numbered neighbours lie on the
// plane before it, its higher-numbered ones after it, wrap neighbours
// included), on the CSR levels host-computed level sets. Each row repeats the
// serial sweep's arithmetic in its entry order, so the sweeps produce the
// serial values. FP64 only.

// Fine-level row in the reference's entry order (faces in face order: per axis
// the lower face first, the wrap face first on the last cell of a periodic
// axis). acc + sign * sum w x_j over the interior faces; a periodic axis of one
// cell (bit a of self_mask) contributes its self face twice.
__device__ __forceinline__ double ref_fine_acc(const Geo& g, uint8_t code, const int q[3], const double nv[6],
                                               double xc, double acc, bool subtract, int self_mask) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double w = g.wi[a];
    if ((self_mask >> a) & 1) {
      acc = subtract ? acc - w * xc : acc + w * xc;
      acc = subtract ? acc - w * xc : acc + w * xc;
      continue;
    }
    const int t = q[a], n = g.n[a];
    const bool P = g.per[a];
    const bool wrap_first = P && t == n - 1;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int side = (k == 0) == wrap_first ? 1 : 0;
      if (!(code & (1u << (2 * a + side)))) continue;
      const bool boundary = side == 0 ? t == 0 : t == n - 1;
      if (boundary && !P) continue;  // outlet: in the diagonal only
      acc = subtract ? acc - w * nv[2 * a + side] : acc + w * nv[2 * a + side];
    }
  }
  return acc;
}
__device__ __forceinline__ double ref_fine_row(const Geo& g, uint8_t code, long long c, const int q[3],
                                               const double* __restrict__ x, double acc, bool subtract, int self_mask) {
  double nv[6];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    nv[2 * a] = x[nb_or_self(g, c, a, q[a], 0)];
    nv[2 * a + 1] = x[nb_or_self(g, c, a, q[a], 1)];
  }
  return ref_fine_acc(g, code, q, nv, x[c], acc, subtract, self_mask);
}

// Gauss-Seidel on hyperplane i + j + k = L: x = (rhs + sum w x_j) / a_ii, which
// is the reference's rhs - sum a_ij x_j with a_ij = -w (Matrix::sweep,
// masked.rs:158-172).
__global__ void __launch_bounds__(BS) k_ref_fine_gs(Geo g, const uint8_t* __restrict__ codes,
                                                    const double* __restrict__ diag, const double* __restrict__ rhs,
                                                    double* __restrict__ x, int L, int j0, int k0, int nj, int nk,
                                                    int self_mask, const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= nj * nk) return;
  const int j = j0 + t % nj, k = k0 + t / nj;
  const int i = L - j - k;
  if (i < 0 || i >= g.n[0]) return;
  const long long c = (long long)i + j * g.sy + k * g.sz;
  const uint8_t code = codes[c];
  if (!(code & FLUID)) return;
  const int q[3] = {i, j, k};
  const double v = ref_fine_row(g, code, c, q, x, rhs[c], false, self_mask);
  const double d = diag[c];
  if (status != ST_RUNNING) return;
  // a sealed singleton (no couplings, no outlet) has no unknown: it stays zero
  // (masked.rs:161-166)
  x[c] = d != 0.0 ? v / d : 0.0;
}

// The same sweep over the tiles of k_ic0_tile: one launch per tile diagonal
// (forward ascending, backward descending), and in a tile one wavefront walks
// its local hyperplanes with a barrier between them. Every neighbour of a tile
// lies in the tile or on a neighbouring diagonal, a wrap neighbour of a
// periodic axis on an earlier diagonal for the sweep that takes it as updated,
// so a cell finds updated (lower index forward, higher backward) the
// neighbours the serial sweep has updated: those inside the tile in LDS, the
// others, and everything not yet updated, as the launch found them. Each cell
// then repeats ref_fine_row's operations (bitwise equal to k_ref_fine_gs).
template <bool FWD, int TX, int TY, int TZ>
__global__ void __launch_bounds__(TX * TY * TZ)
    k_ref_fine_tile(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ diag,
                    const double* __restrict__ rhs, double* __restrict__ x, int self_mask, int P, int J0, int K0,
                    int nJ, int ntx, const PcgState* st) {
  constexpr int NT = TX * TY * TZ;
  __shared__ double xs[NT];
  if (st && st->status != ST_RUNNING) return;
  const int J = J0 + (int)blockIdx.x % nJ, K = K0 + (int)blockIdx.x / nJ;
  const int I = P - J - K;
  if (I < 0 || I >= ntx) return;
  const int ox = I * TX, oy = J * TY, oz = K * TZ;
  const int li = threadIdx.x;
  const int a = li % TX, b = (li / TX) % TY, cz = li / (TX * TY);
  const int q[3] = {ox + a, oy + b, oz + cz};
  const bool live = q[0] < g.n[0] && q[1] < g.n[1] && q[2] < g.n[2];
  double nv[6] = {0, 0, 0, 0, 0, 0}, xc = 0.0, rv = 0.0, d = 0.0;
  int lix[6] = {-1, -1, -1, -1, -1, -1};
  uint8_t code = 0;
  long long c = 0;
  if (live) {
    c = (long long)q[0] + q[1] * g.sy + q[2] * g.sz;
    code = codes[c];
    if (code & FLUID) {
#pragma unroll
      for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const int m = 2 * ax + side;
          const long long nb = nb_or_self(g, c, ax, q[ax], side);
          nv[m] = x[nb];
          if (!(code & (1u << m)) || !(FWD ? nb < c : nb > c)) continue;
          const int ni = (int)(nb % g.n[0]), nj = (int)((nb / g.n[0]) % g.n[1]), nk = (int)(nb / g.sz);
          if (ni >= ox && ni < ox + TX && nj >= oy && nj < oy + TY && nk >= oz && nk < oz + TZ)
            lix[m] = (ni - ox) + TX * ((nj - oy) + TY * (nk - oz));
        }
      xc = x[c];
      rv = rhs[c];
      d = diag[c];
    }
  }
  const int plane = a + b + cz;
  const int S = min(TX, g.n[0] - ox) + min(TY, g.n[1] - oy) + min(TZ, g.n[2] - oz) - 3;
  // the steps exchange values through LDS only; each cell stores its value once
  // after the last step (no global store outstanding at the steps' barriers)
  double v = 0.0;
  for (int step = 0; step <= S; ++step) {
    const int sp = FWD ? step : S - step;
    if (live && plane == sp && (code & FLUID)) {
      double v6[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) v6[m] = lix[m] >= 0 ? xs[lix[m]] : nv[m];
      const double v0 = ref_fine_acc(g, code, q, v6, xc, rv, false, self_mask);
      // a sealed singleton (no couplings, no outlet) has no unknown: it stays zero
      v = d != 0.0 ? v0 / d : 0.0;
      xs[li] = v;
    }
    __syncthreads();
  }
  if (live && (code & FLUID)) x[c] = v;
}

// r = rhs - (a_ii x + sum a_ij x_j) (Matrix::product); zero on solid cells.
__global__ void __launch_bounds__(BS) k_ref_fine_residual(Geo g, const uint8_t* __restrict__ codes,
                                                          const double* __restrict__ diag,
                                                          const double* __restrict__ rhs, const double* __restrict__ x,
                                                          double* __restrict__ r, int self_mask, const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  const long long c = (long long)blockIdx.x * BS + threadIdx.x;
  if (c >= g.N) return;
  const uint8_t code = codes[c];
  double v = 0.0;
  if (code & FLUID) {
    int q[3];
    coords(g, c, q);
    v = rhs[c] - ref_fine_row(g, code, c, q, x, diag[c] * x[c], true, self_mask);
  }
  if (status != ST_RUNNING) return;
  r[c] = v;
}

// Level l -> l + 1: sum over each coarse node's finer nodes in ascending order,
// weighted on the unmasked hierarchy (next.rhs[col] += w r in fine order); the
// coarse iterate starts at zero.
__global__ void __launch_bounds__(BS) k_ref_restrict(int n, const int* __restrict__ child_ptr,
                                                     const int* __restrict__ child, const double* __restrict__ child_w,
                                                     const double* __restrict__ r, double* __restrict__ b,
                                                     double* __restrict__ x, const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  const int J = (int)(blockIdx.x * BS + threadIdx.x);
  if (J >= n) return;
  double sum = 0.0;
  for (int k = child_ptr[J]; k < child_ptr[J + 1]; ++k)
    sum += child_w ? child_w[k] * r[child[k]] : r[child[k]];
  if (status != ST_RUNNING) return;
  b[J] = sum;
  x[J] = 0.0;
}

// Rows of one dependency level of a CSR level's sweep.
__global__ void __launch_bounds__(BS) k_ref_gs(AggDevLevel<double> L, const int* __restrict__ rows, int count,
                                               const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  const int t = (int)(blockIdx.x * BS + threadIdx.x);
  if (t >= count) return;
  const int i = rows[t];
  const int e0 = L.ptr[i], e1 = L.ptr[i + 1];
  const double d = L.diag[i];
  double v = L.b[i];
  for (int e = e0; e < e1; ++e) v -= L.val[e] * L.x[L.col[e]];
  if (status != ST_RUNNING) return;
  L.x[i] = d != 0.0 ? v / d : 0.0;
}

// A whole sweep of a CSR level inside one workgroup: its dependency levels in
// order, a barrier between them (the workgroup's stores are visible to its
// loads after the barrier), each row as in k_ref_gs: b_i - sum a_ij x_j in
// entry order, divided by a_ii. A row's entries are loaded 16 at a time before
// they are summed, so the loads of a chain of dependency levels overlap within
// each level. LDSX: the level's iterate (n <= kRefWgLds) is held in LDS for
// the sweep and written back at its end. For levels whose launches would each
// hold a few rows.
template <bool LDSX>
__global__ void __launch_bounds__(1024) k_ref_gs_wg(AggDevLevel<double> L, const int* __restrict__ rows,
                                                    const int* __restrict__ ptr, int nlev, const PcgState* st) {
  __shared__ double xs[LDSX ? kRefWgLds : 1];
  if (st && st->status != ST_RUNNING) return;
  if (LDSX) {
    for (int i = (int)threadIdx.x; i < L.n; i += (int)blockDim.x) xs[i] = L.x[i];
    __syncthreads();
  }
  for (int dl = 0; dl < nlev; ++dl) {
    const int b = ptr[dl], e = ptr[dl + 1];
    for (int t = b + (int)threadIdx.x; t < e; t += (int)blockDim.x) {
      const int i = rows[t];
      const int e0 = L.ptr[i], e1 = L.ptr[i + 1];
      const double dg = L.diag[i];
      double v = L.b[i];
      for (int k = e0; k < e1; k += 16) {
        double av[16], xv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          av[u] = 0.0;
          xv[u] = 0.0;
          if (k + u < e1) {
            const int col = L.col[k + u];
            av[u] = L.val[k + u];
            xv[u] = LDSX ? xs[col] : L.x[col];
          }
        }
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (k + u < e1) v -= av[u] * xv[u];
      }
      const double out = dg != 0.0 ? v / dg : 0.0;
      if (LDSX)
        xs[i] = out;
      else
        L.x[i] = out;
    }
    __syncthreads();
  }
  if (LDSX)
    for (int i = (int)threadIdx.x; i < L.n; i += (int)blockDim.x) L.x[i] = xs[i];
}

// A uniform int4 loaded through the scalar cache (the constant address space).
__device__ __forceinline__ int4 scalar_load(const int4* p, int i) {
  const __attribute__((address_space(4))) int* q = (const __attribute__((address_space(4))) int*)p;
  return make_int4(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
}

// The operands of one sub-level a thread holds: its entries (column -1: none)
// and, when it has one, its row with the row's entry range, rhs and diagonal.
struct RefSubOps {
  int col[kRefSubEpt];
  double val[kRefSubEpt];
  int4 ri;
  double b, d;
};

// A whole sweep of a CSR level inside one workgroup of NT threads from its
// sweep-ordered copy (RefSweep), a sub-level at a time: the workgroup's
// threads form the products a_ij x_j of every entry of the sub-level at once
// (each rounded, as the serial loop rounds it before subtracting) into LDS;
// after a barrier one thread a row subtracts its products from b_i in entry
// order and divides by a_ii, the serial row's arithmetic in its order, so the
// sweep is bitwise k_ref_gs's. Every load of a sub-level's operands is
// addressed from its meta entry alone (no load waits for another) and is
// issued while the previous sub-level is worked. LDSX: the iterate is held in
// LDS (the barriers wait for LDS only); else it stays in global memory.
template <bool LDSX, int NT>
__global__ void __launch_bounds__(NT) k_ref_gs_pk(AggDevLevel<double> L, RefSweep S, const PcgState* st) {
  extern __shared__ double shm[];
  if (st && st->status != ST_RUNNING) return;
  double* xs = shm;
  double* prod = shm + (LDSX ? L.n : 0);
  const int tid = (int)threadIdx.x;
  // the right-hand side in the sweep's order (read by this workgroup only)
  for (int r = tid; r < L.n; r += NT) S.bs[r] = L.b[S.rinfo[r].x];
  if (LDSX)
    for (int i = tid; i < L.n; i += NT) xs[i] = L.x[i];
  __syncthreads();
  auto sync = [&]() {
    if (LDSX)
      asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    else
      __syncthreads();
  };
  auto fetch = [&](const int4 m, RefSubOps& o) {
    const int r = m.x + tid;
    const bool has = r < m.y;
    o.ri = has ? S.rinfo[r] : make_int4(-1, 0, 0, 0);
    o.b = has ? S.bs[r] : 0.0;
    o.d = has ? S.dsw[r] : 0.0;
#pragma unroll
    for (int u = 0; u < kRefSubEpt; ++u) {
      const int e = m.z + tid + u * NT;
      o.col[u] = e < m.w ? S.col[e] : -1;
      o.val[u] = e < m.w ? S.val[e] : 0.0;
    }
  };
  RefSubOps cur;
  fetch(scalar_load(S.meta, 0), cur);
  int4 mnext = S.nsub > 1 ? scalar_load(S.meta, 1) : make_int4(0, 0, 0, 0);
  for (int s = 0; s < S.nsub; ++s) {
    RefSubOps nxt;
    if (s + 1 < S.nsub) fetch(mnext, nxt);
    const int4 mafter = s + 2 < S.nsub ? scalar_load(S.meta, s + 2) : make_int4(0, 0, 0, 0);
#pragma unroll
    for (int u = 0; u < kRefSubEpt; ++u)
      if (cur.col[u] >= 0) prod[tid + u * NT] = cur.val[u] * (LDSX ? xs[cur.col[u]] : L.x[cur.col[u]]);
    sync();
    if (cur.ri.x >= 0) {
      double v = cur.b;
      int k = cur.ri.y;
      for (; k + 16 <= cur.ri.z; k += 16) {
        double p[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) p[u] = prod[k + u];
#pragma unroll
        for (int u = 0; u < 16; ++u) v -= p[u];
      }
      for (; k < cur.ri.z; ++k) v -= prod[k];
      const double out = cur.d != 0.0 ? v / cur.d : 0.0;
      if (LDSX)
        xs[cur.ri.x] = out;
      else
        L.x[cur.ri.x] = out;
    }
    sync();
    cur = nxt;
    mnext = mafter;
  }
  if (LDSX)
    for (int i = tid; i < L.n; i += NT) L.x[i] = xs[i];
}

__global__ void __launch_bounds__(BS) k_ref_residual(AggDevLevel<double> L, double* __restrict__ r, const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  const int i = (int)(blockIdx.x * BS + threadIdx.x);
  if (i >= L.n) return;
  double v = L.diag[i] * L.x[i];
  for (int e = L.ptr[i]; e < L.ptr[i + 1]; ++e) v += L.val[e] * L.x[L.col[e]];
  const double out = L.b[i] - v;
  if (status != ST_RUNNING) return;
  r[i] = out;
}

// x_i += correction from the next level: its value at the parent (aggregation)
// or -0.0 + sum w x_c over the interpolation row (trilinear).
__global__ void __launch_bounds__(BS) k_ref_prolong(long long n, double* __restrict__ x, const int* __restrict__ parent,
                                                    const int* __restrict__ tptr, const int* __restrict__ tcol,
                                                    const double* __restrict__ tw, const double* __restrict__ xc,
                                                    const PcgState* st) {
  const int status = st ? st->status : ST_RUNNING;
  const long long i = (long long)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  double add;
  if (parent) {
    const int p = parent[i];
    if (p < 0) return;
    add = xc[p];
  } else {
    add = -0.0;
    for (int k = tptr[i]; k < tptr[i + 1]; ++k) add += tw[k] * xc[tcol[k]];
  }
  const double v = x[i] + add;
  if (status != ST_RUNNING) return;
  x[i] = v;
}

// Exact solve of each line of one dependency level (BlockSmoother::sweep,
// multigrid/line.rs:155-190): rhs minus the entries leaving the block, banded
// forward and backward substitution with the reciprocal diagonal, stored back.
// One thread per line; the lines of one level are not coupled to each other.
// The forward and backward substitutions of one banded block (half-bandwidth
// BW, reciprocal diagonal) in the order of k_ref_lines, with the last BW
// solved values kept in registers, so the dependent chain does not wait on LDS.
template <int BW>
__device__ __forceinline__ void line_substitute(int length, double* sv, const double* fa) {
  constexpr int stride = BW + 1;
  double w[BW];  // w[k]: the value k + 1 places back in the sweep direction
#pragma unroll
  for (int k = 0; k < BW; ++k) w[k] = 0.0;
  for (int i = 0; i < length; ++i) {
    const int row = i * stride;
    double value = sv[i];
    // j ascending from max(0, i - BW) to i - 1: w[i - 1 - j]
#pragma unroll
    for (int k = BW - 1; k >= 0; --k)
      if (i - 1 - k >= 0) value -= fa[row + 1 + k] * w[k];
    value *= fa[row];
#pragma unroll
    for (int k = BW - 1; k > 0; --k) w[k] = w[k - 1];
    w[0] = value;
    sv[i] = value;
  }
#pragma unroll
  for (int k = 0; k < BW; ++k) w[k] = 0.0;
  for (int i = length - 1; i >= 0; --i) {
    double value = sv[i];
    // j ascending from i + 1 to min(length, i + BW + 1) - 1: w[j - i - 1]
#pragma unroll
    for (int k = 0; k < BW; ++k)
      if (i + 1 + k < length) value -= fa[(i + 1 + k) * stride + 1 + k] * w[k];
    value *= fa[i * stride];
#pragma unroll
    for (int k = BW - 1; k > 0; --k) w[k] = w[k - 1];
    w[0] = value;
    sv[i] = value;
  }
}

// k_ref_lines with one wavefront per block: the lanes gather the block's
// right-hand side (rhs minus its outside entries) and its banded factor into LDS
// at once, one lane runs the forward and backward substitutions from LDS in the
// same order as k_ref_lines (bit for bit), and the lanes write the block back.
// LDS: length * (bw + 2) doubles (dynamic).
__global__ void __launch_bounds__(64) k_ref_lines_lds(int length, int bw, const int* __restrict__ nodes,
                                                      const double* __restrict__ factor, const int* __restrict__ ooff,
                                                      const int* __restrict__ ocol, const double* __restrict__ oval,
                                                      const int* __restrict__ lines, const double* __restrict__ rhs,
                                                      double* __restrict__ x, const PcgState* st) {
  extern __shared__ double lds[];
  if (st && st->status != ST_RUNNING) return;
  const int stride = bw + 1;
  double* sv = lds;            // length values
  double* fa = lds + length;   // length * stride factor entries
  const long long start = (long long)lines[blockIdx.x] * length;
  const long long base = start * stride;
  for (int i = threadIdx.x; i < length; i += 64) {
    double value = rhs[nodes[start + i]];
    for (int e = ooff[start + i]; e < ooff[start + i + 1]; ++e) value -= oval[e] * x[ocol[e]];
    sv[i] = value;
  }
  for (int i = threadIdx.x; i < length * stride; i += 64) fa[i] = factor[base + i];
  __syncthreads();
  if (threadIdx.x == 0) {
    switch (bw) {
      case 1: line_substitute<1>(length, sv, fa); break;
      case 2: line_substitute<2>(length, sv, fa); break;
      case 3: line_substitute<3>(length, sv, fa); break;
      case 4: line_substitute<4>(length, sv, fa); break;
      default:
        for (int i = 0; i < length; ++i) {
          const int row = i * stride;
          double value = sv[i];
          for (int j = i > bw ? i - bw : 0; j < i; ++j) value -= fa[row + i - j] * sv[j];
          sv[i] = value * fa[row];
        }
        for (int i = length - 1; i >= 0; --i) {
          double value = sv[i];
          const int end = length < i + bw + 1 ? length : i + bw + 1;
          for (int j = i + 1; j < end; ++j) value -= fa[j * stride + j - i] * sv[j];
          sv[i] = value * fa[i * stride];
        }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < length; i += 64) x[nodes[start + i]] = sv[i];
}

__global__ void __launch_bounds__(BS) k_ref_lines(int length, int bw, const int* __restrict__ nodes,
                                                  const double* __restrict__ factor, const int* __restrict__ ooff,
                                                  const int* __restrict__ ocol, const double* __restrict__ oval,
                                                  double* __restrict__ scr, const int* __restrict__ lines, int count,
                                                  const double* __restrict__ rhs, double* __restrict__ x,
                                                  const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
  const int t = (int)(blockIdx.x * BS + threadIdx.x);
  if (t >= count) return;
  const long long start = (long long)lines[t] * length;
  const int stride = bw + 1;
  const long long base = start * stride;
  for (int i = 0; i < length; ++i) {
    double value = rhs[nodes[start + i]];
    for (int e = ooff[start + i]; e < ooff[start + i + 1]; ++e) value -= oval[e] * x[ocol[e]];
    scr[start + i] = value;
  }
  for (int i = 0; i < length; ++i) {
    const long long row = base + (long long)i * stride;
    double value = scr[start + i];
    for (int j = i > bw ? i - bw : 0; j < i; ++j) value -= factor[row + i - j] * scr[start + j];
    scr[start + i] = value * factor[row];
  }
  for (int i = length - 1; i >= 0; --i) {
    double value = scr[start + i];
    const int end = length < i + bw + 1 ? length : i + bw + 1;
    for (int j = i + 1; j < end; ++j) value -= factor[base + (long long)j * stride + j - i] * scr[start + j];
    scr[start + i] = value * factor[base + (long long)i * stride];
  }
  for (int i = 0; i < length; ++i) x[nodes[start + i]] = scr[start + i];
}

// Centering of a coarse level's closed components (masked.rs:16-34) over chunks
// of their member lists: the sum of each chunk, then every chunk's block adds
// its component's chunk sums in order and subtracts the mean from its members.
__global__ void __launch_bounds__(BS) k_center_sum(const double* __restrict__ v, const int* __restrict__ cb,
                                                   const int* __restrict__ members, double* __restrict__ sums,
                                                   const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
  __shared__ double part[BS / 64];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int k = cb[c] + (int)threadIdx.x; k < cb[c + 1]; k += BS) s += v[members[k]];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double total = 0.0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) total += part[w];
    sums[c] = total;
  }
}
__global__ void __launch_bounds__(BS) k_center_sub(double* __restrict__ v, const int* __restrict__ cb,
                                                   const int* __restrict__ chunk_comp, const int* __restrict__ comp_chunk,
                                                   const int* __restrict__ cptr, const int* __restrict__ members,
                                                   const double* __restrict__ sums, const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
  const int c = blockIdx.x, comp = chunk_comp[c];
  double total = 0.0;
  for (int k = comp_chunk[comp]; k < comp_chunk[comp + 1]; ++k) total += sums[k];
  const double mean = total / (double)(cptr[comp + 1] - cptr[comp]);
  for (int k = cb[c] + (int)threadIdx.x; k < cb[c + 1]; k += BS) v[members[k]] -= mean;
}

// ---------------------------------------------------------------- projection
// b = -div(u)/dt and the round-off scale sum(|u_lo| + |u_hi|)/dx/dt, in the
// axis order of Grid::divergence (pressure.rs:440-459). u holds the x-, y- and
// z-face blocks of the reference's face list.
__global__ void __launch_bounds__(BS) k_projection_rhs(Geo g, FaceGeo fg, const uint8_t* __restrict__ codes,
                                                       const double* __restrict__ u, double dt, double* __restrict__ b,
                                                       double* __restrict__ terms) {
  long long c = (long long)blockIdx.x * BS + threadIdx.x;
  if (c >= g.N) return;
  if (!(codes[c] & FLUID)) {
    b[c] = -0.0;
    terms[c] = 0.0;
    return;
  }
  int q[3];
  coords(g, c, q);
  double divergence = 0.0, scale = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int ql[3] = {q[0], q[1], q[2]};
    int qh[3] = {q[0], q[1], q[2]};
    qh[a] = q[a] + 1;
    if (qh[a] == fg.dims[a][a]) qh[a] = 0;  // periodic wrap (a non-periodic block is one longer)
    long long lo = fg.off[a] + ql[0] + (long long)fg.dims[a][0] * (ql[1] + (long long)fg.dims[a][1] * ql[2]);
    long long hi = fg.off[a] + qh[0] + (long long)fg.dims[a][0] * (qh[1] + (long long)fg.dims[a][1] * qh[2]);
    double ul = u[lo], uh = u[hi];
    divergence += (uh - ul) / fg.dx[a];
    scale += (fabs(ul) + fabs(uh)) / fg.dx[a];
  }
  b[c] = -divergence / dt;
  terms[c] = scale / dt;
}

// u -= dt * grad(phi) on interior and outlet faces (pressure.rs:481-496).
__global__ void __launch_bounds__(BS) k_project_faces(Geo g, FaceGeo fg, const uint8_t* __restrict__ codes,
                                                      const double* __restrict__ phi, double dt, double* __restrict__ u) {
  long long f = (long long)blockIdx.x * BS + threadIdx.x;
  if (f >= fg.total) return;
  int a = f < fg.off[1] ? 0 : (f < fg.off[2] ? 1 : 2);
  long long lf = f - fg.off[a];
  int q[3];
  q[0] = (int)(lf % fg.dims[a][0]);
  long long r = lf / fg.dims[a][0];
  q[1] = (int)(r % fg.dims[a][1]);
  q[2] = (int)(r / fg.dims[a][1]);
  double d = 0.0;
  const double dist_in = fg.dx[a], dist_out = 0.5 * fg.dx[a];
  if (q[a] < g.n[a]) {
    long long ch = q[0] + q[1] * g.sy + q[2] * g.sz;
    if (codes[ch] & (1u << (2 * a))) {
      if (q[a] == 0 && !g.per[a]) {
        d = phi[ch] / dist_out;
      } else {
        long long cl = lo_nb(g, ch, a, q[a]);
        d = (phi[ch] - phi[cl]) / dist_in;
      }
    }
  } else {
    int ql[3] = {q[0], q[1], q[2]};
    ql[a] -= 1;
    long long cl = ql[0] + ql[1] * g.sy + ql[2] * g.sz;
    if (codes[cl] & (2u << (2 * a))) d = -phi[cl] / dist_out;
  }
  u[f] -= dt * d;
}

__global__ void __launch_bounds__(BS) k_add(long long N, double* __restrict__ p, const double* __restrict__ q) {
  long long i = (long long)blockIdx.x * BS + threadIdx.x;
  if (i < N) p[i] += q[i];
}

// Whether the operator marches through z on this grid: once every block of the
// fixed grid gets a tile.
bool march_fits(const Geo& g) {
  const MarchShape ms = march_shape(g);
  return (long long)ms.tx * ms.ty * ms.tz >= RP;
}

// ---------------------------------------------------------------- launchers
static inline unsigned blocks_for(long long n) { return (unsigned)((n + BS - 1) / BS); }

void launch_apply(const Geo& g, const uint8_t* codes, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_apply, dim3(blocks_for(g.N)), dim3(BS), 0, s, g, codes, x, y);
}
void launch_pcg_op(const Geo& g, const uint8_t* codes, const double* z, const double* pold, double* pnew, double* ap,
                   const double* mean, const int* cidx, bool closed, PcgState* st, double* part, hipStream_t s,
                   int march, long long zhalf, const float* z32) {
  const MarchShape ms = march_shape(g);
  const bool marching = march == 2 || (march == 1 && ms.tx * ms.ty * ms.tz >= RP);
  if (z32) {
    // z of the mixed-precision cycle, its FP32 values in the natural layout
    const double bytes32 = 29.0 * (double)g.N;  // z 4, p_old 8, code 1 in; p_new, Ap 16 out
    if (marching)
      PROBED_LAUNCH(PROBE_OPERATOR, bytes32, (k_pcg_op_march<false, false, float>), dim3(RP), dim3(RB), 0, s, g, ms,
                    codes, z32, pold, pnew, ap, mean, cidx, st, part, 0LL);
    else
      PROBED_LAUNCH(PROBE_OPERATOR, bytes32, (k_pcg_op<false, false, float>), dim3(RP), dim3(RB), 0, s, g, codes,
                    z32, pold, pnew, ap, mean, cidx, st, part, 0LL);
    return;
  }
  const double bytes = 33.0 * (double)g.N;  // z, p_old, code in; p_new, Ap out
  // marching pays once the grid gives every block of the fixed grid an item
  if (marching) {
    if (closed)
      PROBED_LAUNCH(PROBE_OPERATOR, bytes, (k_pcg_op_march<true, false>), dim3(RP), dim3(RB), 0, s, g, ms, codes, z,
                    pold, pnew, ap, mean, cidx, st, part, 0LL);
    else if (zhalf)
      PROBED_LAUNCH(PROBE_OPERATOR, bytes, (k_pcg_op_march<false, true>), dim3(RP), dim3(RB), 0, s, g, ms, codes, z,
                    pold, pnew, ap, mean, cidx, st, part, zhalf);
    else
      PROBED_LAUNCH(PROBE_OPERATOR, bytes, (k_pcg_op_march<false, false>), dim3(RP), dim3(RB), 0, s, g, ms, codes, z,
                    pold, pnew, ap, mean, cidx, st, part, 0LL);
  } else if (closed) {
    PROBED_LAUNCH(PROBE_OPERATOR, bytes, (k_pcg_op<true, false>), dim3(RP), dim3(RB), 0, s, g, codes, z, pold, pnew, ap,
                  mean, cidx, st, part, 0LL);
  } else if (zhalf) {
    PROBED_LAUNCH(PROBE_OPERATOR, bytes, (k_pcg_op<false, true>), dim3(RP), dim3(RB), 0, s, g, codes, z, pold, pnew, ap,
                  mean, cidx, st, part, zhalf);
  } else {
    PROBED_LAUNCH(PROBE_OPERATOR, bytes, (k_pcg_op<false, false>), dim3(RP), dim3(RB), 0, s, g, codes, z, pold, pnew,
                  ap, mean, cidx, st, part, 0LL);
  }
}
void launch_pcg_axpy(long long N, const double* p, const double* ap, double* x, double* r, PcgState* st, double* part,
                     hipStream_t s) {
  // x, p, r, Ap in; x, r out
  PROBED_LAUNCH(PROBE_AXPY, 48.0 * (double)N, k_pcg_axpy, dim3(RP), dim3(RB), 0, s, N, p, ap, x, r, st, part);
}
void launch_pcg_axpy_rb(const Geo& g, const double* p, const double* ap, double* x, double* r, PcgState* st,
                        double* part, hipStream_t s) {
  const unsigned long long two40 = 1ULL << 40;
  const unsigned long long mnx = (two40 + (unsigned long long)g.n[0] - 1) / (unsigned long long)g.n[0];
  const unsigned long long mny = (two40 + (unsigned long long)g.n[1] - 1) / (unsigned long long)g.n[1];
  PROBED_LAUNCH(PROBE_AXPY, 48.0 * (double)g.N, k_pcg_axpy_rb<double>, dim3(RP), dim3(RB), 0, s, g.N, g.n[0], g.n[1],
                mnx, mny, p, ap, x, r, st, part);
}
void launch_pcg_axpy_rb(const Geo& g, const float* p, const double* ap, double* x, double* r, PcgState* st,
                        double* part, hipStream_t s) {
  const unsigned long long two40 = 1ULL << 40;
  const unsigned long long mnx = (two40 + (unsigned long long)g.n[0] - 1) / (unsigned long long)g.n[0];
  const unsigned long long mny = (two40 + (unsigned long long)g.n[1] - 1) / (unsigned long long)g.n[1];
  // x, r, Ap 8 and p 4 in; x, r out
  PROBED_LAUNCH(PROBE_AXPY, 44.0 * (double)g.N, k_pcg_axpy_rb<float>, dim3(RP), dim3(RB), 0, s, g.N, g.n[0], g.n[1],
                mnx, mny, p, ap, x, r, st, part);
}
template <class TP, class T>
void launch_pcg_axpy_first(const Geo& g, const FineRB<T>& F, const TP* p, const double* ap, double* x, double* r,
                           PcgState* st, double* part, hipStream_t s) {
  const unsigned long long two40 = 1ULL << 40;
  const unsigned long long mnx = (two40 + (unsigned long long)g.n[0] - 1) / (unsigned long long)g.n[0];
  const unsigned long long mny = (two40 + (unsigned long long)g.n[1] - 1) / (unsigned long long)g.n[1];
  // x, r, Ap and p in; x, r out; the red cells' code in and x out; (FP32
  // level) both colours' b out
  const double bytes = (double)g.N * (40.0 + sizeof(TP)) + (double)F.half * (1.0 + sizeof(T)) +
                       (std::is_same<T, double>::value ? 0.0 : (double)g.N * sizeof(T));
  PROBED_LAUNCH(PROBE_AXPY, bytes, (k_pcg_axpy_first<TP, T>), dim3(RP), dim3(RB), 0, s, g, F, mnx, mny, p, ap, x, r,
                st, part);
}
template void launch_pcg_axpy_first<double, double>(const Geo&, const FineRB<double>&, const double*, const double*,
                                                     double*, double*, PcgState*, double*, hipStream_t);
template void launch_pcg_axpy_first<float, float>(const Geo&, const FineRB<float>&, const float*, const double*,
                                                  double*, double*, PcgState*, double*, hipStream_t);
void launch_pcg_op_mixed(const Geo& g, const uint8_t* codes, const float* z, long long zhalf, const float* pold,
                         float* pnew, double* ap, PcgState* st, double* part, hipStream_t s, int march) {
  const MarchShape ms = march_shape(g);
  const bool marching = march == 2 || (march == 1 && ms.tx * ms.ty * ms.tz >= RP);
  const double bytes = 21.0 * (double)g.N;  // z 4, p_old 4, code 1 in; p_new 4, Ap 8 out
  const double* none = nullptr;
  const int* nidx = nullptr;
  if (marching)
    PROBED_LAUNCH(PROBE_OPERATOR, bytes, (k_pcg_op_march<false, true, float, float>), dim3(RP), dim3(RB), 0, s, g, ms,
                  codes, z, pold, pnew, ap, none, nidx, st, part, zhalf);
  else
    PROBED_LAUNCH(PROBE_OPERATOR, bytes, (k_pcg_op<false, true, float, float>), dim3(RP), dim3(RB), 0, s, g, codes, z,
                  pold, pnew, ap, none, nidx, st, part, zhalf);
}
void launch_pcg_true(const Geo& g, const uint8_t* codes, const double* x, const double* bc, PcgState* st, double* part,
                     hipStream_t s) {
  probe_begin(PROBE_TRUE_RESIDUAL, s);
  hipLaunchKernelGGL(k_pcg_true, dim3(RP), dim3(RB), 0, s, g, codes, x, bc, st, part);
  hipLaunchKernelGGL(k_pcg_fin, dim3(1), dim3(RB), 0, s, st, (const double*)part);
  // the residual is recomputed (17 B per cell) only after the recurrent test has
  // passed, which the host does not know at launch: no bytes are counted
  probe_end(PROBE_TRUE_RESIDUAL, s, 0.0);
}
void launch_pcg_rz(long long N, const uint8_t* codes, const double* r, const double* z, const double* mean,
                   const int* cidx, PcgState* st, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_rz, dim3(RP), dim3(RB), 0, s, N, codes, r, z, mean, cidx, st, part);
}
void launch_pcg_init(long long N, const uint8_t* codes, const double* b, const double* mean, const int* cidx, double* r,
                     double* bc, double* x, PcgState* st, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_init, dim3(RP), dim3(RB), 0, s, N, codes, b, mean, cidx, r, bc, x, st, part);
  hipLaunchKernelGGL(k_init_fin, dim3(1), dim3(RB), 0, s, st, (const double*)part);
}
void launch_warm_residual(const Geo& g, const uint8_t* codes, const double* x0, const double* bc, double* x, double* r,
                          const PcgState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_warm_residual, dim3(blocks_for(g.N)), dim3(BS), 0, s, g, codes, x0, bc, x, r, st);
}
void launch_state_init(PcgState* st, double tol, long long max_it, hipStream_t s) {
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, st, tol, max_it);
}
void launch_gauge(long long N, const uint8_t* codes, int n_comp, int n_chunks, const int* members, const int* chunk_begin,
                  const int* comp_first_chunk, const double* inv_count, double inv_count1, const double* src,
                  const double* scale, int e031, double* mean, PcgState* st, double* part, hipStream_t s) {
  if (n_comp == 0) return;
  if (n_comp == 1)
    hipLaunchKernelGGL(k_csum1, dim3(RP), dim3(RB), 0, s, N, codes, src, scale, e031, (const PcgState*)st, part);
  else
    hipLaunchKernelGGL(k_csumN, dim3(n_chunks), dim3(RB), 0, s, members, chunk_begin, src, scale, e031, (const PcgState*)st,
                       part);
  hipLaunchKernelGGL(k_csum_fin, dim3(1), dim3(RB), 0, s, n_comp, comp_first_chunk, inv_count, inv_count1, e031, mean, st,
                     (const double*)part);
}
void launch_center(long long N, const uint8_t* codes, const double* src, const double* mean, const int* cidx, double* dst,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_center, dim3(blocks_for(N)), dim3(BS), 0, s, N, codes, src, mean, cidx, dst);
}
void launch_fill(long long N, double* dst, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(N)), dim3(BS), 0, s, N, dst, v);
}
void launch_add(long long N, double* p, const double* q, hipStream_t s) {
  hipLaunchKernelGGL(k_add, dim3(blocks_for(N)), dim3(BS), 0, s, N, p, q);
}

static void level_box(const Geo& g, int L, int& j0, int& k0, int& nj, int& nk) {
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  int klo = L - (nx - 1) - (ny - 1);
  if (klo < 0) klo = 0;
  int khi = L < nz - 1 ? L : nz - 1;
  int jlo = L - (nx - 1) - khi;
  if (jlo < 0) jlo = 0;
  int jhi = L - klo < ny - 1 ? L - klo : ny - 1;
  j0 = jlo;
  k0 = klo;
  nj = jhi >= jlo ? jhi - jlo + 1 : 0;
  nk = khi >= klo ? khi - klo + 1 : 0;
}

// y: the forward result (z itself in FP64; an FP32 buffer for the mixed sweeps).
template <int TX, int TY, int TZ, class T>
void launch_ic0_tiles_shaped(const Geo& g, const uint8_t* codes, const double* D, const double* r, T* y, double* z,
                             const PcgState* st, hipStream_t s, const double* fill) {
  constexpr int NT = TX * TY * TZ;
  Geo tg{};
  tg.n[0] = (g.n[0] + TX - 1) / TX;
  tg.n[1] = (g.n[1] + TY - 1) / TY;
  tg.n[2] = (g.n[2] + TZ - 1) / TZ;
  const int diagonals = tg.n[0] + tg.n[1] + tg.n[2] - 2;
  for (int P = 0; P < diagonals; ++P) {
    int j0, k0, nj, nk;
    level_box(tg, P, j0, k0, nj, nk);
    if (nj * nk == 0) continue;
    hipLaunchKernelGGL((k_ic0_tile<true, TX, TY, TZ, T, double, T>), dim3(nj * nk), dim3(NT), 0, s, g, codes, D, r, y,
                       P, j0, k0, nj, nk, tg.n[0], st, fill);
  }
  for (int P = diagonals - 1; P >= 0; --P) {
    int j0, k0, nj, nk;
    level_box(tg, P, j0, k0, nj, nk);
    if (nj * nk == 0) continue;
    hipLaunchKernelGGL((k_ic0_tile<false, TX, TY, TZ, T, T, double>), dim3(nj * nk), dim3(NT), 0, s, g, codes, D,
                       (const T*)y, z, P, j0, k0, nj, nk, tg.n[0], st, fill);
  }
}

template <class T>
void launch_ic0_tiles_t(const Geo& g, const uint8_t* codes, const double* D, const double* r, T* y, double* z,
                        const PcgState* st, hipStream_t s, const double* fill) {
  probe_begin(PROBE_IC0, s);
  // an axis of at most 2 cells: flat 16 x 16 tiles across the other two axes
  if (g.n[1] <= 2)
    launch_ic0_tiles_shaped<16, 2, 16, T>(g, codes, D, r, y, z, st, s, fill);
  else if (g.n[2] <= 2)
    launch_ic0_tiles_shaped<16, 16, 2, T>(g, codes, D, r, y, z, st, s, fill);
  else if (g.n[0] <= 2)
    launch_ic0_tiles_shaped<2, 16, 16, T>(g, codes, D, r, y, z, st, s, fill);
  else
    launch_ic0_tiles_shaped<8, 8, 8, T>(g, codes, D, r, y, z, st, s, fill);
  // forward: r 8, D 8, code 1 in, y out; backward: y, D 8, code 1 in, z 8 out
  probe_end(PROBE_IC0, s, (double)g.N * (8 + 8 + 1 + sizeof(T) + sizeof(T) + 8 + 1 + 8));
}

void launch_ic0_tiles(const Geo& g, const uint8_t* codes, const double* D, const double* r, double* z,
                      const PcgState* st, hipStream_t s, const double* fill) {
  launch_ic0_tiles_t<double>(g, codes, D, r, z, z, st, s, fill);
}

void launch_ic0_tiles_mixed(const Geo& g, const uint8_t* codes, const double* D, const double* r, float* y, double* z,
                            const PcgState* st, hipStream_t s, const double* fill) {
  launch_ic0_tiles_t<float>(g, codes, D, r, y, z, st, s, fill);
}

void launch_ic0(const Geo& g, const uint8_t* codes, const double* D, const double* r, double* z, const PcgState* st,
                hipStream_t s, const double* fill) {
  const int levels = g.n[0] + g.n[1] + g.n[2] - 2;
  for (int L = 0; L < levels; ++L) {
    int j0, k0, nj, nk;
    level_box(g, L, j0, k0, nj, nk);
    if (nj * nk == 0) continue;
    hipLaunchKernelGGL(k_ic0_fwd, dim3(blocks_for((long long)nj * nk)), dim3(BS), 0, s, g, codes, D, r, z, L, j0, k0, nj,
                       nk, st, fill);
  }
  for (int L = levels - 1; L >= 0; --L) {
    int j0, k0, nj, nk;
    level_box(g, L, j0, k0, nj, nk);
    if (nj * nk == 0) continue;
    hipLaunchKernelGGL(k_ic0_bwd, dim3(blocks_for((long long)nj * nk)), dim3(BS), 0, s, g, codes, D, z, L, j0, k0, nj, nk,
                       st, fill);
  }
}

template <class T>
static unsigned color_blocks(const Geo& g) {
  long long n = (long long)((g.n[0] + 1) >> 1) * g.n[1] * g.n[2];
  return blocks_for(n);
}

template <class T, class TB>
void launch_fine_gs(const Geo& g0, const uint8_t* codes, const TB* b0, T* x0, int color, const Geo& g1, hipStream_t s) {
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(color_blocks<T>(g0)), dim3(BS), 0, s, g0, codes, b0, x0, color, 0,
                     (const T*)nullptr, g1, (double*)nullptr, 0, (const PcgState*)nullptr);
}
template void launch_fine_gs<double, double>(const Geo&, const uint8_t*, const double*, double*, int, const Geo&, hipStream_t);
template void launch_fine_gs<float, double>(const Geo&, const uint8_t*, const double*, float*, int, const Geo&, hipStream_t);

// One V(2,2) cycle on the aggregation hierarchy: the fine level from the
// coupling codes (red-black, or 2^(1+seams) colour classes across odd periodic
// seams), lv[0..nlev) the coarse levels (lv[nlev-1] solved by the dense
// blocks). The correction of each level is scaled by omega. resident (colour-
// separated fine level with two cells per thread): 1 b0 copied into the fine
// level's b and z left in its x, the last half-sweep forming r . z from them;
// 2 the same with b already there (b0 unused); 3 (an FP32 level) b0 rounded into
// the level's b as without, r . z formed from b0 and z written in FP32 (the
// values of the level's x) to out32 in the natural layout; 4 (an FP32 level,
// two cells per thread) the FP64 right-hand side rz_r in the colour-separated
// layout (b0 unused) rounded into the level's b, r . z formed from rz_r and z
// left in the level's x. first_done (resident 2 and 4): the x/r update that
// formed the right-hand side also did the first half-sweep (the level's b and
// red x are set); the cycle starts with the black half-sweep.
template <class T, class TB>
void launch_agg_vcycle(const Geo& g0, const uint8_t* codes, const TB* b0, T* x0, const int* fparent, int seams,
                       int nlev, const AggDevLevel<T>* lv, const AggCoarsest& D, double* out64, const PcgState* st,
                       double omega, hipStream_t s, const FineRB<T>* rb, int fine_rounds, int coarse_rounds,
                       const double* rz_r, double* rz_part, bool zero_last, int resident, float* out32,
                       bool first_done) {
  // a round is one sweep over every colour class; pre-smoothing runs the classes in
  // ascending order, post-smoothing in descending order (the adjoint)
  const int fr = fine_rounds, cr = coarse_rounds;
  const T cs = (T)omega;
  const unsigned ab = blocks_for(g0.N);
  auto direct = [&](const auto* b, T* x, double* o64) {
    using TI = std::remove_cv_t<std::remove_pointer_t<decltype(b)>>;
    PROBED_LAUNCH(PROBE_COARSEST, 0.0, (k_agg_direct<TI, T>), dim3(D.big_rows + (D.nsmall + 3) / 4), dim3(BS), 0, s,
                  D, b, x, o64, st);
  };
  if (nlev == 0) {
    // the fine level is the coarsest: solid cells stay zero
    hipLaunchKernelGGL((k_fill_t<T>), dim3(ab), dim3(BS), 0, s, g0.N, x0, st);
    if (out64) hipLaunchKernelGGL((k_fill_t<double>), dim3(ab), dim3(BS), 0, s, g0.N, out64, st);
    direct(b0, x0, out64);
    return;
  }
  // algorithmic bytes of a coarse half-sweep of colour c: per row its padded
  // (or CSR) entries, diagonal and right-hand side in and its value out, and the
  // other classes' values read once
  auto coarse_bytes = [&](const AggDevLevel<T>& L, int c) {
    const double rows = (double)(L.cptr[c + 1] - L.cptr[c]);
    const double per_row = L.ecol ? (double)L.K * 4.0 : (double)L.nnz / (double)L.n * (4.0 + sizeof(T)) + 4.0;
    return rows * (per_row + 3.0 * sizeof(T)) + ((double)L.n - rows) * sizeof(T);
  };
  auto relax = [&](const AggDevLevel<T>& L, int c) {
    const int rows = L.cptr[c + 1] - L.cptr[c];
    if (rows <= 0) return;
    PROBED_LAUNCH(PROBE_COARSE_HALF, coarse_bytes(L, c), (k_agg_gs<T>), dim3(blocks_for(rows)), dim3(BS), 0, s, L, c,
                  st, (const T*)nullptr, T(0));
  };
  const Geo none{};
  const T* nox = nullptr;
  const int fine_colours = 2 << __builtin_popcount((unsigned)seams);
  const unsigned fb = color_blocks<T>(g0);
  const long long bins = (long long)((g0.n[0] + 1) >> 1) * ((g0.n[1] + 1) >> 1) * ((g0.n[2] + 1) >> 1);
  if (rb) {
    // colour-separated fine level: scatter + red, black, red, black; restriction
    // pairs of natural cells when every row starts a pair and r allows 16-byte loads
    // r (FP64) and codes in; b of both colours and the red x out (from b kept
    // in the layout: red codes and b in, red x out)
    const double first_bytes =
        resident == 2   ? (double)rb->half * (1 + 2 * sizeof(T))
        : resident == 4 ? (double)g0.N * (8 + sizeof(T)) + (double)rb->half * (1 + sizeof(T))
                        : (double)g0.N * (8 + 1 + sizeof(T)) + (double)rb->half * sizeof(T);
    if (first_done && (resident == 2 || resident == 4)) {
      // the x/r update that formed the right-hand side relaxed the red cells
    } else if (resident == 2)
      PROBED_LAUNCH(PROBE_FINE_FIRST, first_bytes, (k_rb_first_lin<T>), dim3((unsigned)((rb->half / 2 + BS - 1) / BS)),
                    dim3(BS), 0, s, g0, *rb, st);
    else if (resident == 4)
      PROBED_LAUNCH(PROBE_FINE_FIRST, first_bytes, (k_rb_first_cvt<T>),
                    dim3((unsigned)((rb->half / 2 + BS - 1) / BS)), dim3(BS), 0, s, g0, *rb, rz_r, st);
    else if (g0.n[0] % 2 == 0 && ((uintptr_t)b0 & 15) == 0)
      PROBED_LAUNCH(PROBE_FINE_FIRST, first_bytes, (k_rb_first2<T>),
                    dim3(((rb->hx + 63) / 64) * ((g0.n[1] + 3) / 4) * g0.n[2]), dim3(BS), 0, s, g0, codes,
                    (const double*)b0, *rb, st);
    else
      PROBED_LAUNCH(PROBE_FINE_FIRST, first_bytes, (k_rb_first<T>),
                    dim3(((g0.n[0] + 63) / 64) * ((g0.n[1] + 3) / 4) * g0.n[2]), dim3(BS), 0, s, g0, codes,
                    (const double*)b0, *rb, st);
    for (int pass = 1; pass < 2 * fr; ++pass)
      launch_rb_half<T>(g0, *rb, pass & 1, nox, fparent, cs, (double*)nullptr, st, s);
    // red cells: code, b, x; black x; one parent per 2x2x2 block; the first
    // coarse level's b and x written and its diagonal read
    const double restrict_bytes =
        (double)rb->half * (1 + 3 * sizeof(T)) + (double)g0.N / 8 * 4 + (double)lv[0].n * 3 * sizeof(T);
    const unsigned restrict_blocks =
        (unsigned)((((g0.n[0] + 1) / 2 + 63) / 64) * ((g0.n[1] + 1) / 2) * ((g0.n[2] + 1) / 2));
    if (g0.n[0] % 4 == 0 && rb_restrict_pairs())
      PROBED_LAUNCH(PROBE_FINE_RESTRICT, restrict_bytes, (k_rb_restrict2<T>), dim3(restrict_blocks), dim3(RS2), 0, s,
                    g0, *rb, fparent, lv[0].b, lv[0].x, lv[0].diag, lv[0].cptr[1], st);
    else
      PROBED_LAUNCH(PROBE_FINE_RESTRICT, restrict_bytes, (k_rb_restrict<T>), dim3(restrict_blocks), dim3(BS), 0, s,
                    g0, *rb, fparent, lv[0].b, lv[0].x, lv[0].diag, lv[0].cptr[1], st);
  } else if (!seams) {
    for (int pass = 0; pass < 2 * fr; ++pass)
      hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, pass & 1, pass == 0 ? 1 : 0, nox,
                         none, (double*)nullptr, 0, st, cs, fparent, 0);
  } else {
    hipLaunchKernelGGL((k_fill_t<T>), dim3(ab), dim3(BS), 0, s, g0.N, x0, st);
    for (int round = 0; round < fr; ++round)
      for (int c = 0; c < fine_colours; ++c)
        hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(ab), dim3(BS), 0, s, g0, codes, b0, x0, c, 0, nox, none,
                           (double*)nullptr, 1, st, cs, fparent, seams);
  }
  if (!rb)
    hipLaunchKernelGGL((k_agg_fine_restrict<T, TB>), dim3(blocks_for(8 * bins)), dim3(BS), 0, s, g0, codes, b0,
                       (const T*)x0, fparent, lv[0].b, lv[0].x, lv[0].diag, lv[0].cptr[1], st, seams ? -1 : 1);
  const int stop = nlev - 1;
  for (int l = 0; l < stop; ++l) {
    const AggDevLevel<T>& L = lv[l];
    // the first half-sweep (colour 0 from zero) happened in the restriction
    for (int round = 0; round < cr; ++round)
      for (int c = round == 0 ? 1 : 0; c < L.ncol; ++c) relax(L, c);
    // pre-smoothing ended with the last colour class, whose residual is zero to
    // rounding (without rounds, or with zero_last off, nothing is taken as zero)
    const int zero_from = cr >= 1 && zero_last ? L.cptr[L.ncol - 1] : L.n;
    PROBED_LAUNCH(PROBE_COARSE_RESTRICT, 0.0, (k_agg_restrict<T>), dim3((lv[l + 1].n + 63) / 64),
                  dim3(kRestrictThreads), 0, s, L, lv[l + 1], st, zero_from);
  }
  direct((const T*)lv[nlev - 1].b, lv[nlev - 1].x, (double*)nullptr);
  for (int l = stop - 1; l >= 0; --l) {
    const AggDevLevel<T>& L = lv[l];
    const T* xc = lv[l + 1].x;
    if (L.ncol == 2) {
      // prolongation fused into the first half-sweep
      const int rows = L.cptr[2] - L.cptr[1];
      if (rows > 0)
        PROBED_LAUNCH(PROBE_COARSE_HALF, coarse_bytes(L, 1), (k_agg_gs<T>), dim3(blocks_for(rows)), dim3(BS), 0, s, L, 1,
                      st, xc, cs);
      else
        hipLaunchKernelGGL((k_agg_prolong<T>), dim3(blocks_for(L.n)), dim3(BS), 0, s, L, xc, cs, st);
      for (int pass = 1; pass < 2 * cr; ++pass) relax(L, (pass + 1) & 1);
    } else {
      hipLaunchKernelGGL((k_agg_prolong<T>), dim3(blocks_for(L.n)), dim3(BS), 0, s, L, xc, cs, st);
      for (int round = 0; round < cr; ++round)
        for (int c = L.ncol - 1; c >= 0; --c) relax(L, c);
    }
  }
  // fine post-smoothing; in mixed precision (and always in the colour-separated
  // layout) the final half-sweeps also write z in FP64
  if (rb) {
    for (int pass = 0; pass < 2 * fr; ++pass) {
      if (pass == 2 * fr - 1 && pass > 0 && (rz_r || resident)) {
        // the last half-sweep also forms the PCG's r . z partials: the red
        // half-sweep, both colours' z out in FP64 and r in (with r and z kept in
        // the layout: the black b read besides)
        const double rz_bytes = (double)rb->half * (1 + 3 * sizeof(T)) +
                                (resident == 3   ? (double)g0.N * 12
                                 : resident == 4 ? (double)g0.N * 8
                                 : resident      ? (double)rb->half * sizeof(T)
                                                 : (double)g0.N * 16);
        if (resident == 4)
          PROBED_LAUNCH(PROBE_FINE_RZ, rz_bytes, (k_rb_gs_rz2_lin<T, true, double>), dim3(RP), dim3(RB), 0, s, g0,
                        *rb, (double*)nullptr, (float*)nullptr, rz_r, st, rz_part);
        else if (resident == 3 && rb->two_cells == 2)
          PROBED_LAUNCH(PROBE_FINE_RZ, rz_bytes, (k_rb_gs_rz2_lin<T, false>), dim3(RP), dim3(RB), 0, s, g0, *rb,
                        (double*)nullptr, out32, rz_r, st, rz_part);
        else if (resident == 3)
          PROBED_LAUNCH(PROBE_FINE_RZ, rz_bytes, (k_rb_gs_rz2<T, false>), dim3(RP), dim3(RB), 0, s, g0, *rb,
                        (double*)nullptr, out32, rz_r, st, rz_part);
        else if (resident && rb->two_cells == 2)
          PROBED_LAUNCH(PROBE_FINE_RZ, rz_bytes, (k_rb_gs_rz2_lin<T, true>), dim3(RP), dim3(RB), 0, s, g0, *rb,
                        (double*)nullptr, (float*)nullptr, (const double*)nullptr, st, rz_part);
        else if (resident)
          PROBED_LAUNCH(PROBE_FINE_RZ, rz_bytes, (k_rb_gs_rz2<T, true>), dim3(RP), dim3(RB), 0, s, g0, *rb,
                        (double*)nullptr, (float*)nullptr, (const double*)nullptr, st, rz_part);
        else if (rb->two_cells == 2 && g0.n[0] % 4 == 0)
          PROBED_LAUNCH(PROBE_FINE_RZ, rz_bytes, (k_rb_gs_rz2_lin<T, false>), dim3(RP), dim3(RB), 0, s, g0, *rb, out64,
                        (float*)nullptr, rz_r, st, rz_part);
        else if (rb->two_cells && g0.n[0] % 4 == 0)
          PROBED_LAUNCH(PROBE_FINE_RZ, rz_bytes, (k_rb_gs_rz2<T, false>), dim3(RP), dim3(RB), 0, s, g0, *rb, out64,
                        (float*)nullptr, rz_r, st, rz_part);
        else
          PROBED_LAUNCH(PROBE_FINE_RZ, rz_bytes, (k_rb_gs_rz<T>), dim3(RP), dim3(RB), 0, s, g0, *rb, out64, rz_r, st,
                        rz_part);
        continue;
      }
      const T* xcp = pass == 0 ? (const T*)lv[0].x : nox;
      if (xcp && rb->rpar) {
        // the correction added to the red cells first, then a plain half-sweep
        // red x in and out, its parent; the first coarse level's values
        const double prolong_bytes = (double)rb->half * (2 * sizeof(T) + 4) + (double)lv[0].n * sizeof(T);
        if (rb->half % 2 == 0)
          PROBED_LAUNCH(PROBE_FINE_PROLONG, prolong_bytes, (k_rb_prolong2<T>), dim3(blocks_for(rb->half / 2)), dim3(BS),
                        0, s, rb->half, rb->x, rb->rpar, xcp, cs, st);
        else
          PROBED_LAUNCH(PROBE_FINE_PROLONG, prolong_bytes, (k_rb_prolong<T>), dim3(blocks_for(rb->half)), dim3(BS), 0,
                        s, rb->half, rb->x, rb->rpar, xcp, cs, st);
        xcp = nox;
      }
      // z in the natural layout from the last two half-sweeps (the r . z kernel
      // writes both colours itself)
      const bool z_out = pass == 2 * fr - 1 || (pass == 2 * fr - 2 && !(rz_r && fr >= 1));
      launch_rb_half<T>(g0, *rb, (pass + 1) & 1, xcp, fparent, cs, z_out ? out64 : (double*)nullptr, st, s);
    }
  } else if (!seams) {
    for (int pass = 0; pass < 2 * fr; ++pass)
      hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, (pass + 1) & 1, 0,
                         pass == 0 ? (const T*)lv[0].x : nox, none, pass >= 2 * fr - 2 ? out64 : (double*)nullptr, 0,
                         st, cs, fparent, 0);
  } else {
    hipLaunchKernelGGL((k_agg_fine_prolong<T>), dim3(ab), dim3(BS), 0, s, g0.N, fparent, (const T*)lv[0].x, cs, x0, st);
    for (int round = 0; round < fr; ++round)
      for (int c = fine_colours - 1; c >= 0; --c)
        hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(ab), dim3(BS), 0, s, g0, codes, b0, x0, c, 0, nox, none,
                           (round == fr - 1 && c == 0) ? out64 : (double*)nullptr, 1, st, cs, fparent, seams);
  }
}
template void launch_agg_vcycle<double, double>(const Geo&, const uint8_t*, const double*, double*, const int*, int, int,
                                                const AggDevLevel<double>*, const AggCoarsest&, double*,
                                                const PcgState*, double, hipStream_t, const FineRB<double>*, int, int,
                                                const double*, double*, bool, int, float*, bool);
template void launch_agg_vcycle<float, double>(const Geo&, const uint8_t*, const double*, float*, const int*, int, int,
                                               const AggDevLevel<float>*, const AggCoarsest&, double*, const PcgState*,
                                               double, hipStream_t, const FineRB<float>*, int, int, const double*,
                                               double*, bool, int, float*, bool);

template <class T>
void launch_rb_scatter(const Geo& g0, const uint8_t* codes, const double* r, const FineRB<T>& rb, const PcgState* st,
                       hipStream_t s) {
  hipLaunchKernelGGL((k_rb_first2<T>), dim3(((rb.hx + 63) / 64) * ((g0.n[1] + 3) / 4) * g0.n[2]), dim3(BS), 0, s, g0,
                     codes, r, rb, st);
}
template void launch_rb_scatter<double>(const Geo&, const uint8_t*, const double*, const FineRB<double>&,
                                        const PcgState*, hipStream_t);
void launch_rb_first_cvt(const Geo& g0, const FineRB<float>& rb, const double* rres, const PcgState* st,
                         hipStream_t s) {
  hipLaunchKernelGGL((k_rb_first_cvt<float>), dim3((unsigned)((rb.half / 2 + BS - 1) / BS)), dim3(BS), 0, s, g0, rb,
                     rres, st);
}

template <class T>
void launch_rb_gs(const Geo& g0, const FineRB<T>& rb, int colour, hipStream_t s) {
  launch_rb_half<T>(g0, rb, colour, (const T*)nullptr, (const int*)nullptr, T(0), (double*)nullptr,
                    (const PcgState*)nullptr, s);
}
template void launch_rb_gs<double>(const Geo&, const FineRB<double>&, int, hipStream_t);
template void launch_rb_gs<float>(const Geo&, const FineRB<float>&, int, hipStream_t);

// Inverts the deferred coarsest blocks on the device (see k_gj_panel); returns
// false when a pivot is not positive (the block lost definiteness).
bool invert_block(double* Gblock, int m, bool anchored, const double* K, double* work, int* bad, hipStream_t s) {
  const int skip = anchored ? 0 : 1;
  const int k = m - skip;
  double* A = work;                              // k * k
  double* pinv = work + (long long)k * k;        // kGJ * kGJ
  double* rp = pinv + kGJ * kGJ;                 // kGJ * k
  double* cold = rp + (long long)kGJ * k;        // k * kGJ
  double* mu = cold + (long long)k * kGJ;        // m
  if (hipMemcpyAsync(A, K, (size_t)k * k * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) return false;
  if (hipMemsetAsync(bad, 0, sizeof(int), s) != hipSuccess) return false;
  const int tiles = ((k + 63) / 64) * ((k + 63) / 64);
  for (int p0 = 0; p0 < k; p0 += kGJ) {
    const int nb = std::min(kGJ, k - p0);
    hipLaunchKernelGGL(k_gj_panel, dim3(1), dim3(kGJThreads), 0, s, (const double*)A, k, p0, nb, pinv, bad);
    hipLaunchKernelGGL(k_gj_rows, dim3(blocks_for(2LL * nb * k)), dim3(BS), 0, s, (const double*)A, k, p0, nb,
                       (const double*)pinv, rp, cold);
    hipLaunchKernelGGL(k_gj_update, dim3(tiles), dim3(BS), 0, s, A, k, p0, nb, (const double*)rp, (const double*)cold);
  }
  hipLaunchKernelGGL(k_block_embed, dim3(blocks_for((long long)m * m)), dim3(BS), 0, s, Gblock, m, skip,
                     (const double*)A);
  if (!anchored) {
    hipLaunchKernelGGL(k_block_colmean, dim3(blocks_for(m)), dim3(BS), 0, s, (const double*)Gblock, m, mu);
    hipLaunchKernelGGL(k_block_center, dim3(blocks_for((long long)m * m)), dim3(BS), 0, s, Gblock, m, (const double*)mu);
  }
  int h_bad = 0;
  if (hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return false;
  if (hipStreamSynchronize(s) != hipSuccess) return false;
  return h_bad == 0;
}

// One application of the reference's V-cycle: z = M r.
void launch_ref_vcycle(const RefCycle& R, const double* r, double* z, const PcgState* st, hipStream_t s) {
  const Geo& g = R.g0;
  const unsigned ab = blocks_for(g.N);
  const int planes = g.n[0] + g.n[1] + g.n[2] - 2;
  auto tiled = [&](auto tx, auto ty, auto tz, bool forward) {
    constexpr int TX = decltype(tx)::value, TY = decltype(ty)::value, TZ = decltype(tz)::value;
    Geo tg{};
    tg.n[0] = (g.n[0] + TX - 1) / TX;
    tg.n[1] = (g.n[1] + TY - 1) / TY;
    tg.n[2] = (g.n[2] + TZ - 1) / TZ;
    const int diagonals = tg.n[0] + tg.n[1] + tg.n[2] - 2;
    for (int k = 0; k < diagonals; ++k) {
      const int P = forward ? k : diagonals - 1 - k;
      int j0, k0, nj, nk;
      level_box(tg, P, j0, k0, nj, nk);
      if (nj * nk == 0) continue;
      if (forward)
        hipLaunchKernelGGL((k_ref_fine_tile<true, TX, TY, TZ>), dim3(nj * nk), dim3(TX * TY * TZ), 0, s, g, R.codes,
                           R.diag0, (const double*)R.rhs0, z, R.self_mask, P, j0, k0, nj, tg.n[0], st);
      else
        hipLaunchKernelGGL((k_ref_fine_tile<false, TX, TY, TZ>), dim3(nj * nk), dim3(TX * TY * TZ), 0, s, g, R.codes,
                           R.diag0, (const double*)R.rhs0, z, R.self_mask, P, j0, k0, nj, tg.n[0], st);
    }
  };
  using I2 = std::integral_constant<int, 2>;
  using I8 = std::integral_constant<int, 8>;
  using I16 = std::integral_constant<int, 16>;
  auto fine_sweep = [&](bool forward) {
    if (R.tiles) {
      // tiles as the IC(0) sweeps take them: flat ones across an axis of at most 2 cells
      if (g.n[1] <= 2)
        tiled(I16{}, I2{}, I16{}, forward);
      else if (g.n[2] <= 2)
        tiled(I16{}, I16{}, I2{}, forward);
      else if (g.n[0] <= 2)
        tiled(I2{}, I16{}, I16{}, forward);
      else
        tiled(I8{}, I8{}, I8{}, forward);
      return;
    }
    for (int k = 0; k < planes; ++k) {
      const int L = forward ? k : planes - 1 - k;
      int j0, k0, nj, nk;
      level_box(g, L, j0, k0, nj, nk);
      if (nj * nk == 0) continue;
      hipLaunchKernelGGL(k_ref_fine_gs, dim3(blocks_for((long long)nj * nk)), dim3(BS), 0, s, g, R.codes, R.diag0,
                         (const double*)R.rhs0, z, L, j0, k0, nj, nk, R.self_mask, st);
    }
  };
  auto center = [&](const RefCoarse& C, double* v) {
    if (C.nchunks == 0) return;
    hipLaunchKernelGGL(k_center_sum, dim3(C.nchunks), dim3(BS), 0, s, (const double*)v, C.chunk_begin,
                       C.closed_members, C.chunk_sum, st);
    hipLaunchKernelGGL(k_center_sub, dim3(C.nchunks), dim3(BS), 0, s, v, C.chunk_begin, C.chunk_comp, C.comp_chunk,
                       C.closed_ptr, C.closed_members, (const double*)C.chunk_sum, st);
  };
  auto coarse_sweep = [&](const RefCoarse& C, bool forward) {
    const std::vector<int>& ptr = forward ? C.fwd_ptr : C.bwd_ptr;
    const int* rows = forward ? C.fwd_rows : C.bwd_rows;
    const RefSweep& P = forward ? C.fwd_pk : C.bwd_pk;
    if (R.tiles && R.packed && P.nsub > 0) {
      const size_t shm = ((size_t)(P.ldsx ? C.L.n : 0) + (size_t)P.cap) * sizeof(double);
      if (P.ldsx)
        hipLaunchKernelGGL((k_ref_gs_pk<true, kRefSubRows>), dim3(1), dim3(kRefSubRows), shm, s, C.L, P, st);
      else
        hipLaunchKernelGGL((k_ref_gs_pk<false, kRefSubRows>), dim3(1), dim3(kRefSubRows), shm, s, C.L, P, st);
      return;
    }
    if (R.tiles && C.L.n <= kRefWgRows && ptr.size() > 2) {
      if (C.L.n <= kRefWgLds)
        hipLaunchKernelGGL(k_ref_gs_wg<true>, dim3(1), dim3(1024), 0, s, C.L, rows, forward ? C.fwd_dptr : C.bwd_dptr,
                           (int)ptr.size() - 1, st);
      else
        hipLaunchKernelGGL(k_ref_gs_wg<false>, dim3(1), dim3(1024), 0, s, C.L, rows, forward ? C.fwd_dptr : C.bwd_dptr,
                           (int)ptr.size() - 1, st);
      return;
    }
    for (size_t d = 0; d + 1 < ptr.size(); ++d) {
      const int count = ptr[d + 1] - ptr[d];
      if (count > 0)
        hipLaunchKernelGGL(k_ref_gs, dim3(blocks_for(count)), dim3(BS), 0, s, C.L, rows + ptr[d], count, st);
    }
  };
  auto line_sweep = [&](const RefLines& B, const double* rhs, double* x, bool forward) {
    const std::vector<int>& ptr = forward ? B.fwd_ptr : B.bwd_ptr;
    const int* lines = forward ? B.fwd : B.bwd;
    for (size_t d = 0; d + 1 < ptr.size(); ++d) {
      const int count = ptr[d + 1] - ptr[d];
      if (count <= 0) continue;
      const size_t lds = (size_t)B.length * (B.bandwidth + 2) * sizeof(double);
      if (lds <= 48 * 1024)
        hipLaunchKernelGGL(k_ref_lines_lds, dim3(count), dim3(64), lds, s, B.length, B.bandwidth, B.nodes, B.factor,
                           B.ooff, B.ocol, B.oval, lines + ptr[d], rhs, x, st);
      else
        hipLaunchKernelGGL(k_ref_lines, dim3(blocks_for(count)), dim3(BS), 0, s, B.length, B.bandwidth, B.nodes, B.factor,
                           B.ooff, B.ocol, B.oval, B.scratch, lines + ptr[d], count, rhs, x, st);
    }
  };
  const int nlev = (int)R.lv.size();
  // level 0: rhs centered, x = 0, two forward sweeps (one block sweep on an
  // anisotropic level, multigrid.rs:676-688), centered x, centered residual
  R.center0(r, R.rhs0);
  hipLaunchKernelGGL(k_fill_t<double>, dim3(ab), dim3(BS), 0, s, g.N, z, st);
  if (nlev == 0) {
    hipLaunchKernelGGL((k_agg_direct<double, double>), dim3(R.D.big_rows + (R.D.nsmall + 3) / 4), dim3(BS), 0, s, R.D,
                       (const double*)R.rhs0, z, (double*)nullptr, st);
    return;
  }
  if (R.fine_lines.length > 0)
    line_sweep(R.fine_lines, R.rhs0, z, true);
  else
    for (int pass = 0; pass < 2; ++pass) fine_sweep(true);
  R.center0(z, z);
  hipLaunchKernelGGL(k_ref_fine_residual, dim3(ab), dim3(BS), 0, s, g, R.codes, R.diag0, (const double*)R.rhs0,
                     (const double*)z, R.r0, R.self_mask, st);
  R.center0(R.r0, R.r0);
  {
    const RefCoarse& C = R.lv[0];
    hipLaunchKernelGGL(k_ref_restrict, dim3(blocks_for(C.L.n)), dim3(BS), 0, s, C.L.n, C.L.child_ptr, R.fchild,
                       C.child_w, (const double*)R.r0, C.L.b, C.L.x, st);
  }
  for (int l = 0; l + 1 < nlev; ++l) {
    const RefCoarse& C = R.lv[l];
    center(C, C.L.b);
    if (C.lines.length > 0)
      line_sweep(C.lines, C.L.b, C.L.x, true);
    else
      for (int pass = 0; pass < 2; ++pass) coarse_sweep(C, true);
    center(C, C.L.x);
    hipLaunchKernelGGL(k_ref_residual, dim3(blocks_for(C.L.n)), dim3(BS), 0, s, C.L, C.r, st);
    center(C, C.r);
    const RefCoarse& N = R.lv[l + 1];
    hipLaunchKernelGGL(k_ref_restrict, dim3(blocks_for(N.L.n)), dim3(BS), 0, s, N.L.n, N.L.child_ptr, N.L.child,
                       N.child_w, (const double*)C.r, N.L.b, N.L.x, st);
  }
  {
    // coarsest: the dense blocks include the centering of rhs and solution
    const RefCoarse& C = R.lv[nlev - 1];
    hipLaunchKernelGGL((k_agg_direct<double, double>), dim3(R.D.big_rows + (R.D.nsmall + 3) / 4), dim3(BS), 0, s, R.D,
                       (const double*)C.L.b, C.L.x, (double*)nullptr, st);
  }
  for (int l = nlev - 2; l >= 0; --l) {
    const RefCoarse& C = R.lv[l];
    hipLaunchKernelGGL(k_ref_prolong, dim3(blocks_for(C.L.n)), dim3(BS), 0, s, (long long)C.L.n, C.L.x, C.L.parent,
                       C.tptr, C.tcol, C.tw, (const double*)R.lv[l + 1].L.x, st);
    if (C.lines.length > 0)
      line_sweep(C.lines, C.L.b, C.L.x, false);
    else
      for (int pass = 0; pass < 2; ++pass) coarse_sweep(C, false);
    center(C, C.L.x);
  }
  hipLaunchKernelGGL(k_ref_prolong, dim3(ab), dim3(BS), 0, s, g.N, z, R.fparent, R.ftptr, R.ftcol, R.ftw,
                     (const double*)R.lv[0].L.x, st);
  if (R.fine_lines.length > 0)
    line_sweep(R.fine_lines, R.rhs0, z, false);
  else
    for (int pass = 0; pass < 2; ++pass) fine_sweep(false);
  R.center0(z, z);
}

void launch_projection_rhs(const Geo& g, const FaceGeo& fg, const uint8_t* codes, const double* u, double dt, double* b,
                           double* terms, hipStream_t s) {
  hipLaunchKernelGGL(k_projection_rhs, dim3(blocks_for(g.N)), dim3(BS), 0, s, g, fg, codes, u, dt, b, terms);
}
void launch_project_faces(const Geo& g, const FaceGeo& fg, const uint8_t* codes, const double* phi, double dt, double* u,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_project_faces, dim3(blocks_for(fg.total)), dim3(BS), 0, s, g, fg, codes, phi, dt, u);
}

// ---------------------------------------------------------------- batched iterations
// A batched launch carries the arguments of up to kBatchMax single-grid
// launches in its kernel arguments: item i owns the blocks [first[i],
// first[i] + nblk) (first[i] a multiple of 8, so a block's XCD, bid & 7, is the
// one it would have alone) and runs the single-grid kernel's body on them.
template <class A>
struct Batch {
  A a[kBatchMax];
  unsigned first[kBatchMax];
  int n;
};
template <class A>
__device__ __forceinline__ const A& batch_item(const Batch<A>& B, unsigned& bid) {
  int i = 0;
#pragma unroll
  for (int k = 1; k < kBatchMax; ++k)
    if (k < B.n && blockIdx.x >= B.first[k]) i = k;
  bid = blockIdx.x - B.first[i];
  return B.a[i];
}

template <class T>
struct BHalf {
  Geo g;
  FineRB<T> F;
  int colour;
  const PcgState* st;
  unsigned nblk;
};
template <class T>
struct BFineRestrict {
  Geo g;
  FineRB<T> F;
  const int* fparent;
  T* bc;
  T* xc;
  const T* cdiag;
  int red_end;
  const PcgState* st;
  unsigned nblk;
};
template <class T>
struct BAggGs {
  AggDevLevel<T> L;
  int colour;
  const PcgState* st;
  const T* xc;
  T cs;
  unsigned nblk;
};
template <class T>
struct BAggRestrict {
  AggDevLevel<T> F, C;
  const PcgState* st;
  int zero_from;
  unsigned nblk;
};
template <class T>
struct BAggProlong {
  AggDevLevel<T> F;
  const T* xc;
  T cs;
  const PcgState* st;
  unsigned nblk;
};
template <class T>
struct BDirect {
  AggCoarsest D;
  const T* b;
  T* x;
  const PcgState* st;
  unsigned nblk;
};
template <class T>
struct BRbProlong {
  long long half;
  T* xr;
  const int* rpar;
  const T* xc;
  T cs;
  const PcgState* st;
  unsigned nblk;
};
template <class T>
struct BRz {
  Geo g;
  FineRB<T> F;
  const double* r;
  const PcgState* st;
  double* part;
  unsigned nblk;
};
template <class TZ, class TP>
struct BOp {
  Geo g;
  MarchShape ms;
  const uint8_t* codes;
  const TZ* z;
  const TP* pold;
  TP* pnew;
  double* ap;
  PcgState* st;
  double* part;
  long long zhalf;
  unsigned nblk;
};
struct BTrue {
  Geo g;
  const uint8_t* codes;
  const double* x;
  const double* bc;
  PcgState* st;
  double* part;
  unsigned nblk;
};
struct BState {
  const PcgState* src;
  PcgState* dst;
  unsigned nblk;
};

template <class T>
__global__ void __launch_bounds__(BS) kb_rb_gs2_lin(Batch<BHalf<T>> B) {
  unsigned bid;
  const BHalf<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  rb_gs2_lin_body<T>(a.g, a.F, a.colour, (double*)nullptr, a.st, bid, a.nblk);
}
template <class T>
__global__ void __launch_bounds__(BS) kb_rb_restrict(Batch<BFineRestrict<T>> B) {
  unsigned bid;
  const BFineRestrict<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  rb_restrict_body<T>(a.g, a.F, a.fparent, a.bc, a.xc, a.cdiag, a.red_end, a.st, bid);
}
template <class T>
__global__ void __launch_bounds__(RS2) kb_rb_restrict2(Batch<BFineRestrict<T>> B) {
  unsigned bid;
  const BFineRestrict<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  rb_restrict2_body<T>(a.g, a.F, a.fparent, a.bc, a.xc, a.cdiag, a.red_end, a.st, bid);
}
template <class T>
__global__ void __launch_bounds__(BS) kb_agg_gs(Batch<BAggGs<T>> B) {
  unsigned bid;
  const BAggGs<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  agg_gs_body<T>(a.L, a.colour, a.st, a.xc, a.cs, bid);
}
template <class T>
__global__ void __launch_bounds__(kRestrictThreads) kb_agg_restrict(Batch<BAggRestrict<T>> B) {
  unsigned bid;
  const BAggRestrict<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  agg_restrict_body<T>(a.F, a.C, a.st, a.zero_from, bid);
}
template <class T>
__global__ void __launch_bounds__(BS) kb_agg_prolong(Batch<BAggProlong<T>> B) {
  unsigned bid;
  const BAggProlong<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  agg_prolong_body<T>(a.F, a.xc, a.cs, a.st, bid);
}
template <class T>
__global__ void __launch_bounds__(BS) kb_agg_direct(Batch<BDirect<T>> B) {
  unsigned bid;
  const BDirect<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  agg_direct_body<T, T>(a.D, a.b, a.x, (double*)nullptr, a.st, bid);
}
template <class T>
__global__ void __launch_bounds__(BS) kb_rb_prolong(Batch<BRbProlong<T>> B) {
  unsigned bid;
  const BRbProlong<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  rb_prolong_body<T>(a.half, a.xr, a.rpar, a.xc, a.cs, a.st, bid);
}
template <class T>
__global__ void __launch_bounds__(BS) kb_rb_prolong2(Batch<BRbProlong<T>> B) {
  unsigned bid;
  const BRbProlong<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  rb_prolong2_body<T>(a.half, a.xr, a.rpar, a.xc, a.cs, a.st, bid);
}
template <class T, class TR>
__global__ void __launch_bounds__(RB, 8) kb_rb_gs_rz2_lin(Batch<BRz<T>> B) {
  unsigned bid;
  const BRz<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  rb_gs_rz2_lin_body<T, true, TR>(a.g, a.F, (double*)nullptr, (float*)nullptr, a.r, a.st, a.part, bid, a.nblk);
}
template <class TZ, class TP>
__global__ void __launch_bounds__(RB, 8) kb_pcg_op(Batch<BOp<TZ, TP>> B) {
  unsigned bid;
  const BOp<TZ, TP>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  pcg_op_body<false, true, TZ, TP>(a.g, a.codes, a.z, a.pold, a.pnew, a.ap, (const double*)nullptr,
                                   (const int*)nullptr, a.st, a.part, a.zhalf, bid, a.nblk);
}
template <class TZ, class TP>
__global__ void __launch_bounds__(RB, 8) kb_pcg_op_march(Batch<BOp<TZ, TP>> B) {
  unsigned bid;
  const BOp<TZ, TP>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  pcg_op_march_body<false, true, TZ, TP>(a.g, a.ms, a.codes, a.z, a.pold, a.pnew, a.ap, (const double*)nullptr,
                                         (const int*)nullptr, a.st, a.part, a.zhalf, bid, a.nblk);
}
template <class T>
struct BAxpyFirst {
  Geo g;
  FineRB<T> F;
  unsigned long long mnx, mny;
  const T* p;
  const double* ap;
  double* x;
  double* r;
  PcgState* st;
  double* part;
  unsigned nblk;
};
template <class T>
__global__ void __launch_bounds__(RB) kb_pcg_axpy_first(Batch<BAxpyFirst<T>> B) {
  unsigned bid;
  const BAxpyFirst<T>& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  pcg_axpy_first_body<T, T>(a.g, a.F, a.mnx, a.mny, a.p, a.ap, a.x, a.r, a.st, a.part, bid, a.nblk);
}
__global__ void __launch_bounds__(RB, 8) kb_pcg_true(Batch<BTrue> B) {
  unsigned bid;
  const BTrue& a = batch_item(B, bid);
  if (bid >= a.nblk) return;
  pcg_true_body(a.g, a.codes, a.x, a.bc, a.st, a.part, bid, a.nblk);
}
__global__ void __launch_bounds__(RB) kb_pcg_fin(Batch<BTrue> B) {
  unsigned bid;
  const BTrue& a = batch_item(B, bid);
  if (bid != 0) return;
  pcg_fin_body(a.st, a.part);
}
// One wavefront per state: the words of the PcgState copied in order.
__global__ void __launch_bounds__(64) kb_states(Batch<BState> B) {
  unsigned bid;
  const BState& a = batch_item(B, bid);
  if (bid != 0) return;
  static_assert(sizeof(PcgState) % 8 == 0, "copied as 64-bit words");
  const unsigned long long* src = reinterpret_cast<const unsigned long long*>(a.src);
  unsigned long long* dst = reinterpret_cast<unsigned long long*>(a.dst);
  for (unsigned w = threadIdx.x; w < sizeof(PcgState) / 8; w += 64) dst[w] = src[w];
}

// The host side of a batched launch: items whose block count is zero are left
// out, and nothing is launched when no item is left.
template <class A>
struct Batcher {
  Batch<A> B;
  unsigned total = 0;
  Batcher() { B.n = 0; }
  void add(const A& a) {
    if (a.nblk == 0) return;
    B.a[B.n] = a;
    B.first[B.n] = total;
    total += (a.nblk + 7u) & ~7u;
    ++B.n;
  }
};
static_assert(sizeof(Batch<BAggRestrict<double>>) <= 4096, "kernel arguments of a batched launch");
#define BATCH_LAUNCH(K, BT, THREADS, S)                                               \
  do {                                                                                \
    if ((BT).B.n > 0) hipLaunchKernelGGL(K, dim3((BT).total), dim3(THREADS), 0, S, (BT).B); \
  } while (0)

template <class T>
void launch_batch_iteration(const BatchSolve<T>* it, int n, hipStream_t s) {
  if (n <= 0 || n > kBatchMax) return;
  auto pair_blocks = [](const FineRB<T>& F) { return (unsigned)((F.half / 2 + BS - 1) / BS); };
  const int fr = it[0].fine_rounds;
  // the first fine half-sweep from a zero iterate was done by the x/r update
  // that formed the right-hand side (or by the solve's set-up)
  for (int pass = 1; pass < 2 * fr; ++pass) {
    Batcher<BHalf<T>> L;
    for (int i = 0; i < n; ++i) L.add({it[i].g, it[i].F, pass & 1, it[i].st, pair_blocks(it[i].F)});
    BATCH_LAUNCH(kb_rb_gs2_lin<T>, L, BS, s);
  }
  {
    Batcher<BFineRestrict<T>> L;
    for (int i = 0; i < n; ++i) {
      const Geo& g = it[i].g;
      const AggDevLevel<T>& L0 = it[i].lv[0];
      L.add({g, it[i].F, it[i].fparent, L0.b, L0.x, L0.diag, L0.cptr[1], it[i].st,
             (unsigned)((((g.n[0] + 1) / 2 + 63) / 64) * ((g.n[1] + 1) / 2) * ((g.n[2] + 1) / 2))});
    }
    // batched grids have nx a multiple of 4
    if (rb_restrict_pairs())
      BATCH_LAUNCH(kb_rb_restrict2<T>, L, RS2, s);
    else
      BATCH_LAUNCH(kb_rb_restrict<T>, L, BS, s);
  }
  int max_stop = 0;
  for (int i = 0; i < n; ++i) max_stop = std::max(max_stop, it[i].nlev - 1);
  auto rows_of = [](const AggDevLevel<T>& L, int c) { return L.cptr[c + 1] - L.cptr[c]; };
  // coarse pre-smoothing: each level's colour classes in ascending order per
  // round (the first class of the first round was relaxed by the restriction),
  // step k of every solve's sequence in one launch; then the restrictions
  for (int l = 0; l < max_stop; ++l) {
    std::vector<int> seq[kBatchMax];
    size_t steps = 0;
    for (int i = 0; i < n; ++i) {
      if (l >= it[i].nlev - 1) continue;
      const AggDevLevel<T>& L = it[i].lv[l];
      for (int round = 0; round < it[i].coarse_rounds; ++round)
        for (int c = round == 0 ? 1 : 0; c < L.ncol; ++c) seq[i].push_back(c);
      steps = std::max(steps, seq[i].size());
    }
    for (size_t k = 0; k < steps; ++k) {
      Batcher<BAggGs<T>> L;
      for (int i = 0; i < n; ++i) {
        if (k >= seq[i].size()) continue;
        const AggDevLevel<T>& Lv = it[i].lv[l];
        const int c = seq[i][k];
        L.add({Lv, c, it[i].st, (const T*)nullptr, T(0), blocks_for(rows_of(Lv, c))});
      }
      BATCH_LAUNCH(kb_agg_gs<T>, L, BS, s);
    }
    Batcher<BAggRestrict<T>> R;
    for (int i = 0; i < n; ++i) {
      if (l >= it[i].nlev - 1) continue;
      const AggDevLevel<T>& Lv = it[i].lv[l];
      const int zero_from = it[i].coarse_rounds >= 1 && it[i].zero_last ? Lv.cptr[Lv.ncol - 1] : Lv.n;
      R.add({Lv, it[i].lv[l + 1], it[i].st, zero_from, (unsigned)((it[i].lv[l + 1].n + 63) / 64)});
    }
    BATCH_LAUNCH(kb_agg_restrict<T>, R, kRestrictThreads, s);
  }
  {
    Batcher<BDirect<T>> L;
    for (int i = 0; i < n; ++i) {
      const AggCoarsest& D = it[i].D;
      const AggDevLevel<T>& C = it[i].lv[it[i].nlev - 1];
      L.add({D, C.b, C.x, it[i].st, (unsigned)(D.big_rows + (D.nsmall + 3) / 4)});
    }
    BATCH_LAUNCH(kb_agg_direct<T>, L, BS, s);
  }
  // coarse post-smoothing: the prolongation (riding on the first half-sweep of a
  // two-colour level), then the classes in the adjoint order
  struct Op {
    int kind;  // 0 half-sweep, 1 half-sweep with the prolongation, 2 prolongation alone
    int colour;
  };
  for (int l = max_stop - 1; l >= 0; --l) {
    std::vector<Op> seq[kBatchMax];
    size_t steps = 0;
    for (int i = 0; i < n; ++i) {
      if (l >= it[i].nlev - 1) continue;
      const AggDevLevel<T>& L = it[i].lv[l];
      const int cr = it[i].coarse_rounds;
      if (L.ncol == 2) {
        seq[i].push_back(rows_of(L, 1) > 0 ? Op{1, 1} : Op{2, 0});
        for (int pass = 1; pass < 2 * cr; ++pass) seq[i].push_back(Op{0, (pass + 1) & 1});
      } else {
        seq[i].push_back(Op{2, 0});
        for (int round = 0; round < cr; ++round)
          for (int c = L.ncol - 1; c >= 0; --c) seq[i].push_back(Op{0, c});
      }
      steps = std::max(steps, seq[i].size());
    }
    for (size_t k = 0; k < steps; ++k) {
      Batcher<BAggGs<T>> G;
      Batcher<BAggProlong<T>> P;
      for (int i = 0; i < n; ++i) {
        if (k >= seq[i].size()) continue;
        const AggDevLevel<T>& Lv = it[i].lv[l];
        const T* xc = it[i].lv[l + 1].x;
        const Op op = seq[i][k];
        if (op.kind == 2)
          P.add({Lv, xc, it[i].cs, it[i].st, blocks_for(Lv.n)});
        else
          G.add({Lv, op.colour, it[i].st, op.kind == 1 ? xc : (const T*)nullptr, op.kind == 1 ? it[i].cs : T(0),
                 blocks_for(rows_of(Lv, op.colour))});
      }
      BATCH_LAUNCH(kb_agg_prolong<T>, P, BS, s);
      BATCH_LAUNCH(kb_agg_gs<T>, G, BS, s);
    }
  }
  // fine post-smoothing: the correction added to the red cells, black, red, ...,
  // and the last red half-sweep with the partials of r . z
  {
    // batched grids have nx a multiple of 4: an even number of red cells
    Batcher<BRbProlong<T>> L;
    for (int i = 0; i < n; ++i)
      L.add({it[i].F.half, it[i].F.x, it[i].F.rpar, (const T*)it[i].lv[0].x, it[i].cs, it[i].st,
             blocks_for(it[i].F.half / 2)});
    BATCH_LAUNCH(kb_rb_prolong2<T>, L, BS, s);
  }
  for (int pass = 0; pass + 1 < 2 * fr; ++pass) {
    Batcher<BHalf<T>> L;
    for (int i = 0; i < n; ++i) L.add({it[i].g, it[i].F, (pass + 1) & 1, it[i].st, pair_blocks(it[i].F)});
    BATCH_LAUNCH(kb_rb_gs2_lin<T>, L, BS, s);
  }
  {
    // r . z from the FP64 residual: F.b itself (FP64), the FP64 level's b (mixed)
    Batcher<BRz<T>> L;
    for (int i = 0; i < n; ++i) L.add({it[i].g, it[i].F, it[i].rres, it[i].st, it[i].part, (unsigned)RP});
    BATCH_LAUNCH((kb_rb_gs_rz2_lin<T, double>), L, RB, s);
  }
  // p = z + beta p (stored as T), A p, p . Ap; z is the fine level's x
  {
    Batcher<BOp<T, T>> C, M;
    for (int i = 0; i < n; ++i) {
      const BOp<T, T> a{it[i].g, march_shape(it[i].g), it[i].codes, it[i].F.x, (const T*)it[i].pold, (T*)it[i].pnew,
                        it[i].ap, it[i].st, it[i].part, it[i].F.half, (unsigned)RP};
      (it[i].march ? M : C).add(a);
    }
    BATCH_LAUNCH((kb_pcg_op<T, T>), C, RB, s);
    BATCH_LAUNCH((kb_pcg_op_march<T, T>), M, RB, s);
  }
  // alpha, x += alpha p, r -= alpha A p, |r|^2 (r in the FP64 colour-separated
  // arrays), and the next cycle's first half-sweep
  {
    Batcher<BAxpyFirst<T>> L;
    const unsigned long long two40 = 1ULL << 40;
    for (int i = 0; i < n; ++i) {
      const Geo& g = it[i].g;
      const unsigned long long mnx = (two40 + (unsigned long long)g.n[0] - 1) / (unsigned long long)g.n[0];
      const unsigned long long mny = (two40 + (unsigned long long)g.n[1] - 1) / (unsigned long long)g.n[1];
      L.add({g, it[i].F, mnx, mny, (const T*)it[i].pnew, it[i].ap, it[i].x, it[i].rres, it[i].st, it[i].part,
             (unsigned)RP});
    }
    BATCH_LAUNCH(kb_pcg_axpy_first<T>, L, RB, s);
  }
  // the recomputed residual and the iteration's decisions
  {
    Batcher<BTrue> L, F1;
    for (int i = 0; i < n; ++i) {
      L.add({it[i].g, it[i].codes, it[i].x, it[i].bc, it[i].st, it[i].part, (unsigned)RP});
      F1.add({it[i].g, it[i].codes, it[i].x, it[i].bc, it[i].st, it[i].part, 1u});
    }
    BATCH_LAUNCH(kb_pcg_true, L, RB, s);
    BATCH_LAUNCH(kb_pcg_fin, F1, RB, s);
  }
}
template void launch_batch_iteration<double>(const BatchSolve<double>*, int, hipStream_t);
template void launch_batch_iteration<float>(const BatchSolve<float>*, int, hipStream_t);

void launch_batch_states(PcgState* const* st, int n, PcgState* dst, hipStream_t s) {
  Batcher<BState> L;
  for (int i = 0; i < n && i < kBatchMax; ++i) L.add({st[i], dst + i, 1u});
  BATCH_LAUNCH(kb_states, L, 64, s);
}

// ---------------------------------------------------------------- workload helper
// A standard normal deviate per (seed, index), the same on every run: two
// 53-bit uniforms from 64-bit mixes of the seed and the index (the SplitMix64
// finalizer, Steele, Lea and Flood 2014), then the Box-Muller transform.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double normal_at(unsigned long long seed, long long i) {
  const unsigned long long a = mix64(seed ^ mix64(2ull * (unsigned long long)i));
  const unsigned long long b = mix64(seed ^ mix64(2ull * (unsigned long long)i + 1ull));
  const double u1 = ((double)(a >> 11) + 0.5) * 0x1.0p-53, u2 = (double)(b >> 11) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
// Partials of the sum of src^2 over the fluid cells.
__global__ void __launch_bounds__(RB) k_sumsq_fluid(long long N, const uint8_t* __restrict__ codes,
                                                    const double* __restrict__ src, double* part) {
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < N; i += (long long)gridDim.x * RB)
    if (codes[i] & FLUID) acc += src[i] * src[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// dst = src + rel rms xi on the fluid cells (src elsewhere), rms over the fluid cells.
__global__ void __launch_bounds__(RB) k_perturb(long long N, const uint8_t* __restrict__ codes,
                                                const double* __restrict__ src, double* __restrict__ dst, double rel,
                                                unsigned long long seed, const double* part, double fluid_count) {
  const double scale = rel * sqrt(reduce_part(part, RP) / fluid_count);
  for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < N; i += (long long)gridDim.x * RB) {
    const double v = src[i];
    dst[i] = (codes[i] & FLUID) ? v + scale * normal_at(seed, i) : v;
  }
}
void launch_perturb(long long N, const uint8_t* codes, const double* src, double* dst, double rel,
                    unsigned long long seed, double* part, double fluid_count, hipStream_t s) {
  hipLaunchKernelGGL(k_sumsq_fluid, dim3(RP), dim3(RB), 0, s, N, codes, src, part);
  hipLaunchKernelGGL(k_perturb, dim3(RP), dim3(RB), 0, s, N, codes, src, dst, rel, seed, (const double*)part,
                     fluid_count);
}

}  // namespace cfdgpu
