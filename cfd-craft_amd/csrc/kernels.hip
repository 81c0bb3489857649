// CDNA4 (gfx950) kernels of the pressure solve: the masked 7-point operator,
// fused PCG vector updates with deterministic wave64 reductions, segmented
// gauge means, IC(0) hyperplane sweeps, the geometric multigrid V-cycle and
// the projection steps either side of the solve.
//
// Built with -ffp-contract=off: the FP64 operator and IC(0) sweeps repeat the
// reference's per-cell arithmetic (pressure.rs:223-241, 273-285) operation for
// operation, so their results are bitwise equal to the serial CPU code.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

namespace cfdgpu {

constexpr int BS = 256;
constexpr uint8_t FLUID = 1u << 6;
constexpr uint8_t CLOSED = 1u << 7;

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// Sum over a 256-thread block, valid in thread 0. Fixed association order.
__device__ double block_sum(double v) {
  __shared__ double s[BS / 64];
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  return (s[0] + s[1]) + (s[2] + s[3]);
}
// True in every thread of the last block to finish (release/acquire at agent scope).
__device__ bool last_block(unsigned int* ticket) {
  __shared__ unsigned int t;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) t = atomicAdd(ticket, 1u);
  __syncthreads();
  bool last = t == gridDim.x - 1;
  if (last) __threadfence();
  return last;
}
// Deterministic sum of per-block partials inside the last block.
__device__ double sum_partials(const double* part, int n) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += BS) v += __builtin_nontemporal_load(part + i);
  return block_sum(v);
}

// ---------------------------------------------------------------- geometry helpers
__device__ __forceinline__ void coords(const Geo& g, long long c, int q[3]) {
  long long r = c / g.n[0];
  q[0] = (int)(c - r * g.n[0]);
  q[1] = (int)(r % g.n[1]);
  q[2] = (int)(r / g.n[1]);
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
template <class XV>
__device__ __forceinline__ double apply_cell(const Geo& g, uint8_t code, long long c, const int q[3], XV xv) {
  double y = 0.0;
  if (!(code & 0x3f)) return y;
  const double xc = xv(c);
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
          double d = g.wi[a] * (xc - xv(hi_nb(g, c, a, t)));
          y += d;
        }
      } else {
        if (!lo) continue;
        if (t == 0 && !P) {
          y += g.wo[a] * xc;
        } else {
          double d = g.wi[a] * (xv(lo_nb(g, c, a, t)) - xc);
          y -= d;
        }
      }
    }
  }
  return y;
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

// p_new = z_c + beta p (or z_c on the first iteration), Ap = A p_new, and the
// partial sums of p_new . Ap; the last block forms alpha = rz / (p . Ap) and
// raises CFD-E032 (pressure.rs:391-397,426-432).
__global__ void __launch_bounds__(BS)
    k_pcg_op(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ z, const double* __restrict__ pold,
             double* __restrict__ pnew, double* __restrict__ ap, const double* __restrict__ mean,
             const int* __restrict__ cidx, PcgState* st, double* part) {
  if (st->status != ST_RUNNING) return;
  const bool first = st->first;
  const double beta = st->beta;
  auto pv = [&](long long i) {
    uint8_t ci = codes[i];
    double zc = z[i];
    if (ci & CLOSED) zc -= mean[cidx ? cidx[i] : 0];
    return first ? zc : zc + beta * pold[i];
  };
  double acc = 0.0;
  for (long long c = (long long)blockIdx.x * BS + threadIdx.x; c < g.N; c += (long long)gridDim.x * BS) {
    int q[3];
    coords(g, c, q);
    uint8_t code = codes[c];
    double pc = pv(c);
    double y = apply_cell(g, code, c, q, pv);
    pnew[c] = pc;
    ap[c] = y;
    acc += pc * y;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (!last_block(&st->ticket)) return;
  double denom = sum_partials(part, gridDim.x);
  if (threadIdx.x == 0) {
    st->denom = denom;
    if (denom <= 0.0 || !isfinite(denom)) {
      st->status = ST_E032;
    } else {
      st->alpha = st->rz / denom;
    }
    st->first = 0;
    st->ticket = 0;
  }
}

// x += alpha p; r -= alpha Ap; |r|^2 -> recurrent residual test (pressure.rs:397-403).
__global__ void __launch_bounds__(BS) k_pcg_axpy(long long N, const double* __restrict__ p, const double* __restrict__ ap,
                                                 double* __restrict__ x, double* __restrict__ r, PcgState* st,
                                                 double* part) {
  if (st->status != ST_RUNNING) return;
  const double alpha = st->alpha;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < N; i += (long long)gridDim.x * BS) {
    x[i] += alpha * p[i];
    double ri = r[i] - alpha * ap[i];
    r[i] = ri;
    acc += ri * ri;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (!last_block(&st->ticket)) return;
  double rr = sum_partials(part, gridDim.x);
  if (threadIdx.x == 0) {
    double rel = sqrt(rr) / st->norm;
    st->rel = rel;
    if (rel <= st->tol) st->check = 1;
    st->ticket = 0;
  }
}

// Recomputed residual |A x - b_c| / |b_c| and the acceptance gate 10 tol + 1e-12
// (pressure.rs:403-425). Runs only after the recurrent residual crossed tol.
__global__ void __launch_bounds__(BS) k_pcg_true(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ x,
                                                 const double* __restrict__ bc, PcgState* st, double* part) {
  if (st->status != ST_RUNNING || !st->check) return;
  double acc = 0.0;
  for (long long c = (long long)blockIdx.x * BS + threadIdx.x; c < g.N; c += (long long)gridDim.x * BS) {
    int q[3];
    coords(g, c, q);
    double y = apply_cell(g, codes[c], c, q, [&](long long i) { return x[i]; });
    double d = y - bc[c];
    acc += d * d;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (!last_block(&st->ticket)) return;
  double ss = sum_partials(part, gridDim.x);
  if (threadIdx.x == 0) {
    double tr = sqrt(ss) / st->norm;
    st->true_rel = tr;
    st->true_checks += 1;
    if (tr <= st->accept) {
      st->status = ST_CONVERGED;
      st->it += 1;
    } else {
      st->check = 0;
    }
    st->ticket = 0;
  }
}

// rz_next = r . z_c, beta = rz_next / rz; closes one PCG iteration and raises
// CFD-E033 at the limit (pressure.rs:426-435). With init set it forms the
// first r . z instead.
__global__ void __launch_bounds__(BS) k_pcg_rz(long long N, const uint8_t* __restrict__ codes, const double* __restrict__ r,
                                               const double* __restrict__ z, const double* __restrict__ mean,
                                               const int* __restrict__ cidx, PcgState* st, double* part, int init) {
  if (st->status != ST_RUNNING) return;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < N; i += (long long)gridDim.x * BS) {
    uint8_t ci = codes[i];
    double zc = z[i];
    if (ci & CLOSED) zc -= mean[cidx ? cidx[i] : 0];
    acc += r[i] * zc;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (!last_block(&st->ticket)) return;
  double next = sum_partials(part, gridDim.x);
  if (threadIdx.x == 0) {
    if (init) {
      st->rz = next;
      st->first = 1;
    } else {
      st->beta = next / st->rz;
      st->rz = next;
      st->it += 1;
      if (st->it >= st->max_it) st->status = ST_E033;
    }
    st->ticket = 0;
  }
}

// ---------------------------------------------------------------- gauges
// Per closed component: sum of src (and of max(|src|, scale) for the CFD-E031
// compatibility test, pressure.rs:351-362); mean = sum / members.
// One closed component: a grid-stride pass over the cells flagged closed.
__global__ void __launch_bounds__(BS) k_csum1(long long N, const uint8_t* __restrict__ codes, const double* __restrict__ src,
                                              const double* __restrict__ scale, int e031, double inv_count, double* mean,
                                              double* sums, PcgState* st, double* part, unsigned int* ticket) {
  if (st && st->status != ST_RUNNING) return;
  double acc = 0.0, acc2 = 0.0;
  for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < N; i += (long long)gridDim.x * BS) {
    if (!(codes[i] & CLOSED)) continue;
    double v = src[i];
    acc += v;
    if (e031) acc2 += fmax(fabs(v), scale ? scale[i] : 0.0);
  }
  double s = block_sum(acc);
  double s2 = e031 ? block_sum(acc2) : 0.0;
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s;
    part[2 * blockIdx.x + 1] = s2;
  }
  if (!last_block(ticket)) return;
  double v = 0.0, v2 = 0.0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += BS) {
    v += __builtin_nontemporal_load(part + 2 * i);
    v2 += __builtin_nontemporal_load(part + 2 * i + 1);
  }
  double sum = block_sum(v);
  double sum2 = e031 ? block_sum(v2) : 0.0;
  if (threadIdx.x == 0) {
    mean[0] = sum * inv_count;
    sums[0] = sum;
    sums[1] = sum2;
    if (e031 && st && fabs(sum) > 1e-9 * sum2) st->status = ST_E031;
    *ticket = 0;
  }
}

// Several closed components: one block per chunk of a component's member list,
// then the last block sums each component's chunks in order.
__global__ void __launch_bounds__(BS)
    k_csumN(const int* __restrict__ members, const int* __restrict__ chunk_begin, const int* __restrict__ chunk_comp,
            const int* __restrict__ comp_first_chunk, int n_comp, const double* __restrict__ inv_count,
            const double* __restrict__ src, const double* __restrict__ scale, int e031, double* mean, double* sums,
            PcgState* st, double* part, unsigned int* ticket) {
  if (st && st->status != ST_RUNNING) return;
  const int ch = blockIdx.x;
  const int b0 = chunk_begin[ch], b1 = chunk_begin[ch + 1];
  double acc = 0.0, acc2 = 0.0;
  for (int k = b0 + threadIdx.x; k < b1; k += BS) {
    int i = members[k];
    double v = src[i];
    acc += v;
    if (e031) acc2 += fmax(fabs(v), scale ? scale[i] : 0.0);
  }
  double s = block_sum(acc);
  double s2 = e031 ? block_sum(acc2) : 0.0;
  if (threadIdx.x == 0) {
    part[2 * ch] = s;
    part[2 * ch + 1] = s2;
  }
  if (!last_block(ticket)) return;
  bool bad = false;
  for (int comp = threadIdx.x; comp < n_comp; comp += BS) {
    double v = 0.0, v2 = 0.0;
    for (int k = comp_first_chunk[comp]; k < comp_first_chunk[comp + 1]; ++k) {
      v += __builtin_nontemporal_load(part + 2 * k);
      v2 += __builtin_nontemporal_load(part + 2 * k + 1);
    }
    mean[comp] = v * inv_count[comp];
    if (sums) {
      sums[2 * comp] = v;
      sums[2 * comp + 1] = v2;
    }
    if (e031 && fabs(v) > 1e-9 * v2) bad = true;
  }
  (void)chunk_comp;
  __shared__ int any_bad;
  if (threadIdx.x == 0) any_bad = 0;
  __syncthreads();
  if (bad) atomicOr(&any_bad, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (st && any_bad) st->status = ST_E031;
    *ticket = 0;
  }
}

// r = b - mean (closed cells), b_c = r, x = 0, |r| -> norm; a zero norm ends the
// solve with zero iterations (pressure.rs:365-375).
__global__ void __launch_bounds__(BS) k_pcg_init(long long N, const uint8_t* __restrict__ codes, const double* __restrict__ b,
                                                 const double* __restrict__ mean, const int* __restrict__ cidx,
                                                 double* __restrict__ r, double* __restrict__ bc, double* __restrict__ x,
                                                 const double* __restrict__ x0, PcgState* st, double* part) {
  if (st->status != ST_RUNNING) return;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < N; i += (long long)gridDim.x * BS) {
    double v = b[i];
    uint8_t ci = codes[i];
    if (ci & CLOSED) v -= mean[cidx ? cidx[i] : 0];
    r[i] = v;
    bc[i] = v;
    x[i] = 0.0;
    acc += v * v;
  }
  double s = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  if (!last_block(&st->ticket)) return;
  double rr = sum_partials(part, gridDim.x);
  if (threadIdx.x == 0) {
    double norm = sqrt(rr);
    st->norm = norm;
    if (norm < 1e-24) st->status = ST_ZERO;
    st->ticket = 0;
  }
  (void)x0;
}

// Warm start: x = x0, r = b_c - A x0 (b_c already formed), |b_c| from the caller's norm pass.
__global__ void __launch_bounds__(BS) k_warm_residual(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ x0,
                                                      const double* __restrict__ bc, double* __restrict__ x,
                                                      double* __restrict__ r, PcgState* st) {
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
// Forward sweep of the hyperplane i + j + k = L, gather form of Factor::apply
// (pressure.rs:225-232): y = (((r - l y_zlo) - l y_ylo) - l y_xlo) / D with the
// lower neighbours in ascending index order (periodic wraps included) and
// l = a / D[col], the reference's factor entries (no fill without a 3-cell
// periodic axis).
__global__ void __launch_bounds__(BS) k_ic0_fwd(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ D,
                                                const double* __restrict__ r, double* __restrict__ y, int L, int j0,
                                                int k0, int nj, int nk, const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
  int t = blockIdx.x * BS + threadIdx.x;
  if (t >= nj * nk) return;
  int j = j0 + t % nj, k = k0 + t / nj;
  int i = L - j - k;
  if (i < 0 || i >= g.n[0]) return;
  long long c = (long long)i + j * g.sy + k * g.sz;
  const uint8_t code = codes[c];
  const int q[3] = {i, j, k};
  double value = r[c];
#pragma unroll
  for (int a = 2; a >= 0; --a) {
    const int n = g.n[a], tq = q[a];
    const long long s = stride_of(g, a);
    const double w = g.wi[a];
    if (!g.per[a] || n == 1) {
      if (tq > 0 && (code & (1u << (2 * a)))) {
        double l = -w / D[c - s];
        value -= l * y[c - s];
      }
    } else if (n == 2) {
      if (tq == 1 && (code & (1u << (2 * a)))) {
        double l = (-w - w) / D[c - s];
        value -= l * y[c - s];
      }
    } else {
      if (tq == n - 1 && (code & (2u << (2 * a)))) {
        long long nb = c - (long long)(n - 1) * s;
        double l = -w / D[nb];
        value -= l * y[nb];
      }
      if (tq > 0 && (code & (1u << (2 * a)))) {
        double l = -w / D[c - s];
        value -= l * y[c - s];
      }
    }
  }
  y[c] = value / D[c];
}

// Backward sweep (pressure.rs:233-240) in gather form: the reference's scatter
// z[col] -= l * x_i over rows i in descending order becomes, for cell c, the
// subtraction of its upper neighbours' terms in descending index order, then
// the division by D[c]; l = a / D[c].
__global__ void __launch_bounds__(BS) k_ic0_bwd(Geo g, const uint8_t* __restrict__ codes, const double* __restrict__ D,
                                                double* __restrict__ z, int L, int j0, int k0, int nj, int nk,
                                                const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
  int t = blockIdx.x * BS + threadIdx.x;
  if (t >= nj * nk) return;
  int j = j0 + t % nj, k = k0 + t / nj;
  int i = L - j - k;
  if (i < 0 || i >= g.n[0]) return;
  long long c = (long long)i + j * g.sy + k * g.sz;
  const uint8_t code = codes[c];
  const int q[3] = {i, j, k};
  const double dc = D[c];
  double value = z[c];
#pragma unroll
  for (int a = 2; a >= 0; --a) {
    const int n = g.n[a], tq = q[a];
    const long long s = stride_of(g, a);
    const double w = g.wi[a];
    if (!g.per[a] || n == 1) {
      if (tq < n - 1 && (code & (2u << (2 * a)))) {
        double l = -w / dc;
        value -= l * z[c + s];
      }
    } else if (n == 2) {
      if (tq == 0 && (code & (2u << (2 * a)))) {
        double l = (-w - w) / dc;
        value -= l * z[c + s];
      }
    } else {
      if (tq == 0 && (code & (1u << (2 * a)))) {
        long long nb = c + (long long)(n - 1) * s;
        double l = -w / dc;
        value -= l * z[nb];
      }
      if (tq < n - 1 && (code & (2u << (2 * a)))) {
        double l = -w / dc;
        value -= l * z[c + s];
      }
    }
  }
  z[c] = value / dc;
}

// ---------------------------------------------------------------- multigrid (geometric)
// Red-black Gauss-Seidel half-sweep on the fine level from the coupling codes.
// `first`: x starts at zero (neighbours contribute nothing). `xc`: fuse the
// prolongation x += P e of the coarse correction into the neighbour reads (the
// cells of this colour are overwritten, so only neighbours need it). `out64`:
// also store the result in FP64.
template <class T, class TB>
__global__ void __launch_bounds__(BS)
    k_fine_gs(Geo g, const uint8_t* __restrict__ codes, const TB* __restrict__ b, T* __restrict__ x, int color, int first,
              const T* __restrict__ xc, Geo gc, double* __restrict__ out64, int all_cells, const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
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
    const int hx = (g.n[0] + 1) >> 1;
    long long row = t / hx;
    if (row >= (long long)g.n[1] * g.n[2]) return;
    int i2 = (int)(t - row * hx);
    j = (int)(row % g.n[1]);
    k = (int)(row / g.n[1]);
    i = 2 * i2 + ((j + k + color) & 1);
    if (i >= g.n[0]) return;
  }
  const long long c = (long long)i + j * g.sy + k * g.sz;
  if (((i + j + k) & 1) != color) {
    if (out64) out64[c] = (double)x[c];
    return;
  }
  const uint8_t code = codes[c];
  if (!(code & FLUID)) {
    x[c] = T(0);
    if (out64) out64[c] = 0.0;
    return;
  }
  const int q[3] = {i, j, k};
  T diag = T(0);
  T value = (T)b[c];
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
      long long nb = side == 0 ? lo_nb(g, c, a, q[a]) : hi_nb(g, c, a, q[a]);
      T xn = x[nb];
      if (xc) {
        int qn[3] = {i, j, k};
        qn[a] = side == 0 ? (q[a] > 0 ? q[a] - 1 : g.n[a] - 1) : (q[a] < g.n[a] - 1 ? q[a] + 1 : 0);
        xn += xc[(qn[0] >> 1) + (qn[1] >> 1) * gc.sy + (qn[2] >> 1) * gc.sz];
      }
      value += w * xn;
    }
  }
  T v = diag > T(0) ? value / diag : T(0);
  x[c] = v;
  if (out64) out64[c] = (double)v;
}

// Fine residual r = b - A x summed over each 2x2x2 bin: the coarse right-hand side P^T r.
template <class T, class TB>
__global__ void __launch_bounds__(BS) k_fine_restrict(Geo g, const uint8_t* __restrict__ codes, const TB* __restrict__ b,
                                                      const T* __restrict__ x, Geo gc, T* __restrict__ bc,
                                                      const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
  long long C = (long long)blockIdx.x * BS + threadIdx.x;
  if (C >= gc.N) return;
  int Q[3];
  coords(gc, C, Q);
  T sum = T(0);
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int q[3] = {2 * Q[0] + dx, 2 * Q[1] + dy, 2 * Q[2] + dz};
        if (q[0] >= g.n[0] || q[1] >= g.n[1] || q[2] >= g.n[2]) continue;
        const long long c = (long long)q[0] + q[1] * g.sy + q[2] * g.sz;
        const uint8_t code = codes[c];
        if (!(code & FLUID)) continue;
        const T xv = x[c];
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
            long long nb = side == 0 ? lo_nb(g, c, a, q[a]) : hi_nb(g, c, a, q[a]);
            off += (T)g.wi[a] * x[nb];
          }
        sum += ((T)b[c] - (ax * xv - off));
      }
  bc[C] = sum;
}

// Coarse-level red-black half-sweep with explicit per-face weights.
template <class T>
__global__ void __launch_bounds__(BS) k_coarse_gs(DevLevel<T> L, int color, int first, const T* __restrict__ xc, Geo gc,
                                                  const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
  const Geo& g = L.g;
  long long t = (long long)blockIdx.x * BS + threadIdx.x;
  const int hx = (g.n[0] + 1) >> 1;
  long long row = t / hx;
  if (row >= (long long)g.n[1] * g.n[2]) return;
  int i2 = (int)(t - row * hx);
  int j = (int)(row % g.n[1]);
  int k = (int)(row / g.n[1]);
  int i = 2 * i2 + ((j + k + color) & 1);
  if (i >= g.n[0]) return;
  const long long c = (long long)i + j * g.sy + k * g.sz;
  const T d = L.diag[c];
  if (!(d > T(0))) {
    L.x[c] = T(0);
    return;
  }
  T value = L.b[c];
  if (!first) {
    const int q[3] = {i, j, k};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int n = g.n[a];
      if (q[a] < n - 1 || g.per[a]) {
        T w = L.wp[a][c];
        if (w != T(0)) {
          long long nb = hi_nb(g, c, a, q[a]);
          T xn = L.x[nb];
          if (xc) {
            int qn[3] = {i, j, k};
            qn[a] = q[a] < n - 1 ? q[a] + 1 : 0;
            xn += xc[(qn[0] >> 1) + (qn[1] >> 1) * gc.sy + (qn[2] >> 1) * gc.sz];
          }
          value += w * xn;
        }
      }
      if (q[a] > 0 || g.per[a]) {
        long long nb = lo_nb(g, c, a, q[a]);
        T w = L.wp[a][nb];
        if (w != T(0)) {
          T xn = L.x[nb];
          if (xc) {
            int qn[3] = {i, j, k};
            qn[a] = q[a] > 0 ? q[a] - 1 : n - 1;
            xn += xc[(qn[0] >> 1) + (qn[1] >> 1) * gc.sy + (qn[2] >> 1) * gc.sz];
          }
          value += w * xn;
        }
      }
    }
  }
  L.x[c] = value / d;
}

template <class T>
__global__ void __launch_bounds__(BS) k_coarse_restrict(DevLevel<T> L, Geo gc, T* __restrict__ bc, const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
  const Geo& g = L.g;
  long long C = (long long)blockIdx.x * BS + threadIdx.x;
  if (C >= gc.N) return;
  int Q[3];
  coords(gc, C, Q);
  T sum = T(0);
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int q[3] = {2 * Q[0] + dx, 2 * Q[1] + dy, 2 * Q[2] + dz};
        if (q[0] >= g.n[0] || q[1] >= g.n[1] || q[2] >= g.n[2]) continue;
        const long long c = (long long)q[0] + q[1] * g.sy + q[2] * g.sz;
        const T d = L.diag[c];
        if (!(d > T(0))) continue;
        T off = T(0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const int n = g.n[a];
          if (q[a] < n - 1 || g.per[a]) {
            T w = L.wp[a][c];
            if (w != T(0)) off += w * L.x[hi_nb(g, c, a, q[a])];
          }
          if (q[a] > 0 || g.per[a]) {
            long long nb = lo_nb(g, c, a, q[a]);
            T w = L.wp[a][nb];
            if (w != T(0)) off += w * L.x[nb];
          }
        }
        sum += L.b[c] - (d * L.x[c] - off);
      }
  bc[C] = sum;
}

// Coarsest level: x = G b with the dense symmetric pseudo-inverse (FP64), one wave per row.
template <class TI, class TO>
__global__ void __launch_bounds__(BS) k_direct(int n, const double* __restrict__ G, const TI* __restrict__ b,
                                               TO* __restrict__ x, double* __restrict__ out64, const PcgState* st) {
  if (st && st->status != ST_RUNNING) return;
  int row = blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  double acc = 0.0;
  if (row < n)
    for (int j = lane; j < n; j += 64) acc += G[(long long)row * n + j] * (double)b[j];
  acc = wave_sum(acc);
  if (row < n && lane == 0) {
    x[row] = (TO)acc;
    if (out64) out64[row] = (double)(TO)acc;
  }
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

// ---------------------------------------------------------------- launchers
static inline unsigned blocks_for(long long n) { return (unsigned)((n + BS - 1) / BS); }

void launch_apply(const Geo& g, const uint8_t* codes, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_apply, dim3(blocks_for(g.N)), dim3(BS), 0, s, g, codes, x, y);
}
void launch_pcg_op(const Geo& g, int nb, const uint8_t* codes, const double* z, const double* pold, double* pnew, double* ap,
                   const double* mean, const int* cidx, PcgState* st, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_op, dim3(nb), dim3(BS), 0, s, g, codes, z, pold, pnew, ap, mean, cidx, st, part);
}
void launch_pcg_axpy(long long N, int nb, const double* p, const double* ap, double* x, double* r, PcgState* st,
                     double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_axpy, dim3(nb), dim3(BS), 0, s, N, p, ap, x, r, st, part);
}
void launch_pcg_true(const Geo& g, int nb, const uint8_t* codes, const double* x, const double* bc, PcgState* st,
                     double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_true, dim3(nb), dim3(BS), 0, s, g, codes, x, bc, st, part);
}
void launch_pcg_rz(long long N, int nb, const uint8_t* codes, const double* r, const double* z, const double* mean,
                   const int* cidx, PcgState* st, double* part, int init, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_rz, dim3(nb), dim3(BS), 0, s, N, codes, r, z, mean, cidx, st, part, init);
}
void launch_pcg_init(long long N, int nb, const uint8_t* codes, const double* b, const double* mean, const int* cidx,
                     double* r, double* bc, double* x, PcgState* st, double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_init, dim3(nb), dim3(BS), 0, s, N, codes, b, mean, cidx, r, bc, x, (const double*)nullptr, st,
                     part);
}
void launch_warm_residual(const Geo& g, const uint8_t* codes, const double* x0, const double* bc, double* x, double* r,
                          PcgState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_warm_residual, dim3(blocks_for(g.N)), dim3(BS), 0, s, g, codes, x0, bc, x, r, st);
}
void launch_state_init(PcgState* st, double tol, long long max_it, hipStream_t s) {
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, st, tol, max_it);
}
void launch_csum1(long long N, int nb, const uint8_t* codes, const double* src, const double* scale, int e031,
                  double inv_count, double* mean, double* sums, PcgState* st, double* part, unsigned int* ticket,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_csum1, dim3(nb), dim3(BS), 0, s, N, codes, src, scale, e031, inv_count, mean, sums, st, part,
                     ticket);
}
void launch_csumN(int n_chunks, const int* members, const int* chunk_begin, const int* comp_first_chunk, int n_comp,
                  const double* inv_count, const double* src, const double* scale, int e031, double* mean, double* sums,
                  PcgState* st, double* part, unsigned int* ticket, hipStream_t s) {
  hipLaunchKernelGGL(k_csumN, dim3(n_chunks), dim3(BS), 0, s, members, chunk_begin, (const int*)nullptr, comp_first_chunk,
                     n_comp, inv_count, src, scale, e031, mean, sums, st, part, ticket);
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

void launch_ic0(const Geo& g, const uint8_t* codes, const double* D, const double* r, double* z, const PcgState* st,
                hipStream_t s) {
  const int levels = g.n[0] + g.n[1] + g.n[2] - 2;
  for (int L = 0; L < levels; ++L) {
    int j0, k0, nj, nk;
    level_box(g, L, j0, k0, nj, nk);
    if (nj * nk == 0) continue;
    hipLaunchKernelGGL(k_ic0_fwd, dim3(blocks_for((long long)nj * nk)), dim3(BS), 0, s, g, codes, D, r, z, L, j0, k0, nj,
                       nk, st);
  }
  for (int L = levels - 1; L >= 0; --L) {
    int j0, k0, nj, nk;
    level_box(g, L, j0, k0, nj, nk);
    if (nj * nk == 0) continue;
    hipLaunchKernelGGL(k_ic0_bwd, dim3(blocks_for((long long)nj * nk)), dim3(BS), 0, s, g, codes, D, z, L, j0, k0, nj, nk,
                       st);
  }
}

template <class T>
static unsigned color_blocks(const Geo& g) {
  long long n = (long long)((g.n[0] + 1) >> 1) * g.n[1] * g.n[2];
  return blocks_for(n);
}

template <class T, class TB>
void launch_vcycle(const Geo& g0, const uint8_t* codes, const TB* b0, T* x0, int nlev, DevLevel<T>* lv, int direct_n,
                   const double* G, double* out64, const PcgState* st, hipStream_t s) {
  // Level 0 is the fine grid; lv[0..nlev) are the coarse levels, lv[nlev-1] solved directly.
  const unsigned fb = color_blocks<T>(g0);
  if (nlev == 0) {
    // No coarse level: the fine grid itself is tiny and solved directly (T == double here).
    hipLaunchKernelGGL((k_direct<TB, T>), dim3((direct_n + 3) / 4), dim3(BS), 0, s, direct_n, G, b0, x0, out64, st);
    return;
  }
  const Geo& g1 = lv[0].g;
  // pre-smoothing on the fine level
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, 0, 1, (const T*)nullptr, g1,
                     (double*)nullptr, 0, st);
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, 1, 0, (const T*)nullptr, g1,
                     (double*)nullptr, 0, st);
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, 0, 0, (const T*)nullptr, g1,
                     (double*)nullptr, 0, st);
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, 1, 0, (const T*)nullptr, g1,
                     (double*)nullptr, 0, st);
  hipLaunchKernelGGL((k_fine_restrict<T, TB>), dim3(blocks_for(g1.N)), dim3(BS), 0, s, g0, codes, b0, (const T*)x0, g1,
                     lv[0].b, st);
  for (int l = 0; l + 1 < nlev; ++l) {
    const unsigned cb = color_blocks<T>(lv[l].g);
    const Geo& gn = lv[l + 1].g;
    hipLaunchKernelGGL((k_coarse_gs<T>), dim3(cb), dim3(BS), 0, s, lv[l], 0, 1, (const T*)nullptr, gn, st);
    hipLaunchKernelGGL((k_coarse_gs<T>), dim3(cb), dim3(BS), 0, s, lv[l], 1, 0, (const T*)nullptr, gn, st);
    hipLaunchKernelGGL((k_coarse_gs<T>), dim3(cb), dim3(BS), 0, s, lv[l], 0, 0, (const T*)nullptr, gn, st);
    hipLaunchKernelGGL((k_coarse_gs<T>), dim3(cb), dim3(BS), 0, s, lv[l], 1, 0, (const T*)nullptr, gn, st);
    hipLaunchKernelGGL((k_coarse_restrict<T>), dim3(blocks_for(gn.N)), dim3(BS), 0, s, lv[l], gn, lv[l + 1].b, st);
  }
  hipLaunchKernelGGL((k_direct<T, T>), dim3((direct_n + 3) / 4), dim3(BS), 0, s, direct_n, G, (const T*)lv[nlev - 1].b,
                     lv[nlev - 1].x, (double*)nullptr, st);
  for (int l = nlev - 2; l >= 0; --l) {
    const unsigned cb = color_blocks<T>(lv[l].g);
    const Geo& gn = lv[l + 1].g;
    hipLaunchKernelGGL((k_coarse_gs<T>), dim3(cb), dim3(BS), 0, s, lv[l], 1, 0, (const T*)lv[l + 1].x, gn, st);
    hipLaunchKernelGGL((k_coarse_gs<T>), dim3(cb), dim3(BS), 0, s, lv[l], 0, 0, (const T*)nullptr, gn, st);
    hipLaunchKernelGGL((k_coarse_gs<T>), dim3(cb), dim3(BS), 0, s, lv[l], 1, 0, (const T*)nullptr, gn, st);
    hipLaunchKernelGGL((k_coarse_gs<T>), dim3(cb), dim3(BS), 0, s, lv[l], 0, 0, (const T*)nullptr, gn, st);
  }
  // post-smoothing on the fine level; in mixed precision the last two half-sweeps also write z in FP64
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, 1, 0, (const T*)lv[0].x, g1,
                     (double*)nullptr, 0, st);
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, 0, 0, (const T*)nullptr, g1,
                     (double*)nullptr, 0, st);
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, 1, 0, (const T*)nullptr, g1,
                     out64, 0, st);
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(fb), dim3(BS), 0, s, g0, codes, b0, x0, 0, 0, (const T*)nullptr, g1, out64,
                     0, st);
}

template <class T, class TB>
void launch_fine_gs(const Geo& g0, const uint8_t* codes, const TB* b0, T* x0, int color, const Geo& g1, hipStream_t s) {
  hipLaunchKernelGGL((k_fine_gs<T, TB>), dim3(color_blocks<T>(g0)), dim3(BS), 0, s, g0, codes, b0, x0, color, 0,
                     (const T*)nullptr, g1, (double*)nullptr, 0, (const PcgState*)nullptr);
}
template void launch_fine_gs<double, double>(const Geo&, const uint8_t*, const double*, double*, int, const Geo&, hipStream_t);
template void launch_fine_gs<float, double>(const Geo&, const uint8_t*, const double*, float*, int, const Geo&, hipStream_t);

template void launch_vcycle<double, double>(const Geo&, const uint8_t*, const double*, double*, int, DevLevel<double>*, int,
                                            const double*, double*, const PcgState*, hipStream_t);
template void launch_vcycle<float, double>(const Geo&, const uint8_t*, const double*, float*, int, DevLevel<float>*, int,
                                           const double*, double*, const PcgState*, hipStream_t);

void launch_projection_rhs(const Geo& g, const FaceGeo& fg, const uint8_t* codes, const double* u, double dt, double* b,
                           double* terms, hipStream_t s) {
  hipLaunchKernelGGL(k_projection_rhs, dim3(blocks_for(g.N)), dim3(BS), 0, s, g, fg, codes, u, dt, b, terms);
}
void launch_project_faces(const Geo& g, const FaceGeo& fg, const uint8_t* codes, const double* phi, double dt, double* u,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_project_faces, dim3(blocks_for(fg.total)), dim3(BS), 0, s, g, fg, codes, phi, dt, u);
}

}  // namespace cfdgpu
