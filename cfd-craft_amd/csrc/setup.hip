// Device set-up of a grid's topology: the mask rasterization (geometry.rs:113-193),
// the per-cell face codes (grid.rs:341-429) and the fluid components with their
// anchoring (grid.rs:430-462) on the GPU, giving exactly the HostGrid that
// build_host_grid gives on the host.
//
// Components come from a lock-free union-find over the coupled faces in which a
// root is always linked below a smaller index (the hooking scheme of ECL-CC,
// Jaiganesh and Burtscher, 2018), so every component ends with its smallest cell
// as root. The reference numbers components by breadth-first search from the
// lowest unassigned cell, which is the order of their smallest cells: the roots,
// counted in index order, give the same numbering.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

namespace cfdgpu {
namespace {

constexpr int TB = 256;     // threads per block of the per-cell kernels
constexpr int SB = 1024;    // cells per block of the root numbering

__device__ __forceinline__ uint8_t side(int axis, int s) { return (uint8_t)(1u << (2 * axis + s)); }

struct RasterBlock {
  int r0[3], r1[3];  // cell ranges whose centers lie within the bounds (block_ranges)
  double min_m[3], max_m[3];
  int nv;            // polygon vertices (0: axis-aligned box)
  int voff;          // first vertex in the vertex array
};

struct Dims {
  int n[3];
  int per[3];
  double dx[3];
  long long N;
};

__device__ __forceinline__ void cell_coords(const Dims& d, long long c, int q[3]) {
  q[0] = (int)(c % d.n[0]);
  const long long r = c / d.n[0];
  q[1] = (int)(r % d.n[1]);
  q[2] = (int)(r / d.n[1]);
}

// Polygon membership by even-odd crossings (geometry.rs:13-26), the host's arithmetic.
__device__ bool in_polygon(const double* p, int n, double x, double y) {
  bool inside = false;
  for (int i = 0; i < n; ++i) {
    const double* a = p + 2 * i;
    const double* b = p + 2 * ((i + 1) % n);
    if ((a[1] > y) != (b[1] > y)) {
      const double* lo = a[1] < b[1] ? a : b;
      const double* hi = a[1] < b[1] ? b : a;
      if (x < lo[0] + (y - lo[1]) * (hi[0] - lo[0]) / (hi[1] - lo[1])) inside = !inside;
    }
  }
  return inside;
}

// Solid where some block contains the cell: within the block's cell ranges, the
// center (x, y of the cell, z of the range's first cell, as the host tests one
// center per column) within its bounds and, for a prism, inside its footprint.
__global__ void __launch_bounds__(TB) k_raster(Dims d, const RasterBlock* __restrict__ blocks, int nb,
                                               const double* __restrict__ verts, uint8_t* __restrict__ fluid,
                                               int* __restrict__ hit) {
  const long long c = (long long)blockIdx.x * TB + threadIdx.x;
  if (c >= d.N) return;
  int q[3];
  cell_coords(d, c, q);
  uint8_t f = 1;
  for (int b = 0; b < nb; ++b) {
    const RasterBlock& B = blocks[b];
    if (q[0] < B.r0[0] || q[0] >= B.r1[0] || q[1] < B.r0[1] || q[1] >= B.r1[1] || q[2] < B.r0[2] || q[2] >= B.r1[2])
      continue;
    const double p[3] = {((double)q[0] + 0.5) * d.dx[0], ((double)q[1] + 0.5) * d.dx[1],
                         ((double)B.r0[2] + 0.5) * d.dx[2]};
    bool inside = true;
    for (int a = 0; a < 3; ++a) inside = inside && p[a] >= B.min_m[a] && p[a] < B.max_m[a];
    if (inside && B.nv > 0) inside = in_polygon(verts + 2 * B.voff, B.nv, p[0], p[1]);
    if (!inside) continue;
    f = 0;
    hit[b] = 1;
  }
  fluid[c] = f;
}

// The neighbour across side `side` of axis a, wrapping on a periodic axis; -1 at
// a domain face.
__device__ __forceinline__ long long neighbour(const Dims& d, long long c, const int q[3], int a, int side) {
  const long long stride = a == 0 ? 1 : a == 1 ? (long long)d.n[0] : (long long)d.n[0] * d.n[1];
  const int n = d.n[a];
  if (side == 0) return q[a] > 0 ? c - stride : (d.per[a] ? c + (long long)(n - 1) * stride : -1);
  return q[a] + 1 < n ? c + stride : (d.per[a] ? c - (long long)(n - 1) * stride : -1);
}

// Face codes (grid.rs:341-429): a side couples when its face is Interior (a fluid
// neighbour other than the cell itself) or an Outlet (an open domain face the
// flow leaves through; `outlet` holds that per axis, side and z layer).
__global__ void __launch_bounds__(TB) k_codes(Dims d, const uint8_t* __restrict__ fluid,
                                              const uint8_t* __restrict__ outlet, uint8_t* __restrict__ codes) {
  const long long c = (long long)blockIdx.x * TB + threadIdx.x;
  if (c >= d.N) return;
  if (!fluid[c]) {
    codes[c] = 0;
    return;
  }
  int q[3];
  cell_coords(d, c, q);
  uint8_t code = kFluid;
  for (int a = 0; a < 3; ++a)
    for (int sd = 0; sd < 2; ++sd) {
      const long long other = neighbour(d, c, q, a, sd);
      bool coupled;
      if (other >= 0)
        coupled = other != c && fluid[other];
      else
        coupled = outlet[(2 * a + sd) * d.n[2] + q[2]] != 0;
      if (coupled) code |= side(a, sd);
    }
  codes[c] = code;
}

__device__ __forceinline__ int load_parent(int* parent, int x) {
  return __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The root of x. Links only ever point to smaller indices, so the walk strictly
// decreases and ends; it shortcuts the entries it passes (to an ancestor, which
// keeps every path intact).
__device__ int find_root(int* parent, int x) {
  int cur = load_parent(parent, x);
  if (cur == x) return x;
  int prev = x;
  for (;;) {
    const int next = load_parent(parent, cur);
    if (next >= cur) return cur;
    __hip_atomic_store(parent + prev, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    prev = cur;
    cur = next;
  }
}

__global__ void __launch_bounds__(TB) k_uf_init(long long N, int* __restrict__ parent) {
  const long long c = (long long)blockIdx.x * TB + threadIdx.x;
  if (c < N) parent[c] = (int)c;
}

// Unites each fluid cell with its coupled upper neighbours (every interior face
// once, periodic wraps included). A failed link means the root was linked
// meanwhile to a smaller index: the loop continues from there, so its value
// strictly decreases and the loop ends.
__global__ void __launch_bounds__(TB) k_uf_hook(Dims d, const uint8_t* __restrict__ codes, int* parent) {
  const long long c = (long long)blockIdx.x * TB + threadIdx.x;
  if (c >= d.N) return;
  const uint8_t code = codes[c];
  if (!(code & kFluid)) return;
  int q[3];
  cell_coords(d, c, q);
  for (int a = 0; a < 3; ++a) {
    if (!(code & side(a, 1))) continue;
    const long long other = neighbour(d, c, q, a, 1);
    if (other < 0) continue;  // an outlet
    int u = find_root(parent, (int)c), v = find_root(parent, (int)other);
    while (u != v) {
      if (u < v) {
        const int t = u;
        u = v;
        v = t;
      }
      // link the larger root u below v
      const int seen = atomicCAS(parent + u, u, v);
      if (seen == u) break;
      u = find_root(parent, seen);
      v = find_root(parent, v);
    }
  }
}

// The final root of every cell into `root` (a read-only walk: once hooking is
// done no link changes, and writing into `parent` here could race with another
// thread's walk).
__global__ void __launch_bounds__(TB) k_uf_roots(long long N, const int* __restrict__ parent, int* __restrict__ root) {
  const long long c = (long long)blockIdx.x * TB + threadIdx.x;
  if (c >= N) return;
  int x = (int)c, p = parent[x];
  while (p != x) {
    x = p;
    p = parent[x];
  }
  root[c] = x;
}

// Roots (fluid cells that are their component's smallest cell) per block of SB cells.
__global__ void __launch_bounds__(SB) k_root_count(long long N, const uint8_t* __restrict__ fluid,
                                                   const int* __restrict__ root, int* __restrict__ counts) {
  const long long c = (long long)blockIdx.x * SB + threadIdx.x;
  const bool is_root = c < N && fluid[c] && root[c] == (int)c;
  const int n = __syncthreads_count(is_root);
  if (threadIdx.x == 0) counts[blockIdx.x] = n;
}

// The component number of each root: its rank among the roots in index order.
__global__ void __launch_bounds__(SB) k_root_ids(long long N, const uint8_t* __restrict__ fluid,
                                                 const int* __restrict__ root, const int* __restrict__ offsets,
                                                 int* __restrict__ id) {
  __shared__ int wave_total[SB / 64];
  const long long c = (long long)blockIdx.x * SB + threadIdx.x;
  const bool is_root = c < N && fluid[c] && root[c] == (int)c;
  const unsigned long long mask = __ballot(is_root);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int before = __popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) wave_total[wave] = __popcll(mask);
  __syncthreads();
  int base = offsets[blockIdx.x];
  for (int w = 0; w < wave; ++w) base += wave_total[w];
  if (is_root) id[c] = base + before;
}

// Component of every cell (-1 solid) and the anchored flag of every component:
// a coupled side at a non-periodic domain face is an outlet.
__global__ void __launch_bounds__(TB) k_assign(Dims d, const uint8_t* __restrict__ codes, const int* __restrict__ root,
                                               const int* __restrict__ id, int* __restrict__ comp,
                                               uint8_t* __restrict__ anchored) {
  const long long c = (long long)blockIdx.x * TB + threadIdx.x;
  if (c >= d.N) return;
  const uint8_t code = codes[c];
  if (!(code & kFluid)) {
    comp[c] = -1;
    return;
  }
  const int k = id[root[c]];
  comp[c] = k;
  int q[3];
  cell_coords(d, c, q);
  for (int a = 0; a < 3; ++a)
    for (int sd = 0; sd < 2; ++sd)
      if ((code & side(a, sd)) && neighbour(d, c, q, a, sd) < 0) anchored[k] = 1;
}

struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int alloc(T** out, size_t n) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) {
      set_error(std::string("device set-up: hipMalloc: ") + hipGetErrorString(e));
      return CFD_ENOMEM;
    }
    ptrs.push_back(p);
    *out = (T*)p;
    return 0;
  }
};

#define SETUP_TRY(expr)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      set_error(std::string("device set-up: ") + #expr + ": " + hipGetErrorString(e_)); \
      return CFD_EDEVICE;                                                               \
    }                                                                                   \
  } while (0)

unsigned blocks(long long n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

int build_device_grid(const cfd_case& c, HostGrid& g) {
  int rc = grid_header(c, g);
  if (rc) return rc;
  const long long N = g.count;
  if (N > 0x7fffffffLL) {
    set_error("device set-up: more cells than 32-bit component labels hold");
    return CFD_EUNSUPPORTED;
  }
  Dims d{};
  for (int a = 0; a < 3; ++a) {
    d.n[a] = (int)g.n[a];
    d.per[a] = g.periodic[a] ? 1 : 0;
    d.dx[a] = g.dx[a];
  }
  d.N = N;
  // the blocks' cell ranges, from the host's search over cell centers
  std::vector<RasterBlock> rb((size_t)c.n_blocks);
  std::vector<double> verts;
  for (int b = 0; b < c.n_blocks; ++b) {
    const cfd_block& B = c.blocks[b];
    int64_t r0[3], r1[3];
    if (!block_ranges(B, g.n, g.dx, r0, r1)) return block_error(c, b);
    RasterBlock& R = rb[(size_t)b];
    for (int a = 0; a < 3; ++a) {
      R.r0[a] = (int)r0[a];
      R.r1[a] = (int)r1[a];
      R.min_m[a] = B.min_m[a];
      R.max_m[a] = B.max_m[a];
    }
    R.nv = B.n_vertices;
    R.voff = (int)(verts.size() / 2);
    if (B.n_vertices > 0) verts.insert(verts.end(), B.footprint_xy, B.footprint_xy + 2 * B.n_vertices);
  }
  std::vector<uint8_t> outlet;
  open_outlets(c, g, outlet);

  Scratch S;
  RasterBlock* d_blocks;
  double* d_verts;
  int *d_hit, *d_parent, *d_root, *d_id, *d_comp, *d_counts;
  uint8_t *d_fluid, *d_codes, *d_outlet, *d_anch;
  const unsigned nsb = blocks(N, SB);
  if ((rc = S.alloc(&d_blocks, rb.size())) || (rc = S.alloc(&d_verts, verts.size())) ||
      (rc = S.alloc(&d_hit, rb.size())) || (rc = S.alloc(&d_parent, (size_t)N)) || (rc = S.alloc(&d_root, (size_t)N)) ||
      (rc = S.alloc(&d_id, (size_t)N)) ||
      (rc = S.alloc(&d_comp, (size_t)N)) || (rc = S.alloc(&d_counts, nsb)) || (rc = S.alloc(&d_fluid, (size_t)N)) ||
      (rc = S.alloc(&d_codes, (size_t)N)) || (rc = S.alloc(&d_outlet, outlet.size())))
    return rc;
  hipStream_t s = nullptr;
  SETUP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  if (!rb.empty()) {
    SETUP_TRY(hipMemcpyAsync(d_blocks, rb.data(), rb.size() * sizeof(RasterBlock), hipMemcpyHostToDevice, s));
    SETUP_TRY(hipMemsetAsync(d_hit, 0, rb.size() * sizeof(int), s));
  }
  if (!verts.empty())
    SETUP_TRY(hipMemcpyAsync(d_verts, verts.data(), verts.size() * sizeof(double), hipMemcpyHostToDevice, s));
  SETUP_TRY(hipMemcpyAsync(d_outlet, outlet.data(), outlet.size(), hipMemcpyHostToDevice, s));
  const unsigned nb = blocks(N, TB);
  hipLaunchKernelGGL(k_raster, dim3(nb), dim3(TB), 0, s, d, d_blocks, c.n_blocks, d_verts, d_fluid, d_hit);
  // CFD-E028 for the first block that contains no cell center
  std::vector<int> hit(rb.size(), 0);
  if (!rb.empty()) SETUP_TRY(hipMemcpyAsync(hit.data(), d_hit, hit.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  g.fluid.resize((size_t)N);
  SETUP_TRY(hipMemcpyAsync(g.fluid.data(), d_fluid, (size_t)N, hipMemcpyDeviceToHost, s));
  SETUP_TRY(hipStreamSynchronize(s));
  for (int b = 0; b < c.n_blocks; ++b)
    if (!hit[(size_t)b]) return block_error(c, b);
  if ((rc = count_fluid(g))) return rc;
  hipLaunchKernelGGL(k_codes, dim3(nb), dim3(TB), 0, s, d, d_fluid, d_outlet, d_codes);
  hipLaunchKernelGGL(k_uf_init, dim3(nb), dim3(TB), 0, s, N, d_parent);
  hipLaunchKernelGGL(k_uf_hook, dim3(nb), dim3(TB), 0, s, d, d_codes, d_parent);
  hipLaunchKernelGGL(k_uf_roots, dim3(nb), dim3(TB), 0, s, N, d_parent, d_root);
  hipLaunchKernelGGL(k_root_count, dim3(nsb), dim3(SB), 0, s, N, d_fluid, d_root, d_counts);
  std::vector<int> counts(nsb);
  SETUP_TRY(hipMemcpyAsync(counts.data(), d_counts, nsb * sizeof(int), hipMemcpyDeviceToHost, s));
  SETUP_TRY(hipStreamSynchronize(s));
  int ncomp = 0;
  for (unsigned b = 0; b < nsb; ++b) {
    const int n = counts[b];
    counts[b] = ncomp;
    ncomp += n;
  }
  if ((rc = S.alloc(&d_anch, (size_t)ncomp))) return rc;
  SETUP_TRY(hipMemcpyAsync(d_counts, counts.data(), nsb * sizeof(int), hipMemcpyHostToDevice, s));
  SETUP_TRY(hipMemsetAsync(d_anch, 0, (size_t)ncomp, s));
  hipLaunchKernelGGL(k_root_ids, dim3(nsb), dim3(SB), 0, s, N, d_fluid, d_root, d_counts, d_id);
  hipLaunchKernelGGL(k_assign, dim3(nb), dim3(TB), 0, s, d, d_codes, d_root, d_id, d_comp, d_anch);
  SETUP_TRY(hipGetLastError());
  g.codes.resize((size_t)N);
  g.component.resize((size_t)N);
  g.anchored.assign((size_t)ncomp, 0);
  SETUP_TRY(hipMemcpyAsync(g.codes.data(), d_codes, (size_t)N, hipMemcpyDeviceToHost, s));
  SETUP_TRY(hipMemcpyAsync(g.component.data(), d_comp, (size_t)N * sizeof(int), hipMemcpyDeviceToHost, s));
  if (ncomp > 0) SETUP_TRY(hipMemcpyAsync(g.anchored.data(), d_anch, (size_t)ncomp, hipMemcpyDeviceToHost, s));
  SETUP_TRY(hipStreamSynchronize(s));
  closed_lists(g);
  return 0;
}

}  // namespace cfdgpu
