// C-ABI of the cfdgpu library: grid contexts, the device-resident PCG driver
// (pressure.rs:342-436), projection (pressure.rs:440-502) and device memory
// helpers. One HIP stream per grid; the PCG scalars live on the device and the
// host reads them once per segment of 8 iterations, when it also polls the
// caller's control callback (pressure.rs:387-389). Each segment is a captured
// hipGraph replayed until the device status leaves RUNNING.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <tuple>
#include <thread>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <unordered_map>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "cfdgpu.h"
#include "device.hpp"
#include "internal.hpp"
#include "kernels.hpp"

namespace cfdgpu {

thread_local std::string g_error;
void set_error(const std::string& text) { g_error = text; }

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);     \
      return CFD_EDEVICE;                                                               \
    }                                                                                   \
  } while (0)

constexpr int kSegment = 8;          // PCG iterations per host poll (pressure.rs:387)
constexpr int kRing = 4;             // host slots for state read-backs
constexpr int kAhead = 2;            // segments queued beyond the last read-back
// The coarse-grid correction is scaled by kOmega. A 2x2x2 piecewise-constant
// Galerkin operator is about twice as stiff as the rediscretized one in 3-D, so
// the plain correction falls short by about half; omega = 1.9 cut PCG
// iterations on the urban block from 36 to 11 (2.0 is the optimum on every
// tested grid; 1.9 keeps a margin below the loss of definiteness past 2).
// CFD_MG_OMEGA overrides.
constexpr double kOmega = 1.9;
// Aggregation hierarchy of the red-black cycle: coarsening stops once every
// component has at most kAggDense nodes, and that level is solved exactly by a
// dense inverse per component (one launch instead of the ten of each further
// level; CFD_MG_DENSE overrides). Blocks above kHostBlock nodes are inverted
// on the device at set-up.
constexpr int64_t kAggDense = 2048;
constexpr int64_t kHostBlock = 128;
constexpr int kMaxReduceBlocks = 2048;

// Device bytes allocated while a grid is being set up (its HBM footprint).
thread_local size_t* tl_bytes = nullptr;

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    n = count;
    if (count == 0) return 0;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
      return CFD_ENOMEM;
    }
    if (tl_bytes) *tl_bytes += count * sizeof(T);
    return 0;
  }
  int upload(const T* h, size_t count) {
    int rc = alloc(count);
    if (rc) return rc;
    if (count) HIP_TRY(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
    return 0;
  }
};

// Aggregation hierarchy on the device. The integer arrays (CSR structure,
// parents, children, coarsest blocks) serve both precisions; values, right-hand
// sides and iterates exist per precision.
struct DevAggShared {
  std::vector<std::unique_ptr<DevBuf<int>>> ints;
  DevBuf<int> fparent;  // per fine cell: node of the first coarse level (-1 solid)
  DevBuf<int> bptr, rows;
  DevBuf<long long> goff;
  DevBuf<double> G;
  AggCoarsest D{};
  DevBuf<int> big_block, big_first, small_block;
  // colour-separated fine level (grids without odd periodic seams)
  DevBuf<uint8_t> rb_codes;
  DevBuf<int> rb_rpar;  // parent of each red cell in the colour-separated order
  DevBuf<int> rb_bpar;  // common parent of each 2x2x2 block of cells (kMixedParent: several)
  DevBuf<double> rb_b64, rb_x64;
  DevBuf<float> rb_b32, rb_x32;
  FineRB<double> rb64{};
  FineRB<float> rb32{};
  bool has_rb = false;
  int seams = 0;       // periodic axes of odd length on the fine level
  int nlev = 0;        // coarse levels
  std::vector<AggDevLevel<double>> shape;  // integer pointers and colour ranges per coarse level
  // per coarse level: the distinct off-diagonal values its padded rows index
  // (empty without padded rows)
  std::vector<std::vector<double>> ell_values;
  bool built = false;
};
// The reference's cycle on the device (FP64).
struct DevRef {
  AggHierarchy h;
  std::vector<std::unique_ptr<DevBuf<int>>> ints;
  std::vector<std::unique_ptr<DevBuf<double>>> dbls;
  DevBuf<int> bptr, rows, small_block, big_block, big_first;
  DevBuf<long long> goff;
  DevBuf<double> G;
  RefCycle cyc;
  bool built = false;
};

template <class T>
struct DevAgg {
  std::vector<AggDevLevel<T>> levels;
  std::vector<std::unique_ptr<DevBuf<T>>> store;
};

}  // namespace cfdgpu

using namespace cfdgpu;

struct cfd_grid {
  HostGrid host;
  bool has_hier = false;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Geo geo;
  FaceGeo fgeo;
  int nb = 1;  // reduction grid
  cfd_options opt{CFD_PRECISION_FP64, 0, kSegment, CFD_CYCLE_RB};
  // topology
  DevBuf<uint8_t> codes;
  DevBuf<int> cidx;
  DevBuf<double> D;
  DevBuf<double> ic0_fill;  // factor entries with their dot terms (a periodic axis of 3 cells), else empty
  int n_closed = 0;
  double inv_count1 = 0.0;
  DevBuf<int> members, chunk_begin, comp_first_chunk;
  DevBuf<double> inv_count;
  int n_chunks = 0;
  // PCG state and vectors
  DevBuf<double> x, r, bc, z, p0, p1, ap, b, scale, xo, phi, part;
  DevBuf<double> mean_b, mean_z, mean_x, sums;
  DevBuf<unsigned int> tickets;
  DevBuf<PcgState> st;
  PcgState* h_st = nullptr;  // kRing read-back slots
  hipEvent_t evr[kRing] = {};  // recorded after the read-back into each slot
  // multigrid
  DevBuf<float> x0f;     // FP32 fine iterate of the mixed-precision cycles
  DevBuf<float> z32;     // z of the mixed-precision red-black cycle, FP32 values in the natural layout
  DevBuf<float> ic0_y32;  // FP32 forward result of the mixed-precision IC(0) application
  // aggregation hierarchy (CFD_CYCLE_RB)
  AggHierarchy agg;
  DevAggShared ash;
  DevAgg<double> a64;
  DevAgg<float> a32;
  int64_t agg_dense = kAggDense;
  bool rb_layout = true;  // CFD_MG_RB_LAYOUT=0: the fine level in the natural layout
  int rb_pairs = -1;      // CFD_MG_RB_PAIRS: two fine cells per thread (-1 by size, 0 never, 1 when hx is even)
  bool force_pairs = false;  // two fine cells per thread whatever the size (grids of a batching server)
  bool rb_lin = true;     // CFD_MG_RB_LIN=0: the two-cell half-sweep in tiles instead of linear pairs
  bool ic0_tiles = true;  // CFD_IC0_TILES=0: one launch per hyperplane
  // a hierarchy built since the last solve of its cycle (MultigridWork::hierarchy_builds)
  bool unreported_build[4] = {false, false, false, false};
  bool cycle_built[4] = {false, false, false, false};
  double build_seconds[4] = {0, 0, 0, 0};
  // the reference's cycle (CFD_CYCLE_REFERENCE) and its multi-colour variant (CFD_CYCLE_ZEBRA)
  DevRef ref, zeb;
  DevBuf<double> ref_mean;
  // captured iteration segments, by (length, parity of the first iteration, precision)
  std::map<int, hipGraphExec_t> seg;
  // iterations of the last converged solve, by (tol, precision, cycle, warm start):
  // the next such solve queues that many iterations without waiting for a read-back
  std::map<std::tuple<double, int, int, int>, int64_t> learned;
  // resident solution of the previous solve (warm start) and the HBM footprint
  DevBuf<double> warm;
  bool warm_valid = false;
  size_t footprint = 0;
  // the solve in flight (asynchronous state machine driven by job_poll)
  struct Job {
    int phase = 0;  // 0 idle, 2 iterating, 3 final kernels, 4 done
    int64_t enq = 0;  // iterations queued
    std::deque<std::pair<int, int64_t>> inflight;  // (slot, iterations queued) of pending read-backs
    int slot = 0;
    const double* b = nullptr;
    const double* scale = nullptr;
    double* x = nullptr;
    double tol = 0.0;
    int64_t max = 0;
    int rc = 0;
    bool graphs = true;  // false: iterations launched kernel by kernel (nothing instantiated)
    int64_t learned = 0;  // iterations of the last converged solve with these settings (0: none)
    std::string error;
    cfd_linear_info info{};
  } job;
  hipEvent_t evs = nullptr;  // recorded after each state read-back
  // CFD_GPU_GRAPHS=0 launches the segments kernel by kernel (for profilers that do not follow graphs).
  bool use_graphs = true;
  LaunchProbe probe;  // event pool of cfd_profile_solve
  double omega = kOmega;
  // red-black cycle: smoothing rounds (sweeps over all colours) before and after
  // the correction on the fine level and on the coarse levels (CFD_MG_SWEEPS=f,c)
  int fine_rounds = 2, coarse_rounds = 2;
  // the coarse restriction takes the residual of the class relaxed last as zero
  // (CFD_MG_ZERO_LAST=0: every residual is formed)
  bool zero_last = true;
  // stencil kernels marching through z on large grids (CFD_MARCH: 0 never, 1 by size, 2 always)
  int march = 1;
  // the reference and zebra cycles sweep the fine level over tiles and small
  // coarse levels in one workgroup (CFD_REF_TILES=0: a launch per hyperplane
  // and per dependency level)
  bool ref_tiles = true;
  // ... and sweep coarse levels of at most 65536 rows from sweep-ordered copies
  // of their rows (CFD_REF_PACKED=0: from the CSR by dependency level)
  bool ref_packed = true;
  // the red-black cycle in FP64 keeps the PCG's r and z in the fine level's
  // colour-separated layout where it can (CFD_RB_RESIDENT=0: never)
  bool rb_resident = true;
  ~cfd_grid() {
    for (auto& kv : seg)
      if (kv.second) (void)hipGraphExecDestroy(kv.second);
    if (h_st) (void)hipHostFree(h_st);
    for (hipEvent_t e : evr)
      if (e) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (evs) (void)hipEventDestroy(evs);
    for (hipEvent_t e : probe.events) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Set-up stage times on stderr (CFD_SETUP_TIMES=1): the seconds since the
// previous mark of the same thread.
struct StageClock {
  bool on = std::getenv("CFD_SETUP_TIMES") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void mark(const char* what) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[setup] %-28s %8.1f ms\n", what, 1e3 * std::chrono::duration<double>(now - t).count());
    t = now;
  }
};
thread_local StageClock t_clock;

int fill_geo(cfd_grid* g) {
  const HostGrid& h = g->host;
  Geo& G = g->geo;
  for (int a = 0; a < 3; ++a) {
    G.n[a] = (int)h.n[a];
    G.per[a] = h.periodic[a] ? 1 : 0;
    G.wi[a] = h.w_int[a];
    G.wo[a] = h.w_out[a];
  }
  G.N = h.count;
  G.sy = h.n[0];
  G.sz = h.n[0] * h.n[1];
  FaceGeo& F = g->fgeo;
  long long off = 0;
  for (int a = 0; a < 3; ++a) {
    for (int d = 0; d < 3; ++d) F.dims[a][d] = (int)h.n[d];
    if (!h.periodic[a]) F.dims[a][a] += 1;
    F.off[a] = off;
    off += (long long)F.dims[a][0] * F.dims[a][1] * F.dims[a][2];
    F.dx[a] = h.dx[a];
  }
  F.total = off;
  long long nb = (h.count + 255) / 256;
  g->nb = (int)std::max<long long>(1, std::min<long long>(nb, kMaxReduceBlocks));
  return 0;
}

template <class T>
int upload_agg_values(cfd_grid* g, DevAgg<T>& da) {
  const AggHierarchy& H = g->agg;
  da.levels.clear();
  da.store.clear();
  auto up = [&](const std::vector<double>& v, T** out) -> int {
    std::vector<T> t(v.begin(), v.end());
    da.store.emplace_back(new DevBuf<T>());
    int rc = da.store.back()->upload(t.data(), t.size());
    *out = da.store.back()->p;
    return rc;
  };
  for (int l = 0; l < g->ash.nlev; ++l) {
    const AggLevel& L = H.levels[(size_t)l + 1];
    const AggDevLevel<double>& sh = g->ash.shape[(size_t)l];
    AggDevLevel<T> d{};
    d.n = sh.n;
    d.ncol = sh.ncol;
    d.nnz = sh.nnz;
    for (int c = 0; c <= kMaxColours; ++c) d.cptr[c] = sh.cptr[c];
    d.ptr = sh.ptr;
    d.col = sh.col;
    d.parent = sh.parent;
    d.child_ptr = sh.child_ptr;
    d.child = sh.child;
    T *val = nullptr, *diag = nullptr, *b = nullptr, *x = nullptr;
    int rc;
    // a level in packed rows reads its values from the table (its CSR values are
    // not read on the device)
    if ((!sh.ecol && (rc = up(L.val, &val))) || (rc = up(L.diag, &diag))) return rc;
    std::vector<double> zero((size_t)L.n, 0.0);
    if ((rc = up(zero, &b)) || (rc = up(zero, &x))) return rc;
    d.val = val;
    d.diag = diag;
    d.b = b;
    d.x = x;
    d.K = sh.K;
    d.ecol = sh.ecol;
    d.etab = nullptr;
    if (sh.ecol) {
      T* tab = nullptr;
      if ((rc = up(g->ash.ell_values[(size_t)l], &tab))) return rc;
      d.etab = tab;
    }
    da.levels.push_back(d);
  }
  return 0;
}

// Builds the aggregation hierarchy (host) and its device arrays.
int build_agg(cfd_grid* g) {
  if (g->ash.built) return 0;
  const HostGrid& h = g->host;
  t_clock.mark("before hierarchy");
  // (the colouring below forms the children lists in its numbering)
  int rc = build_agg_hierarchy(h, g->agg, g->agg_dense, kHostBlock, true, false);
  if (rc) return rc;
  t_clock.mark("host hierarchy");
  if (g->agg.fine_shift > 1) {
    // the fine-level restrictions sum each 2x2x2 bin's residuals per aggregate
    set_error("multigrid: no fluid cells of the grid share a 2x2x2 bin with a face neighbour; the red-black "
              "cycle cannot aggregate the fine level (CFD_CYCLE_REFERENCE can)");
    return CFD_EUNSUPPORTED;
  }
  colour_agg_hierarchy(g->agg);
  t_clock.mark("colouring");
  AggHierarchy& H = g->agg;
  DevAggShared& A = g->ash;
  A.nlev = (int)H.levels.size() - 1;
  A.seams = 0;
  for (int a = 0; a < 3; ++a)
    if (h.periodic[a] && h.n[a] >= 3 && (h.n[a] & 1)) A.seams |= 1 << a;
  if (A.nlev > 0) {
    // every entry written below: no zero fill of the host array
    std::unique_ptr<int[]> fp(new int[(size_t)h.count]);
    parallel_ranges(h.count, [&](int64_t c0, int64_t c1) {
      for (int64_t c = c0; c < c1; ++c) {
        const int32_t node = H.node_of_cell[(size_t)c];
        fp[(size_t)c] = node >= 0 ? H.levels[0].parent[(size_t)node] : -1;
      }
    });
    if ((rc = A.fparent.upload(fp.get(), (size_t)h.count))) return rc;
  }
  auto upi = [&](const int32_t* v, size_t n, const int** out) -> int {
    A.ints.emplace_back(new DevBuf<int>());
    int e = A.ints.back()->upload(v, n);
    *out = A.ints.back()->p;
    return e;
  };
  auto upi64 = [&](const std::vector<int64_t>& v, const int** out) -> int {
    std::vector<int32_t> t(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
      if (v[i] > INT32_MAX) {
        set_error("multigrid: a level has more than 2^31 entries");
        return CFD_EUNSUPPORTED;
      }
      t[i] = (int32_t)v[i];
    }
    return upi(t.data(), t.size(), out);
  };
  t_clock.mark("fine parents");
  A.shape.clear();
  A.ell_values.clear();
  for (int l = 1; l <= A.nlev; ++l) {
    const AggLevel& L = H.levels[(size_t)l];
    AggDevLevel<double> sh{};
    sh.n = (int)L.n;
    sh.ncol = (int)L.colour_ptr.size() - 1;
    sh.nnz = (int)L.ptr.back();
    if (sh.ncol > kMaxColours) {
      set_error("multigrid: a coarse level needs more colour classes than supported");
      return CFD_EUNSUPPORTED;
    }
    for (int c = 0; c <= sh.ncol; ++c) sh.cptr[c] = (int)L.colour_ptr[(size_t)c];
    if ((rc = upi64(L.ptr, &sh.ptr)) || (rc = upi(L.col.data(), L.col.size(), &sh.col)) ||
        (rc = upi64(L.child_ptr, &sh.child_ptr)) || (rc = upi(L.child.data(), L.child.size(), &sh.child)))
      return rc;
    if (!L.parent.empty() && (rc = upi(L.parent.data(), L.parent.size(), &sh.parent))) return rc;
    // padded rows when every row has at most kEllWidth entries, the columns fit
    // 24 bits and the off-diagonal values take at most 256 distinct bit patterns:
    // each entry then packs its column and the index of its value
    int64_t K = 0;
    for (int64_t i = 0; i < L.n; ++i) K = std::max<int64_t>(K, L.ptr[(size_t)i + 1] - L.ptr[(size_t)i]);
    sh.K = 0;
    sh.ecol = nullptr;
    std::vector<double> table;
    if (K > 0 && K <= kEllWidth && L.n < (1LL << 24)) {
      std::unordered_map<uint64_t, int32_t> index;
      auto slot = [&](double v) -> int32_t {
        uint64_t key;
        std::memcpy(&key, &v, sizeof key);
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        if (table.size() == 256) return -1;
        index.emplace(key, (int32_t)table.size());
        table.push_back(v);
        return (int32_t)table.size() - 1;
      };
      const int32_t pad = slot(0.0);
      // the distinct values in order of first appearance (runs of one value
      // skip the lookup), then the packed entries over threads (read-only map)
      bool fits = true;
      {
        uint64_t last = 0;
        std::memcpy(&last, &L.val[0], sizeof last);
        fits = slot(L.val[0]) >= 0;
        for (size_t e = 1; e < L.val.size() && fits; ++e) {
          uint64_t key;
          std::memcpy(&key, &L.val[e], sizeof key);
          if (key == last) continue;
          last = key;
          fits = slot(L.val[e]) >= 0;
        }
      }
      // every entry written below: no zero fill of the host array
      std::unique_ptr<int32_t[]> ec(new int32_t[fits ? (size_t)K * (size_t)L.n : 1]);
      if (fits)
        parallel_ranges(L.n, [&](int64_t i0, int64_t i1) {
          for (int64_t i = i0; i < i1; ++i)
            for (int64_t k = 0; k < K; ++k) {
              const int64_t e = L.ptr[(size_t)i] + k;
              const bool entry = e < L.ptr[(size_t)i + 1];
              int32_t v = pad;
              if (entry) {
                uint64_t key;
                std::memcpy(&key, &L.val[(size_t)e], sizeof key);
                v = index.find(key)->second;
              }
              const uint32_t w = (uint32_t)(entry ? L.col[(size_t)e] : (int32_t)i) | ((uint32_t)v << 24);
              ec[(size_t)k * (size_t)L.n + (size_t)i] = (int32_t)w;
            }
        });
      if (fits) {
        const int* ep = nullptr;
        if ((rc = upi(ec.get(), (size_t)K * (size_t)L.n, &ep))) return rc;
        sh.ecol = reinterpret_cast<const unsigned*>(ep);
        sh.K = (int)K;
      } else {
        table.clear();
      }
    }
    A.shape.push_back(sh);
    A.ell_values.push_back(std::move(table));
  }
  t_clock.mark("coarse levels (ELL)");
  // coarsest blocks (on the fine level itself the rows are cells)
  std::vector<int32_t> rows = H.block_rows;
  if (A.nlev == 0)
    for (auto& r : rows) r = (int32_t)H.cell_of_node[(size_t)r];
  std::vector<int32_t> bptr(H.block_ptr.begin(), H.block_ptr.end());
  std::vector<long long> goff(H.block_goff.begin(), H.block_goff.end());
  if ((rc = A.bptr.upload(bptr.data(), bptr.size())) || (rc = A.rows.upload(rows.data(), rows.size())) ||
      (rc = A.goff.upload(goff.data(), goff.size())) || (rc = A.G.upload(H.block_G.data(), H.block_G.size())))
    return rc;
  A.D.nblocks = (int)H.block_ptr.size() - 1;
  A.D.bptr = A.bptr.p;
  A.D.rows = A.rows.p;
  A.D.goff = A.goff.p;
  A.D.G = A.G.p;
  {
    std::vector<int> bb, bf{0}, sb;
    for (int k = 0; k < A.D.nblocks; ++k) {
      const int64_t m = H.block_ptr[(size_t)k + 1] - H.block_ptr[(size_t)k];
      if (m > kSmallBlock) {
        bb.push_back(k);
        bf.push_back(bf.back() + (int)m);
      } else {
        sb.push_back(k);
      }
    }
    A.D.nbig = (int)bb.size();
    A.D.big_rows = bf.back();
    A.D.nsmall = (int)sb.size();
    if ((rc = A.big_first.upload(bf.data(), bf.size()))) return rc;
    if (!bb.empty() && (rc = A.big_block.upload(bb.data(), bb.size()))) return rc;
    if (!sb.empty() && (rc = A.small_block.upload(sb.data(), sb.size()))) return rc;
    A.D.big_block = A.big_block.p;
    A.D.big_first = A.big_first.p;
    A.D.small_block = A.small_block.p;
  }
  if (!H.deferred.empty()) {
    int64_t most = 0;
    for (const auto& d : H.deferred) most = std::max(most, d.m);
    // set-up scratch, not part of the grid's resident footprint
    size_t* counting = tl_bytes;
    tl_bytes = nullptr;
    DevBuf<double> work;
    DevBuf<int> bad;
    rc = work.alloc((size_t)(most * most + 64 * 64 + 129 * most));  // matrix, pivot block, panels, means
    if (!rc) rc = bad.alloc(1);
    tl_bytes = counting;
    if (rc) return rc;
    for (const auto& d : H.deferred) {
      if (!invert_block(A.G.p + H.block_goff[(size_t)d.block], (int)d.m, d.anchored, d.K.data(), work.p, bad.p,
                        g->stream)) {
        set_error("CFD-E035: numerics.pressure_preconditioner: masked coarse Cholesky lost positive definiteness");
        return CFD_E035;
      }
    }
    H.deferred.clear();
  }
  t_clock.mark("coarsest blocks");
  if ((rc = upload_agg_values<double>(g, g->a64))) return rc;
  if ((rc = upload_agg_values<float>(g, g->a32))) return rc;
  t_clock.mark("level values");
  if (A.seams == 0 && A.nlev > 0 && g->rb_layout) {
    const int hx = (int)((h.n[0] + 1) / 2);
    const long long half = (long long)hx * h.n[1] * h.n[2];
    std::vector<uint8_t> rc8((size_t)(2 * half), 0);
    std::vector<int> rpar((size_t)half, -1);
    const int64_t bx = (h.n[0] + 1) / 2, by = (h.n[1] + 1) / 2, bz = (h.n[2] + 1) / 2;
    std::vector<int> bpar((size_t)(bx * by * bz), 0);
    std::vector<uint8_t> bseen(bpar.size(), 0);
    const std::vector<int32_t>& parent0 = H.levels[0].parent;
    // layers in pairs (a 2x2x2 block's cells lie in one pair), in parallel
    parallel_ranges(bz, [&](int64_t z0, int64_t z1) {
      for (int64_t k = 2 * z0; k < std::min<int64_t>(h.n[2], 2 * z1); ++k)
        for (int64_t j = 0, c = k * h.n[0] * h.n[1]; j < h.n[1]; ++j)
          for (int64_t i = 0; i < h.n[0]; ++i, ++c) {
            const int colour = (int)((i + j + k) & 1);
            const size_t u = (size_t)((i >> 1) + (long long)hx * (j + h.n[1] * k));
            rc8[(size_t)colour * (size_t)half + u] = h.codes[(size_t)c];
            const int32_t node = H.node_of_cell[(size_t)c];
            const int par = node >= 0 ? parent0[(size_t)node] : -1;
            if (colour == 0 && node >= 0) rpar[u] = par;
            const size_t blk = (size_t)((i >> 1) + bx * ((j >> 1) + by * (k >> 1)));
            if (!bseen[blk]) {
              bseen[blk] = 1;
              bpar[blk] = par;
            } else if (bpar[blk] != par) {
              bpar[blk] = kMixedParent;
            }
          }
    });
    if ((rc = A.rb_codes.upload(rc8.data(), rc8.size())) || (rc = A.rb_rpar.upload(rpar.data(), rpar.size())) ||
        (rc = A.rb_bpar.upload(bpar.data(), bpar.size())) ||
        (rc = A.rb_b64.alloc((size_t)(2 * half))) ||
        (rc = A.rb_x64.alloc((size_t)(2 * half))) || (rc = A.rb_b32.alloc((size_t)(2 * half))) ||
        (rc = A.rb_x32.alloc((size_t)(2 * half))))
      return rc;
    HIP_TRY(hipMemset(A.rb_b64.p, 0, (size_t)(2 * half) * sizeof(double)));
    HIP_TRY(hipMemset(A.rb_x64.p, 0, (size_t)(2 * half) * sizeof(double)));
    HIP_TRY(hipMemset(A.rb_b32.p, 0, (size_t)(2 * half) * sizeof(float)));
    HIP_TRY(hipMemset(A.rb_x32.p, 0, (size_t)(2 * half) * sizeof(float)));
    // two cells per thread (nx a multiple of 4: every pair of a colour row is
    // aligned and complete) when a colour holds at least 3 * 2^18 cells: below that
    // one cell per thread fills the GPU in one round of workgroups and is faster
    // (1M cells: 4.9 against 6.4 us per half-sweep); above it the pairs keep more
    // bytes in flight (2M cells: 10.5 against 8.8 us, 3.5M: 14.2 against 11.4,
    // 8.4M: 23.0 against 18.2). CFD_MG_RB_PAIRS=0 / 1 forces.
    const int two = h.n[0] % 4 == 0 && (g->rb_pairs == 1 || (g->rb_pairs < 0 && half >= 3 * (1LL << 18)));
    // two cells per thread: consecutive threads on consecutive pairs (CFD_MG_RB_LIN=0: tiles of 128 x 4)
    const int pairs_mode = two ? (g->rb_lin ? 2 : 1) : 0;
    A.rb64 = FineRB<double>{half, hx, A.rb_codes.p, A.rb_b64.p, A.rb_x64.p, pairs_mode, A.rb_rpar.p, A.rb_bpar.p};
    A.rb32 = FineRB<float>{half, hx, A.rb_codes.p, A.rb_b32.p, A.rb_x32.p, pairs_mode, A.rb_rpar.p, A.rb_bpar.p};
    const unsigned long long two40 = 1ULL << 40;
    for (unsigned long long* m : {&A.rb64.mhx, &A.rb32.mhx}) *m = (two40 + (unsigned long long)hx - 1) / (unsigned long long)hx;
    for (unsigned long long* m : {&A.rb64.mny, &A.rb32.mny})
      *m = (two40 + (unsigned long long)h.n[1] - 1) / (unsigned long long)h.n[1];
    A.has_rb = true;
    t_clock.mark("fine colour layout");
  }
  // the host copies of the coarse matrices are no longer needed
  for (auto& L : H.levels) {
    std::vector<int32_t>().swap(L.col);
    std::vector<double>().swap(L.val);
    std::vector<int64_t>().swap(L.child_ptr);
    std::vector<int32_t>().swap(L.child);
  }
  A.built = true;
  return 0;
}

void enqueue_gauge(cfd_grid* g, const double* src, const double* scale, int e031, double* mean, PcgState* st);

// Builds the reference's hierarchy (host) and the device arrays of its cycle.
int build_ref(cfd_grid* g, DevRef& R, bool zebra) {
  if (R.built) return 0;
  const HostGrid& hg = g->host;
  AggHierarchy& H = R.h;
  int rc = build_ref_hierarchy(hg, H, zebra);
  if (rc) return rc;
  auto upi = [&](const std::vector<int32_t>& v) -> const int* {
    R.ints.emplace_back(new DevBuf<int>());
    if (R.ints.back()->upload(v.data(), v.size())) return nullptr;
    return R.ints.back()->p;
  };
  auto upi64 = [&](const std::vector<int64_t>& v) -> const int* {
    std::vector<int32_t> t(v.begin(), v.end());
    return upi(t);
  };
  auto upd = [&](const std::vector<double>& v) -> double* {
    R.dbls.emplace_back(new DevBuf<double>());
    if (R.dbls.back()->upload(v.data(), v.size())) return nullptr;
    return R.dbls.back()->p;
  };
  // the sweep of a coarse level in one workgroup (k_ref_gs_pk): its rows in the
  // sweep's order, each with its entries, in sub-levels of at most kRefSubRows
  // rows and `cap` entries; the iterate in LDS when it fits beside a product
  // buffer of 2048 entries or more. Not built (nsub 0) for a row longer than cap.
  int lds_max = 0;
  HIP_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, g->device));
  auto pack_sweep = [&](const AggLevel& L, const std::vector<int64_t>& ptr, const std::vector<int32_t>& order,
                        double* bs, RefSweep& P) -> int {
    P = RefSweep();
    const int64_t n = L.n, words = lds_max / (int64_t)sizeof(double);
    const bool ldsx = n + 2048 <= words;
    const int64_t cap = std::min<int64_t>((int64_t)kRefSubEpt * kRefSubRows, ldsx ? words - n : words);
    std::vector<int4> meta, rinfo((size_t)n);
    std::vector<int32_t> col;
    std::vector<double> val, dsw((size_t)n);
    col.reserve(L.col.size());
    val.reserve(L.val.size());
    int64_t most = 0;
    for (size_t d = 0; d + 1 < ptr.size(); ++d) {
      int4 m = make_int4((int)ptr[d], (int)ptr[d], (int)col.size(), (int)col.size());
      auto close = [&]() {
        if (m.y == m.x) return;
        meta.push_back(m);
        most = std::max<int64_t>(most, m.w - m.z);
      };
      for (int64_t t = ptr[d]; t < ptr[d + 1]; ++t) {
        const int32_t i = order[(size_t)t];
        const int64_t e0 = L.ptr[(size_t)i], e1 = L.ptr[(size_t)i + 1];
        if (e1 - e0 > cap) return 0;
        if (m.y - m.x == kRefSubRows || (m.w - m.z) + (e1 - e0) > cap) {
          close();
          m = make_int4((int)t, (int)t, (int)col.size(), (int)col.size());
        }
        rinfo[(size_t)t] = make_int4(i, m.w - m.z, m.w - m.z + (int)(e1 - e0), 0);
        dsw[(size_t)t] = L.diag[(size_t)i];
        for (int64_t e = e0; e < e1; ++e) {
          col.push_back(L.col[(size_t)e]);
          val.push_back(L.val[(size_t)e]);
        }
        m.y += 1;
        m.w += (int)(e1 - e0);
      }
      close();
    }
    if (col.empty()) col.push_back(0), val.push_back(0.0);  // a level without couplings
    R.ints.emplace_back(new DevBuf<int>());
    if (R.ints.back()->upload((const int*)meta.data(), meta.size() * 4)) return CFD_ENOMEM;
    P.meta = (const int4*)R.ints.back()->p;
    R.ints.emplace_back(new DevBuf<int>());
    if (R.ints.back()->upload((const int*)rinfo.data(), rinfo.size() * 4)) return CFD_ENOMEM;
    P.rinfo = (const int4*)R.ints.back()->p;
    P.dsw = upd(dsw);
    P.col = upi(col);
    P.val = upd(val);
    P.bs = bs;
    if (!P.dsw || !P.col || !P.val) return CFD_ENOMEM;
    P.nsub = (int)meta.size();
    P.cap = (int)std::max<int64_t>(most, 1);
    P.ldsx = ldsx;
    return 0;
  };
  const int64_t N = hg.count;
  RefCycle& C = R.cyc;
  C.g0 = g->geo;
  C.codes = g->codes.p;
  C.self_mask = 0;
  if (!H.masked)
    for (int a = 0; a < 3; ++a)
      if (hg.periodic[a] && hg.n[a] == 1) C.self_mask |= 1 << a;
  {
    std::vector<double> d0((size_t)N, 0.0);
    for (int64_t c = 0; c < N; ++c)
      if (H.node_of_cell[(size_t)c] >= 0) d0[(size_t)c] = H.levels[0].diag[(size_t)H.node_of_cell[(size_t)c]];
    C.diag0 = upd(d0);
    std::vector<double> zero((size_t)N, 0.0);
    C.rhs0 = upd(zero);
    C.r0 = upd(zero);
    if (!C.diag0 || !C.rhs0 || !C.r0) return CFD_ENOMEM;
  }
  // block smoothers: the level's lines in node numbers (cells on the fine level)
  auto lines_of = [&](const AggLevel& L, bool fine, RefLines& B) -> int {
    B = RefLines();
    if (L.line_length == 0) return 0;
    B.length = (int)L.line_length;
    B.bandwidth = (int)L.line_bandwidth;
    std::vector<int32_t> nodes = L.line_nodes, ocol = L.line_ocol;
    if (fine) {
      for (auto& v : nodes) v = (int32_t)H.cell_of_node[(size_t)v];
      for (auto& v : ocol) v = (int32_t)H.cell_of_node[(size_t)v];
    }
    B.nodes = upi(nodes);
    B.ocol = upi(ocol);
    B.ooff = upi64(L.line_ooff);
    B.oval = upd(L.line_oval);
    B.factor = upd(L.line_factor);
    std::vector<double> zero(L.line_nodes.size(), 0.0);
    B.scratch = upd(zero);
    B.fwd = upi(L.line_fwd);
    B.bwd = upi(L.line_bwd);
    B.fwd_ptr.assign(L.line_fwd_ptr.begin(), L.line_fwd_ptr.end());
    B.bwd_ptr.assign(L.line_bwd_ptr.begin(), L.line_bwd_ptr.end());
    return (!B.nodes || !B.ooff || !B.factor || !B.scratch || !B.fwd || !B.bwd) ? CFD_ENOMEM : 0;
  };
  if ((rc = lines_of(H.levels[0], true, C.fine_lines))) return rc;
  const int nlev = (int)H.levels.size() - 1;
  C.fparent = nullptr;
  C.ftptr = nullptr;
  C.ftcol = nullptr;
  C.ftw = nullptr;
  C.fchild = nullptr;
  if (nlev > 0) {
    const AggLevel& L0 = H.levels[0];
    if (H.masked) {
      std::vector<int32_t> fp((size_t)N, -1);
      for (int64_t c = 0; c < N; ++c)
        if (H.node_of_cell[(size_t)c] >= 0) fp[(size_t)c] = L0.parent[(size_t)H.node_of_cell[(size_t)c]];
      C.fparent = upi(fp);
    } else {
      C.ftptr = upi64(L0.tptr);
      C.ftcol = upi(L0.tcol);
      C.ftw = upd(L0.tw);
    }
    std::vector<int32_t> fc(H.levels[1].child.size());
    for (size_t k = 0; k < fc.size(); ++k) fc[k] = (int32_t)H.cell_of_node[(size_t)H.levels[1].child[k]];
    C.fchild = upi(fc);
  }
  C.lv.clear();
  for (int l = 1; l <= nlev; ++l) {
    const AggLevel& L = H.levels[(size_t)l];
    RefCoarse rc2{};
    AggDevLevel<double>& d = rc2.L;
    d.n = (int)L.n;
    d.ncol = 0;
    d.ptr = upi64(L.ptr);
    d.col = upi(L.col);
    d.val = upd(L.val);
    d.diag = upd(L.diag);
    std::vector<double> zero((size_t)L.n, 0.0);
    d.b = upd(zero);
    d.x = upd(zero);
    rc2.r = upd(zero);
    d.parent = L.parent.empty() ? nullptr : upi(L.parent);
    d.child_ptr = upi64(L.child_ptr);
    d.child = upi(L.child);
    rc2.child_w = L.child_w.empty() ? nullptr : upd(L.child_w);
    rc2.tptr = L.tptr.empty() ? nullptr : upi64(L.tptr);
    rc2.tcol = L.tcol.empty() ? nullptr : upi(L.tcol);
    rc2.tw = L.tw.empty() ? nullptr : upd(L.tw);
    rc2.nclosed = (int)L.closed_ptr.size() - 1;
    rc2.closed_ptr = rc2.nclosed > 0 ? upi64(L.closed_ptr) : nullptr;
    rc2.closed_members = rc2.nclosed > 0 ? upi(L.closed_members) : nullptr;
    {
      std::vector<int32_t> cb, cc, kc{0};
      for (int k = 0; k < rc2.nclosed; ++k) {
        for (int64_t m = L.closed_ptr[(size_t)k]; m < L.closed_ptr[(size_t)k + 1]; m += 2048) {
          cb.push_back((int32_t)m);
          cc.push_back(k);
        }
        kc.push_back((int32_t)cb.size());
      }
      rc2.nchunks = (int)cb.size();
      cb.push_back(rc2.nclosed > 0 ? (int32_t)L.closed_ptr.back() : 0);
      rc2.chunk_begin = rc2.nchunks > 0 ? upi(cb) : nullptr;
      rc2.chunk_comp = rc2.nchunks > 0 ? upi(cc) : nullptr;
      rc2.comp_chunk = rc2.nchunks > 0 ? upi(kc) : nullptr;
      rc2.chunk_sum = rc2.nchunks > 0 ? upd(std::vector<double>((size_t)rc2.nchunks, 0.0)) : nullptr;
    }
    rc2.fwd_rows = upi(L.fwd_rows);
    rc2.bwd_rows = upi(L.bwd_rows);
    rc2.fwd_ptr.assign(L.fwd_ptr.begin(), L.fwd_ptr.end());
    rc2.bwd_ptr.assign(L.bwd_ptr.begin(), L.bwd_ptr.end());
    rc2.fwd_dptr = upi64(L.fwd_ptr);
    rc2.bwd_dptr = upi64(L.bwd_ptr);
    if (g->ref_packed && L.n <= kRefWgRows && L.fwd_ptr.size() > 2) {
      double* bs = upd(zero);
      if (!bs) return CFD_ENOMEM;
      if ((rc = pack_sweep(L, L.fwd_ptr, L.fwd_rows, bs, rc2.fwd_pk)) ||
          (rc = pack_sweep(L, L.bwd_ptr, L.bwd_rows, bs, rc2.bwd_pk)))
        return rc;
    }
    if ((rc = lines_of(L, false, rc2.lines))) return rc;
    if (!d.x || !d.b || !rc2.r || !d.diag) return CFD_ENOMEM;
    C.lv.push_back(std::move(rc2));
  }
  // coarsest blocks (on the fine level itself the rows are cells)
  std::vector<int32_t> rows = H.block_rows;
  if (nlev == 0)
    for (auto& r : rows) r = (int32_t)H.cell_of_node[(size_t)r];
  std::vector<int32_t> bptr(H.block_ptr.begin(), H.block_ptr.end());
  std::vector<long long> goff(H.block_goff.begin(), H.block_goff.end());
  std::vector<int> bb, bf{0}, sb;
  for (size_t k = 0; k + 1 < H.block_ptr.size(); ++k) {
    const int64_t m = H.block_ptr[k + 1] - H.block_ptr[k];
    if (m > kSmallBlock) {
      bb.push_back((int)k);
      bf.push_back(bf.back() + (int)m);
    } else {
      sb.push_back((int)k);
    }
  }
  if ((rc = R.bptr.upload(bptr.data(), bptr.size())) || (rc = R.rows.upload(rows.data(), rows.size())) ||
      (rc = R.goff.upload(goff.data(), goff.size())) || (rc = R.G.upload(H.block_G.data(), H.block_G.size())) ||
      (rc = R.big_first.upload(bf.data(), bf.size())))
    return rc;
  if (!bb.empty() && (rc = R.big_block.upload(bb.data(), bb.size()))) return rc;
  if (!sb.empty() && (rc = R.small_block.upload(sb.data(), sb.size()))) return rc;
  AggCoarsest& D = C.D;
  D.nblocks = (int)H.block_ptr.size() - 1;
  D.bptr = R.bptr.p;
  D.rows = R.rows.p;
  D.goff = R.goff.p;
  D.G = R.G.p;
  D.nbig = (int)bb.size();
  D.big_rows = bf.back();
  D.big_block = R.big_block.p;
  D.big_first = R.big_first.p;
  D.nsmall = (int)sb.size();
  D.small_block = R.small_block.p;
  if (!g->ref_mean.p && (rc = g->ref_mean.alloc((size_t)std::max(3, g->n_closed)))) return rc;
  C.tiles = g->ref_tiles;
  C.packed = g->ref_packed;
  C.center0 = [g](const double* src, double* dst) {
    if (g->n_closed == 0) {
      if (src != dst)
        (void)hipMemcpyAsync(dst, src, (size_t)g->geo.N * sizeof(double), hipMemcpyDeviceToDevice, g->stream);
      return;
    }
    const int* cidx = g->n_closed > 1 ? g->cidx.p : nullptr;
    enqueue_gauge(g, src, nullptr, 0, g->ref_mean.p, nullptr);
    launch_center(g->geo.N, g->codes.p, src, g->ref_mean.p, cidx, dst, g->stream);
  };
  R.built = true;
  return 0;
}

// The aggregation or unmasked hierarchy of the grid's current (non-geometric) cycle.
const AggHierarchy& hierarchy_of(const cfd_grid* g) {
  return g->opt.cycle == CFD_CYCLE_REFERENCE ? g->ref.h : g->opt.cycle == CFD_CYCLE_ZEBRA ? g->zeb.h : g->agg;
}

int ensure_cycle(cfd_grid* g, int cycle) {
  if (cycle != CFD_CYCLE_RB && cycle != CFD_CYCLE_REFERENCE && cycle != CFD_CYCLE_ZEBRA) {
    set_error("cfd_options.cycle must be CFD_CYCLE_RB, CFD_CYCLE_REFERENCE or CFD_CYCLE_ZEBRA");
    return CFD_EINVAL;
  }
  if (g->cycle_built[cycle]) return 0;
  const auto t0 = std::chrono::steady_clock::now();
  int rc = cycle == CFD_CYCLE_RB ? build_agg(g)
                                 : build_ref(g, cycle == CFD_CYCLE_ZEBRA ? g->zeb : g->ref, cycle == CFD_CYCLE_ZEBRA);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  g->build_seconds[cycle] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  g->cycle_built[cycle] = true;
  g->unreported_build[cycle] = true;
  return 0;
}

int setup_device_inner(cfd_grid* g);
int setup_device(cfd_grid* g) {
  size_t bytes = 0;
  tl_bytes = &bytes;
  int rc = setup_device_inner(g);
  tl_bytes = nullptr;
  g->footprint = bytes;
  return rc;
}
int setup_device_inner(cfd_grid* g) {
  HostGrid& h = g->host;
  HIP_TRY(hipGetDevice(&g->device));
  HIP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&g->ev0));
  HIP_TRY(hipEventCreate(&g->ev1));
  HIP_TRY(hipEventCreateWithFlags(&g->evs, hipEventDisableTiming));
  HIP_TRY(hipHostMalloc((void**)&g->h_st, kRing * sizeof(PcgState), hipHostMallocDefault));
  for (hipEvent_t& e : g->evr) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  fill_geo(g);
  if (const char* env = std::getenv("CFD_GPU_GRAPHS")) g->use_graphs = std::string(env) != "0";
  if (const char* env = std::getenv("CFD_MG_DENSE")) g->agg_dense = std::max<int64_t>(1, std::atoll(env));
  if (const char* env = std::getenv("CFD_MG_RB_LAYOUT")) g->rb_layout = std::string(env) != "0";
  if (const char* env = std::getenv("CFD_MG_RB_PAIRS")) g->rb_pairs = std::atoi(env) ? 1 : 0;
  if (g->force_pairs) g->rb_pairs = 1;
  if (const char* env = std::getenv("CFD_MG_RB_LIN")) g->rb_lin = std::string(env) != "0";
  if (const char* env = std::getenv("CFD_IC0_TILES")) g->ic0_tiles = std::string(env) != "0";
  if (const char* env = std::getenv("CFD_MG_OMEGA")) g->omega = std::atof(env);
  if (const char* env = std::getenv("CFD_MG_ZERO_LAST")) g->zero_last = std::string(env) != "0";
  if (const char* env = std::getenv("CFD_MARCH")) g->march = std::atoi(env);
  if (const char* env = std::getenv("CFD_RB_RESIDENT")) g->rb_resident = std::string(env) != "0";
  if (const char* env = std::getenv("CFD_REF_TILES")) g->ref_tiles = std::string(env) != "0";
  if (const char* env = std::getenv("CFD_REF_PACKED")) g->ref_packed = std::string(env) != "0";
  if (const char* env = std::getenv("CFD_MG_SWEEPS")) {
    int f = 2, c = 2;
    if (std::sscanf(env, "%d,%d", &f, &c) == 2 && f >= 1 && f <= 8 && c >= 1 && c <= 8) {
      g->fine_rounds = f;
      g->coarse_rounds = c;
    }
  }
  size_t N = (size_t)h.count;
  int rc;
  if ((rc = g->codes.upload(h.codes.data(), N))) return rc;
  g->n_closed = h.n_closed();
  if (g->n_closed == 1) g->inv_count1 = 1.0 / (double)h.closed_offsets[1];
  if (g->n_closed > 1) {
    if ((rc = g->cidx.upload(h.closed_index.data(), N))) return rc;
    // chunks of at most max(2048, size/64) members, never spanning two components
    std::vector<int> begin, first;
    std::vector<double> inv;
    for (int cl = 0; cl < g->n_closed; ++cl) {
      int64_t b0 = h.closed_offsets[(size_t)cl], b1 = h.closed_offsets[(size_t)cl + 1];
      int64_t size = b1 - b0;
      int64_t ch = std::max<int64_t>(2048, (size + 63) / 64);
      first.push_back((int)begin.size());
      for (int64_t s = b0; s < b1; s += ch) begin.push_back((int)s);
      inv.push_back(1.0 / (double)size);
    }
    first.push_back((int)begin.size());
    g->n_chunks = (int)begin.size();
    begin.push_back((int)h.closed_offsets.back());
    if ((rc = g->members.upload(h.closed_members.data(), h.closed_members.size()))) return rc;
    if ((rc = g->chunk_begin.upload(begin.data(), begin.size()))) return rc;
    if ((rc = g->comp_first_chunk.upload(first.data(), first.size()))) return rc;
    if ((rc = g->inv_count.upload(inv.data(), inv.size()))) return rc;
  }
  size_t nm = (size_t)std::max(3, g->n_closed);
  for (DevBuf<double>* v :
       {&g->x, &g->r, &g->bc, &g->z, &g->p0, &g->p1, &g->ap, &g->b, &g->scale, &g->xo, &g->phi, &g->warm})
    if ((rc = v->alloc(N))) return rc;
  for (DevBuf<double>* v : {&g->mean_b, &g->mean_z, &g->mean_x})
    if ((rc = v->alloc(nm))) return rc;
  if ((rc = g->sums.alloc(2 * nm))) return rc;
  if ((rc = g->part.alloc((size_t)(P_CS + 2 * std::max(kReduceGrid, g->n_chunks + 1))))) return rc;
  HIP_TRY(hipMemset(g->part.p, 0, g->part.n * sizeof(double)));
  if ((rc = g->tickets.alloc(8))) return rc;
  HIP_TRY(hipMemset(g->tickets.p, 0, 8 * sizeof(unsigned int)));
  if ((rc = g->st.alloc(1))) return rc;
  HIP_TRY(hipMemset(g->st.p, 0, sizeof(PcgState)));
  if (h.preconditioner == CFD_IC0) {
    build_ic0_diagonal(h);
    if ((rc = g->D.upload(h.ic0_diag.data(), N))) return rc;
    if (!h.ic0_fill_coef.empty() && (rc = g->ic0_fill.upload(h.ic0_fill_coef.data(), h.ic0_fill_coef.size())))
      return rc;
    if ((rc = g->ic0_y32.alloc(N))) return rc;
  } else {
    g->has_hier = true;
    if ((rc = g->x0f.alloc(N)) || (rc = g->z32.alloc(N))) return rc;
    int cycle = CFD_CYCLE_RB;
    if (const char* env = std::getenv("CFD_MG_CYCLE")) cycle = std::atoi(env);
    if ((rc = ensure_cycle(g, cycle))) return rc;
    g->opt.cycle = cycle;
  }
  HIP_TRY(hipDeviceSynchronize());
  return 0;
}

// Per closed component: mean of src into `mean` (and the E031 test when e031 is set).
void enqueue_gauge(cfd_grid* g, const double* src, const double* scale, int e031, double* mean, PcgState* st) {
  launch_gauge(g->geo.N, g->codes.p, g->n_closed, g->n_chunks, g->members.p, g->chunk_begin.p, g->comp_first_chunk.p,
               g->inv_count.p, g->inv_count1, src, scale, e031, mean, st, g->part.p, g->stream);
}
// The red-black cycle's last fine half-sweep computes the PCG's r . z partials
// (colour-separated fine level, no closed component to center z in).
bool rz_fused(const cfd_grid* g) {
  return g->host.preconditioner != CFD_IC0 && g->opt.cycle == CFD_CYCLE_RB && g->ash.has_rb && g->n_closed == 0 &&
         g->fine_rounds >= 1;
}

// The solve keeps r in the fine level's right-hand side and z in its x (both in
// the colour-separated layout, FP64) instead of natural-layout vectors: the
// fused r . z half-sweep, two cells per thread (nx a multiple of 4) and the
// divisions of the colour-separated PCG update exact (N nx < 2^40).
bool resident(const cfd_grid* g) {
  return g->rb_resident && rz_fused(g) && g->opt.precision == CFD_PRECISION_FP64 && g->ash.rb64.two_cells >= 1 &&
         g->geo.n[0] % 4 == 0 && g->geo.N < (1LL << 31) && (double)g->geo.N * g->geo.n[0] < 1099511627776.0;
}
// The mixed-precision solve on such a grid keeps z as the FP32 values it is
// (written once in the natural layout, read so by the operator) instead of an
// FP64 copy.
bool resident32(const cfd_grid* g) {
  return g->rb_resident && rz_fused(g) && g->opt.precision == CFD_PRECISION_MIXED && g->ash.rb32.two_cells >= 1 &&
         g->geo.n[0] % 4 == 0 && g->z32.p;
}
// ... and with consecutive pairs of cells per thread it keeps its FP64 residual
// in the FP64 level's colour-separated b arrays, z in the FP32 level's x arrays
// and p in FP32 (no natural-layout r, z or FP64 p per iteration).
bool resident_mixed(const cfd_grid* g) {
  return resident32(g) && g->ash.rb32.two_cells == 2 && g->geo.N < (1LL << 31) &&
         (double)g->geo.N * g->geo.n[0] < 1099511627776.0;
}
// The mode enqueue_precondition takes in a solve (job set-up: first).
int resident_mode(const cfd_grid* g, bool first) {
  return resident(g) ? (first ? 1 : 2) : resident_mixed(g) ? 4 : resident32(g) ? 3 : 0;
}

// z = M r. resident (see resident()): 0 r and z natural; 1 r natural, copied
// into the fine level's right-hand side, z left in its x; 2 r already there;
// 3 (mixed) r natural, z written as FP32 to z32; 4 (mixed) r in the FP64
// level's b arrays, z left in the FP32 level's x. first_done (2, 4): the x/r
// update did the cycle's first half-sweep.
void enqueue_precondition(cfd_grid* g, const double* r, double* z, PcgState* st, int resident = 0,
                          bool first_done = false) {
  if (g->host.preconditioner == CFD_IC0) {
    // mixed precision: both sweeps in FP32 over the FP64 pivots (tiled only)
    const double* fill = g->ic0_fill.p;
    if (g->opt.precision == CFD_PRECISION_MIXED)
      launch_ic0_tiles_mixed(g->geo, g->codes.p, g->D.p, r, g->ic0_y32.p, z, st, g->stream, fill);
    else if (g->ic0_tiles)
      launch_ic0_tiles(g->geo, g->codes.p, g->D.p, r, z, st, g->stream, fill);
    else
      launch_ic0(g->geo, g->codes.p, g->D.p, r, z, st, g->stream, fill);
    return;
  }
  if (g->opt.cycle == CFD_CYCLE_REFERENCE || g->opt.cycle == CFD_CYCLE_ZEBRA) {
    launch_ref_vcycle((g->opt.cycle == CFD_CYCLE_ZEBRA ? g->zeb : g->ref).cyc, r, z, st, g->stream);
    return;
  }
  const DevAggShared& A = g->ash;
  // without closed components z needs no centering: the last fine half-sweep
  // forms the r . z partials as well (no separate k_pcg_rz)
  const double* rz_r = rz_fused(g) ? r : nullptr;
  if (g->opt.precision == CFD_PRECISION_MIXED)
    launch_agg_vcycle<float, double>(g->geo, g->codes.p, r, g->x0f.p, A.fparent.p, A.seams, A.nlev,
                                     g->a32.levels.data(), A.D, resident >= 3 ? nullptr : z, st, g->omega, g->stream,
                                     A.has_rb ? &A.rb32 : nullptr, g->fine_rounds, g->coarse_rounds,
                                     resident == 4 ? A.rb64.b : rz_r, g->part.p, g->zero_last,
                                     resident >= 3 ? resident : 0, resident == 3 ? g->z32.p : nullptr, first_done);
  else
    launch_agg_vcycle<double, double>(g->geo, g->codes.p, r, z, A.fparent.p, A.seams, A.nlev, g->a64.levels.data(),
                                      A.D, A.has_rb && !resident ? z : nullptr, st, g->omega, g->stream,
                                      A.has_rb ? &A.rb64 : nullptr, g->fine_rounds, g->coarse_rounds,
                                      resident ? nullptr : rz_r, g->part.p, g->zero_last, resident, nullptr, first_done);
}

// One PCG iteration (pressure.rs:386-434): the preconditioner applied to the
// residual the previous iteration left (`precondition`; the job's set-up applies
// it before the first iteration), then the search direction, its operator
// product, the update of x and r, and the convergence test. The preconditioner
// leads the iteration so that the iteration that converges queues no cycle
// after its test. p alternates between p0 and p1.
void enqueue_iteration(cfd_grid* g, int parity, bool precondition) {
  PcgState* st = g->st.p;
  double* pold = parity ? g->p1.p : g->p0.p;
  double* pnew = parity ? g->p0.p : g->p1.p;
  const double* mz = g->mean_z.p;
  const int* cidx = g->n_closed > 1 ? g->cidx.p : nullptr;
  const bool res = resident(g), res32 = resident32(g), resm = resident_mixed(g);
  if (precondition) {
    // the previous iteration's x/r update did the cycle's first half-sweep
    enqueue_precondition(g, g->r.p, g->z.p, st, resident_mode(g, false), res || resm);
    enqueue_gauge(g, g->z.p, nullptr, 0, g->mean_z.p, st);
    if (!rz_fused(g)) launch_pcg_rz(g->geo.N, g->codes.p, g->r.p, g->z.p, mz, cidx, st, g->part.p, g->stream);
  }
  if (resm) {
    // p in FP32 in the space of the FP64 vectors p0 / p1
    const FineRB<float>& F = g->ash.rb32;
    launch_pcg_op_mixed(g->geo, g->codes.p, F.x, F.half, (const float*)pold, (float*)pnew, g->ap.p, st, g->part.p,
                        g->stream, g->march);
    launch_pcg_axpy_first<float, float>(g->geo, F, (const float*)pnew, g->ap.p, g->x.p, g->ash.rb64.b, st, g->part.p,
                                        g->stream);
  } else if (res) {
    const FineRB<double>& F = g->ash.rb64;
    launch_pcg_op(g->geo, g->codes.p, F.x, pold, pnew, g->ap.p, mz, cidx, false, st, g->part.p, g->stream, g->march,
                  F.half);
    launch_pcg_axpy_first<double, double>(g->geo, F, pnew, g->ap.p, g->x.p, F.b, st, g->part.p, g->stream);
  } else if (res32) {
    launch_pcg_op(g->geo, g->codes.p, nullptr, pold, pnew, g->ap.p, mz, cidx, false, st, g->part.p, g->stream,
                  g->march, 0, g->z32.p);
    launch_pcg_axpy(g->geo.N, pnew, g->ap.p, g->x.p, g->r.p, st, g->part.p, g->stream);
  } else {
    launch_pcg_op(g->geo, g->codes.p, g->z.p, pold, pnew, g->ap.p, mz, cidx, g->n_closed > 0, st, g->part.p,
                  g->stream, g->march);
    launch_pcg_axpy(g->geo.N, pnew, g->ap.p, g->x.p, g->r.p, st, g->part.p, g->stream);
  }
  launch_pcg_true(g->geo, g->codes.p, g->x.p, g->bc.p, st, g->part.p, g->stream);
}

// The captured graph of `len` iterations whose first one has parity `parity`
// (`first`: it is the job's first iteration, whose preconditioner the set-up
// applied).
int segment_graph(cfd_grid* g, int len, int parity, bool first, hipGraphExec_t* out) {
  const int key = (len << 3) | (first ? 4 : 0) | (parity << 1) | (g->opt.precision == CFD_PRECISION_MIXED ? 1 : 0);
  auto it = g->seg.find(key);
  if (it != g->seg.end()) {
    *out = it->second;
    return 0;
  }
  hipGraph_t graph;
  HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < len; ++i) enqueue_iteration(g, (parity + i) & 1, !first || i > 0);
  HIP_TRY(hipStreamEndCapture(g->stream, &graph));
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return CFD_EDEVICE;
  }
  g->seg[key] = exec;
  *out = exec;
  return 0;
}

std::tuple<double, int, int, int> learn_key(const cfd_grid* g, double tol) {
  return {tol, g->opt.precision, g->opt.cycle, g->opt.warm_start ? 1 : 0};
}

int finish_status(cfd_grid* g, const PcgState& s, int64_t max_iterations) {
  switch (s.status) {
    case ST_CONVERGED:
    case ST_ZERO:
      return CFD_OK;
    case ST_E031:
      set_error("CFD-E031: incompatible flux in a closed fluid component");
      return CFD_E031;
    case ST_E032:
      set_error("CFD-E032: pressure CG lost positive definiteness");
      return CFD_E032;
    case ST_E033: {
      char buf[96];
      std::snprintf(buf, sizeof buf, "CFD-E033: pressure iteration limit %lld", (long long)max_iterations);
      set_error(buf);
      return CFD_E033;
    }
    default:
      set_error("pressure solve ended in an unknown device state");
      return CFD_EDEVICE;
  }
  (void)g;
}

int read_state(cfd_grid* g) {
  HIP_TRY(hipMemcpyAsync(g->h_st, g->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return 0;
}

// Asynchronous solve: job_begin enqueues the set-up kernels; job_poll advances
// the job whenever its last read-back of the device state has landed, so one
// host thread can drive the solves of many grids (each on its own stream).
enum { JOB_PENDING = 0, JOB_DONE = 1 };

void fill_info(cfd_grid* g, const PcgState& s, cfd_linear_info& local);

// Queues `len` iterations and a read-back of the state after them into the next
// host slot.
int enqueue_chunk(cfd_grid* g, int len) {
  cfd_grid::Job& J = g->job;
  const int parity = (int)(J.enq & 1);
  if (g->use_graphs && J.graphs && g->opt.cycle != CFD_CYCLE_REFERENCE) {
    hipGraphExec_t seg = nullptr;
    int rc = segment_graph(g, len, parity, J.enq == 0, &seg);
    if (rc) return rc;
    HIP_TRY(hipGraphLaunch(seg, g->stream));
  } else {
    for (int i = 0; i < len; ++i) {
      if (t_probe) t_probe->tag = J.enq + i + 1;
      enqueue_iteration(g, (parity + i) & 1, J.enq + i > 0);
    }
  }
  J.enq += len;
  const int slot = J.slot;
  J.slot = (J.slot + 1) % kRing;
  HIP_TRY(hipMemcpyAsync(g->h_st + slot, g->st.p, sizeof(PcgState), hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipEventRecord(g->evr[slot], g->stream));
  J.inflight.emplace_back(slot, J.enq);
  return 0;
}

// Queues segments of at most kSegment iterations until `target` iterations are
// queued (at most the iteration limit, and at most kAhead segments past the last
// read-back); at least one segment when none is pending.
int queue_ahead(cfd_grid* g, int64_t target) {
  cfd_grid::Job& J = g->job;
  target = std::min(target, J.max);
  if (target <= J.enq && J.inflight.empty()) target = std::min(J.enq + 1, J.max);
  while (J.enq < target && (int)J.inflight.size() < kAhead) {
    int rc = enqueue_chunk(g, (int)std::min<int64_t>(kSegment, target - J.enq));
    if (rc) return rc;
  }
  return 0;
}

// The iterations a running solve is expected to need, from a read-back state: the
// mean reduction of the recurrent residual per iteration so far, extrapolated to
// the tolerance.
int64_t expected_iterations(const PcgState& s, double tol, int64_t queued) {
  if (s.it <= 0 || !(s.rel > 0.0) || !(s.rel < 1.0)) return queued + kSegment;
  if (s.rel <= tol) return s.it + 2;  // the recomputed residual missed its gate
  const double need = std::log(tol / s.rel) / (std::log(s.rel) / (double)s.it);
  if (!(need < 1e6)) return queued + kSegment;
  return s.it + (int64_t)std::ceil(need);
}

int job_begin(cfd_grid* g, const double* d_b, const double* d_scale, double tol, int64_t max_iterations, double* d_x,
              cfd_poll_fn poll = nullptr, void* user = nullptr, bool graphs = true) {
  if (!(tol > 0.0) || max_iterations <= 0) {
    set_error("pressure tolerance must be positive and the iteration limit nonzero (CFD-E008/E009)");
    return CFD_EINVAL;
  }
  HIP_TRY(hipSetDevice(g->device));
  cfd_grid::Job& J = g->job;
  J = cfd_grid::Job();
  J.graphs = graphs;
  J.b = d_b;
  J.scale = d_scale;
  J.x = d_x;
  J.tol = tol;
  J.max = max_iterations;
  PcgState* st = g->st.p;
  HIP_TRY(hipEventRecord(g->ev0, g->stream));
  launch_state_init(st, tol, max_iterations, g->stream);
  const int* cidx = g->n_closed > 1 ? g->cidx.p : nullptr;
  enqueue_gauge(g, d_b, d_scale, 1, g->mean_b.p, st);
  launch_pcg_init(g->geo.N, g->codes.p, d_b, g->mean_b.p, cidx, g->r.p, g->bc.p, g->x.p, st, g->part.p, g->stream);
  if (g->opt.warm_start && g->warm_valid)
    launch_warm_residual(g->geo, g->codes.p, g->warm.p, g->bc.p, g->x.p, g->r.p, st, g->stream);
  if (t_probe) t_probe->tag = 0;
  if (resident_mixed(g)) launch_rb_scatter<double>(g->geo, g->codes.p, g->r.p, g->ash.rb64, st, g->stream);
  enqueue_precondition(g, g->r.p, g->z.p, st, resident_mode(g, true));
  enqueue_gauge(g, g->z.p, nullptr, 0, g->mean_z.p, st);
  if (!rz_fused(g))
    launch_pcg_rz(g->geo.N, g->codes.p, g->r.p, g->z.p, g->mean_z.p, cidx, st, g->part.p, g->stream);
  // Control::check before the first iteration (pressure.rs:387-389)
  // (a zero or incompatible right-hand side ends the solve before that check)
  if (int p = poll ? poll(user) : 0; p == 1 || p == 2) {
    int rc = read_state(g);
    if (rc) return rc;
    if (g->h_st[0].status == ST_RUNNING) {
      fill_info(g, g->h_st[0], J.info);
      J.rc = p == 1 ? CFD_CANCELLED : CFD_WALL_TIME;
      J.error = p == 1 ? "cancelled" : "wall_time_limit";
      J.phase = 4;
      return 0;
    }
  }
  // the first iterations go into the queue at once: as many as the last such solve
  // took, else one segment; kernels queued past the end of a solve do nothing
  auto known = g->learned.find(learn_key(g, tol));
  J.learned = known != g->learned.end() ? known->second : 0;
  int rc = queue_ahead(g, J.learned > 0 ? J.learned : kSegment);
  if (rc) return rc;
  J.phase = 2;
  return 0;
}

// LinearInfo and the attempted work (pressure.rs:350-410, multigrid.rs:646-726) of
// the job whose last read-back state is `s`.
void fill_info(cfd_grid* g, const PcgState& s, cfd_linear_info& local) {
  std::memset(&local, 0, sizeof local);
  local.iterations = s.status == ST_ZERO ? 0 : s.it;
  local.relative_residual = s.status == ST_ZERO ? 0.0 : s.true_rel;
  local.recurrent_residual = s.rel;
  local.true_residual_checks = s.true_checks;
  local.operator_applications = local.iterations + s.true_checks;
  local.preconditioner_applications = s.status == ST_ZERO ? 0 : (s.status == ST_CONVERGED ? s.it : s.it + 1);
  if (g->has_hier) {
    // MultigridWork (multigrid.rs:667-726): per cycle and nonterminal level, one
    // residual, restriction and prolongation per cell and 2 x 2 point sweeps, or
    // one block sweep each way on a block-smoothed level
    int64_t cells = 0, block_cells = 0, block_lines = 0;
    const AggHierarchy& H = hierarchy_of(g);
    for (size_t l = 0; l + 1 < H.levels.size(); ++l) {
      cells += H.levels[l].n;
      if (H.levels[l].line_length > 0) {
        block_cells += H.levels[l].n;
        block_lines += (int64_t)H.levels[l].line_nodes.size() / H.levels[l].line_length;
      }
    }
    cfd_mg_stats ms;
    if (cfd_multigrid_stats(g, &ms) == CFD_OK) {
      local.hierarchy_levels = ms.levels;
      for (int l = 0; l < ms.levels; ++l) {
        local.hierarchy_cells += ms.level_cells[l];
        local.hierarchy_nonzeros += ms.level_nonzeros[l];
      }
      local.hierarchy_bytes = ms.allocated_bytes;
    }
    local.cycles = local.preconditioner_applications;
    local.smoothing_cell_updates = local.cycles * (4 * (cells - block_cells) + 2 * block_cells);
    if (g->opt.cycle == CFD_CYCLE_RB && cells > 0) {
      // the red-black cycle's rounds per side: fine level and coarse levels
      const int64_t n0 = g->agg.levels[0].n;
      local.smoothing_cell_updates = local.cycles * 2 * (g->fine_rounds * n0 + g->coarse_rounds * (cells - n0));
    }
    local.residual_cell_updates = local.cycles * cells;
    local.restriction_cell_updates = local.residual_cell_updates;
    local.prolongation_cell_updates = local.residual_cell_updates;
    local.coarse_solves = local.cycles;
    local.block_smoothing_cell_updates = local.cycles * 2 * block_cells;
    local.block_solves = local.cycles * 2 * block_lines;
    local.block_level_cell_visits = local.cycles * block_cells;
    if (g->unreported_build[g->opt.cycle]) {
      local.hierarchy_builds = 1;
      local.hierarchy_build_seconds = g->build_seconds[g->opt.cycle];
      g->unreported_build[g->opt.cycle] = false;
    }
  }
}

// Advances the job; `block` waits for the pending read-back. Returns JOB_PENDING,
// JOB_DONE (result in g->job.rc / info) or an error code of the device runtime.
int job_poll(cfd_grid* g, bool block, cfd_poll_fn poll, void* user) {
  cfd_grid::Job& J = g->job;
  if (J.phase == 4 || J.phase == 0) return JOB_DONE;
  if (J.phase == 2 && J.inflight.empty()) {
    set_error("pressure solve: no pending read-back of the solver state");
    return CFD_EDEVICE;
  }
  hipEvent_t ev = J.phase == 3 ? g->evs : g->evr[J.inflight.front().first];
  if (block) {
    HIP_TRY(hipEventSynchronize(ev));
  } else {
    hipError_t q = hipEventQuery(ev);
    if (q == hipErrorNotReady) return JOB_PENDING;
    if (q != hipSuccess) {
      set_error(std::string("HIP error: ") + hipGetErrorString(q));
      return CFD_EDEVICE;
    }
  }
  if (J.phase == 3) {
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, g->ev0, g->ev1));
    J.info.seconds = ms * 1e-3;
    J.phase = 4;
    return JOB_DONE;
  }
  const PcgState s = g->h_st[J.inflight.front().first];
  J.inflight.pop_front();
  if (s.status == ST_RUNNING) {
    if (poll) {
      int p = poll(user);
      if (p == 1 || p == 2) {
        fill_info(g, s, J.info);  // the attempted work counts (pressure.rs:387-389)
        J.rc = p == 1 ? CFD_CANCELLED : CFD_WALL_TIME;
        J.error = p == 1 ? "cancelled" : "wall_time_limit";
        J.phase = 4;
        return JOB_DONE;
      }
    }
    // no further than the count the last such solve converged in until a
    // read-back shows that count reached
    int64_t target = expected_iterations(s, J.tol, J.enq);
    if (J.learned > s.it) target = std::min(target, J.learned);
    int rc = queue_ahead(g, target);
    if (rc) return rc;
    return JOB_PENDING;
  }
  if (s.status == ST_CONVERGED) g->learned[learn_key(g, J.tol)] = s.it;
  fill_info(g, s, J.info);
  int rc = finish_status(g, s, J.max);
  if (rc) {
    J.rc = rc;
    J.error = g_error;
    J.phase = 4;
    return JOB_DONE;
  }
  const int* cidx = g->n_closed > 1 ? g->cidx.p : nullptr;
  if (s.status == ST_ZERO) {
    launch_fill(g->geo.N, J.x, 0.0, g->stream);
  } else {
    enqueue_gauge(g, g->x.p, nullptr, 0, g->mean_x.p, nullptr);
    launch_center(g->geo.N, g->codes.p, g->x.p, g->mean_x.p, cidx, J.x, g->stream);
  }
  if (g->opt.warm_start) {
    HIP_TRY(hipMemcpyAsync(g->warm.p, J.x, (size_t)g->geo.N * sizeof(double), hipMemcpyDeviceToDevice, g->stream));
    g->warm_valid = true;
  }
  HIP_TRY(hipEventRecord(g->ev1, g->stream));
  HIP_TRY(hipEventRecord(g->evs, g->stream));
  J.phase = 3;
  return JOB_PENDING;
}

int solve_device(cfd_grid* g, const double* d_b, const double* d_scale, double tol, int64_t max_iterations,
                 cfd_poll_fn poll, void* user, double* d_x, cfd_linear_info* info, bool graphs = true) {
  int rc = job_begin(g, d_b, d_scale, tol, max_iterations, d_x, poll, user, graphs);
  if (rc) return rc;
  for (;;) {
    rc = job_poll(g, true, poll, user);
    if (rc == JOB_DONE) break;
    if (rc != JOB_PENDING) return rc;
  }
  if (info) *info = g->job.info;
  if (g->job.rc) {
    // a cancelled job may still have kernels queued: let them finish before returning
    HIP_TRY(hipStreamSynchronize(g->stream));
    set_error(g->job.error);
    return g->job.rc;
  }
  return CFD_OK;
}

// A grid context from a case; force_pairs: the fine red-black level takes two
// cells per thread whatever its size (the layout of the batched iterations).
int create_grid(const cfd_case* c, cfd_grid** out, bool force_pairs) {
  if (!c || !out) {
    set_error("cfd_grid_create: null argument");
    return CFD_EINVAL;
  }
  *out = nullptr;
  if (c->preconditioner != CFD_IC0 && c->preconditioner != CFD_MULTIGRID) {
    set_error("numerics.pressure_preconditioner must be ic0 or multigrid");
    return CFD_EINVAL;
  }
  std::unique_ptr<cfd_grid> g(new cfd_grid());
  g->force_pairs = force_pairs;
  t_clock.t = std::chrono::steady_clock::now();
  // rasterization, face codes and components on the device (CFD_GPU_SETUP=0: on the host)
  const char* env = std::getenv("CFD_GPU_SETUP");
  int rc = env && std::string(env) == "0" ? build_host_grid(*c, g->host) : build_device_grid(*c, g->host);
  if (rc) return rc;
  t_clock.mark("topology");
  rc = setup_device(g.get());
  if (rc) return rc;
  t_clock.mark("end");
  *out = g.release();
  return CFD_OK;
}

}  // namespace

// ---------------------------------------------------------------- C ABI
extern "C" {

const char* cfd_last_error(void) { return g_error.c_str(); }

int cfd_grid_create(const cfd_case* c, cfd_grid** out) { return create_grid(c, out, false); }

int cfd_grid_create_desc(const cfd_grid_desc* d, cfd_grid** out) {
  if (!d || !out) {
    set_error("cfd_grid_create_desc: null argument");
    return CFD_EINVAL;
  }
  *out = nullptr;
  std::unique_ptr<cfd_grid> g(new cfd_grid());
  int rc = grid_from_desc(*d, g->host, true);
  if (rc) return rc;
  rc = setup_device(g.get());
  if (rc) return rc;
  *out = g.release();
  return CFD_OK;
}

void cfd_grid_destroy(cfd_grid* g) { delete g; }

// Signature::matches over two host grids (multigrid.rs:158-222).
static bool same_operator(const HostGrid& h, const HostGrid& o) {
  auto bits = [](double v) {
    uint64_t u;
    std::memcpy(&u, &v, sizeof u);
    return u;
  };
  bool eq = h.preconditioner == o.preconditioner;
  for (int a = 0; a < 3 && eq; ++a)
    eq = h.n[a] == o.n[a] && h.periodic[a] == o.periodic[a] && bits(h.dx[a]) == bits(o.dx[a]) &&
         bits(h.w_int[a]) == bits(o.w_int[a]) && bits(h.w_out[a]) == bits(o.w_out[a]);
  // the codes carry every interior and outlet face (a coupled side) and, with the
  // fluid bit and the closed bit, the fluid mask and the anchoring of each cell's
  // component; components are numbered by their lowest cell in both builds
  eq = eq && h.fluid == o.fluid && h.anchored == o.anchored && h.codes == o.codes;
  return eq;
}

int cfd_grid_matches(const cfd_grid* g, const cfd_case* c, int* same) {
  if (!g || !c || !same) {
    set_error("cfd_grid_matches: null argument");
    return CFD_EINVAL;
  }
  *same = 0;
  HostGrid h;
  if (int rc = build_host_grid(*c, h)) return rc;
  *same = same_operator(h, g->host) ? 1 : 0;
  return CFD_OK;
}

int cfd_grid_matches_desc(const cfd_grid* g, const cfd_grid_desc* d, int* same) {
  if (!g || !d || !same) {
    set_error("cfd_grid_matches_desc: null argument");
    return CFD_EINVAL;
  }
  *same = 0;
  HostGrid h;
  if (int rc = grid_from_desc(*d, h, false)) return rc;
  uint64_t a = 0, b = 0;
  std::memcpy(&a, &h.volume, sizeof a);
  std::memcpy(&b, &g->host.volume, sizeof b);
  *same = a == b && same_operator(h, g->host) ? 1 : 0;
  return CFD_OK;
}

static void host_counts(const HostGrid& h, int64_t counts[8]) {
  int64_t anchored = 0;
  for (uint8_t a : h.anchored) anchored += a;
  counts[0] = h.count;
  counts[1] = h.face_count[0] + h.face_count[1] + h.face_count[2];
  counts[2] = h.fluid_count;
  counts[3] = (int64_t)h.anchored.size();
  counts[4] = anchored;
  counts[5] = h.face_count[0];
  counts[6] = h.face_count[1];
  counts[7] = h.face_count[2];
}

// cfd_case_topology on the host or the device.
static int case_topology(const cfd_case* c, bool device, int64_t counts[8], uint8_t* codes, int64_t* component,
                  uint8_t* anchored, int64_t anchored_cap, double* ic0_diagonal) {
  if (!c) {
    set_error("cfd_case_topology: null case");
    return CFD_EINVAL;
  }
  HostGrid h;
  int rc = device ? build_device_grid(*c, h) : build_host_grid(*c, h);
  if (rc) return rc;
  if (counts) host_counts(h, counts);
  if (codes) std::memcpy(codes, h.codes.data(), h.codes.size());
  if (component)
    for (int64_t i = 0; i < h.count; ++i) component[i] = h.component[(size_t)i];
  if (anchored)
    for (size_t k = 0; k < h.anchored.size() && (int64_t)k < anchored_cap; ++k) anchored[k] = h.anchored[k];
  if (ic0_diagonal) {
    build_ic0_diagonal(h);
    std::memcpy(ic0_diagonal, h.ic0_diag.data(), h.ic0_diag.size() * sizeof(double));
  }
  return CFD_OK;
}

int cfd_case_topology(const cfd_case* c, int64_t counts[8], uint8_t* codes, int64_t* component, uint8_t* anchored,
                      int64_t anchored_cap, double* ic0_diagonal) {
  return case_topology(c, false, counts, codes, component, anchored, anchored_cap, ic0_diagonal);
}

int cfd_case_topology_device(const cfd_case* c, int64_t counts[8], uint8_t* codes, int64_t* component,
                             uint8_t* anchored, int64_t anchored_cap) {
  return case_topology(c, true, counts, codes, component, anchored, anchored_cap, nullptr);
}

int cfd_case_multigrid_stats(const cfd_case* c, int32_t cycle, cfd_mg_stats* s) {
  if (!c || !s) {
    set_error("cfd_case_multigrid_stats: null argument");
    return CFD_EINVAL;
  }
  std::memset(s, 0, sizeof *s);
  if (cycle != CFD_CYCLE_RB && cycle != CFD_CYCLE_REFERENCE && cycle != CFD_CYCLE_ZEBRA) {
    set_error("cfd_case_multigrid_stats: cycle must be CFD_CYCLE_RB, CFD_CYCLE_REFERENCE or CFD_CYCLE_ZEBRA");
    return CFD_EINVAL;
  }
  HostGrid h;
  int rc = build_host_grid(*c, h);
  if (rc) return rc;
  AggHierarchy H;
  if (cycle == CFD_CYCLE_RB)
    rc = build_agg_hierarchy(h, H, kAggDense, kHostBlock, true);  // sizes only: no large block is inverted
  else
    rc = build_ref_hierarchy(h, H, cycle == CFD_CYCLE_ZEBRA);
  if (rc) return rc;
  int L = 0;
  for (const AggLevel& lv : H.levels) {
    if (L >= 32) break;
    s->level_cells[L] = lv.n;
    s->level_nonzeros[L] = L == 0 ? lv.n + H.level0_entries : (int64_t)lv.col.size() + lv.n;
    ++L;
  }
  s->levels = L;
  agg_hierarchy_shape(H, s);
  s->digest = agg_hierarchy_digest(H);
  return CFD_OK;
}

int cfd_grid_counts(const cfd_grid* g, int64_t counts[8]) {
  host_counts(g->host, counts);
  return CFD_OK;
}

int cfd_grid_codes(const cfd_grid* g, uint8_t* codes) {
  std::memcpy(codes, g->host.codes.data(), g->host.codes.size());
  return CFD_OK;
}

int cfd_grid_components(const cfd_grid* g, int64_t* component, uint8_t* anchored) {
  const HostGrid& h = g->host;
  for (int64_t i = 0; i < h.count; ++i) component[i] = h.component[(size_t)i];
  for (size_t c = 0; c < h.anchored.size(); ++c) anchored[c] = h.anchored[c];
  return CFD_OK;
}

int cfd_grid_set_options(cfd_grid* g, const cfd_options* o) {
  if (o->precision != CFD_PRECISION_FP64 && o->precision != CFD_PRECISION_MIXED) {
    set_error("cfd_options.precision must be FP64 or MIXED");
    return CFD_EINVAL;
  }
  if (g->has_hier && (o->cycle == CFD_CYCLE_REFERENCE || o->cycle == CFD_CYCLE_ZEBRA) &&
      o->precision != CFD_PRECISION_FP64) {
    set_error("cfd_options: the reference and zebra cycles run in FP64");
    return CFD_EINVAL;
  }
  if (g->has_hier && o->cycle != g->opt.cycle) {
    HIP_TRY(hipSetDevice(g->device));
    size_t bytes = 0;
    tl_bytes = &bytes;
    int rc = ensure_cycle(g, o->cycle);
    tl_bytes = nullptr;
    g->footprint += bytes;
    if (rc) return rc;
    // the captured iteration segments launch the previous cycle's kernels
    for (auto& kv : g->seg)
      if (kv.second) (void)hipGraphExecDestroy(kv.second);
    g->seg.clear();
  }
  const int cycle = g->has_hier ? o->cycle : g->opt.cycle;
  g->opt = *o;
  g->opt.cycle = cycle;
  return CFD_OK;
}

int cfd_ic0_diagonal(const cfd_grid* g, double* diag) {
  if (g->host.ic0_diag.empty()) {
    set_error("the grid's preconditioner is not ic0");
    return CFD_EINVAL;
  }
  std::memcpy(diag, g->host.ic0_diag.data(), g->host.ic0_diag.size() * sizeof(double));
  return CFD_OK;
}

int cfd_multigrid_stats(const cfd_grid* g, cfd_mg_stats* s) {
  std::memset(s, 0, sizeof *s);
  if (!g->has_hier) {
    set_error("the grid's preconditioner is not multigrid");
    return CFD_EINVAL;
  }
  const HostGrid& h = g->host;
  {
    // MultigridStats of the aggregation hierarchy (masked.rs:286-290): unknowns and
    // stored entries (off-diagonal plus diagonal) per level
    const AggHierarchy& H = hierarchy_of(g);
    int L = 0;
    int64_t bytes = 0;
    for (const AggLevel& lv : H.levels) {
      if (L >= 32) break;
      s->level_cells[L] = lv.n;
      s->level_nonzeros[L] = L == 0 ? h.fluid_count + (int64_t)H.level0_entries : (int64_t)lv.ptr.back() + lv.n;
      if (L > 0) bytes += lv.n * (int64_t)(3 * sizeof(int) + 4 * sizeof(double) + 4 * sizeof(float)) +
                         (int64_t)lv.ptr.back() * (int64_t)(sizeof(int) + sizeof(double) + sizeof(float));
      ++L;
    }
    s->levels = L;
    agg_hierarchy_shape(H, s);
    s->allocated_bytes = bytes + h.count * (int64_t)(sizeof(int) + sizeof(float)) +
                         (int64_t)H.block_G.size() * (int64_t)sizeof(double);
    for (int64_t k = 0; k < s->n_block_smoothers; ++k) s->allocated_bytes += s->block_smoothers[k].allocated_bytes;
    return CFD_OK;
  }
}

int cfd_apply_device(cfd_grid* g, const double* d_x, double* d_y) {
  HIP_TRY(hipSetDevice(g->device));
  launch_apply(g->geo, g->codes.p, d_x, d_y, g->stream);
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

int cfd_apply(cfd_grid* g, const double* x, double* y) {
  HIP_TRY(hipSetDevice(g->device));
  size_t bytes = (size_t)g->geo.N * sizeof(double);
  HIP_TRY(hipMemcpyAsync(g->b.p, x, bytes, hipMemcpyHostToDevice, g->stream));
  launch_apply(g->geo, g->codes.p, g->b.p, g->xo.p, g->stream);
  HIP_TRY(hipMemcpyAsync(y, g->xo.p, bytes, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

int cfd_center_nullspaces(cfd_grid* g, double* x) {
  HIP_TRY(hipSetDevice(g->device));
  size_t bytes = (size_t)g->geo.N * sizeof(double);
  HIP_TRY(hipMemcpyAsync(g->b.p, x, bytes, hipMemcpyHostToDevice, g->stream));
  const int* cidx = g->n_closed > 1 ? g->cidx.p : nullptr;
  enqueue_gauge(g, g->b.p, nullptr, 0, g->mean_x.p, nullptr);
  launch_center(g->geo.N, g->codes.p, g->b.p, g->mean_x.p, cidx, g->xo.p, g->stream);
  HIP_TRY(hipMemcpyAsync(x, g->xo.p, bytes, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

int cfd_precondition_device(cfd_grid* g, const double* d_r, double* d_z) {
  HIP_TRY(hipSetDevice(g->device));
  launch_state_init(g->st.p, 1.0, 1, g->stream);
  enqueue_precondition(g, d_r, d_z, g->st.p);
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

int cfd_precondition(cfd_grid* g, const double* r, double* z) {
  HIP_TRY(hipSetDevice(g->device));
  size_t bytes = (size_t)g->geo.N * sizeof(double);
  HIP_TRY(hipMemcpyAsync(g->b.p, r, bytes, hipMemcpyHostToDevice, g->stream));
  launch_state_init(g->st.p, 1.0, 1, g->stream);
  enqueue_precondition(g, g->b.p, g->xo.p, g->st.p);
  HIP_TRY(hipMemcpyAsync(z, g->xo.p, bytes, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

int cfd_solve_scaled_device(cfd_grid* g, const double* d_b, const double* d_scale, double tol, int64_t max_iterations,
                            cfd_poll_fn poll, void* poll_user, double* d_x, cfd_linear_info* info) {
  return solve_device(g, d_b, d_scale, tol, max_iterations, poll, poll_user, d_x, info);
}

int cfd_solve_scaled(cfd_grid* g, const double* b, const double* roundoff_scale, double tol, int64_t max_iterations,
                     cfd_poll_fn poll, void* poll_user, double* x, cfd_linear_info* info) {
  HIP_TRY(hipSetDevice(g->device));
  size_t bytes = (size_t)g->geo.N * sizeof(double);
  HIP_TRY(hipMemcpyAsync(g->b.p, b, bytes, hipMemcpyHostToDevice, g->stream));
  if (roundoff_scale) HIP_TRY(hipMemcpyAsync(g->scale.p, roundoff_scale, bytes, hipMemcpyHostToDevice, g->stream));
  int rc = solve_device(g, g->b.p, roundoff_scale ? g->scale.p : nullptr, tol, max_iterations, poll, poll_user, g->xo.p,
                        info);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(x, g->xo.p, bytes, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

int cfd_projection_rhs(cfd_grid* g, const double* u, double dt, double* b, double* terms) {
  HIP_TRY(hipSetDevice(g->device));
  size_t bytes = (size_t)g->geo.N * sizeof(double);
  size_t fbytes = (size_t)g->fgeo.total * sizeof(double);
  DevBuf<double> du;
  int rc = du.alloc((size_t)g->fgeo.total);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(du.p, u, fbytes, hipMemcpyHostToDevice, g->stream));
  launch_projection_rhs(g->geo, g->fgeo, g->codes.p, du.p, dt, g->b.p, g->scale.p, g->stream);
  HIP_TRY(hipMemcpyAsync(b, g->b.p, bytes, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(terms, g->scale.p, bytes, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

}  // extern "C"

namespace {
// pressure::project on device buffers u (faces) and p (cells); graphs false
// launches the iterations kernel by kernel.
int project_device(cfd_grid* g, double* d_u, double* d_p, double dt, double tol, int64_t max_iterations,
                   cfd_poll_fn poll, void* poll_user, cfd_linear_info* info, bool graphs) {
  HIP_TRY(hipSetDevice(g->device));
  launch_projection_rhs(g->geo, g->fgeo, g->codes.p, d_u, dt, g->b.p, g->scale.p, g->stream);
  int rc = solve_device(g, g->b.p, g->scale.p, tol, max_iterations, poll, poll_user, g->phi.p, info, graphs);
  if (rc) return rc;
  launch_project_faces(g->geo, g->fgeo, g->codes.p, g->phi.p, dt, d_u, g->stream);
  launch_add(g->geo.N, d_p, g->phi.p, g->stream);
  const int* cidx = g->n_closed > 1 ? g->cidx.p : nullptr;
  enqueue_gauge(g, d_p, nullptr, 0, g->mean_x.p, nullptr);
  launch_center(g->geo.N, g->codes.p, d_p, g->mean_x.p, cidx, d_p, g->stream);
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

// Staging memory of cfd_project_static: one array per GPU (every device that
// loads the library gets its own copy of a __device__ variable, held until the
// process ends): the faces and then the cells of one grid, kStaticDoubles
// values at most (48 MiB; about 1.4 M cells).
constexpr size_t kStaticDoubles = size_t(6) << 20;
static __device__ __align__(256) double scratch[kStaticDoubles];
constexpr int kMaxDevices = 64;
// the device address of each GPU's array, from hipGetSymbolAddress with that
// GPU current (the host-side name `scratch` is not a device address)
double* scratch_ptr[kMaxDevices] = {};

int static_scratch(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("cfd_project_static: device ordinal out of range");
    return CFD_EINVAL;
  }
  if (!scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("cfd_project_static: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return CFD_EDEVICE;
    }
    scratch_ptr[device] = (double*)p;
  }
  *out = scratch_ptr[device];
  return CFD_OK;
}
}  // namespace

extern "C" {

int cfd_project_device(cfd_grid* g, double* d_u, double* d_p, double dt, double tol, int64_t max_iterations,
                       cfd_poll_fn poll, void* poll_user, cfd_linear_info* info) {
  return project_device(g, d_u, d_p, dt, tol, max_iterations, poll, poll_user, info, true);
}

int cfd_project_static(cfd_grid* g, double* u, double* p, double dt, double tol, int64_t max_iterations,
                       cfd_poll_fn poll, void* poll_user, cfd_linear_info* info) {
  if (!g || !u || !p) {
    set_error("cfd_project_static: null argument");
    return CFD_EINVAL;
  }
  const size_t faces = (size_t)g->fgeo.total, cells = (size_t)g->geo.N;
  if (faces + cells > kStaticDoubles) {
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "cfd_project_static: the grid's %zu faces and %zu cells need %zu doubles of staging; the static "
                  "scratch holds %zu (use cfd_project)",
                  faces, cells, faces + cells, kStaticDoubles);
    set_error(buf);
    return CFD_EINVAL;
  }
  HIP_TRY(hipSetDevice(g->device));
  double* base = nullptr;
  if (int rc = static_scratch(g->device, &base)) return rc;
  double* du = base;
  double* dp = base + faces;
  HIP_TRY(hipMemcpyAsync(du, u, faces * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(dp, p, cells * sizeof(double), hipMemcpyHostToDevice, g->stream));
  int rc = project_device(g, du, dp, dt, tol, max_iterations, poll, poll_user, info, false);
  if (rc) return rc;  // u and p untouched on error
  HIP_TRY(hipMemcpyAsync(u, du, faces * sizeof(double), hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(p, dp, cells * sizeof(double), hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

int cfd_project(cfd_grid* g, double* u, double* p, double dt, double tol, int64_t max_iterations, cfd_poll_fn poll,
                void* poll_user, cfd_linear_info* info) {
  HIP_TRY(hipSetDevice(g->device));
  size_t bytes = (size_t)g->geo.N * sizeof(double);
  size_t fbytes = (size_t)g->fgeo.total * sizeof(double);
  DevBuf<double> du, dp;
  int rc = du.alloc((size_t)g->fgeo.total);
  if (!rc) rc = dp.alloc((size_t)g->geo.N);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(du.p, u, fbytes, hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipMemcpyAsync(dp.p, p, bytes, hipMemcpyHostToDevice, g->stream));
  rc = cfd_project_device(g, du.p, dp.p, dt, tol, max_iterations, poll, poll_user, info);
  if (rc) return rc;  // u and p untouched on error
  HIP_TRY(hipMemcpyAsync(u, du.p, fbytes, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipMemcpyAsync(p, dp.p, bytes, hipMemcpyDeviceToHost, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

int cfd_profile_kernel(cfd_grid* g, int kernel, int reps, double* avg_seconds, double* bytes_per_launch) {
  if (reps < 1 || kernel < 0 || kernel > 5) {
    set_error("cfd_profile_kernel: kernel must be 0..5 and reps positive");
    return CFD_EINVAL;
  }
  if ((kernel == 1 || kernel == 4) && !g->has_hier) {
    set_error("cfd_profile_kernel: the grid has no multigrid hierarchy");
    return CFD_EINVAL;
  }
  if (kernel == 3 && g->host.preconditioner != CFD_IC0) {
    set_error("cfd_profile_kernel: the grid's preconditioner is not ic0");
    return CFD_EINVAL;
  }
  HIP_TRY(hipSetDevice(g->device));
  const long long N = g->geo.N;
  PcgState* st = g->st.p;
  // Work vectors with finite values; the state stays RUNNING with alpha = beta = 0.
  launch_fill(N, g->z.p, 1.0, g->stream);
  launch_fill(N, g->p0.p, 0.5, g->stream);
  launch_fill(N, g->r.p, 0.25, g->stream);
  launch_fill(N, g->x.p, 0.0, g->stream);
  launch_state_init(st, 1.0, 1LL << 60, g->stream);
  const double* mz = g->mean_z.p;
  const int* cidx = g->n_closed > 1 ? g->cidx.p : nullptr;
  if (g->n_closed > 0) launch_fill(std::max(1, g->n_closed), g->mean_z.p, 0.0, g->stream);
  if (kernel == 5) {
    // a later iteration: p = z + beta p_old with beta = (sum of the r . z partials) / rz = 1
    launch_fill(kReduceGrid, g->part.p + P_RZ, 1.0, g->stream);
    PcgState h{};
    h.rz = (double)kReduceGrid;
    h.tol = 1.0;
    h.max_it = 1LL << 60;
    h.status = ST_RUNNING;
    h.first = 0;
    HIP_TRY(hipMemcpyAsync(st, &h, sizeof h, hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
  }
  const bool mixed = g->opt.precision == CFD_PRECISION_MIXED;
  auto once = [&]() {
    switch (kernel) {
      case 0:
      case 5:
        launch_pcg_op(g->geo, g->codes.p, g->z.p, g->p0.p, g->p1.p, g->ap.p, mz, cidx, g->n_closed > 0, st, g->part.p,
                      g->stream, g->march);
        break;
      case 1:
        if (g->opt.cycle == CFD_CYCLE_RB && g->ash.has_rb) {
          if (mixed)
            launch_rb_gs<float>(g->geo, g->ash.rb32, 1, g->stream);
          else
            launch_rb_gs<double>(g->geo, g->ash.rb64, 1, g->stream);
        } else if (mixed) {
          launch_fine_gs<float, double>(g->geo, g->codes.p, g->r.p, g->x0f.p, 1, Geo{}, g->stream);
        } else {
          launch_fine_gs<double, double>(g->geo, g->codes.p, g->r.p, g->z.p, 1, Geo{}, g->stream);
        }
        break;
      case 2:
        launch_pcg_axpy(N, g->p0.p, g->ap.p, g->x.p, g->r.p, st, g->part.p, g->stream);
        break;
      default:
        enqueue_precondition(g, g->r.p, g->z.p, st);
        break;
    }
  };
  once();  // warm
  HIP_TRY(hipEventRecord(g->ev0, g->stream));
  for (int i = 0; i < reps; ++i) once();
  HIP_TRY(hipEventRecord(g->ev1, g->stream));
  HIP_TRY(hipStreamSynchronize(g->stream));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, g->ev0, g->ev1));
  *avg_seconds = (double)ms * 1e-3 / reps;
  const double n = (double)N;
  switch (kernel) {
    case 0:  // first iteration: z, code in; p_new, Ap out
      *bytes_per_launch = n * (8 + 1 + 8 + 8);
      break;
    case 5:  // z, p_old, code in; p_new, Ap out
      *bytes_per_launch = n * (8 + 8 + 1 + 8 + 8);
      break;
    case 1:
      if (g->opt.cycle == CFD_CYCLE_RB && g->ash.has_rb) {
        // colour-separated: the swept half's code, b and new x, the other colour's x
        const double sz = mixed ? 4.0 : 8.0;
        *bytes_per_launch = (double)g->ash.rb64.half * (1.0 + 3.0 * sz);
      } else {
        // natural layout: half the cells; neighbour x, own b (the FP64 r), own code in; own x out
        *bytes_per_launch = mixed ? 0.5 * n * (4 + 8 + 1 + 4) : 0.5 * n * (8 + 8 + 1 + 8);
      }
      break;
    case 2:  // x, p, r, Ap in; x, r out
      *bytes_per_launch = n * 48;
      break;
    case 3:  // forward r, D, code in, y out; backward y, D, code in, z out
      *bytes_per_launch = n * 50;
      break;
    default: {  // fine level: 8 half-sweeps + residual/restriction; coarse levels add about 1/7 more
      double fine = mixed ? 8 * 0.5 * (4 + 8 + 1 + 4) + (4 + 8 + 1 + 0.5) : 8 * 0.5 * (8 + 8 + 1 + 8) + (8 + 8 + 1 + 1);
      *bytes_per_launch = n * fine * 8.0 / 7.0;
    }
  }
  return CFD_OK;
}

int cfd_profile_solve(cfd_grid* g, const double* d_b, const double* d_scale, double tol, int64_t max_iterations,
                      double* d_x, int kernel, cfd_kernel_probe* out, cfd_linear_info* info) {
  if (!g || !out || kernel < 1 || kernel >= PROBE_KERNELS) {
    set_error("cfd_profile_solve: kernel must be one of CFD_PROBE_*");
    return CFD_EINVAL;
  }
  HIP_TRY(hipSetDevice(g->device));
  LaunchProbe& P = g->probe;
  if (P.events.empty()) {
    P.events.resize(8192, nullptr);
    for (hipEvent_t& e : P.events) HIP_TRY(hipEventCreate(&e));
  }
  P.kernel = kernel;
  P.used = 0;
  P.tag = 0;
  P.bytes.clear();
  P.tags.clear();
  t_probe = &P;
  cfd_linear_info local{};
  int rc = solve_device(g, d_b, d_scale, tol, max_iterations, nullptr, nullptr, d_x, &local, false);
  t_probe = nullptr;
  if (info) *info = local;
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  // launches queued past the converging iteration end at once (the device state
  // is no longer RUNNING): only the solve's own work is counted, the launches of
  // the set-up (tag 0) and of iterations 1..it
  const long long it = local.iterations;
  std::memset(out, 0, sizeof *out);
  for (size_t k = 0; 2 * k + 1 < P.used; ++k) {
    const long long tag = P.tags[k];
    if (tag > it) {
      ++out->dropped;
      continue;
    }
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, P.events[2 * k], P.events[2 * k + 1]));
    out->launches += 1;
    out->seconds += (double)ms * 1e-3;
    out->bytes += P.bytes[k];
  }
  out->truncated = P.used + 2 > P.events.size() ? 1 : 0;
  return CFD_OK;
}

int cfd_grid_perturb(cfd_grid* g, const double* d_src, double* d_dst, double rel, uint64_t seed) {
  if (!g || !d_src || !d_dst) {
    set_error("cfd_grid_perturb: null argument");
    return CFD_EINVAL;
  }
  HIP_TRY(hipSetDevice(g->device));
  launch_perturb(g->geo.N, g->codes.p, d_src, d_dst, rel, seed, g->part.p, (double)std::max<int64_t>(1, g->host.fluid_count),
                 g->stream);
  HIP_TRY(hipStreamSynchronize(g->stream));
  return CFD_OK;
}

int cfd_device_alloc(size_t bytes, void** d_ptr) {
  HIP_TRY(hipMalloc(d_ptr, bytes));
  return CFD_OK;
}
int cfd_device_free(void* d_ptr) {
  HIP_TRY(hipFree(d_ptr));
  return CFD_OK;
}
int cfd_copy_to_device(void* d_dst, const void* h_src, size_t bytes) {
  HIP_TRY(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  return CFD_OK;
}
int cfd_copy_to_host(void* h_dst, const void* d_src, size_t bytes) {
  HIP_TRY(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
  return CFD_OK;
}
int cfd_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return CFD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- serving
// Continuous batching of independent pressure solves: each request names a
// resident grid; up to max_active requests run at once, and a request that
// arrives while others are on the GPU starts as soon as a slot and its grid are
// free. Requests whose grid and precision allow it (red-black cycle,
// colour-separated fine level with two cells per thread, no closed component)
// join a batch of their precision: the iterations of all its solves go out as
// one launch per stage (launch_batch_iteration), and a solve joins or leaves
// the batch at any iteration. Other requests run on their grid's own stream.
// Grids are admitted against an HBM budget by their measured footprint.
namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
constexpr int kStepRing = 4;   // host slots for the states read back after each batched step
// batched steps queued beyond the last one seen finished (CFD_SERVE_AHEAD): one,
// so that a solve that converged leaves its batch at the next step; the other
// lanes keep the GPU busy while the host turns a lane's step around (2: 1042 /s
// against 1120 /s with one, two lanes on the served mix)
constexpr int kStepsAhead = 1;
}  // namespace

struct cfd_server {
  struct Req {
    cfd_grid* g = nullptr;
    const double* b = nullptr;
    const double* scale = nullptr;
    double* x = nullptr;
    double tol = 0.0;
    int64_t max = 0;
    int precision = 0;
    int grid_precision = 0;  // the grid's own setting, restored when the request ends
    double t_submit = 0.0, t_start = 0.0, t_end = 0.0;
    // 0 queued, 1 running, 2 done; a batched request is driven by its lane's
    // thread, which publishes 2 (release) after its results
    std::atomic<int> state{0};
    int lane = -1;    // batched: the lane it runs in
    int64_t enq = 0;  // batched: iterations queued (lane thread)
    int rc = 0;
    std::string error;
    cfd_linear_info info{};
  };
  // A lane: a stream running batches (one per precision) step after step, driven
  // by a host thread of its own (the kernel launches of a step take about 3 us
  // each on the host). Lanes run concurrently, so one lane's latency-bound coarse
  // levels overlap another's streaming fine-level kernels.
  struct Step {
    int slot = 0;
    std::vector<Req*> items;  // the solves whose states the step reads back, in slot order
  };
  struct Lane {
    hipStream_t bs = nullptr;
    std::vector<Req*> group[2];  // solves in the FP64 and the mixed-precision batch (lane thread)
    std::vector<Req*> finishing;  // solves whose final kernels are queued (lane thread)
    std::deque<Step> steps;
    int next_slot = 0;
    // kStepRing x 2 kBatchMax states in coherent pinned host memory, written by
    // the step's last kernel (no copy call per step)
    PcgState* h_states = nullptr;
    PcgState* hd_states = nullptr;  // the same memory as the device addresses it
    hipEvent_t step_ev[kStepRing] = {};
    std::vector<Req*> joining;  // handed over by the driving thread (under the server's lock)
    std::atomic<bool> has_joining{false};
    int load = 0;               // solves assigned and not yet collected (driving thread)
    int failed = 0;             // a device error stopped the lane (its code)
    std::thread worker;
  };
  // A queued admission (cfd_server_admit_async): the case copied, its grid built
  // by the server's builder thread once the budget has room for it.
  struct Admission {
    std::vector<cfd_block> blocks;
    std::vector<std::vector<double>> footprints;
    cfd_case c{};
    size_t need = 0;  // measured footprint of an earlier build that did not fit (0: estimate)
    int state = 0;    // 0 waiting, 1 building, 2 admitted, 3 failed
    cfd_grid* g = nullptr;
    int rc = 0;
    std::string error;
  };
  int max_active = 8;
  int device = 0;
  size_t budget = 0, used = 0;  // used: footprints of the admitted grids (under am)
  uint64_t next = 1;
  std::unordered_map<uint64_t, Req> reqs;  // node-based: a Req keeps its address while others come and go
  std::deque<uint64_t> queued;
  std::vector<uint64_t> running;  // on their grids' streams
  std::vector<uint64_t> batched;  // handed to lanes, not yet collected
  std::unordered_map<cfd_grid*, int> busy;
  std::unordered_map<cfd_grid*, size_t> grids;  // admitted grids and their footprints
  int n_lanes = 3;  // 0: no batching
  int ahead = kStepsAhead;
  std::vector<std::unique_ptr<Lane>> lanes;
  std::mutex m;
  std::condition_variable cv;
  bool stop = false;
  // admissions, grids, used and reserved are guarded by am
  std::mutex am;
  std::condition_variable acv;
  std::map<uint64_t, Admission> admissions;  // by ticket: builds start in ticket order
  uint64_t next_admission = 1;
  size_t reserved = 0;         // bytes held for the admissions being built
  double bytes_per_cell = 0;   // the largest measured footprint per cell so far
  int n_builders = 2;          // builder threads (CFD_SERVE_BUILDERS)
  std::vector<std::thread> builders;
  bool builder_stop = false;
  ~cfd_server() {
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    cv.notify_all();
    {
      std::lock_guard<std::mutex> lk(am);
      builder_stop = true;
    }
    acv.notify_all();
    for (std::thread& t : builders)
      if (t.joinable()) t.join();
    // grids admitted from the queue and never collected
    for (auto& kv : admissions)
      if (kv.second.state == 2 && kv.second.g) cfd_grid_destroy(kv.second.g);
    for (auto& L : lanes)
      if (L->worker.joinable()) L->worker.join();
    for (auto& L : lanes) {
      if (L->bs) (void)hipStreamSynchronize(L->bs);
      for (hipEvent_t e : L->step_ev)
        if (e) (void)hipEventDestroy(e);
      if (L->h_states) (void)hipHostFree(L->h_states);
      if (L->bs) (void)hipStreamDestroy(L->bs);
    }
  }
};

namespace {
// The request's grid and precision run in a batch: the red-black cycle on the
// colour-separated fine level with two cells per thread (nx a multiple of 4), no
// closed component, and the FP64 solve's r and z kept in that level's arrays
// (or the mixed solve's z as FP32 values).
bool batchable(const cfd_grid* g, int precision) {
  if (!g->has_hier || g->host.preconditioner == CFD_IC0 || g->opt.cycle != CFD_CYCLE_RB) return false;
  const DevAggShared& A = g->ash;
  if (!A.has_rb || A.nlev < 1 || A.seams != 0 || g->n_closed != 0 || g->fine_rounds < 1 || g->coarse_rounds < 1)
    return false;
  if (g->geo.n[0] % 4 != 0 || g->geo.N >= (1LL << 31) || (double)g->geo.N * g->geo.n[0] >= 1099511627776.0)
    return false;
  if (precision == CFD_PRECISION_MIXED) return A.rb32.two_cells == 2 && g->rb_resident;
  return A.rb64.two_cells == 2;
}

// The solve of grid g as the batched launches see it; pold / pnew by the parity
// of its next iteration.
template <class T>
BatchSolve<T> batch_solve(cfd_grid* g, const FineRB<T>& F, const std::vector<AggDevLevel<T>>& lv, int64_t enq) {
  BatchSolve<T> b;
  const DevAggShared& A = g->ash;
  b.g = g->geo;
  b.codes = g->codes.p;
  b.F = F;
  b.fparent = A.fparent.p;
  b.nlev = A.nlev;
  b.lv = lv.data();
  b.D = A.D;
  b.st = g->st.p;
  b.part = g->part.p;
  b.x = g->x.p;
  b.bc = g->bc.p;
  b.ap = g->ap.p;
  const bool odd = enq & 1;
  b.pold = odd ? g->p1.p : g->p0.p;
  b.pnew = odd ? g->p0.p : g->p1.p;
  b.rres = A.rb64.b;
  b.cs = (T)g->omega;
  b.zero_last = g->zero_last;
  b.fine_rounds = g->fine_rounds;
  b.coarse_rounds = g->coarse_rounds;
  b.march = g->march == 2 || (g->march == 1 && march_fits(g->geo));
  return b;
}

// A request joins a batch of its precision (lane thread): the set-up kernels of
// solve_scaled (pressure.rs:342-375) on the lane's stream, and the residual
// scattered into the FP64 fine level's colour-separated arrays, where the
// batched iterations keep it. Its first batched iteration applies the
// preconditioner to that residual.
int batch_begin(cfd_server::Lane& L, cfd_server::Req& R) {
  cfd_grid* g = R.g;
  hipStream_t st = L.bs;
  PcgState* sp = g->st.p;
  HIP_TRY(hipEventRecord(g->ev0, st));
  launch_state_init(sp, R.tol, R.max, st);
  launch_pcg_init(g->geo.N, g->codes.p, R.b, g->mean_b.p, nullptr, g->r.p, g->bc.p, g->x.p, sp, g->part.p, st);
  if (g->opt.warm_start && g->warm_valid)
    launch_warm_residual(g->geo, g->codes.p, g->warm.p, g->bc.p, g->x.p, g->r.p, sp, st);
  // (FP64: the scatter is the cycle's first half-sweep; mixed: from the FP64
  // residual into the FP32 level)
  launch_rb_scatter<double>(g->geo, g->codes.p, g->r.p, g->ash.rb64, sp, st);
  if (R.precision == CFD_PRECISION_MIXED) launch_rb_first_cvt(g->geo, g->ash.rb32, g->ash.rb64.b, sp, st);
  R.enq = 0;
  return 0;
}

// One batched step of a lane: an iteration of every solve in each of its
// batches, then their states written into the lane's next host slot.
int batch_step(cfd_server::Lane& L) {
  cfd_server::Step step;
  step.slot = L.next_slot;
  L.next_slot = (L.next_slot + 1) % kStepRing;
  std::vector<PcgState*> sts;
  for (int grp = 0; grp < 2; ++grp) {
    std::vector<cfd_server::Req*>& rs = L.group[grp];
    if (rs.empty()) continue;
    if (grp == 0) {
      std::vector<BatchSolve<double>> items;
      for (cfd_server::Req* R : rs) items.push_back(batch_solve<double>(R->g, R->g->ash.rb64, R->g->a64.levels, R->enq));
      launch_batch_iteration<double>(items.data(), (int)items.size(), L.bs);
    } else {
      std::vector<BatchSolve<float>> items;
      for (cfd_server::Req* R : rs) items.push_back(batch_solve<float>(R->g, R->g->ash.rb32, R->g->a32.levels, R->enq));
      launch_batch_iteration<float>(items.data(), (int)items.size(), L.bs);
    }
    for (cfd_server::Req* R : rs) {
      R->enq += 1;
      step.items.push_back(R);
      sts.push_back(R->g->st.p);
    }
  }
  PcgState* d = L.hd_states + (size_t)step.slot * 2 * kBatchMax;
  for (size_t k = 0; k < sts.size(); k += kBatchMax)
    launch_batch_states(sts.data() + k, (int)std::min<size_t>(kBatchMax, sts.size() - k), d + k, L.bs);
  HIP_TRY(hipEventRecord(L.step_ev[step.slot], L.bs));
  L.steps.push_back(std::move(step));
  return 0;
}

// A batched solve whose state shows it ended (lane thread): its LinearInfo and
// code, and the final kernels of solve_scaled (x centered into the caller's
// buffer, the warm start kept) on the lane's stream. The request ends once that
// stream has passed them, and with them every queued step that still names its
// grid.
int batch_finish(cfd_server::Lane& L, cfd_server::Req& R, const PcgState& st) {
  cfd_grid* g = R.g;
  std::vector<cfd_server::Req*>& grp = L.group[R.precision == CFD_PRECISION_MIXED ? 1 : 0];
  grp.erase(std::find(grp.begin(), grp.end(), &R));
  if (st.status == ST_CONVERGED) g->learned[learn_key(g, R.tol)] = st.it;
  fill_info(g, st, R.info);
  R.rc = finish_status(g, st, R.max);
  R.error = R.rc ? g_error : std::string();
  if (!R.rc) {
    if (st.status == ST_ZERO) {
      launch_fill(g->geo.N, R.x, 0.0, L.bs);
    } else {
      launch_center(g->geo.N, g->codes.p, g->x.p, g->mean_x.p, nullptr, R.x, L.bs);
    }
    if (g->opt.warm_start) {
      HIP_TRY(hipMemcpyAsync(g->warm.p, R.x, (size_t)g->geo.N * sizeof(double), hipMemcpyDeviceToDevice, L.bs));
      g->warm_valid = true;
    }
  }
  HIP_TRY(hipEventRecord(g->ev1, L.bs));
  HIP_TRY(hipEventRecord(g->evs, L.bs));
  L.finishing.push_back(&R);
  return 0;
}

// One pass of a lane: joining solves set up, steps whose states landed read
// (solves that ended leave their batch), finished final kernels complete their
// requests, and steps queued while a batch has a solve, at most `ahead` past
// the last step seen finished. Returns whether anything happened.
int lane_pass(cfd_server* s, cfd_server::Lane& L, bool& moved) {
  std::vector<cfd_server::Req*> join;
  if (L.has_joining.load(std::memory_order_acquire)) {
    std::lock_guard<std::mutex> lk(s->m);
    join.swap(L.joining);
    L.has_joining.store(false, std::memory_order_relaxed);
  }
  for (cfd_server::Req* R : join) {
    int rc = batch_begin(L, *R);
    if (rc) return rc;
    L.group[R->precision == CFD_PRECISION_MIXED ? 1 : 0].push_back(R);
    moved = true;
  }
  while (!L.steps.empty()) {
    const cfd_server::Step& step = L.steps.front();
    hipError_t q = hipEventQuery(L.step_ev[step.slot]);
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) {
      set_error(std::string("HIP error: ") + hipGetErrorString(q));
      return CFD_EDEVICE;
    }
    const PcgState* h = L.h_states + (size_t)step.slot * 2 * kBatchMax;
    for (size_t k = 0; k < step.items.size(); ++k) {
      cfd_server::Req& R = *step.items[k];
      const bool live = std::find(L.group[0].begin(), L.group[0].end(), &R) != L.group[0].end() ||
                        std::find(L.group[1].begin(), L.group[1].end(), &R) != L.group[1].end();
      if (live && h[k].status != ST_RUNNING) {
        int rc = batch_finish(L, R, h[k]);
        if (rc) return rc;
      }
    }
    L.steps.pop_front();
    moved = true;
  }
  for (size_t i = 0; i < L.finishing.size();) {
    cfd_server::Req& R = *L.finishing[i];
    // a step still to be read names the solve: it is handed back after that step
    bool named = false;
    for (const cfd_server::Step& st : L.steps)
      named = named || std::find(st.items.begin(), st.items.end(), &R) != st.items.end();
    hipError_t q = named ? hipErrorNotReady : hipEventQuery(R.g->evs);
    if (q == hipErrorNotReady) {
      ++i;
      continue;
    }
    if (q != hipSuccess) {
      set_error(std::string("HIP error: ") + hipGetErrorString(q));
      return CFD_EDEVICE;
    }
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, R.g->ev0, R.g->ev1));
    R.info.seconds = ms * 1e-3;
    R.t_end = now_s();
    R.state.store(2, std::memory_order_release);
    L.finishing[i] = L.finishing.back();
    L.finishing.pop_back();
    moved = true;
  }
  while ((!L.group[0].empty() || !L.group[1].empty()) && (int)L.steps.size() < s->ahead) {
    int rc = batch_step(L);
    if (rc) return rc;
    moved = true;
  }
  return 0;
}

// A lane's thread: passes while the lane has work (taking the server's lock
// only to pick up joining solves), else it waits for a joining solve (or the
// server's end).
void lane_worker(cfd_server* s, cfd_server::Lane* L, int device) {
  (void)hipSetDevice(device);
  for (;;) {
    const bool idle = L->group[0].empty() && L->group[1].empty() && L->steps.empty() && L->finishing.empty();
    if (idle && !L->has_joining.load(std::memory_order_acquire)) {
      std::unique_lock<std::mutex> lk(s->m);
      s->cv.wait(lk, [&] { return s->stop || !L->joining.empty(); });
      if (s->stop && L->joining.empty()) return;
    }
    bool moved = false;
    const int rc = lane_pass(s, *L, moved);
    if (rc) {
      // the lane stops; its solves end with the device error
      const std::string error = g_error;
      std::vector<cfd_server::Req*> all;
      for (int grp = 0; grp < 2; ++grp) all.insert(all.end(), L->group[grp].begin(), L->group[grp].end());
      all.insert(all.end(), L->finishing.begin(), L->finishing.end());
      {
        std::lock_guard<std::mutex> lk(s->m);
        all.insert(all.end(), L->joining.begin(), L->joining.end());
        L->joining.clear();
        L->failed = rc;
      }
      for (cfd_server::Req* R : all) {
        R->rc = rc;
        R->error = error;
        R->t_end = now_s();
        R->state.store(2, std::memory_order_release);
      }
      L->group[0].clear();
      L->group[1].clear();
      L->finishing.clear();
      L->steps.clear();
      continue;
    }
    if (!moved) std::this_thread::yield();
  }
}

int ensure_lanes(cfd_server* s) {
  if ((int)s->lanes.size() == s->n_lanes) return 0;
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  for (int i = (int)s->lanes.size(); i < s->n_lanes; ++i) {
    std::unique_ptr<cfd_server::Lane> L(new cfd_server::Lane());
    HIP_TRY(hipStreamCreateWithFlags(&L->bs, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc((void**)&L->h_states, (size_t)kStepRing * 2 * kBatchMax * sizeof(PcgState),
                          hipHostMallocMapped | hipHostMallocCoherent));
    HIP_TRY(hipHostGetDevicePointer((void**)&L->hd_states, L->h_states, 0));
    for (hipEvent_t& e : L->step_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    L->worker = std::thread(lane_worker, s, L.get(), device);
    s->lanes.push_back(std::move(L));
  }
  return 0;
}

void end_request(cfd_server* s, cfd_server::Req& R, int rc, const std::string& error) {
  R.t_end = now_s();
  R.rc = rc;
  R.error = error;
  R.state.store(2, std::memory_order_relaxed);
  R.g->opt.precision = R.grid_precision;
  s->busy.erase(R.g);
}

// The lane a batched request of precision group grp joins: the one with the
// least load that has room in that group (-1: none). Load counts the solves
// assigned and not yet collected (at most kBatchMax per precision, shared by
// both groups here).
int pick_lane(const cfd_server* s) {
  int best = -1;
  for (size_t i = 0; i < s->lanes.size(); ++i) {
    const cfd_server::Lane& L = *s->lanes[i];
    if (L.failed || L.load >= kBatchMax) continue;
    if (best < 0 || L.load < s->lanes[(size_t)best]->load) best = (int)i;
  }
  return best;
}

int active_count(const cfd_server* s) { return (int)(s->running.size() + s->batched.size()); }

// Drives the requests on their grids' own streams one step, collects batched
// requests that their lanes finished and starts queued requests whose grid is
// idle (batched ones are handed to a lane's thread).
int server_progress(cfd_server* s) {
  std::lock_guard<std::mutex> lk(s->m);
  for (size_t i = 0; i < s->batched.size();) {
    cfd_server::Req& R = s->reqs[s->batched[i]];
    if (R.state.load(std::memory_order_acquire) != 2) {
      ++i;
      continue;
    }
    s->lanes[(size_t)R.lane]->load -= 1;
    R.g->opt.precision = R.grid_precision;
    s->busy.erase(R.g);
    s->batched[i] = s->batched.back();
    s->batched.pop_back();
  }
  for (size_t i = 0; i < s->running.size();) {
    cfd_server::Req& R = s->reqs[s->running[i]];
    int jr = job_poll(R.g, false, nullptr, nullptr);
    if (jr == JOB_PENDING) {
      ++i;
      continue;
    }
    if (jr == JOB_DONE) {
      R.info = R.g->job.info;
      end_request(s, R, R.g->job.rc, R.g->job.error);
    } else {
      end_request(s, R, jr, g_error);
    }
    s->running[i] = s->running.back();
    s->running.pop_back();
  }
  bool handed = false;
  for (size_t i = 0; i < s->queued.size() && active_count(s) < s->max_active;) {
    const uint64_t id = s->queued[i];
    cfd_server::Req& R = s->reqs[id];
    if (s->busy.count(R.g)) {
      ++i;
      continue;
    }
    const bool batch = s->n_lanes > 0 && batchable(R.g, R.precision);
    int lane = -1;
    if (batch) {
      if (int rc = ensure_lanes(s)) return -rc;
      lane = pick_lane(s);
      if (lane < 0) {
        ++i;
        continue;
      }
    }
    // the request's precision for the duration of the job (the iteration graphs
    // and learned counts are keyed by it); the grid's own setting comes back after
    R.grid_precision = R.g->opt.precision;
    R.g->opt.precision = R.precision;
    int rc = 0;
    if (!batch) rc = job_begin(R.g, R.b, R.scale, R.tol, R.max, R.x);
    R.t_start = now_s();
    s->queued.erase(s->queued.begin() + (long)i);
    if (rc) {
      R.g->opt.precision = R.grid_precision;
      R.rc = rc;
      R.error = g_error;
      R.t_end = R.t_start;
      R.state.store(2, std::memory_order_relaxed);
      continue;
    }
    R.state.store(1, std::memory_order_relaxed);
    s->busy[R.g] = 1;
    if (batch) {
      R.lane = lane;
      cfd_server::Lane& L = *s->lanes[(size_t)lane];
      L.load += 1;
      L.joining.push_back(&R);
      L.has_joining.store(true, std::memory_order_release);
      s->batched.push_back(id);
      handed = true;
    } else {
      s->running.push_back(id);
    }
  }
  if (handed) s->cv.notify_all();
  return (int)(active_count(s) + s->queued.size());
}
}  // namespace

namespace {
// An admitted grid counted against the budget (the caller holds am).
void admit_locked(cfd_server* s, cfd_grid* g) {
  s->used += g->footprint;
  s->grids[g] = g->footprint;
  const double cells = (double)g->geo.N;
  if (cells > 0) s->bytes_per_cell = std::max(s->bytes_per_cell, (double)g->footprint / cells);
}

// Bytes a waiting admission is expected to take: the measured footprint of an
// earlier build that did not fit, else its cells times the largest footprint
// per cell measured so far (2 % over), or 256 B a cell before any grid.
size_t admission_need(const cfd_server& s, const cfd_server::Admission& A) {
  if (A.need) return A.need;
  const double cells = (double)A.c.cells[0] * (double)A.c.cells[1] * (double)A.c.cells[2];
  const double per = s.bytes_per_cell > 0 ? 1.02 * s.bytes_per_cell : 256.0;
  return (size_t)(cells * per);
}

// A builder thread: queued admissions taken in ticket order, each built once the
// budget has room for it beside the admitted grids and the builds under way (an
// empty server builds whatever comes next, so a grid larger than the whole
// budget ends with CFD_ENOMEM instead of waiting forever). The grid set-up runs
// outside the lock, beside the solves of the lanes and the other builder's set-up
// (two builders: the host stages of one set-up overlap the other's device stages).
void builder_loop(cfd_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->am);
  for (;;) {
    cfd_server::Admission* A = nullptr;
    size_t need = 0;
    s->acv.wait(lk, [&] {
      if (s->builder_stop) return true;
      A = nullptr;
      for (auto& kv : s->admissions)
        if (kv.second.state == 0) {
          A = &kv.second;
          break;
        }
      if (!A) return false;
      need = admission_need(*s, *A);
      return s->used + s->reserved + need <= s->budget || (s->used == 0 && s->reserved == 0);
    });
    if (s->builder_stop) return;
    A->state = 1;
    s->reserved += need;
    const bool pairs = s->n_lanes > 0;
    lk.unlock();
    cfd_grid* g = nullptr;
    const int rc = create_grid(&A->c, &g, pairs);
    const std::string error = rc ? g_error : std::string();
    lk.lock();
    s->reserved -= need;
    if (rc) {
      A->state = 3;
      A->rc = rc;
      A->error = error;
      continue;
    }
    if (s->used + s->reserved + g->footprint > s->budget) {
      if (s->used == 0 && s->reserved == 0) {
        char buf[200];
        std::snprintf(buf, sizeof buf, "HBM admission: the grid needs %zu bytes, more than the whole budget of %zu",
                      g->footprint, s->budget);
        A->state = 3;
        A->rc = CFD_ENOMEM;
        A->error = buf;
      } else {
        // waits for room for its measured footprint
        A->state = 0;
        A->need = g->footprint;
      }
      lk.unlock();
      cfd_grid_destroy(g);
      lk.lock();
      s->acv.notify_all();
      continue;
    }
    admit_locked(s, g);
    A->g = g;
    A->state = 2;
    // a smaller measured footprint than the estimate leaves room for another build
    s->acv.notify_all();
  }
}
}  // namespace

extern "C" {

int cfd_server_create(int max_active, size_t hbm_budget_bytes, cfd_server** out) {
  if (!out || max_active < 1) {
    set_error("cfd_server_create: max_active must be positive");
    return CFD_EINVAL;
  }
  cfd_server* s = new cfd_server();
  s->max_active = max_active;
  if (hbm_budget_bytes == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) hbm_budget_bytes = free_b / 10 * 9;
  }
  s->budget = hbm_budget_bytes;
  if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
  if (const char* env = std::getenv("CFD_SERVE_BUILDERS")) s->n_builders = std::max(1, std::min(4, std::atoi(env)));
  if (const char* env = std::getenv("CFD_SERVE_AHEAD")) s->ahead = std::max(1, std::min(kStepRing - 1, std::atoi(env)));
  *out = s;
  return CFD_OK;
}

int cfd_server_set_batching(cfd_server* s, int lanes) {
  if (!s || lanes < 0 || lanes > 8) {
    set_error("cfd_server_set_batching: lanes must be 0 .. 8");
    return CFD_EINVAL;
  }
  std::lock_guard<std::mutex> lk(s->m);
  if (!s->lanes.empty() && lanes != (int)s->lanes.size()) {
    set_error("cfd_server_set_batching: the lanes are set before the first request");
    return CFD_EINVAL;
  }
  s->n_lanes = lanes;
  return CFD_OK;
}

void cfd_server_destroy(cfd_server* s) {
  if (!s) return;
  while (server_progress(s) > 0) std::this_thread::yield();
  delete s;
}

int cfd_server_admit(cfd_server* s, const cfd_case* c, cfd_grid** out) {
  cfd_grid* g = nullptr;
  // a batching server's grids take two fine cells per thread (the batched layout)
  int rc = create_grid(c, &g, s->n_lanes > 0);
  if (rc) return rc;
  std::unique_lock<std::mutex> lk(s->am);
  if (s->used + s->reserved + g->footprint > s->budget) {
    char buf[200];
    std::snprintf(buf, sizeof buf, "HBM admission refused: the grid needs %zu bytes, %zu of %zu are in use", g->footprint,
                  s->used + s->reserved, s->budget);
    lk.unlock();
    cfd_grid_destroy(g);
    set_error(buf);
    return CFD_ENOMEM;
  }
  admit_locked(s, g);
  *out = g;
  return CFD_OK;
}

int cfd_server_admit_async(cfd_server* s, const cfd_case* c, uint64_t* ticket) {
  if (!s || !c || !ticket || c->n_blocks < 0 || (c->n_blocks > 0 && !c->blocks)) {
    set_error("cfd_server_admit_async: null argument or block list");
    return CFD_EINVAL;
  }
  for (int32_t i = 0; i < c->n_blocks; ++i)
    if (c->blocks[i].n_vertices < 0 || (c->blocks[i].n_vertices > 0 && !c->blocks[i].footprint_xy)) {
      set_error("cfd_server_admit_async: a prism without its footprint");
      return CFD_EINVAL;
    }
  std::lock_guard<std::mutex> lk(s->am);
  const uint64_t id = s->next_admission++;
  cfd_server::Admission& A = s->admissions[id];
  // the case is copied: the caller's arrays need not outlive this call
  A.c = *c;
  A.blocks.assign(c->blocks, c->blocks + c->n_blocks);
  A.footprints.resize(A.blocks.size());
  for (size_t i = 0; i < A.blocks.size(); ++i) {
    cfd_block& b = A.blocks[i];
    if (b.n_vertices > 0) {
      A.footprints[i].assign(b.footprint_xy, b.footprint_xy + 2 * (size_t)b.n_vertices);
      b.footprint_xy = A.footprints[i].data();
    } else {
      b.footprint_xy = nullptr;
    }
  }
  A.c.blocks = A.blocks.empty() ? nullptr : A.blocks.data();
  if (s->builders.empty())
    for (int i = 0; i < s->n_builders; ++i) s->builders.emplace_back(builder_loop, s);
  s->acv.notify_all();
  *ticket = id;
  return CFD_OK;
}

int cfd_server_admitted(cfd_server* s, uint64_t ticket, int* done, cfd_grid** out) {
  if (!s || !done || !out) {
    set_error("cfd_server_admitted: null argument");
    return CFD_EINVAL;
  }
  *done = 0;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(s->am);
  auto it = s->admissions.find(ticket);
  if (it == s->admissions.end()) {
    set_error("cfd_server_admitted: unknown admission ticket");
    return CFD_EINVAL;
  }
  cfd_server::Admission& A = it->second;
  if (A.state < 2) return CFD_OK;
  *done = 1;
  const int rc = A.rc;
  if (A.state == 2)
    *out = A.g;
  else
    set_error(A.error);
  s->admissions.erase(it);
  return rc;
}

int cfd_server_release(cfd_server* s, cfd_grid* g) {
  size_t footprint = 0;
  {
    std::lock_guard<std::mutex> lk(s->am);
    auto it = s->grids.find(g);
    if (it == s->grids.end()) {
      set_error("cfd_server_release: the grid was not admitted by this server");
      return CFD_EINVAL;
    }
    footprint = it->second;
  }
  while (s->busy.count(g)) {
    if (server_progress(s) < 0) return CFD_EDEVICE;
    std::this_thread::yield();
  }
  for (const uint64_t id : s->queued)
    if (s->reqs[id].g == g) {
      set_error("cfd_server_release: requests on this grid are still queued");
      return CFD_EINVAL;
    }
  // batched steps queued before the last solve on the grid ended still name it
  for (auto& L : s->lanes)
    if (L->bs) HIP_TRY(hipStreamSynchronize(L->bs));
  // the memory is freed before a queued admission may take its place
  cfd_grid_destroy(g);
  {
    std::lock_guard<std::mutex> lk(s->am);
    s->used -= footprint;
    s->grids.erase(g);
  }
  s->acv.notify_all();
  return CFD_OK;
}

int cfd_server_usage(const cfd_server* s, size_t* used, size_t* budget) {
  std::lock_guard<std::mutex> lk(const_cast<cfd_server*>(s)->am);
  if (used) *used = s->used;
  if (budget) *budget = s->budget;
  return CFD_OK;
}

size_t cfd_grid_footprint(const cfd_grid* g) { return g->footprint; }

int cfd_server_submit(cfd_server* s, cfd_grid* g, const double* d_b, const double* d_scale, double tol, int64_t max_iterations,
                      int precision, double* d_x, uint64_t* ticket) {
  if (!(tol > 0.0) || max_iterations <= 0 || (precision != CFD_PRECISION_FP64 && precision != CFD_PRECISION_MIXED)) {
    set_error("cfd_server_submit: tolerance, iteration limit or precision out of range");
    return CFD_EINVAL;
  }
  if (precision == CFD_PRECISION_MIXED && g->has_hier &&
      (g->opt.cycle == CFD_CYCLE_REFERENCE || g->opt.cycle == CFD_CYCLE_ZEBRA)) {
    set_error("cfd_server_submit: the reference and zebra cycles run in FP64 only");
    return CFD_EINVAL;
  }
  std::unique_lock<std::mutex> lk(s->m);
  const uint64_t id = s->next++;
  cfd_server::Req& R = s->reqs[id];
  R.g = g;
  R.b = d_b;
  R.scale = d_scale;
  R.x = d_x;
  R.tol = tol;
  R.max = max_iterations;
  R.precision = precision;
  R.t_submit = now_s();
  s->queued.push_back(id);
  lk.unlock();
  if (ticket) *ticket = id;
  if (server_progress(s) < 0) return CFD_EDEVICE;
  return CFD_OK;
}

int cfd_server_progress(cfd_server* s) { return server_progress(s); }

int cfd_server_wait(cfd_server* s, uint64_t ticket) {
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) {
    set_error("cfd_server_wait: unknown ticket");
    return CFD_EINVAL;
  }
  // collected: done, and (batched) handed back by its lane
  auto collected = [&] {
    std::lock_guard<std::mutex> lk(s->m);
    return it->second.state.load(std::memory_order_acquire) == 2 &&
           std::find(s->batched.begin(), s->batched.end(), ticket) == s->batched.end();
  };
  while (!collected()) {
    if (server_progress(s) < 0) return CFD_EDEVICE;
    std::this_thread::yield();
  }
  return CFD_OK;
}

int cfd_server_finished(cfd_server* s, uint64_t ticket, int* done) {
  if (!s || !done) {
    set_error("cfd_server_finished: null argument");
    return CFD_EINVAL;
  }
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) {
    set_error("cfd_server_finished: unknown ticket");
    return CFD_EINVAL;
  }
  std::lock_guard<std::mutex> lk(s->m);
  *done = it->second.state.load(std::memory_order_acquire) == 2 &&
          std::find(s->batched.begin(), s->batched.end(), ticket) == s->batched.end();
  return CFD_OK;
}

int cfd_server_drain(cfd_server* s) {
  for (;;) {
    const int n = server_progress(s);
    if (n < 0) return CFD_EDEVICE;
    if (n == 0) break;
    std::this_thread::yield();
  }
  return CFD_OK;
}

int cfd_server_result(cfd_server* s, uint64_t ticket, cfd_linear_info* info, double* latency_s, double* queued_s) {
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) {
    set_error("cfd_server_result: unknown ticket");
    return CFD_EINVAL;
  }
  cfd_server::Req& R = it->second;
  {
    std::lock_guard<std::mutex> lk(s->m);
    if (R.state.load(std::memory_order_acquire) != 2 ||
        std::find(s->batched.begin(), s->batched.end(), ticket) != s->batched.end()) {
      set_error("cfd_server_result: the request has not finished");
      return CFD_EINVAL;
    }
  }
  if (info) *info = R.info;
  if (latency_s) *latency_s = R.t_end - R.t_submit;
  if (queued_s) *queued_s = R.t_start - R.t_submit;
  const int rc = R.rc;
  if (rc) set_error(R.error);
  s->reqs.erase(it);
  return rc;
}

}  // extern "C"
