// Host-side launchers of the kernels in kernels.hip.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include "device.hpp"

namespace cfdgpu {

// Launch probe: during an instrumented solve the host names one kernel class,
// and every launch of that class is bracketed by a pair of HIP events on its
// stream, with the launch's algorithmic bytes (each array crossing HBM once).
// Inactive (null) otherwise: the launch sites then only test a pointer.
enum ProbeKernel : int {
  PROBE_FINE_HALF = 1,        // fine red-black half-sweep (k_rb_gs, k_rb_gs2)
  PROBE_COARSE_HALF = 2,      // coarse half-sweep of one colour class (k_agg_gs)
  PROBE_OPERATOR = 3,         // p update, A p and p . Ap (k_pcg_op)
  PROBE_AXPY = 4,             // x and r update with |r|^2 (k_pcg_axpy)
  PROBE_FINE_RESTRICT = 5,    // fine residual, restriction, first coarse half-sweep (k_rb_restrict)
  PROBE_FINE_RZ = 6,          // last fine half-sweep with z and r . z (k_rb_gs_rz, k_rb_gs_rz2)
  PROBE_FINE_FIRST = 7,       // first fine half-sweep from zero (k_rb_first, k_rb_first2)
  PROBE_COARSE_RESTRICT = 8,  // coarse restriction (k_agg_restrict)
  PROBE_COARSEST = 9,         // dense coarsest solve (k_agg_direct)
  PROBE_FINE_PROLONG = 10,    // coarse correction of the red cells (k_rb_prolong)
  PROBE_IC0 = 11,             // one IC(0) application (both tiled sweeps)
  PROBE_TRUE_RESIDUAL = 12,   // recomputed residual and the iteration's decisions (k_pcg_true, k_pcg_fin)
  PROBE_KERNELS = 13
};
struct LaunchProbe {
  int kernel = 0;
  long long tag = 0;               // the PCG iteration being enqueued (0: the set-up before the first)
  std::vector<hipEvent_t> events;  // begin, end, begin, end, ...
  std::vector<double> bytes;       // per completed pair
  std::vector<long long> tags;     // per completed pair
  size_t used = 0;
};
extern thread_local LaunchProbe* t_probe;
inline void probe_begin(int kernel, hipStream_t s) {
  LaunchProbe* p = t_probe;
  if (p && p->kernel == kernel && (p->used & 1) == 0 && p->used + 2 <= p->events.size())
    (void)hipEventRecord(p->events[p->used++], s);
}
// The probe when it times `kernel` and has a pair of events left (a launch of
// one kernel then records them as its own start and end: hipExtLaunchKernelGGL
// takes them into the dispatch, so they time the kernel's execution alone).
inline LaunchProbe* probe_slot(int kernel) {
  LaunchProbe* p = t_probe;
  return p && p->kernel == kernel && (p->used & 1) == 0 && p->used + 2 <= p->events.size() ? p : nullptr;
}
inline void probe_done(LaunchProbe* p, double bytes) {
  p->used += 2;
  p->bytes.push_back(bytes);
  p->tags.push_back(p->tag);
}
#define PROBED_LAUNCH(ID, BYTES, K, G, B, SHM, S, ...)                                                           \
  do {                                                                                                           \
    if (LaunchProbe* probe_ = probe_slot(ID)) {                                                                  \
      hipExtLaunchKernelGGL(K, G, B, SHM, S, probe_->events[probe_->used], probe_->events[probe_->used + 1], 0, \
                            __VA_ARGS__);                                                                        \
      probe_done(probe_, BYTES);                                                                                 \
    } else {                                                                                                     \
      hipLaunchKernelGGL(K, G, B, SHM, S, __VA_ARGS__);                                                          \
    }                                                                                                            \
  } while (0)
// A pair of events around several launches of a class (a V-cycle's IC(0)
// tiles, the residual kernel and its decision kernel).
inline void probe_end(int kernel, hipStream_t s, double bytes) {
  LaunchProbe* p = t_probe;
  if (p && p->kernel == kernel && (p->used & 1) == 1) {
    (void)hipEventRecord(p->events[p->used++], s);
    p->bytes.push_back(bytes);
    p->tags.push_back(p->tag);
  }
}

void launch_apply(const Geo& g, const uint8_t* codes, const double* x, double* y, hipStream_t s);
// march: 0 cell by cell; 1 marching through z once the grid has an item per
// block of the fixed grid; 2 always marching
// zhalf > 0: z in the fine red-black level's colour-separated layout (cell c at
// c / 2 of its colour's half of zhalf entries; nx even, no closed component);
// z32: z as FP32 values in the natural layout (no closed component)
// Whether the PCG operator kernel marches through z on this grid (a tile of
// 64 x 16 columns over 16 layers, or over 8, 4 ... down to CFD_MARCH_MZ layers,
// for every block of the fixed reduction grid).
bool march_fits(const Geo& g);
void launch_pcg_op(const Geo& g, const uint8_t* codes, const double* z, const double* pold, double* pnew, double* ap,
                   const double* mean, const int* cidx, bool closed, PcgState* st, double* part, hipStream_t s,
                   int march = 1, long long zhalf = 0, const float* z32 = nullptr);
void launch_pcg_axpy(long long N, const double* p, const double* ap, double* x, double* r, PcgState* st, double* part,
                     hipStream_t s);
// the same with r in the colour-separated layout (nx even, N nx < 2^40)
void launch_pcg_axpy_rb(const Geo& g, const double* p, const double* ap, double* x, double* r, PcgState* st,
                        double* part, hipStream_t s);
// ... with p in FP32 (the mixed-precision solve)
void launch_pcg_axpy_rb(const Geo& g, const float* p, const double* ap, double* x, double* r, PcgState* st,
                        double* part, hipStream_t s);
// The same x/r update (r in the FP64 colour-separated arrays, p stored as TP)
// with the next V-cycle's first half-sweep from a zero iterate on the fine
// level F done on the way (T float: r rounded into F's b arrays first).
template <class TP, class T>
void launch_pcg_axpy_first(const Geo& g, const FineRB<T>& F, const TP* p, const double* ap, double* x, double* r,
                           PcgState* st, double* part, hipStream_t s);
// The operator kernel of the mixed-precision solve: z the FP32 values of the
// fine level's x (colour-separated, zhalf entries per colour), p stored in FP32
// (each value rounded once; A p and p . Ap formed from the rounded values).
void launch_pcg_op_mixed(const Geo& g, const uint8_t* codes, const float* z, long long zhalf, const float* pold,
                         float* pnew, double* ap, PcgState* st, double* part, hipStream_t s, int march = 1);
// the recomputed-residual kernel and the one-block decision kernel of an iteration
void launch_pcg_true(const Geo& g, const uint8_t* codes, const double* x, const double* bc, PcgState* st, double* part,
                     hipStream_t s);
void launch_pcg_rz(long long N, const uint8_t* codes, const double* r, const double* z, const double* mean,
                   const int* cidx, PcgState* st, double* part, hipStream_t s);
void launch_pcg_init(long long N, const uint8_t* codes, const double* b, const double* mean, const int* cidx, double* r,
                     double* bc, double* x, PcgState* st, double* part, hipStream_t s);
void launch_warm_residual(const Geo& g, const uint8_t* codes, const double* x0, const double* bc, double* x, double* r,
                          const PcgState* st, hipStream_t s);
void launch_state_init(PcgState* st, double tol, long long max_it, hipStream_t s);
// component means of src (and the CFD-E031 test when e031 is set) into mean
void launch_gauge(long long N, const uint8_t* codes, int n_comp, int n_chunks, const int* members, const int* chunk_begin,
                  const int* comp_first_chunk, const double* inv_count, double inv_count1, const double* src,
                  const double* scale, int e031, double* mean, PcgState* st, double* part, hipStream_t s);
void launch_center(long long N, const uint8_t* codes, const double* src, const double* mean, const int* cidx, double* dst,
                   hipStream_t s);
void launch_fill(long long N, double* dst, double v, hipStream_t s);
void launch_add(long long N, double* p, const double* q, hipStream_t s);
void launch_ic0(const Geo& g, const uint8_t* codes, const double* D, const double* r, double* z, const PcgState* st,
                hipStream_t s, const double* fill = nullptr);
// The same sweeps as a tiled wavefront (one launch per tile diagonal).
void launch_ic0_tiles(const Geo& g, const uint8_t* codes, const double* D, const double* r, double* z,
                      const PcgState* st, hipStream_t s, const double* fill = nullptr);
// The same sweeps in FP32 (the mixed-precision preconditioner): the factor
// entries formed in FP64 and rounded, the forward result held in y (FP32), z
// written in FP64.
void launch_ic0_tiles_mixed(const Geo& g, const uint8_t* codes, const double* D, const double* r, float* y, double* z,
                            const PcgState* st, hipStream_t s, const double* fill = nullptr);
template <class T, class TB>
void launch_agg_vcycle(const Geo& g0, const uint8_t* codes, const TB* b0, T* x0, const int* fparent, int seams,
                       int nlev, const AggDevLevel<T>* lv, const AggCoarsest& D, double* out64, const PcgState* st,
                       double omega, hipStream_t s, const FineRB<T>* rb = nullptr, int fine_rounds = 2,
                       int coarse_rounds = 2, const double* rz_r = nullptr, double* rz_part = nullptr,
                       bool zero_last = true, int resident = 0, float* out32 = nullptr, bool first_done = false);
// One colour-separated fine half-sweep (timing hook).
template <class T>
void launch_rb_gs(const Geo& g0, const FineRB<T>& rb, int colour, hipStream_t s);

// Batched iterations: one PCG iteration (red-black V(2,2) preconditioner led,
// then p, A p, the x/r update with the next cycle's first half-sweep, and the
// convergence decision) of up to
// kBatchMax independent solves, each on its own grid, issued as one launch per
// stage of the iteration; every launch gives each solve the blocks its own
// single-grid launch would have (from a multiple of 8, so XCD placement is
// unchanged), so each solve computes exactly what it computes alone. The fine
// level takes two cells per thread (nx a multiple of 4); r stays in the FP64
// colour-separated b arrays and z in the fine level's x. T double: the FP64
// solve; T float: the mixed-precision solve (FP32 cycle, p in FP32).
constexpr int kBatchMax = 8;
template <class T>
struct BatchSolve {
  Geo g;
  const uint8_t* codes;
  FineRB<T> F;
  const int* fparent;
  int nlev;                   // coarse levels, the last solved by the dense blocks
  const AggDevLevel<T>* lv;   // host array of nlev levels
  AggCoarsest D;
  PcgState* st;
  double* part;
  double *x, *bc, *ap;
  void *pold, *pnew;  // p, stored as T
  double* rres;       // the FP64 residual, colour-separated (F.b for T double)
  T cs;             // coarse correction scale
  bool zero_last;   // the restriction takes the class relaxed last as zero residual
  int fine_rounds, coarse_rounds;
  bool march;       // the operator marches through z
};
template <class T>
void launch_batch_iteration(const BatchSolve<T>* items, int n, hipStream_t s);
// r (natural layout) scattered into the fine level's colour-separated b arrays,
// and the red x set to b / a_ii (the first half-sweep's values).
template <class T>
void launch_rb_scatter(const Geo& g0, const uint8_t* codes, const double* r, const FineRB<T>& rb, const PcgState* st,
                       hipStream_t s);
// The mixed-precision first half-sweep: the FP64 residual rres (colour-separated)
// rounded into the FP32 level's b arrays and its red x set to b / a_ii.
void launch_rb_first_cvt(const Geo& g0, const FineRB<float>& rb, const double* rres, const PcgState* st,
                         hipStream_t s);
// Workload helper: dst = src + rel rms(src) xi on the fluid cells, xi a seeded
// standard normal field, rms over the fluid cells (part: RP partials).
void launch_perturb(long long N, const uint8_t* codes, const double* src, double* dst, double rel,
                    unsigned long long seed, double* part, double fluid_count, hipStream_t s);
// The PcgState of each solve copied to dst[i] (device memory or coherent pinned
// host memory).
void launch_batch_states(PcgState* const* st, int n, PcgState* dst, hipStream_t s);
// The reference's cycle: fine level from the codes (diagonal diag0 in the
// reference's summation order), coarse CSR levels in the reference's numbering.
// Block smoother of one level of the reference cycle (multigrid/line.rs): lines
// of `length` nodes, banded factor (reciprocal diagonal) with half-bandwidth
// `bandwidth`, entries leaving each node's block, and the lines grouped by
// dependency level (host offsets) for the forward and backward sweeps.
struct RefLines {
  int length = 0, bandwidth = 0;
  const int* nodes = nullptr;
  const double* factor = nullptr;
  const int* ooff = nullptr;
  const int* ocol = nullptr;
  const double* oval = nullptr;
  double* scratch = nullptr;
  const int* fwd = nullptr;
  const int* bwd = nullptr;
  std::vector<int> fwd_ptr, bwd_ptr;
};
// One direction of a CSR level's sweep inside one workgroup with the rows in
// the sweep's order and their off-diagonal entries copied in that order.
// Sub-level s (one dependency level or a part of one: at most kRefSubRows rows
// and `cap` entries) is meta[s] = {first position, end position, first entry,
// end entry}; position r of the order is rinfo[r] = {row, first entry, end
// entry (both counted from the sub-level's first), 0} with the row's diagonal
// at dsw[r]; bs holds the right-hand side in the order (written by the sweep).
struct RefSweep {
  const int4* meta = nullptr;
  const int4* rinfo = nullptr;
  const double* dsw = nullptr;
  const int* col = nullptr;
  const double* val = nullptr;
  double* bs = nullptr;
  int nsub = 0;
  int cap = 0;        // entries of the largest sub-level (the products' LDS buffer)
  bool ldsx = false;  // the level's iterate held in LDS for the sweep
};
constexpr int kRefSubRows = 1024;  // the workgroup: one thread a row of a sub-level
constexpr int kRefSubEpt = 8;      // entries a thread of a sub-level: cap <= 8 * 1024

struct RefCoarse {
  AggDevLevel<double> L;  // n, CSR, b, x, parent (aggregation) and children (restriction into this level)
  RefLines lines;         // anisotropic levels: block smoother instead of point sweeps
  double* r;              // residual scratch
  const double* child_w;  // trilinear restriction weights into this level (null: aggregation)
  const int* tptr;        // trilinear interpolation rows to the next level (null: aggregation)
  const int* tcol;
  const double* tw;
  const int* closed_ptr;  // closed components of this level (centering)
  const int* closed_members;
  int nclosed;
  // the closed members in chunks of at most 2048, never spanning two components:
  // chunk c covers closed_members[chunk_begin[c] .. chunk_begin[c + 1]) of
  // component chunk_comp[c]; component k owns chunks comp_chunk[k] .. comp_chunk[k + 1]
  const int* chunk_begin;
  const int* chunk_comp;
  const int* comp_chunk;
  int nchunks;
  double* chunk_sum;
  const int* fwd_rows;  // sweep schedules: rows of dependency level d at [ptr[d], ptr[d+1])
  const int* bwd_rows;
  std::vector<int> fwd_ptr, bwd_ptr;
  const int* fwd_dptr;  // device copies of fwd_ptr / bwd_ptr
  const int* bwd_dptr;
  RefSweep fwd_pk, bwd_pk;  // both sweeps in one workgroup (nsub == 0: not built)
};
// levels of at most this many rows sweep inside one workgroup (k_ref_gs_wg),
// with the iterate in LDS up to kRefWgLds rows
constexpr int kRefWgRows = 65536;
constexpr int kRefWgLds = 6144;
struct RefCycle {
  Geo g0;
  const uint8_t* codes;
  const double* diag0;
  double* rhs0;  // centered copy of the fine right-hand side
  double* r0;    // fine residual
  int self_mask;  // periodic axes of one cell (self faces, trilinear hierarchy)
  const int* fchild;   // children of the first coarse level as fine cells
  const int* fparent;  // fine cell -> first coarse node (aggregation)
  const int* ftptr;    // fine interpolation rows (trilinear)
  const int* ftcol;
  const double* ftw;
  RefLines fine_lines;  // block smoother of the fine level (anisotropic grids)
  std::vector<RefCoarse> lv;
  AggCoarsest D;
  std::function<void(const double*, double*)> center0;  // fine-level centering (closed components)
  // fine sweeps over tiles and small coarse levels in one workgroup per sweep
  // (false: one launch per hyperplane and per dependency level)
  bool tiles = true;
  // coarse sweeps from the sweep-ordered copies where built (k_ref_gs_pk)
  bool packed = true;
};
void launch_ref_vcycle(const RefCycle& R, const double* r, double* z, const PcgState* st, hipStream_t s);

// Device inversion of a big coarsest block (host matrix K, k = m - !anchored);
// work holds k * k + 2 k + m doubles. False if a pivot was not positive.
bool invert_block(double* Gblock, int m, bool anchored, const double* K, double* work, int* bad, hipStream_t s);
template <class T, class TB>
void launch_fine_gs(const Geo& g0, const uint8_t* codes, const TB* b0, T* x0, int color, const Geo& g1, hipStream_t s);
void launch_projection_rhs(const Geo& g, const FaceGeo& fg, const uint8_t* codes, const double* u, double dt, double* b,
                           double* terms, hipStream_t s);
void launch_project_faces(const Geo& g, const FaceGeo& fg, const uint8_t* codes, const double* phi, double dt, double* u,
                          hipStream_t s);

}  // namespace cfdgpu
