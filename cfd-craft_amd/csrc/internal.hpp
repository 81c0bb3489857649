// Internal host-side structures of the cfdgpu library.
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "cfdgpu.h"

namespace cfdgpu {

// Coupling-code bits (one byte per cell, the only per-cell topology on the device).
constexpr uint8_t kFluid = 1u << 6;
constexpr uint8_t kClosed = 1u << 7;
inline uint8_t side_bit(int axis, int side) { return (uint8_t)(1u << (2 * axis + side)); }

void set_error(const std::string& text);

// Runs f(begin, end) over [0, n) split among the host's threads (at most 16).
template <class F>
void parallel_ranges(int64_t n, F f) {
  const int64_t threads = std::max<int64_t>(1, std::min<int64_t>({16, (int64_t)std::thread::hardware_concurrency(), n}));
  if (threads == 1) {
    f((int64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (int64_t t = 0; t < threads; ++t)
    pool.emplace_back([&, t] { f(n * t / threads, n * (t + 1) / threads); });
  for (auto& th : pool) th.join();
}

// Runs each callable on a thread of its own (the last on the calling thread)
// and waits for all: large zero-filled allocations fault their pages in
// parallel instead of one after another.
template <class F, class... R>
void concurrently(F&& f, R&&... rest) {
  std::vector<std::thread> pool;
  (pool.emplace_back(std::forward<R>(rest)), ...);
  f();
  for (auto& th : pool) th.join();
}

// Host description of a uniform staggered grid as the pressure solve sees it:
// the restatement of Grid::build (grid.rs:246-464) in per-cell form.
struct HostGrid {
  int64_t n[3];
  int64_t count;
  double dx[3];
  double volume;
  int boundary[3];
  bool periodic[3];
  int preconditioner;
  // Interior and outlet face weights area / (distance * volume) per axis (pressure.rs:257).
  double w_int[3], w_out[3];
  std::vector<uint8_t> fluid;
  std::vector<uint8_t> codes;
  int64_t fluid_count = 0;
  int64_t face_count[3];
  // Fluid components in the reference's numbering (BFS from the lowest index, grid.rs:430-462).
  std::vector<int32_t> component;  // -1 for solid
  std::vector<uint8_t> anchored;   // per component
  // Closed (non-anchored) components: index into the closed list per component, and member lists.
  std::vector<int32_t> closed_of_comp;           // -1 if anchored
  std::vector<int32_t> closed_index;             // per cell: closed index or -1
  std::vector<int64_t> closed_offsets;           // members CSR (cells sorted ascending)
  std::vector<int32_t> closed_members;
  int n_closed() const { return (int)closed_offsets.size() - 1; }
  // IC(0): diagonal of the shifted factor (pressure.rs:191-214); ic0_fill marks a
  // topology whose factor has fill coupling (a periodic axis of 3 cells).
  std::vector<double> ic0_diag;
  bool ic0_fill = false;
  // with fill coupling: per cell 6 forward and 6 backward factor entries in the
  // device sweeps' slot order (empty otherwise)
  std::vector<double> ic0_fill_coef;
};

// Builds the grid; returns a CFD error code (0 on success) and sets the error text.
int build_host_grid(const cfd_case& c, HostGrid& g);
// The same grid with rasterization, face codes and components on the device.
int build_device_grid(const cfd_case& c, HostGrid& g);
// Stages both builds share: validation, sizes and weights; the cell ranges a
// block covers (false when empty); the CFD-E028 error of block b; the fluid
// count with CFD-E030; per (axis, side, k) whether an open domain face is an
// outlet (Flow::vector, case.rs:345-367); the closed-component lists.
int grid_header(const cfd_case& c, HostGrid& g);
bool block_ranges(const cfd_block& b, const int64_t n[3], const double dx[3], int64_t r0[3], int64_t r1[3]);
int block_error(const cfd_case& c, int b);
int count_fluid(HostGrid& g);
void open_outlets(const cfd_case& c, const HostGrid& g, std::vector<uint8_t>& out);
void closed_lists(HostGrid& g);
// Components and anchoring from the codes (grid.rs:430-462).
void find_components(HostGrid& g);
// The grid of a caller's built Grid; strict: CFD_EUNSUPPORTED for weights that
// vary within an axis and face kind (else those weights become NaN).
int grid_from_desc(const cfd_grid_desc& d, HostGrid& g, bool strict);
void build_ic0_diagonal(HostGrid& g);
// The factor's dot terms of a grid with fill coupling (called by build_ic0_diagonal).
void build_ic0_fill(HostGrid& g, const std::vector<double>& raw, double maxdiag);


// Aggregation hierarchy (multigrid/masked.rs): level 0 = fluid cells in index
// order, coarser levels = connected groups of 2x2x2 bins of coordinates.
struct AggLevel {
  int64_t n = 0;
  std::vector<int64_t> ptr;  // off-diagonal entries per row (CSR)
  std::vector<int32_t> col;
  std::vector<double> val;     // -w
  std::vector<double> diag;    // anchor + sum of w in entry order
  std::vector<double> anchor;  // outlet weights
  std::vector<int32_t> coords;  // 3 per node: bin coordinates
  std::vector<int64_t> closed_ptr;  // closed components, members ascending
  std::vector<int32_t> closed_members;
  std::vector<int32_t> parent;  // node of the next level (empty at the coarsest)
  std::vector<int64_t> child_ptr;  // finer nodes of each node (levels >= 1)
  std::vector<int32_t> child;
  std::vector<int64_t> colour_ptr;  // colour-sorted layout: independent sets
  // unmasked hierarchy (multigrid.rs:291-433): interpolation rows to the next
  // level (columns ascending) and, on the next level, the weight of each child
  std::vector<int64_t> tptr;
  std::vector<int32_t> tcol;
  std::vector<double> tw;
  std::vector<double> child_w;
  // lexicographic sweeps by dependency level: forward rows grouped by the
  // longest chain of lower-numbered neighbours, backward by higher-numbered ones
  // (or, in a zebra hierarchy, by colour class: ascending, then descending)
  std::vector<int64_t> fwd_ptr, bwd_ptr;
  std::vector<int32_t> fwd_rows, bwd_rows;
  // block smoother of an anisotropic level (multigrid/line.rs): lines (or small
  // planes) of `length` nodes in `nodes` order, banded factor of half-bandwidth
  // `bandwidth` (reciprocal diagonal), the entries outside each node's block,
  // and the lines grouped by dependency level for the forward and backward sweeps
  int64_t line_length = 0, line_bandwidth = 0;
  std::vector<int> block_axes;  // axes the blocks span, ascending
  int64_t dims[3] = {0, 0, 0};  // cell dimensions (unmasked hierarchy)
  std::vector<int32_t> line_nodes, line_ocol;
  std::vector<int64_t> line_ooff;
  std::vector<double> line_oval, line_factor;
  std::vector<int64_t> line_fwd_ptr, line_bwd_ptr;
  std::vector<int32_t> line_fwd, line_bwd;
  // edge list (a, b, w) while the next level is built
  std::vector<int32_t> ea, eb;
  std::vector<double> ew;
};
struct AggHierarchy {
  std::vector<int32_t> node_of_cell;  // level-0 node of a cell, -1 for solid
  std::vector<int64_t> cell_of_node;
  int64_t level0_entries = 0;  // off-diagonal entries of level 0 (two per interior face)
  int fine_shift = 1;          // largest bin shift of the first aggregation (1: bins within 2x2x2)
  std::vector<AggLevel> levels;
  // coarsest level: per component, its members and a dense m x m block
  std::vector<int64_t> block_ptr, block_goff;
  std::vector<int32_t> block_rows;
  std::vector<double> block_G;
  // blocks too large to invert on the host: the component's matrix (without the
  // first member of a closed component), inverted on the device
  struct Deferred {
    int64_t block = 0, m = 0;
    bool anchored = true;
    std::vector<double> K;
  };
  std::vector<Deferred> deferred;
  bool permuted = false;
  bool masked = true;  // false: the unmasked trilinear hierarchy of a solid-free grid
};
// MultigridStats shape of an aggregation or unmasked hierarchy: level dimensions
// (unmasked only) and the block smoothers (multigrid.rs:610-617, line.rs:200-209).
void agg_hierarchy_shape(const AggHierarchy& h, cfd_mg_stats* s);
// FNV-1a digest of the hierarchy's host arrays (indices, values, coarsest blocks).
uint64_t agg_hierarchy_digest(const AggHierarchy& h);
// Coarsening stops when every component has at most coarse_cells nodes (the
// reference: 32); blocks above host_block_max nodes are left to the device.
// semicoarsen: bin an axis only while its coupling is within kSemiRatio of the
// strongest axis (the red-black cycle; the reference bins every axis).
// children: form each coarse node's list of finer nodes (false when a
// renumbering that forms them follows).
constexpr double kSemiRatio = 25.0;
int build_agg_hierarchy(const HostGrid& g, AggHierarchy& h, int64_t coarse_cells = 32,
                        int64_t host_block_max = INT64_MAX, bool semicoarsen = false, bool children = true);
// The hierarchy the reference builds for a grid (multigrid.rs:530-645): masked
// aggregation when the grid has solids or several components, else trilinear
// interpolation with Galerkin coarse operators (with line or small-plane block
// smoothers and semicoarsening on anisotropic grids, multigrid.rs:33-47,573-611);
// with the dependency levels of the lexicographic sweeps on levels >= 1.
// zebra: colour classes instead of dependency levels (coupled rows, and coupled
// lines or planes, differ in colour), for a multi-colour smoother that needs a
// few launches per sweep instead of one per dependency level.
int build_ref_hierarchy(const HostGrid& g, AggHierarchy& h, bool zebra = false);
// Renumbers levels >= 1 so that each colour class (an independent set) is contiguous.
void colour_agg_hierarchy(AggHierarchy& h);
void agg_hierarchy_stats(const AggHierarchy& h, std::vector<int64_t>& cells, std::vector<int64_t>& nonzeros);

}  // namespace cfdgpu
