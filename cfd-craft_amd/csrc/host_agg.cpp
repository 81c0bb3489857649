// Host set-up of the aggregation multigrid hierarchy: the reference's masked
// hierarchy (multigrid/masked.rs:181-291) built from the per-cell coupling
// codes instead of a face list.
//
// Level 0 holds the fluid cells in index order. Each coarser level groups the
// nodes of a 2x2x2 bin of coordinates into aggregates, one per connected group
// inside the bin (masked.rs:127-156), numbered by their lowest member; the
// coarse couplings sum the finer ones between distinct aggregates in the finer
// level's edge order, and anchors (outlet weights) add up. Coarsening stops
// when every connected component has at most 32 nodes; those are solved
// densely (masked.rs:21-51).
//
// The device cycle uses this hierarchy in two layouts: the reference's node
// numbering (the lexicographic cycle) and a colour-sorted numbering in which
// the nodes of one independent set are contiguous (the red-black cycle).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <numeric>
#include <mutex>
#include <thread>

#include "internal.hpp"

namespace cfdgpu {
namespace {



struct UnionFind {
  std::vector<int32_t> p;
  explicit UnionFind(int64_t n) : p((size_t)n) { std::iota(p.begin(), p.end(), 0); }
  int32_t find(int32_t x) {
    while (p[(size_t)x] != x) {
      p[(size_t)x] = p[(size_t)p[(size_t)x]];
      x = p[(size_t)x];
    }
    return x;
  }
  void unite(int32_t a, int32_t b) {
    a = find(a);
    b = find(b);
    if (a == b) return;
    if (a < b)
      p[(size_t)b] = a;
    else
      p[(size_t)a] = b;
  }
};

// Visits the faces of the grid in the reference's order (x faces, then y, then
// z; each block x-fastest over its face dimensions, grid.rs:341-429): `edge(lo,
// hi, w)` for an interior face joining two distinct fluid cells, `outlet(c, w)`
// for an outlet face of cell c.
// With self_loops, the face of a periodic axis of one cell, which joins a cell
// to itself, is visited as an edge too (the unmasked hierarchy keeps it,
// multigrid.rs:568-575; it cancels in products but not in a sweep).
template <class E, class O>
void for_each_face(const HostGrid& g, E edge, O outlet, bool self_loops = false) {
  const int64_t* n = g.n;
  const int64_t st[3] = {1, n[0], n[0] * n[1]};
  for (int a = 0; a < 3; ++a) {
    int64_t fd[3] = {n[0], n[1], n[2]};
    fd[a] = g.periodic[a] ? n[a] : n[a] + 1;
    const uint8_t lo_bit = side_bit(a, 0), hi_bit = side_bit(a, 1);
    for (int64_t k = 0; k < fd[2]; ++k)
      for (int64_t j = 0; j < fd[1]; ++j)
        for (int64_t i = 0; i < fd[0]; ++i) {
          const int64_t q[3] = {i, j, k};
          const int64_t f = q[a];
          const int64_t base = i * (a == 0 ? 0 : 1) + j * st[1] * (a == 1 ? 0 : 1) + k * st[2] * (a == 2 ? 0 : 1);
          const int64_t hi = f < n[a] ? base + f * st[a] : -1;
          int64_t lo = -1;
          if (f > 0)
            lo = base + (f - 1) * st[a];
          else if (g.periodic[a])
            lo = base + (n[a] - 1) * st[a];
          if (lo >= 0 && hi >= 0) {
            if (lo != hi && (g.codes[(size_t)lo] & hi_bit)) edge(lo, hi, g.w_int[a]);
            if (lo == hi && self_loops && g.fluid[(size_t)lo]) edge(lo, hi, g.w_int[a]);
          } else if (hi >= 0) {
            if (g.codes[(size_t)hi] & lo_bit) outlet(hi, g.w_out[a]);
          } else if (lo >= 0) {
            if (g.codes[(size_t)lo] & hi_bit) outlet(lo, g.w_out[a]);
          }
        }
  }
}

// Connected components of a level by breadth-first search from the lowest
// unvisited node (masked.rs:90-126); members are stored in ascending order.
void level_components(const AggLevel& L, std::vector<std::vector<int32_t>>& comps, std::vector<uint8_t>& anchored) {
  // Components by union-find over threads: a root is only ever linked under a
  // smaller root, so each component ends rooted at its smallest node, and the
  // components numbered by that node in ascending order are the ones a
  // breadth-first search from each unvisited node in order would number.
  const int64_t n = L.n;
  std::vector<std::atomic<int32_t>> up((size_t)n);
  parallel_ranges(n, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) up[(size_t)i].store((int32_t)i, std::memory_order_relaxed);
  });
  auto find = [&](int32_t x) {
    for (;;) {
      int32_t p = up[(size_t)x].load(std::memory_order_relaxed);
      if (p == x) return x;
      const int32_t q = up[(size_t)p].load(std::memory_order_relaxed);
      if (q != p) up[(size_t)x].compare_exchange_weak(p, q, std::memory_order_relaxed);
      x = q;
    }
  };
  parallel_ranges(n, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i)
      for (int64_t e = L.ptr[(size_t)i]; e < L.ptr[(size_t)i + 1]; ++e) {
        int32_t a = (int32_t)i, b = L.col[(size_t)e];
        if (b <= a) continue;  // each coupling once
        for (;;) {
          a = find(a);
          b = find(b);
          if (a == b) break;
          if (a > b) std::swap(a, b);
          int32_t expect = b;
          if (up[(size_t)b].compare_exchange_strong(expect, a, std::memory_order_relaxed)) break;
        }
      }
  });
  std::vector<int32_t> comp((size_t)n);
  parallel_ranges(n, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) comp[(size_t)i] = find((int32_t)i);
  });
  // number the roots in ascending order; members ascending
  std::vector<int32_t> id((size_t)n, -1);
  std::vector<int64_t> size;
  for (int64_t i = 0; i < n; ++i)
    if (comp[(size_t)i] == (int32_t)i) {
      id[(size_t)i] = (int32_t)size.size();
      size.push_back(0);
      anchored.push_back(0);
    }
  for (int64_t i = 0; i < n; ++i) {
    const int32_t r = id[(size_t)comp[(size_t)i]];
    comp[(size_t)i] = r;
    size[(size_t)r] += 1;
    if (L.anchor[(size_t)i] > 0.0) anchored[(size_t)r] = 1;
  }
  comps.resize(size.size());
  for (size_t r = 0; r < size.size(); ++r) comps[r].reserve((size_t)size[r]);
  for (int64_t i = 0; i < n; ++i) comps[(size_t)comp[(size_t)i]].push_back((int32_t)i);
}

// CSR of a level from its edge list (a < b, sorted by (a, b)) and anchors, in
// the entry order Matrix::build gives (multigrid.rs:225-262): row r lists the
// edges (x, r) by x, then (r, y) by y, so columns ascend; the diagonal is the
// anchor plus the weights in that order.
void csr_from_edges(AggLevel& L) {
  const int64_t n = L.n;
  L.ptr.assign((size_t)n + 1, 0);
  for (size_t e = 0; e < L.ea.size(); ++e) {
    L.ptr[(size_t)L.ea[e] + 1] += 1;
    L.ptr[(size_t)L.eb[e] + 1] += 1;
  }
  for (int64_t i = 0; i < n; ++i) L.ptr[(size_t)i + 1] += L.ptr[(size_t)i];
  std::vector<int64_t> next;
  concurrently([&] { L.col.assign((size_t)L.ptr[(size_t)n], 0); }, [&] { L.val.assign((size_t)L.ptr[(size_t)n], 0.0); },
               [&] { next.assign(L.ptr.begin(), L.ptr.end() - 1); }, [&] { L.diag = L.anchor; });
  for (size_t e = 0; e < L.ea.size(); ++e) {
    const int32_t a = L.ea[e], b = L.eb[e];
    const double w = L.ew[e];
    L.diag[(size_t)a] += w;
    L.col[(size_t)next[(size_t)a]] = b;
    L.val[(size_t)next[(size_t)a]++] = -w;
    L.diag[(size_t)b] += w;
    L.col[(size_t)next[(size_t)b]] = a;
    L.val[(size_t)next[(size_t)b]++] = -w;
  }
}

// CSR of level 0 with its entries in face order (each face adds an entry to
// both of its rows, Matrix::build); the diagonal is already set.
void csr_level0(AggLevel& L) {
  const int64_t n = L.n;
  L.ptr.assign((size_t)n + 1, 0);
  for (size_t e = 0; e < L.ea.size(); ++e) {
    L.ptr[(size_t)L.ea[e] + 1] += 1;
    L.ptr[(size_t)L.eb[e] + 1] += 1;
  }
  for (int64_t i = 0; i < n; ++i) L.ptr[(size_t)i + 1] += L.ptr[(size_t)i];
  L.col.assign((size_t)L.ptr[(size_t)n], 0);
  L.val.assign((size_t)L.ptr[(size_t)n], 0.0);
  std::vector<int64_t> next(L.ptr.begin(), L.ptr.end() - 1);
  for (size_t e = 0; e < L.ea.size(); ++e) {
    const int32_t rows[2] = {L.ea[e], L.eb[e]};
    for (int s = 0; s < 2; ++s) {
      const int32_t r = rows[s];
      L.col[(size_t)next[(size_t)r]] = rows[1 - s];
      L.val[(size_t)next[(size_t)r]++] = -L.ew[e];
    }
  }
}

// Children lists: the finer nodes of each aggregate, ascending.
void build_children(AggHierarchy& h) {
  // children of each node ascending: counted and placed over threads (atomic
  // counters), then each node's few children sorted
  for (size_t l = 1; l < h.levels.size(); ++l) {
    AggLevel& C = h.levels[l];
    const AggLevel& F = h.levels[l - 1];
    std::vector<std::atomic<int32_t>> cnt((size_t)C.n);
    parallel_ranges(C.n, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) cnt[(size_t)i].store(0, std::memory_order_relaxed);
    });
    parallel_ranges(F.n, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) cnt[(size_t)F.parent[(size_t)i]].fetch_add(1, std::memory_order_relaxed);
    });
    C.child_ptr.assign((size_t)C.n + 1, 0);
    for (int64_t i = 0; i < C.n; ++i)
      C.child_ptr[(size_t)i + 1] = C.child_ptr[(size_t)i] + cnt[(size_t)i].load(std::memory_order_relaxed);
    C.child.resize((size_t)F.n);
    parallel_ranges(C.n, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) cnt[(size_t)i].store((int32_t)C.child_ptr[(size_t)i], std::memory_order_relaxed);
    });
    parallel_ranges(F.n, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i)
        C.child[(size_t)cnt[(size_t)F.parent[(size_t)i]].fetch_add(1, std::memory_order_relaxed)] = (int32_t)i;
    });
    parallel_ranges(C.n, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i)
        std::sort(C.child.begin() + C.child_ptr[(size_t)i], C.child.begin() + C.child_ptr[(size_t)i + 1]);
    });
  }
}

// Coarse pair sums of the trilinear Galerkin product, keyed by (min, max) and
// accumulated in arrival order, emitted by row with columns ascending.
struct EdgeSums {
  std::vector<std::vector<std::pair<int32_t, double>>> rows;
  explicit EdgeSums(int64_t n) : rows((size_t)n) {}
  void emit(AggLevel& C) {
    for (size_t a = 0; a < rows.size(); ++a) {
      auto& r = rows[a];
      std::sort(r.begin(), r.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      for (auto& e : r) {
        C.ea.push_back((int32_t)a);
        C.eb.push_back(e.first);
        C.ew.push_back(e.second);
      }
      std::vector<std::pair<int32_t, double>>().swap(r);
    }
  }
};

// Sums fine couplings between distinct aggregates into the coarse edge list,
// keyed by (min, max) and accumulated in the order the fine edges come; the
// coarse edges are emitted by row, columns ascending. `edges(f)` calls
// f(a, b, w) for every fine edge in order and is run twice: once to size each
// coarse row's slots (its count of contributions bounds its distinct columns),
// once to add each contribution to the entry of its column in the row, or to
// the row's next slot for a new column.
template <class Edges>
void coarse_edges(int64_t m, const std::vector<int32_t>& parent, Edges edges, AggLevel& C) {
  std::vector<int64_t> start((size_t)m + 1, 0);
  edges([&](int32_t a, int32_t b, double) {
    const int32_t i = parent[(size_t)a], j = parent[(size_t)b];
    if (i != j) start[(size_t)std::min(i, j) + 1] += 1;
  });
  for (int64_t r = 0; r < m; ++r) start[(size_t)r + 1] += start[(size_t)r];
  std::vector<int32_t> col((size_t)start[(size_t)m]);
  std::vector<double> val((size_t)start[(size_t)m]);
  std::vector<int64_t> end(start.begin(), start.end() - 1);
  edges([&](int32_t a, int32_t b, double w) {
    const int32_t i = parent[(size_t)a], j = parent[(size_t)b];
    if (i == j) return;
    const int32_t r = std::min(i, j), c = std::max(i, j);
    int64_t e = start[(size_t)r];
    const int64_t stop = end[(size_t)r];
    while (e < stop && col[(size_t)e] != c) ++e;
    if (e < stop) {
      val[(size_t)e] += w;
    } else {
      col[(size_t)e] = c;
      val[(size_t)e] = w;
      end[(size_t)r] = e + 1;
    }
  });
  int64_t total = 0;
  for (int64_t r = 0; r < m; ++r) total += end[(size_t)r] - start[(size_t)r];
  C.ea.reserve((size_t)total);
  C.eb.reserve((size_t)total);
  C.ew.reserve((size_t)total);
  for (int64_t r = 0; r < m; ++r) {
    // the row's distinct columns in ascending order (a row holds a handful)
    const int64_t b0 = start[(size_t)r], b1 = end[(size_t)r];
    for (int64_t x = b0 + 1; x < b1; ++x)
      for (int64_t y = x; y > b0 && col[(size_t)y - 1] > col[(size_t)y]; --y) {
        std::swap(col[(size_t)y - 1], col[(size_t)y]);
        std::swap(val[(size_t)y - 1], val[(size_t)y]);
      }
    for (int64_t e = b0; e < b1; ++e) {
      C.ea.push_back((int32_t)r);
      C.eb.push_back(col[(size_t)e]);
      C.ew.push_back(val[(size_t)e]);
    }
  }
}

// Groups the nodes of level L into aggregates: bins of coordinates >> shift[d]
// per axis, split into connected groups, numbered by their lowest member.
// `edges(f)` calls f(a, b) for every edge of the level (any order: only the
// partition matters).
template <class Edges>
int64_t aggregate(int64_t n, const std::vector<int32_t>& coords, const int shift[3], Edges edges,
                  std::vector<int32_t>& parent, std::vector<int32_t>& coarse_coords) {
  UnionFind uf(n);
  auto same_bin = [&](int32_t i, int32_t j) {
    for (int d = 0; d < 3; ++d)
      if ((coords[3 * (size_t)i + d] >> shift[d]) != (coords[3 * (size_t)j + d] >> shift[d])) return false;
    return true;
  };
  edges([&](int32_t i, int32_t j) {
    if (same_bin(i, j)) uf.unite(i, j);
  });
  parent.assign((size_t)n, -1);
  std::vector<int32_t> id_of_root((size_t)n, -1);
  coarse_coords.clear();
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t r = uf.find((int32_t)i);
    if (id_of_root[(size_t)r] < 0) {
      id_of_root[(size_t)r] = (int32_t)m++;
      for (int d = 0; d < 3; ++d) coarse_coords.push_back(coords[3 * (size_t)i + d] >> shift[d]);
    }
    parent[(size_t)i] = id_of_root[(size_t)r];
  }
  return m;
}

// The component's matrix without its first member when it is closed (the
// gauge that Direct::build fixes, masked.rs:21-51): k x k, row-major.
std::vector<double> assemble_block(const AggLevel& L, const std::vector<int32_t>& members, bool anchored) {
  const int64_t m = (int64_t)members.size();
  const int64_t skip = anchored ? 0 : 1;
  const int64_t k = m - skip;
  std::vector<double> K((size_t)(k * k), 0.0);
  for (int64_t r = 0; r < k; ++r) {
    const int32_t row = members[(size_t)(r + skip)];
    K[(size_t)(r * k + r)] = L.diag[(size_t)row];
    for (int64_t e = L.ptr[(size_t)row]; e < L.ptr[(size_t)row + 1]; ++e) {
      auto it = std::lower_bound(members.begin() + skip, members.end(), L.col[(size_t)e]);
      if (it != members.end() && *it == L.col[(size_t)e])
        K[(size_t)(r * k + (it - members.begin() - skip))] += L.val[(size_t)e];
    }
  }
  return K;
}

// Dense block of the coarsest level for one component: the inverse of the
// component's matrix (anchored) or, for a closed component, the map
// rhs -> centered solution with the first member held at zero (masked.rs:21-51,
// 82-89, 330-336): G = C E K^-1 E^T C.
int coarsest_block(const AggLevel& L, const std::vector<int32_t>& members, bool anchored, std::vector<double>& G) {
  const int64_t m = (int64_t)members.size();
  const int64_t skip = anchored ? 0 : 1;
  const int64_t k = m - skip;
  std::vector<double> K = assemble_block(L, members, anchored);
  // Cholesky
  for (int64_t i = 0; i < k; ++i)
    for (int64_t j = 0; j <= i; ++j) {
      double v = K[(size_t)(i * k + j)];
      for (int64_t t = 0; t < j; ++t) v -= K[(size_t)(i * k + t)] * K[(size_t)(j * k + t)];
      if (i == j) {
        if (!std::isfinite(v) || v <= 0.0) {
          set_error("CFD-E035: numerics.pressure_preconditioner: masked coarse Cholesky lost positive definiteness");
          return CFD_E035;
        }
        K[(size_t)(i * k + i)] = std::sqrt(v);
      } else {
        K[(size_t)(i * k + j)] = v / K[(size_t)(j * k + j)];
      }
    }
  G.assign((size_t)(m * m), 0.0);
  std::vector<double> e((size_t)k);
  for (int64_t c = 0; c < k; ++c) {
    std::fill(e.begin(), e.end(), 0.0);
    e[(size_t)c] = 1.0;
    for (int64_t i = 0; i < k; ++i) {
      double v = e[(size_t)i];
      for (int64_t t = 0; t < i; ++t) v -= K[(size_t)(i * k + t)] * e[(size_t)t];
      e[(size_t)i] = v / K[(size_t)(i * k + i)];
    }
    for (int64_t i = k - 1; i >= 0; --i) {
      double v = e[(size_t)i];
      for (int64_t t = i + 1; t < k; ++t) v -= K[(size_t)(t * k + i)] * e[(size_t)t];
      e[(size_t)i] = v / K[(size_t)(i * k + i)];
    }
    for (int64_t i = 0; i < k; ++i) G[(size_t)((i + skip) * m + c + skip)] = e[(size_t)i];
  }
  if (!anchored) {
    const double inv = 1.0 / (double)m;
    for (int pass = 0; pass < 2; ++pass)
      for (int64_t o = 0; o < m; ++o) {
        double mean = 0.0;
        for (int64_t i = 0; i < m; ++i) mean += pass == 0 ? G[(size_t)(i * m + o)] : G[(size_t)(o * m + i)];
        mean *= inv;
        for (int64_t i = 0; i < m; ++i) (pass == 0 ? G[(size_t)(i * m + o)] : G[(size_t)(o * m + i)]) -= mean;
      }
  }
  return 0;
}

// The first aggregation of level 0 with bins of 2x2x2 cells, straight from the
// codes, and the coarse edge list it gives: the same partition, numbering and
// sums as aggregate() and coarse_edges() over level 0's face-ordered edge list.
// Within a bin the fluid cells, in node order, are joined along their coupled
// faces; an aggregate is numbered by its lowest member. Two aggregates in
// different bins meet only on the interface of their two bins, whose faces all
// belong to one axis, so the sum of each coarse edge is formed over that
// interface's faces in face order. Returns false (nothing built) where a
// periodic axis has exactly two bins (a bin pair would meet twice) or where the
// bins reduce nothing.
bool bins_level0(const HostGrid& g, const std::vector<int32_t>& node, const AggLevel& L0, std::vector<int32_t>& parent,
                 std::vector<int32_t>& coarse_coords, AggLevel& C) {
  const int64_t* n = g.n;
  const int64_t st[3] = {1, n[0], n[0] * n[1]};
  int64_t nb[3];
  for (int a = 0; a < 3; ++a) {
    nb[a] = (n[a] + 1) / 2;
    if (g.periodic[a] && nb[a] == 2) return false;
  }
  // lowest member of each node's aggregate (bins of one layer per task: they
  // hold disjoint nodes)
  std::vector<int32_t> low;
  concurrently([&] { low.resize((size_t)L0.n); }, [&] { parent.assign((size_t)L0.n, -1); });
  parallel_ranges(nb[2], [&](int64_t z0, int64_t z1) {
  for (int64_t bz = z0; bz < z1; ++bz)
    for (int64_t by = 0; by < nb[1]; ++by)
      for (int64_t bx = 0; bx < nb[0]; ++bx) {
        int loc[2][2][2];
        int64_t cell[8];
        int up[8];
        int cnt = 0;
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
              loc[dz][dy][dx] = -1;
              const int64_t x = 2 * bx + dx, y = 2 * by + dy, z = 2 * bz + dz;
              if (x >= n[0] || y >= n[1] || z >= n[2]) continue;
              const int64_t c = x + y * st[1] + z * st[2];
              if (!g.fluid[(size_t)c]) continue;
              loc[dz][dy][dx] = cnt;
              cell[cnt] = c;
              up[cnt] = cnt;
              ++cnt;
            }
        if (cnt == 0) continue;
        auto find = [&](int u) {
          while (up[u] != u) u = up[u];
          return u;
        };
        auto unite = [&](int a, int b) {
          a = find(a);
          b = find(b);
          if (a < b)
            up[b] = a;
          else if (b < a)
            up[a] = b;
        };
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
              const int u = loc[dz][dy][dx];
              if (u < 0) continue;
              const uint8_t code = g.codes[(size_t)cell[u]];
              if (dx == 0 && loc[dz][dy][1] >= 0 && (code & side_bit(0, 1))) unite(u, loc[dz][dy][1]);
              if (dy == 0 && loc[dz][1][dx] >= 0 && (code & side_bit(1, 1))) unite(u, loc[dz][1][dx]);
              if (dz == 0 && loc[1][dy][dx] >= 0 && (code & side_bit(2, 1))) unite(u, loc[1][dy][dx]);
            }
        // the root of a group is its smallest local index, the lowest node
        for (int u = 0; u < cnt; ++u) low[(size_t)node[(size_t)cell[u]]] = node[(size_t)cell[find(u)]];
      }
  });
  // aggregates numbered by their lowest node: the roots counted per chunk,
  // numbered from the chunk's offset, then every other node takes its root's
  constexpr int64_t kChunks = 64;
  std::vector<int64_t> roots(kChunks + 1, 0);
  parallel_ranges(kChunks, [&](int64_t t0, int64_t t1) {
    for (int64_t t = t0; t < t1; ++t) {
      int64_t r = 0;
      for (int64_t i = L0.n * t / kChunks; i < L0.n * (t + 1) / kChunks; ++i) r += low[(size_t)i] == (int32_t)i;
      roots[(size_t)t + 1] = r;
    }
  });
  for (int64_t t = 0; t < kChunks; ++t) roots[(size_t)t + 1] += roots[(size_t)t];
  const int64_t m = roots[kChunks];
  if (m == L0.n) return false;
  coarse_coords.assign(3 * (size_t)m, 0);
  parallel_ranges(kChunks, [&](int64_t t0, int64_t t1) {
    for (int64_t t = t0; t < t1; ++t) {
      int64_t next = roots[(size_t)t];
      for (int64_t i = L0.n * t / kChunks; i < L0.n * (t + 1) / kChunks; ++i)
        if (low[(size_t)i] == (int32_t)i) {
          parent[(size_t)i] = (int32_t)next;
          for (int d = 0; d < 3; ++d) coarse_coords[3 * (size_t)next + d] = L0.coords[3 * (size_t)i + d] >> 1;
          ++next;
        }
    }
  });
  parallel_ranges(L0.n, [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i)
      if (low[(size_t)i] != (int32_t)i) parent[(size_t)i] = parent[(size_t)low[(size_t)i]];
  });
  // coarse edges: each interface between a bin and its upper neighbour along an
  // axis (wrapping on a periodic axis of three or more bins), its faces in face
  // order (the two other axes, the higher one outer)
  struct Entry {
    int32_t r, c;
    double w;
  };
  // (layers of bins per task, each with its own list: every pair is formed on
  // one interface, so the lists only need joining and sorting)
  std::vector<std::vector<Entry>> lists(16);
  std::vector<int64_t> count((size_t)m + 1, 0);
  std::mutex join;
  size_t next_list = 0;
  parallel_ranges(nb[2], [&](int64_t z0, int64_t z1) {
  std::vector<Entry> list;
  for (int a = 0; a < 3; ++a) {
    if (nb[a] == 1) continue;
    const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
    const double w = g.w_int[a];
    for (int64_t bz = z0; bz < z1; ++bz)
      for (int64_t by = 0; by < nb[1]; ++by)
        for (int64_t bx = 0; bx < nb[0]; ++bx) {
          const int64_t B[3] = {bx, by, bz};
          int64_t next = B[a] + 1;
          if (next == nb[a]) {
            if (!g.periodic[a]) continue;
            next = 0;
          }
          const int64_t t_lo = std::min(2 * B[a] + 1, n[a] - 1), t_hi = 2 * next;
          Entry e[4];
          int ne = 0;
          for (int s2 = 0; s2 < 2; ++s2) {
            const int64_t q2 = 2 * B[b2] + s2;
            if (q2 >= n[b2]) continue;
            for (int s1 = 0; s1 < 2; ++s1) {
              const int64_t q1 = 2 * B[b1] + s1;
              if (q1 >= n[b1]) continue;
              const int64_t lo = t_lo * st[a] + q1 * st[b1] + q2 * st[b2];
              const int64_t hi = t_hi * st[a] + q1 * st[b1] + q2 * st[b2];
              if (!g.fluid[(size_t)lo] || !(g.codes[(size_t)lo] & side_bit(a, 1))) continue;
              const int32_t P = parent[(size_t)node[(size_t)lo]], Q = parent[(size_t)node[(size_t)hi]];
              const int32_t r = std::min(P, Q), c = std::max(P, Q);
              int k = 0;
              while (k < ne && !(e[k].r == r && e[k].c == c)) ++k;
              if (k < ne) {
                e[k].w += w;
              } else {
                e[ne++] = Entry{r, c, w};
              }
            }
          }
          for (int k = 0; k < ne; ++k) list.push_back(e[k]);
        }
  }
  std::lock_guard<std::mutex> lock(join);
  lists[next_list++] = std::move(list);
  });
  // by row, columns ascending (each pair comes from one interface, so the
  // order in which a row's entries are placed does not matter once the row is
  // sorted): counted and placed over the lists' threads with atomic counters
  size_t total = 0;
  for (const auto& l : lists) total += l.size();
  std::vector<std::atomic<int64_t>> at((size_t)m);
  parallel_ranges(m, [&](int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; ++r) at[(size_t)r].store(0, std::memory_order_relaxed);
  });
  parallel_ranges((int64_t)lists.size(), [&](int64_t l0, int64_t l1) {
    for (int64_t l = l0; l < l1; ++l)
      for (const Entry& e : lists[(size_t)l]) at[(size_t)e.r].fetch_add(1, std::memory_order_relaxed);
  });
  for (int64_t r = 0; r < m; ++r) count[(size_t)r + 1] = count[(size_t)r] + at[(size_t)r].load(std::memory_order_relaxed);
  parallel_ranges(m, [&](int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; ++r) at[(size_t)r].store(count[(size_t)r], std::memory_order_relaxed);
  });
  // every entry placed below: no zero fill
  std::unique_ptr<Entry[]> sorted_mem(new Entry[total ? total : 1]);
  Entry* sorted_p = sorted_mem.get();
  parallel_ranges((int64_t)lists.size(), [&](int64_t l0, int64_t l1) {
    for (int64_t l = l0; l < l1; ++l) {
      for (const Entry& e : lists[(size_t)l]) sorted_p[at[(size_t)e.r].fetch_add(1, std::memory_order_relaxed)] = e;
      std::vector<Entry>().swap(lists[(size_t)l]);
    }
  });
  struct View {
    Entry* p;
    size_t n;
    size_t size() const { return n; }
    Entry& operator[](size_t i) const { return p[i]; }
  } sorted{sorted_p, total};
  concurrently([&] { C.ea.resize(sorted.size()); }, [&] { C.eb.resize(sorted.size()); },
               [&] { C.ew.resize(sorted.size()); });
  // rows are independent: sorted and copied over threads
  parallel_ranges(m, [&](int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; ++r) {
      const int64_t b0 = count[(size_t)r], b1e = count[(size_t)r + 1];
      for (int64_t x = b0 + 1; x < b1e; ++x)
        for (int64_t y = x; y > b0 && sorted[(size_t)y - 1].c > sorted[(size_t)y].c; --y)
          std::swap(sorted[(size_t)y - 1], sorted[(size_t)y]);
      for (int64_t e = b0; e < b1e; ++e) {
        C.ea[(size_t)e] = sorted[(size_t)e].r;
        C.eb[(size_t)e] = sorted[(size_t)e].c;
        C.ew[(size_t)e] = sorted[(size_t)e].w;
      }
    }
  });
  return true;
}

}  // namespace

namespace {
// Stage times of the hierarchy build on stderr (CFD_SETUP_TIMES=1).
struct AggClock {
  bool on = std::getenv("CFD_SETUP_TIMES") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void mark(const char* what, int64_t level = -1) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[setup]   %-24s %3lld %8.1f ms\n", what, (long long)level,
                 1e3 * std::chrono::duration<double>(now - t).count());
    t = now;
  }
};
}  // namespace

int build_agg_hierarchy(const HostGrid& g, AggHierarchy& h, int64_t coarse_cells, int64_t host_block_max,
                        bool semicoarsen, bool children) {
  AggClock clk;
  h = AggHierarchy();
  if (g.fluid_count == 0) {
    set_error("CFD-E034: numerics.pressure_preconditioner: multigrid requires fluid cells");
    return CFD_E034;
  }
  const int64_t* n = g.n;
  // Level 0: fluid cells in index order (numbered from per-layer counts, the
  // layers filled in parallel).
  AggLevel L0;
  const int64_t layer = n[0] * n[1];
  std::vector<int64_t> first((size_t)n[2] + 1, 0);
  concurrently(
      [&] {
        parallel_ranges(n[2], [&](int64_t k0, int64_t k1) {
          for (int64_t k = k0; k < k1; ++k) {
            int64_t m = 0;
            for (int64_t c = k * layer; c < (k + 1) * layer; ++c) m += g.fluid[(size_t)c] ? 1 : 0;
            first[(size_t)k + 1] = m;
          }
        });
      },
      [&] { h.node_of_cell.assign((size_t)g.count, -1); });
  for (int64_t k = 0; k < n[2]; ++k) first[(size_t)k + 1] += first[(size_t)k];
  L0.n = first[(size_t)n[2]];
  concurrently([&] { h.cell_of_node.resize((size_t)L0.n); }, [&] { L0.coords.resize(3 * (size_t)L0.n); },
               [&] { L0.anchor.assign((size_t)L0.n, 0.0); }, [&] { L0.diag.assign((size_t)L0.n, 0.0); });
  // Per node, in face order (x faces, then y, then z; every face of an axis
  // carries that axis's weight): the anchor sums its outlet sides (a coupled
  // side on a non-periodic domain face), the diagonal is the anchor plus the
  // weights of its interior sides; each interior face has two such sides.
  std::vector<int64_t> sides_of((size_t)n[2], 0);
  parallel_ranges(n[2], [&](int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; ++k) {
      int64_t v = first[(size_t)k], sides = 0;
      for (int64_t j = 0, c = k * layer; j < n[1]; ++j)
        for (int64_t i = 0; i < n[0]; ++i, ++c) {
          if (!g.fluid[(size_t)c]) continue;
          h.node_of_cell[(size_t)c] = (int32_t)v;
          h.cell_of_node[(size_t)v] = c;
          L0.coords[3 * (size_t)v] = (int32_t)i;
          L0.coords[3 * (size_t)v + 1] = (int32_t)j;
          L0.coords[3 * (size_t)v + 2] = (int32_t)k;
          const uint8_t code = g.codes[(size_t)c];
          const int64_t q[3] = {i, j, k};
          int inner[3] = {0, 0, 0};
          double anchor = 0.0;
          for (int a = 0; a < 3; ++a)
            for (int side = 0; side < 2; ++side) {
              if (!(code & side_bit(a, side))) continue;
              if (g.periodic[a] || (side == 0 ? q[a] > 0 : q[a] + 1 < n[a]))
                ++inner[a];
              else
                anchor += g.w_out[a];
            }
          double d = anchor;
          for (int a = 0; a < 3; ++a)
            for (int t = 0; t < inner[a]; ++t) d += g.w_int[a];
          L0.anchor[(size_t)v] = anchor;
          L0.diag[(size_t)v] = d;
          sides += inner[0] + inner[1] + inner[2];
          ++v;
        }
      sides_of[(size_t)k] = sides;
    }
  });
  const auto& node = h.node_of_cell;
  std::vector<int64_t> comp_size(g.anchored.size(), 0);
  for (int64_t c = 0; c < g.count; ++c)
    if (g.component[(size_t)c] >= 0) comp_size[(size_t)g.component[(size_t)c]] += 1;
  int64_t sides0 = 0;
  for (int64_t k = 0; k < n[2]; ++k) sides0 += sides_of[(size_t)k];
  const int64_t faces0 = sides0 / 2;
  h.level0_entries = 2 * faces0;
  h.levels.push_back(std::move(L0));
  clk.mark("level 0 nodes, diagonal");
  // level 0's edge list in face order, formed only where it is read: the CSR of
  // a coarsest level 0 and the general aggregation (bins other than 2x2x2)
  const auto list_level0 = [&](AggLevel& L) {
    if (!L.ea.empty() || faces0 == 0) return;
    L.ea.resize((size_t)faces0);
    L.eb.resize((size_t)faces0);
    L.ew.resize((size_t)faces0);
    size_t e = 0;
    for_each_face(
        g,
        [&](int64_t lo, int64_t hi, double w) {
          L.ea[e] = node[(size_t)lo];
          L.eb[e] = node[(size_t)hi];
          L.ew[e++] = w;
        },
        [](int64_t, double) {});
  };
  const bool general = std::getenv("CFD_AGG_GENERAL") != nullptr;
  // coupling strength per axis of the current level's matrix: a coarse face sums
  // the fine faces of its bin, so w_a grows by the bin's extent on the other axes
  double w[3] = {g.w_int[0], g.w_int[1], g.w_int[2]};
  for (;;) {
    AggLevel& L = h.levels.back();
    const bool fine = h.levels.size() == 1;
    std::vector<std::vector<int32_t>> comps;
    std::vector<uint8_t> anch;
    if (fine) {
      // the grid's own components (the same BFS numbering), in node numbers
      comps.resize(g.anchored.size());
      for (size_t r = 0; r < comps.size(); ++r) comps[r].reserve((size_t)comp_size[r]);
      for (int64_t c = 0; c < g.count; ++c)
        if (g.component[(size_t)c] >= 0) comps[(size_t)g.component[(size_t)c]].push_back(node[(size_t)c]);
      anch.assign(g.anchored.begin(), g.anchored.end());
      bool small = true;
      for (auto& m : comps) small &= (int64_t)m.size() <= coarse_cells;
      if (small) {
        list_level0(L);
        csr_level0(L);
      }
    } else {
      csr_from_edges(L);
      clk.mark("csr", (int64_t)h.levels.size() - 1);
      level_components(L, comps, anch);
    }
    clk.mark("components", (int64_t)h.levels.size() - 1);

    L.closed_ptr.assign(1, 0);
    bool small = true;
    for (size_t r = 0; r < comps.size(); ++r) {
      small &= (int64_t)comps[r].size() <= coarse_cells;
      if (!anch[r]) {
        L.closed_members.insert(L.closed_members.end(), comps[r].begin(), comps[r].end());
        L.closed_ptr.push_back((int64_t)L.closed_members.size());
      }
    }
    if (small) {
      // Coarsest: one dense block per component.
      h.block_ptr.assign(1, 0);
      h.block_goff.assign(1, 0);
      for (size_t r = 0; r < comps.size(); ++r) {
        const int64_t m = (int64_t)comps[r].size();
        if (m <= host_block_max) {
          std::vector<double> G;
          int e = coarsest_block(L, comps[r], anch[r] != 0, G);
          if (e) return e;
          h.block_G.insert(h.block_G.end(), G.begin(), G.end());
        } else {
          // inverted on the device; its slot stays zero until then
          AggHierarchy::Deferred d;
          d.block = (int64_t)r;
          d.m = m;
          d.anchored = anch[r] != 0;
          d.K = assemble_block(L, comps[r], d.anchored);
          h.deferred.push_back(std::move(d));
          h.block_G.resize(h.block_G.size() + (size_t)(m * m), 0.0);
        }
        h.block_rows.insert(h.block_rows.end(), comps[r].begin(), comps[r].end());
        h.block_ptr.push_back((int64_t)h.block_rows.size());
        h.block_goff.push_back((int64_t)h.block_G.size());
      }
      break;
    }
    // Aggregate; grow the bins while nothing reduces (masked.rs:262-273). With
    // semicoarsening an axis is binned only while its coupling is within
    // kSemiRatio of the strongest axis that still spans more than one node, so
    // weakly coupled axes (thin or stretched cells) keep their nodes apart.
    bool coarsen[3] = {true, true, true};
    if (semicoarsen) {
      int32_t ext[3] = {0, 0, 0};
      std::mutex mx;
      parallel_ranges(L.n, [&](int64_t i0, int64_t i1) {
        int32_t e[3] = {0, 0, 0};
        for (int64_t i = i0; i < i1; ++i)
          for (int d = 0; d < 3; ++d) e[d] = std::max(e[d], L.coords[3 * (size_t)i + d] + 1);
        std::lock_guard<std::mutex> lk(mx);
        for (int d = 0; d < 3; ++d) ext[d] = std::max(ext[d], e[d]);
      });
      double wmax = 0.0;
      for (int d = 0; d < 3; ++d)
        if (ext[d] > 1) wmax = std::max(wmax, w[d]);
      for (int d = 0; d < 3; ++d) coarsen[d] = ext[d] > 1 && w[d] * kSemiRatio >= wmax;
      if (!coarsen[0] && !coarsen[1] && !coarsen[2]) coarsen[0] = coarsen[1] = coarsen[2] = true;
    }
    std::vector<int32_t> parent, cc;
    int64_t m = 0;
    int s = 1;
    int shift[3];
    AggLevel C;
    bool binned = false;
    if (fine && coarsen[0] && coarsen[1] && coarsen[2] && !general) {
      clk.mark("before bins", 0);
      binned = bins_level0(g, node, L, parent, cc, C);
      clk.mark("bins level 0", 0);
      if (binned) {
        m = (int64_t)cc.size() / 3;
        shift[0] = shift[1] = shift[2] = 1;
      }
    }
    if (!binned && fine) list_level0(L);
    // the level's edges in order
    const auto weighted = [&](auto&& f) {
      for (size_t e = 0; e < L.ea.size(); ++e) f(L.ea[e], L.eb[e], L.ew[e]);
    };
    const auto edges = [&](auto&& f) { weighted([&](int32_t a, int32_t b, double) { f(a, b); }); };
    while (!binned) {
      for (int d = 0; d < 3; ++d) shift[d] = coarsen[d] ? s : 0;
      m = aggregate(L.n, L.coords, shift, edges, parent, cc);
      if (m < L.n) break;
      bool all_zero = true;
      for (size_t e = 0; e < L.coords.size(); ++e) all_zero &= (L.coords[e] >> shift[e % 3]) == 0;
      if (all_zero) {
        if (!(coarsen[0] && coarsen[1] && coarsen[2])) {
          // the binned axes are exhausted: bin every axis
          coarsen[0] = coarsen[1] = coarsen[2] = true;
          s = 1;
          continue;
        }
        set_error("CFD-E035: numerics.pressure_preconditioner: masked aggregation failed to reduce a connected component");
        return CFD_E035;
      }
      ++s;
    }
    if (fine) h.fine_shift = std::max(shift[0], std::max(shift[1], shift[2]));
    {
      double f[3], wn[3];
      for (int d = 0; d < 3; ++d) f[d] = (double)(1 << shift[d]);
      for (int d = 0; d < 3; ++d) wn[d] = w[d] * f[(d + 1) % 3] * f[(d + 2) % 3];
      for (int d = 0; d < 3; ++d) w[d] = wn[d];
    }
    C.n = m;
    C.coords = std::move(cc);
    C.anchor.assign((size_t)m, 0.0);
    for (int64_t i = 0; i < L.n; ++i) C.anchor[(size_t)parent[(size_t)i]] += L.anchor[(size_t)i];
    clk.mark("aggregate", (int64_t)h.levels.size() - 1);
    if (!binned) coarse_edges(m, parent, weighted, C);
    clk.mark("coarse edges", (int64_t)h.levels.size() - 1);
    L.parent = std::move(parent);
    // the finer edge list is not needed once the next level exists
    std::vector<int32_t>().swap(L.ea);
    std::vector<int32_t>().swap(L.eb);
    std::vector<double>().swap(L.ew);
    h.levels.push_back(std::move(C));
  }
  std::vector<int32_t>().swap(h.levels.back().ea);
  std::vector<int32_t>().swap(h.levels.back().eb);
  std::vector<double>().swap(h.levels.back().ew);
  if (children) build_children(h);
  clk.mark("children");
  return 0;
}

namespace {

// Interpolation weights of a fine cell centre x from the coarse centres along
// one axis (multigrid.rs:296-320): the two bracketing centres, wrapped on a
// periodic axis and clamped otherwise; zero weights are dropped.
void axis_weights(double x, const std::vector<double>& centers, bool periodic, double extent,
                  std::vector<std::pair<int64_t, double>>& out) {
  out.clear();
  const int64_t n = (int64_t)centers.size();
  if (n == 1) {
    out.emplace_back(0, 1.0);
    return;
  }
  const int64_t hi0 = std::lower_bound(centers.begin(), centers.end(), x) - centers.begin();
  int64_t lo, hi;
  double left, right;
  if (hi0 == 0) {
    if (!periodic) {
      out.emplace_back(0, 1.0);
      return;
    }
    lo = n - 1;
    hi = 0;
    left = centers[(size_t)(n - 1)] - extent;
    right = centers[0];
  } else if (hi0 == n) {
    if (!periodic) {
      out.emplace_back(n - 1, 1.0);
      return;
    }
    lo = n - 1;
    hi = 0;
    left = centers[(size_t)(n - 1)];
    right = centers[0] + extent;
  } else {
    lo = hi0 - 1;
    hi = hi0;
    left = centers[(size_t)lo];
    right = centers[(size_t)hi];
  }
  const double t = (x - left) / (right - left);
  if (1.0 - t != 0.0) out.emplace_back(lo, 1.0 - t);
  if (t != 0.0) out.emplace_back(hi, t);
}

// Trilinear interpolation rows of one level (multigrid.rs:330-371): axes of more
// than two cells are paired, the coarse centre being the mean of the pair.
void transfer_rows(const int64_t n[3], const std::vector<double> centers[3], const bool periodic[3],
                   const double extent[3], std::vector<double> coarse[3], AggLevel& L, const bool coarsen[3]) {
  int64_t nc[3];
  for (int a = 0; a < 3; ++a) {
    coarse[a].clear();
    const size_t step = coarsen[a] ? 2 : 1;
    for (size_t s0 = 0; s0 < centers[a].size(); s0 += step) {
      const size_t e = std::min(centers[a].size(), s0 + step);
      double sum = -0.0;
      for (size_t k = s0; k < e; ++k) sum += centers[a][k];
      coarse[a].push_back(sum / (double)(e - s0));
    }
    nc[a] = (int64_t)coarse[a].size();
  }
  std::vector<std::vector<std::pair<int64_t, double>>> w[3];
  std::vector<std::pair<int64_t, double>> tmp;
  for (int a = 0; a < 3; ++a)
    for (double x : centers[a]) {
      axis_weights(x, coarse[a], periodic[a], extent[a], tmp);
      w[a].push_back(tmp);
    }
  const int64_t total = n[0] * n[1] * n[2];
  L.tptr.assign(1, 0);
  L.tcol.clear();
  L.tw.clear();
  std::vector<std::pair<int64_t, double>> row;
  for (int64_t i = 0; i < total; ++i) {
    const int64_t q[3] = {i % n[0], i / n[0] % n[1], i / (n[0] * n[1])};
    row.clear();
    for (auto& z : w[2][(size_t)q[2]])
      for (auto& y : w[1][(size_t)q[1]])
        for (auto& x : w[0][(size_t)q[0]])
          row.emplace_back((z.first * nc[1] + y.first) * nc[0] + x.first, x.second * y.second * z.second);
    std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto& e : row) {
      L.tcol.push_back((int32_t)e.first);
      L.tw.push_back(e.second);
    }
    L.tptr.push_back((int64_t)L.tcol.size());
  }
}

// Galerkin coarse operator from edge energies (multigrid.rs:379-429): each fine
// edge w (x_i - x_j)^2 with P rows p_i, p_j adds w d_a d_b to the coarse pair
// (a, b) for the entries d of p_i - p_j; each anchor w x_i^2 adds w p_a p_b and
// w p_a to the coarse anchor. Pairs are accumulated in that order and emitted
// sorted, as edges (a, b, -sum).
void galerkin(const AggLevel& F, int64_t nc, AggLevel& C) {
  EdgeSums sums(nc);
  std::vector<std::pair<int32_t, double>> delta;
  auto pairs = [&](double w) {
    for (size_t at = 0; at < delta.size(); ++at)
      for (size_t k = at + 1; k < delta.size(); ++k) {
        // keys ascend with position, so the pair is already (min, max)
        auto& r = sums.rows[(size_t)delta[at].first];
        const int32_t b = delta[k].first;
        const double v = w * delta[at].second * delta[k].second;
        bool found = false;
        for (auto& e : r)
          if (e.first == b) {
            e.second += v;
            found = true;
            break;
          }
        if (!found) r.emplace_back(b, v);
      }
  };
  for (size_t e = 0; e < F.ea.size(); ++e) {
    delta.clear();
    int64_t a0 = F.tptr[(size_t)F.ea[e]], a1 = F.tptr[(size_t)F.ea[e] + 1];
    int64_t b0 = F.tptr[(size_t)F.eb[e]], b1 = F.tptr[(size_t)F.eb[e] + 1];
    while (a0 < a1 || b0 < b1) {
      const int64_t ca = a0 < a1 ? F.tcol[(size_t)a0] : INT64_MAX;
      const int64_t cb = b0 < b1 ? F.tcol[(size_t)b0] : INT64_MAX;
      int64_t col;
      double value;
      if (ca == cb) {
        col = ca;
        value = F.tw[(size_t)a0++] - F.tw[(size_t)b0++];
      } else if (ca < cb) {
        col = ca;
        value = F.tw[(size_t)a0++];
      } else {
        col = cb;
        value = -F.tw[(size_t)b0++];
      }
      if (value != 0.0) delta.emplace_back((int32_t)col, value);
    }
    pairs(F.ew[e]);
  }
  C.anchor.assign((size_t)nc, 0.0);
  for (int64_t i = 0; i < F.n; ++i) {
    const double w = F.anchor[(size_t)i];
    if (w == 0.0) continue;
    delta.clear();
    for (int64_t at = F.tptr[(size_t)i]; at < F.tptr[(size_t)i + 1]; ++at)
      delta.emplace_back(F.tcol[(size_t)at], F.tw[(size_t)at]);
    for (size_t at = 0; at < delta.size(); ++at) C.anchor[(size_t)delta[at].first] += w * delta[at].second;
    pairs(w);
  }
  C.ea.clear();
  C.eb.clear();
  C.ew.clear();
  for (size_t a = 0; a < sums.rows.size(); ++a) {
    auto& r = sums.rows[a];
    std::sort(r.begin(), r.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    for (auto& e : r)
      if (e.second != 0.0) {
        C.ea.push_back((int32_t)a);
        C.eb.push_back(e.first);
        C.ew.push_back(-e.second);
      }
  }
}

// Anisotropic spacing (multigrid.rs:33-47): when the axis weights 1/dx^2
// spread by more than 25 over the axes of more than two cells, the axes within a
// factor 4 of the largest weight are relaxed by line blocks; none otherwise.
std::vector<int> line_axes(const HostGrid& g) {
  double w[3];
  for (int a = 0; a < 3; ++a) {
    w[a] = 1.0 / (g.dx[a] * g.dx[a]);
    if (!std::isfinite(w[a]) || w[a] <= 0.0) return {};
  }
  double lo = INFINITY, hi = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (g.n[a] > 2) lo = std::fmin(lo, w[a]);
    hi = std::fmax(hi, w[a]);
  }
  if (hi / lo <= 25.0) return {};
  std::vector<int> out;
  for (int a = 0; a < 3; ++a)
    if (w[a] >= hi / 4.0) out.push_back(a);
  return out;
}

// Block smoother of one level (multigrid/line.rs:38-154): the nodes of each
// line in order (a periodic cycle folded as 0, n-1, 1, n-2, ... to keep the band
// narrow), the entries leaving the block, and the banded Cholesky factor with
// reciprocal diagonal; then the lines' dependency levels for parallel sweeps.
int build_lines(AggLevel& L, const int64_t n[3], const std::vector<int>& axes, const bool periodic[3], bool zebra) {
  const int64_t count = L.n;
  int64_t length = 1;
  for (int a : axes) length *= n[a];
  L.line_length = length;
  L.line_nodes.clear();
  for (int64_t i = 0; i < count; ++i) {
    int64_t q[3] = {i % n[0], i / n[0] % n[1], i / (n[0] * n[1])};
    bool base = true;
    for (int a : axes) base &= q[a] == 0;
    if (!base) continue;
    for (int64_t at = 0; at < length; ++at) {
      int64_t rest = at;
      for (int a : axes) {
        const int64_t k = rest % n[a];
        rest /= n[a];
        q[a] = !periodic[a] ? k : (k % 2 == 0 ? k / 2 : n[a] - 1 - k / 2);
      }
      L.line_nodes.push_back((int32_t)((q[2] * n[1] + q[1]) * n[0] + q[0]));
    }
  }
  std::vector<int64_t> pos((size_t)count);
  for (size_t at = 0; at < L.line_nodes.size(); ++at) pos[(size_t)L.line_nodes[at]] = (int64_t)at;
  int64_t bw = 0;
  for (int64_t i = 0; i < count; ++i)
    for (int64_t e = L.ptr[(size_t)i]; e < L.ptr[(size_t)i + 1]; ++e) {
      const int64_t pi = pos[(size_t)i], pj = pos[(size_t)L.col[(size_t)e]];
      if (pi / length == pj / length) bw = std::max(bw, pi > pj ? pi - pj : pj - pi);
    }
  L.line_bandwidth = bw;
  const int64_t stride = bw + 1;
  L.line_ooff.assign(1, 0);
  L.line_ocol.clear();
  L.line_oval.clear();
  for (int64_t at = 0; at < count; ++at) {
    const int32_t cell = L.line_nodes[(size_t)at];
    for (int64_t e = L.ptr[(size_t)cell]; e < L.ptr[(size_t)cell + 1]; ++e) {
      const int32_t c = L.col[(size_t)e];
      if (pos[(size_t)c] / length != at / length) {
        L.line_ocol.push_back(c);
        L.line_oval.push_back(L.val[(size_t)e]);
      }
    }
    L.line_ooff.push_back((int64_t)L.line_ocol.size());
  }
  std::vector<double>& F = L.line_factor;
  F.assign((size_t)(count * stride), 0.0);
  for (int64_t at = 0; at < count; ++at) {
    const int32_t cell = L.line_nodes[(size_t)at];
    F[(size_t)(at * stride)] = L.diag[(size_t)cell];
    for (int64_t e = L.ptr[(size_t)cell]; e < L.ptr[(size_t)cell + 1]; ++e) {
      const int64_t c = pos[(size_t)L.col[(size_t)e]];
      if (c / length == at / length && c < at) F[(size_t)(at * stride + at - c)] += L.val[(size_t)e];
    }
  }
  for (int64_t line = 0; line < count / length; ++line) {
    const int64_t base = line * length * stride;
    for (int64_t i = 0; i < length; ++i)
      for (int64_t j = std::max<int64_t>(0, i - bw); j <= i; ++j) {
        const int64_t row = base + i * stride, other = base + j * stride;
        double v = F[(size_t)(row + i - j)];
        for (int64_t k = std::max<int64_t>(0, i - bw); k < j; ++k) v -= F[(size_t)(row + i - k)] * F[(size_t)(other + j - k)];
        if (i == j) {
          if (!std::isfinite(v) || v <= 0.0) {
            set_error("CFD-E036: numerics.pressure_preconditioner: multigrid line Cholesky lost positive definiteness");
            return CFD_E036;
          }
          F[(size_t)row] = std::sqrt(v);
        } else {
          F[(size_t)(row + i - j)] = v / F[(size_t)other];
        }
      }
  }
  for (int64_t row = 0; row < count; ++row) F[(size_t)(row * stride)] = 1.0 / F[(size_t)(row * stride)];
  // dependency levels of the block Gauss-Seidel sweep over lines
  const int64_t nl = count / length;
  std::vector<int32_t> depth((size_t)nl, 0);
  auto group = [&](std::vector<int64_t>& ptr, std::vector<int32_t>& lines) {
    int32_t most = 0;
    for (int32_t d : depth) most = std::max(most, d);
    ptr.assign((size_t)most + 2, 0);
    for (int32_t d : depth) ptr[(size_t)d + 1] += 1;
    for (int32_t d = 0; d <= most; ++d) ptr[(size_t)d + 1] += ptr[(size_t)d];
    lines.assign((size_t)nl, 0);
    std::vector<int64_t> at(ptr.begin(), ptr.end() - 1);
    for (int64_t l = 0; l < nl; ++l) lines[(size_t)at[(size_t)depth[(size_t)l]]++] = (int32_t)l;
  };
  if (zebra) {
    // colour classes instead: lines coupled by an outside entry differ in colour
    // (greedy in line order); forward sweeps run the classes in ascending order,
    // backward sweeps in descending order
    std::vector<int32_t> colour((size_t)nl, -1);
    std::vector<char> used;
    for (int64_t l = 0; l < nl; ++l) {
      used.assign(used.size(), 0);
      for (int64_t at = l * length; at < (l + 1) * length; ++at)
        for (int64_t e = L.line_ooff[(size_t)at]; e < L.line_ooff[(size_t)at + 1]; ++e) {
          const int32_t c = colour[(size_t)(pos[(size_t)L.line_ocol[(size_t)e]] / length)];
          if (c < 0) continue;
          if ((size_t)c >= used.size()) used.resize((size_t)c + 1, 0);
          used[(size_t)c] = 1;
        }
      int32_t c = 0;
      while ((size_t)c < used.size() && used[(size_t)c]) ++c;
      colour[(size_t)l] = c;
    }
    depth = colour;
    group(L.line_fwd_ptr, L.line_fwd);
    int32_t most = 0;
    for (int32_t c : colour) most = std::max(most, c);
    for (int64_t l = 0; l < nl; ++l) depth[(size_t)l] = most - colour[(size_t)l];
    group(L.line_bwd_ptr, L.line_bwd);
    return 0;
  }
  for (int pass = 0; pass < 2; ++pass) {
    const bool forward = pass == 0;
    std::fill(depth.begin(), depth.end(), 0);
    for (int64_t k = 0; k < nl; ++k) {
      const int64_t l = forward ? k : nl - 1 - k;
      int32_t d = 0;
      for (int64_t at = l * length; at < (l + 1) * length; ++at)
        for (int64_t e = L.line_ooff[(size_t)at]; e < L.line_ooff[(size_t)at + 1]; ++e) {
          const int64_t other = pos[(size_t)L.line_ocol[(size_t)e]] / length;
          if (forward ? other < l : other > l) d = std::max(d, depth[(size_t)other] + 1);
        }
      depth[(size_t)l] = d;
    }
    if (forward)
      group(L.line_fwd_ptr, L.line_fwd);
    else
      group(L.line_bwd_ptr, L.line_bwd);
  }
  return 0;
}

int build_unmasked(const HostGrid& g, AggHierarchy& h, bool zebra) {
  h.masked = false;
  const int64_t N = g.count;
  h.node_of_cell.resize((size_t)N);
  h.cell_of_node.resize((size_t)N);
  for (int64_t c = 0; c < N; ++c) {
    h.node_of_cell[(size_t)c] = (int32_t)c;
    h.cell_of_node[(size_t)c] = c;
  }
  const bool anchored = g.anchored[0] != 0;
  AggLevel L0;
  L0.n = N;
  L0.anchor.assign((size_t)N, 0.0);
  for_each_face(
      g, [&](int64_t lo, int64_t hi, double w) {
        L0.ea.push_back((int32_t)lo);
        L0.eb.push_back((int32_t)hi);
        L0.ew.push_back(w);
      },
      [&](int64_t c, double w) { L0.anchor[(size_t)c] += w; }, true);
  h.level0_entries = 2 * (int64_t)L0.ea.size();
  L0.diag = L0.anchor;
  for (size_t e = 0; e < L0.ea.size(); ++e) {
    L0.diag[(size_t)L0.ea[e]] += L0.ew[e];
    L0.diag[(size_t)L0.eb[e]] += L0.ew[e];
  }
  csr_level0(L0);
  int64_t n[3] = {g.n[0], g.n[1], g.n[2]};
  std::vector<double> centers[3];
  double extent[3];
  bool periodic[3];
  for (int a = 0; a < 3; ++a) {
    for (int64_t i = 0; i < n[a]; ++i) centers[a].push_back((double)i + 0.5);
    extent[a] = (double)n[a];
    periodic[a] = g.periodic[a] && n[a] >= 2;
  }
  h.levels.push_back(std::move(L0));
  const std::vector<int> axes = line_axes(g);
  for (;;) {
    AggLevel& L = h.levels.back();
    for (int a = 0; a < 3; ++a) L.dims[a] = n[a];
    L.closed_ptr.assign(1, 0);
    if (!anchored) {
      for (int64_t i = 0; i < L.n; ++i) L.closed_members.push_back((int32_t)i);
      L.closed_ptr.push_back(L.n);
    }
    if (L.n <= 32) {
      std::vector<int32_t> members((size_t)L.n);
      std::iota(members.begin(), members.end(), 0);
      std::vector<double> G;
      int e = coarsest_block(L, members, anchored, G);
      if (e) {
        set_error("CFD-E035: numerics.pressure_preconditioner: multigrid coarse Cholesky lost positive definiteness");
        return CFD_E035;
      }
      h.block_ptr = {0, L.n};
      h.block_goff = {0, (int64_t)G.size()};
      h.block_rows = members;
      h.block_G = std::move(G);
      break;
    }
    std::vector<double> coarse[3];
    bool coarsen[3];
    for (int a = 0; a < 3; ++a) coarsen[a] = n[a] > 2;
    if (!axes.empty()) {
      // multigrid.rs:582-611: the weak axes wait until the stiff ones have
      // collapsed; one block group per level: the stiff axes together when their
      // span is at most 16 nodes, else the stiffest axis alone
      bool stiff_can_coarsen = false;
      for (int a : axes) stiff_can_coarsen |= n[a] > 2;
      for (int a = 0; a < 3; ++a)
        coarsen[a] = n[a] > 2 && (!stiff_can_coarsen || std::find(axes.begin(), axes.end(), a) != axes.end());
      int64_t span = 1;
      for (int a : axes) span *= n[a];
      std::vector<int> grp;
      if (axes.size() > 1 && span <= 16) {
        grp = axes;
      } else {
        int best = axes[0];
        for (int a : axes)
          if (g.dx[a] < g.dx[best]) best = a;
        grp = {best};
      }
      int e = build_lines(L, n, grp, periodic, zebra);
      if (e) return e;
      L.block_axes = grp;
      std::sort(L.block_axes.begin(), L.block_axes.end());
    }
    transfer_rows(n, centers, periodic, extent, coarse, L, coarsen);
    for (int a = 0; a < 3; ++a) n[a] = (int64_t)coarse[a].size();
    AggLevel C;
    C.n = n[0] * n[1] * n[2];
    galerkin(L, C.n, C);
    csr_from_edges(C);
    for (int a = 0; a < 3; ++a) centers[a] = coarse[a];
    std::vector<int32_t>().swap(L.ea);
    std::vector<int32_t>().swap(L.eb);
    std::vector<double>().swap(L.ew);
    h.levels.push_back(std::move(C));
  }
  // children of each coarse node: the finer nodes whose row has it, ascending, with weights
  for (size_t l = 1; l < h.levels.size(); ++l) {
    AggLevel& C = h.levels[l];
    const AggLevel& F = h.levels[l - 1];
    C.child_ptr.assign((size_t)C.n + 1, 0);
    for (int32_t c : F.tcol) C.child_ptr[(size_t)c + 1] += 1;
    for (int64_t i = 0; i < C.n; ++i) C.child_ptr[(size_t)i + 1] += C.child_ptr[(size_t)i];
    C.child.assign(F.tcol.size(), 0);
    C.child_w.assign(F.tcol.size(), 0.0);
    std::vector<int64_t> at(C.child_ptr.begin(), C.child_ptr.end() - 1);
    for (int64_t i = 0; i < F.n; ++i)
      for (int64_t k = F.tptr[(size_t)i]; k < F.tptr[(size_t)i + 1]; ++k) {
        const int64_t slot = at[(size_t)F.tcol[(size_t)k]]++;
        C.child[(size_t)slot] = (int32_t)i;
        C.child_w[(size_t)slot] = F.tw[(size_t)k];
      }
  }
  std::vector<int32_t>().swap(h.levels.back().ea);
  std::vector<int32_t>().swap(h.levels.back().eb);
  std::vector<double>().swap(h.levels.back().ew);
  return 0;
}

// Dependency levels of the lexicographic sweeps of one CSR level.
void sweep_schedule(AggLevel& L, bool zebra) {
  const int64_t n = L.n;
  std::vector<int32_t> depth((size_t)n, 0);
  auto group = [&](std::vector<int64_t>& ptr, std::vector<int32_t>& rows) {
    int32_t most = 0;
    for (int32_t d : depth) most = std::max(most, d);
    ptr.assign((size_t)most + 2, 0);
    for (int32_t d : depth) ptr[(size_t)d + 1] += 1;
    for (int32_t d = 0; d <= most; ++d) ptr[(size_t)d + 1] += ptr[(size_t)d];
    rows.assign((size_t)n, 0);
    std::vector<int64_t> at(ptr.begin(), ptr.end() - 1);
    for (int64_t i = 0; i < n; ++i) rows[(size_t)at[(size_t)depth[(size_t)i]]++] = (int32_t)i;
  };
  if (zebra) {
    // colour classes (greedy in row order): coupled rows differ in colour
    std::vector<char> used;
    std::fill(depth.begin(), depth.end(), -1);
    int32_t most = 0;
    for (int64_t i = 0; i < n; ++i) {
      used.assign(used.size(), 0);
      for (int64_t e = L.ptr[(size_t)i]; e < L.ptr[(size_t)i + 1]; ++e) {
        const int32_t c = depth[(size_t)L.col[(size_t)e]];
        if (c < 0) continue;
        if ((size_t)c >= used.size()) used.resize((size_t)c + 1, 0);
        used[(size_t)c] = 1;
      }
      int32_t c = 0;
      while ((size_t)c < used.size() && used[(size_t)c]) ++c;
      depth[(size_t)i] = c;
      most = std::max(most, c);
    }
    group(L.fwd_ptr, L.fwd_rows);
    for (int64_t i = 0; i < n; ++i) depth[(size_t)i] = most - depth[(size_t)i];
    group(L.bwd_ptr, L.bwd_rows);
    return;
  }
  for (int64_t i = 0; i < n; ++i) {
    int32_t d = 0;
    for (int64_t e = L.ptr[(size_t)i]; e < L.ptr[(size_t)i + 1]; ++e)
      if (L.col[(size_t)e] < i) d = std::max(d, depth[(size_t)L.col[(size_t)e]] + 1);
    depth[(size_t)i] = d;
  }
  group(L.fwd_ptr, L.fwd_rows);
  for (int64_t i = n - 1; i >= 0; --i) {
    int32_t d = 0;
    for (int64_t e = L.ptr[(size_t)i]; e < L.ptr[(size_t)i + 1]; ++e)
      if (L.col[(size_t)e] > i) d = std::max(d, depth[(size_t)L.col[(size_t)e]] + 1);
    depth[(size_t)i] = d;
  }
  group(L.bwd_ptr, L.bwd_rows);
}

}  // namespace

int build_ref_hierarchy(const HostGrid& g, AggHierarchy& h, bool zebra) {
  h = AggHierarchy();
  if (g.fluid_count == 0) {
    set_error("CFD-E034: numerics.pressure_preconditioner: multigrid requires fluid cells");
    return CFD_E034;
  }
  int rc;
  if (g.fluid_count < g.count || g.anchored.size() != 1) {
    rc = build_agg_hierarchy(g, h, 32);
  } else {
    rc = build_unmasked(g, h, zebra);
  }
  if (rc) return rc;
  for (size_t l = 1; l < h.levels.size(); ++l) sweep_schedule(h.levels[l], zebra);
  return 0;
}

void colour_agg_hierarchy(AggHierarchy& h) {
  const size_t nl = h.levels.size();
  h.permuted = true;
  for (size_t l = 1; l < nl; ++l) {
    AggLevel& L = h.levels[l];
    const int64_t n = L.n;
    // Parity of the bin coordinates separates coupled aggregates (they lie in
    // face-adjacent bins) except across an odd periodic seam; fall back to a
    // greedy colouring in node order where it fails.
    std::vector<int32_t> colour((size_t)n);
    parallel_ranges(n, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i)
        colour[(size_t)i] =
            (L.coords[3 * (size_t)i] + L.coords[3 * (size_t)i + 1] + L.coords[3 * (size_t)i + 2]) & 1;
    });
    std::atomic<bool> clash{false};
    parallel_ranges(n, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1 && !clash.load(std::memory_order_relaxed); ++i)
        for (int64_t e = L.ptr[(size_t)i]; e < L.ptr[(size_t)i + 1]; ++e)
          if (colour[(size_t)L.col[(size_t)e]] == colour[(size_t)i]) {
            clash.store(true, std::memory_order_relaxed);
            break;
          }
    });
    bool ok = !clash.load();
    int ncol = 2;
    if (!ok) {
      ncol = 1;
      std::fill(colour.begin(), colour.end(), -1);
      std::vector<uint8_t> used;
      for (int64_t i = 0; i < n; ++i) {
        used.assign((size_t)ncol + 1, 0);
        for (int64_t e = L.ptr[(size_t)i]; e < L.ptr[(size_t)i + 1]; ++e) {
          const int32_t c = colour[(size_t)L.col[(size_t)e]];
          if (c >= 0) used[(size_t)c] = 1;
        }
        int c = 0;
        while (used[(size_t)c]) ++c;
        colour[(size_t)i] = c;
        ncol = std::max(ncol, c + 1);
      }
    }
    // new numbering: colour classes in order, ascending old index inside a class
    std::vector<int32_t> perm((size_t)n);
    L.colour_ptr.assign((size_t)ncol + 1, 0);
    for (int64_t i = 0; i < n; ++i) L.colour_ptr[(size_t)colour[(size_t)i] + 1] += 1;
    for (int c = 0; c < ncol; ++c) L.colour_ptr[(size_t)c + 1] += L.colour_ptr[(size_t)c];
    {
      std::vector<int64_t> at(L.colour_ptr.begin(), L.colour_ptr.end() - 1);
      for (int64_t i = 0; i < n; ++i) perm[(size_t)i] = (int32_t)at[(size_t)colour[(size_t)i]]++;
    }
    // permute this level's rows, columns, diagonal, anchors, coordinates and parent
    std::vector<int64_t> ptr((size_t)n + 1, 0);
    std::vector<int32_t> inv((size_t)n);
    parallel_ranges(n, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) inv[(size_t)perm[(size_t)i]] = (int32_t)i;
    });
    for (int64_t r = 0; r < n; ++r) {
      const int32_t o = inv[(size_t)r];
      ptr[(size_t)r + 1] = ptr[(size_t)r] + (L.ptr[(size_t)o + 1] - L.ptr[(size_t)o]);
    }
    std::vector<int32_t> col, coords, parent;
    std::vector<double> val, diag, anchor;
    concurrently([&] { col.resize((size_t)ptr[(size_t)n]); }, [&] { val.resize((size_t)ptr[(size_t)n]); },
                 [&] { diag.resize((size_t)n); }, [&] { anchor.resize((size_t)n); },
                 [&] { coords.resize((size_t)(3 * n)); },
                 [&] { parent.resize(L.parent.empty() ? 0 : (size_t)n); });
    parallel_ranges(n, [&](int64_t r0, int64_t r1) {
      for (int64_t r = r0; r < r1; ++r) {
        const int32_t o = inv[(size_t)r];
        int64_t at = ptr[(size_t)r];
        for (int64_t e = L.ptr[(size_t)o]; e < L.ptr[(size_t)o + 1]; ++e, ++at) {
          col[(size_t)at] = perm[(size_t)L.col[(size_t)e]];
          val[(size_t)at] = L.val[(size_t)e];
        }
        diag[(size_t)r] = L.diag[(size_t)o];
        anchor[(size_t)r] = L.anchor[(size_t)o];
        for (int d = 0; d < 3; ++d) coords[3 * (size_t)r + d] = L.coords[3 * (size_t)o + d];
        if (!parent.empty()) parent[(size_t)r] = L.parent[(size_t)o];
      }
    });
    L.ptr.swap(ptr);
    L.col.swap(col);
    L.val.swap(val);
    L.diag.swap(diag);
    L.anchor.swap(anchor);
    L.coords.swap(coords);
    L.parent.swap(parent);
    for (auto& m : L.closed_members) m = perm[(size_t)m];
    for (size_t c = 0; c + 1 < L.closed_ptr.size(); ++c)
      std::sort(L.closed_members.begin() + L.closed_ptr[c], L.closed_members.begin() + L.closed_ptr[c + 1]);
    // the finer level's parents point at the new numbers
    std::vector<int32_t>& fpar = h.levels[l - 1].parent;
    parallel_ranges((int64_t)fpar.size(), [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) fpar[(size_t)i] = perm[(size_t)fpar[(size_t)i]];
    });
    if (l + 1 == nl)
      for (auto& r : h.block_rows) r = perm[(size_t)r];
  }
  build_children(h);
}

// FNV-1a over the hierarchy's host arrays: a digest of every value and index
// the device copies, for checks that two builds are the same bit for bit.
uint64_t agg_hierarchy_digest(const AggHierarchy& h) {
  uint64_t x = 1469598103934665603ULL;
  auto bytes = [&](const void* p, size_t len) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < len; ++i) {
      x ^= b[i];
      x *= 1099511628211ULL;
    }
  };
  auto vec = [&](const auto& v) {
    const uint64_t len = v.size();
    bytes(&len, sizeof len);
    if (len) bytes(v.data(), v.size() * sizeof(v[0]));
  };
  vec(h.node_of_cell);
  bytes(&h.level0_entries, sizeof h.level0_entries);
  for (const AggLevel& L : h.levels) {
    bytes(&L.n, sizeof L.n);
    vec(L.ptr);
    vec(L.col);
    vec(L.val);
    vec(L.diag);
    vec(L.anchor);
    vec(L.coords);
    vec(L.parent);
    vec(L.closed_ptr);
    vec(L.closed_members);
  }
  vec(h.block_ptr);
  vec(h.block_rows);
  vec(h.block_G);
  for (const auto& d : h.deferred) vec(d.K);
  return x;
}

void agg_hierarchy_stats(const AggHierarchy& h, std::vector<int64_t>& cells, std::vector<int64_t>& nonzeros) {
  cells.clear();
  nonzeros.clear();
  for (const auto& L : h.levels) {
    cells.push_back(L.n);
    nonzeros.push_back((int64_t)L.col.size() + L.n);
  }
}

void agg_hierarchy_shape(const AggHierarchy& h, cfd_mg_stats* s) {
  s->n_block_smoothers = 0;
  for (size_t l = 0; l < h.levels.size() && l < 32; ++l) {
    const AggLevel& L = h.levels[l];
    for (int a = 0; a < 3; ++a) s->level_dims[l][a] = h.masked ? 0 : L.dims[a];
    if (L.line_length <= 0 || s->n_block_smoothers >= 32) continue;
    cfd_block_smoother_stats& b = s->block_smoothers[s->n_block_smoothers++];
    b.level = (int64_t)l;
    b.n_axes = (int64_t)L.block_axes.size();
    for (size_t k = 0; k < L.block_axes.size() && k < 3; ++k) b.axes[k] = L.block_axes[k];
    b.block_unknowns = L.line_length;
    b.block_count = (int64_t)L.line_nodes.size() / L.line_length;
    b.half_bandwidth = L.line_bandwidth;
    b.allocated_bytes = (int64_t)(L.line_nodes.size() * 4 + L.line_ooff.size() * 8 + L.line_ocol.size() * 4 +
                                  L.line_oval.size() * 8 + L.line_factor.size() * 8 +
                                  (L.line_fwd_ptr.size() + L.line_bwd_ptr.size()) * 8 +
                                  (L.line_fwd.size() + L.line_bwd.size()) * 4);
  }
}

}  // namespace cfdgpu
