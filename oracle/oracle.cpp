// CPU oracle of the pressure-projection path: a serial FP64 restatement of the
// reference's arithmetic, for parity checks and as the timed CPU baseline.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py; the GPU product never links or calls it.
//
// Built with -O2 -ffp-contract=off so that no a - b*c is contracted into an
// FMA, as in the reference (Rust never contracts). Every floating-point
// expression keeps the reference's operand order; sums start from -0.0, the
// neutral element of Rust's `f64: Sum` since Rust 1.81.
//
// Parity is pinned (tests/test_oracle*.py) by the reference's published
// iteration counts (docs/evidence/pressure-multigrid.md:34-38,
// docs/evidence/native-cpu-performance.md:210-211), its hierarchy sizes
// (pressure-multigrid.md:129-142), closed-form stencils, dense and pivoted
// reference solves on tiny grids, the Galerkin / symmetry / connectivity
// invariants of its own unit tests, and the 18-shape anisotropy suite
// (docs/evidence/anisotropic-pressure.md:43-60: IC(0) and multigrid iterations,
// the latter with the line and small-plane block smoothers of multigrid/line.rs).
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace {

using std::size_t;
using std::vector;
constexpr size_t NONE = (size_t)-1;
thread_local std::string g_error;

int fail(int code, const std::string& text) {
  g_error = text;
  return code;
}

double neg_zero_sum() { return -0.0; }

struct Face {
  int axis;
  size_t coord[3];
  double center_z;
  size_t lo, hi;  // NONE for solid / outside
  int kind;
  double area, distance;
};

// ---------------------------------------------------------------- geometry.rs
// geometry.rs:13-26 even-odd crossing test
bool poly_contains(const double* p, int n, double x, double y) {
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
// case.rs:436-443 Block::contains
bool block_contains(const orc_block& b, const double p[3]) {
  for (int a = 0; a < 3; ++a)
    if (!(p[a] >= b.min_m[a] && p[a] < b.max_m[a])) return false;
  if (b.n_vertices > 0) return poly_contains(b.footprint_xy, b.n_vertices, p[0], p[1]);
  return true;
}
double cell_center(size_t q, double dx) { return ((double)q + 0.5) * dx; }
// geometry.rs:96-107
template <class F>
size_t partition_point(size_t n, F pred) {
  size_t lo = 0, hi = n;
  while (lo < hi) {
    size_t mid = lo + (hi - lo) / 2;
    if (pred(mid))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}
// geometry.rs:113-165 Raster
struct Raster {
  size_t r0[3], r1[3];
  vector<char> columns;
  Raster(const orc_block& b, const size_t n[3], const double dx[3]) {
    for (int a = 0; a < 3; ++a) {
      double lo = b.min_m[a];
      size_t first = partition_point(n[a], [&](size_t q) { return cell_center(q, dx[a]) < lo || std::isnan(lo); });
      size_t end = partition_point(n[a], [&](size_t q) { return cell_center(q, dx[a]) < b.max_m[a]; });
      r0[a] = first;
      r1[a] = std::max(end, first);
    }
    bool empty = false;
    for (int a = 0; a < 3; ++a) empty |= r0[a] == r1[a];
    if (!empty) {
      double z = cell_center(r0[2], dx[2]);
      for (size_t j = r0[1]; j < r1[1]; ++j)
        for (size_t i = r0[0]; i < r1[0]; ++i) {
          double p[3] = {cell_center(i, dx[0]), cell_center(j, dx[1]), z};
          columns.push_back(block_contains(b, p) ? 1 : 0);
        }
    }
  }
  template <class F>
  void for_each(const size_t n[3], F f) const {
    if (columns.empty()) return;
    size_t wx = r1[0] - r0[0];
    for (size_t k = r0[2]; k < r1[2]; ++k)
      for (size_t j = r0[1]; j < r1[1]; ++j)
        for (size_t i = r0[0]; i < r1[0]; ++i)
          if (columns[(j - r0[1]) * wx + (i - r0[0])]) f((k * n[1] + j) * n[0] + i);
  }
  size_t count() const {
    size_t c = 0;
    for (char v : columns) c += v ? 1 : 0;
    return c * (r1[2] - r0[2]);
  }
};

// ---------------------------------------------------------------- case.rs
// case.rs:345-367 Flow::vector
void flow_vector(const orc_case& c, double z, double v[3]) {
  if (c.flow_kind == ORC_FLOW_NEUTRAL_RANS) {
    double z0 = c.roughness_m;
    double u = c.speed_m_s * std::log((std::max(z, 0.0) + z0) / z0) / std::log((c.reference_height_m + z0) / z0);
    const double rads_per_deg = M_PI / 180.0;  // f64::to_radians
    double t = c.direction_deg_from * rads_per_deg;
    v[0] = -u * std::sin(t);
    v[1] = -u * std::cos(t);
    v[2] = 0.0;
  } else {
    for (int a = 0; a < 3; ++a) v[a] = c.velocity_m_s[a];
  }
}

// ---------------------------------------------------------------- pressure.rs Factor
struct Factor {
  vector<size_t> row_offsets, columns;
  vector<double> lower, diagonal;
};

// ---------------------------------------------------------------- multigrid.rs
constexpr size_t COARSE_CELLS = 32;
constexpr int SWEEPS = 2;
struct Edge {
  size_t i, j;
  double w;
};
struct Matrix {  // multigrid.rs:225-289
  vector<size_t> offsets, columns;
  vector<double> values, diagonal;
  static Matrix build(const vector<Edge>& edges, const vector<double>& anchors) {
    Matrix m;
    size_t n = anchors.size();
    m.offsets.assign(n + 1, 0);
    for (auto& e : edges) {
      m.offsets[e.i + 1] += 1;
      m.offsets[e.j + 1] += 1;
    }
    for (size_t i = 0; i < n; ++i) m.offsets[i + 1] += m.offsets[i];
    m.columns.assign(m.offsets[n], 0);
    m.values.assign(m.offsets[n], 0.0);
    vector<size_t> next(m.offsets.begin(), m.offsets.begin() + n);
    m.diagonal = anchors;
    for (auto& e : edges) {
      size_t rc[2][2] = {{e.i, e.j}, {e.j, e.i}};
      for (auto& p : rc) {
        size_t row = p[0], col = p[1];
        m.diagonal[row] += e.w;
        size_t at = next[row];
        m.columns[at] = col;
        m.values[at] = -e.w;
        next[row] += 1;
      }
    }
    return m;
  }
  void product(const vector<double>& x, vector<double>& y) const {
    for (size_t i = 0; i < x.size(); ++i) {
      double value = diagonal[i] * x[i];
      for (size_t at = offsets[i]; at < offsets[i + 1]; ++at) value += values[at] * x[columns[at]];
      y[i] = value;
    }
  }
  void sweep(const vector<double>& rhs, vector<double>& x, bool backward) const {
    size_t n = x.size();
    for (size_t at = 0; at < n; ++at) {
      size_t i = backward ? n - 1 - at : at;
      double value = rhs[i];
      for (size_t e = offsets[i]; e < offsets[i + 1]; ++e) value -= values[e] * x[columns[e]];
      x[i] = value / diagonal[i];
    }
  }
};

struct Transfer {  // multigrid.rs:291-433
  vector<size_t> offsets, columns;
  vector<double> weights;
  bool empty() const { return offsets.empty(); }
};
// multigrid.rs:296-320
vector<std::pair<size_t, double>> axis_weights(double x, const vector<double>& centers, bool periodic, double extent) {
  size_t n = centers.size();
  if (n == 1) return {{0, 1.0}};
  size_t hi = partition_point(n, [&](size_t k) { return centers[k] < x; });
  size_t lo;
  double left, right;
  if (hi == 0) {
    if (!periodic) return {{0, 1.0}};
    lo = n - 1;
    hi = 0;
    left = centers[n - 1] - extent;
    right = centers[0];
  } else if (hi == n) {
    if (!periodic) return {{n - 1, 1.0}};
    lo = n - 1;
    hi = 0;
    left = centers[n - 1];
    right = centers[0] + extent;
  } else {
    lo = hi - 1;
    left = centers[hi - 1];
    right = centers[hi];
  }
  double t = (x - left) / (right - left);
  vector<std::pair<size_t, double>> out;
  if (1.0 - t != 0.0) out.push_back({lo, 1.0 - t});
  if (t != 0.0) out.push_back({hi, t});
  return out;
}
// multigrid.rs:322-371 (Transfer::build: every axis of more than two cells;
// build_selected: the axes the caller selects)
Transfer transfer_build(const size_t n[3], const vector<double> centers[3], const bool periodic[3], const double extent[3],
                        vector<double> coarse[3], const bool* selected = nullptr) {
  bool coarsen[3];
  for (int a = 0; a < 3; ++a) coarsen[a] = selected ? selected[a] : n[a] > 2;
  size_t nc[3];
  for (int a = 0; a < 3; ++a) {
    coarse[a].clear();
    size_t step = coarsen[a] ? 2 : 1;
    for (size_t s = 0; s < centers[a].size(); s += step) {
      size_t e = std::min(centers[a].size(), s + step);
      double sum = neg_zero_sum();
      for (size_t k = s; k < e; ++k) sum += centers[a][k];
      coarse[a].push_back(sum / (double)(e - s));
    }
    nc[a] = coarse[a].size();
  }
  vector<vector<std::pair<size_t, double>>> weights[3];
  for (int a = 0; a < 3; ++a)
    for (double x : centers[a]) weights[a].push_back(axis_weights(x, coarse[a], periodic[a], extent[a]));
  Transfer p;
  p.offsets.push_back(0);
  size_t total = n[0] * n[1] * n[2];
  vector<std::pair<size_t, double>> row;
  for (size_t i = 0; i < total; ++i) {
    size_t q[3] = {i % n[0], i / n[0] % n[1], i / (n[0] * n[1])};
    row.clear();
    for (auto& z : weights[2][q[2]])
      for (auto& y : weights[1][q[1]])
        for (auto& x : weights[0][q[0]])
          row.push_back({(z.first * nc[1] + y.first) * nc[0] + x.first, x.second * y.second * z.second});
    std::sort(row.begin(), row.end(), [](auto& a, auto& b) { return a.first < b.first; });
    for (auto& e : row) {
      p.columns.push_back(e.first);
      p.weights.push_back(e.second);
    }
    p.offsets.push_back(p.columns.size());
  }
  return p;
}
// multigrid.rs:379-429
void galerkin(const Transfer& t, const vector<Edge>& edges, const vector<double>& anchors, size_t nc, vector<Edge>& out_edges,
              vector<double>& out_anchors) {
  std::map<std::pair<size_t, size_t>, double> off;
  vector<std::pair<size_t, double>> delta;
  for (auto& e : edges) {
    delta.clear();
    size_t a0 = t.offsets[e.i], a1 = t.offsets[e.i + 1];
    size_t b0 = t.offsets[e.j], b1 = t.offsets[e.j + 1];
    while (a0 < a1 || b0 < b1) {
      size_t a = a0 < a1 ? t.columns[a0] : NONE;
      size_t b = b0 < b1 ? t.columns[b0] : NONE;
      size_t col;
      double value;
      if (a == b) {
        col = a;
        value = t.weights[a0] - t.weights[b0];
        ++a0;
        ++b0;
      } else if (a < b) {
        col = a;
        value = t.weights[a0];
        ++a0;
      } else {
        col = b;
        value = -t.weights[b0];
        ++b0;
      }
      if (value != 0.0) delta.push_back({col, value});
    }
    for (size_t at = 0; at < delta.size(); ++at)
      for (size_t k = at + 1; k < delta.size(); ++k) {
        auto key = std::make_pair(delta[at].first, delta[k].first);
        auto it = off.find(key);
        if (it == off.end()) it = off.emplace(key, 0.0).first;
        it->second += e.w * delta[at].second * delta[k].second;
      }
  }
  out_anchors.assign(nc, 0.0);
  for (size_t i = 0; i < anchors.size(); ++i) {
    double w = anchors[i];
    if (w == 0.0) continue;
    delta.clear();
    for (size_t at = t.offsets[i]; at < t.offsets[i + 1]; ++at) delta.push_back({t.columns[at], t.weights[at]});
    for (size_t at = 0; at < delta.size(); ++at) {
      out_anchors[delta[at].first] += w * delta[at].second;
      for (size_t k = at + 1; k < delta.size(); ++k) {
        auto key = std::make_pair(delta[at].first, delta[k].first);
        auto it = off.find(key);
        if (it == off.end()) it = off.emplace(key, 0.0).first;
        it->second += w * delta[at].second * delta[k].second;
      }
    }
  }
  out_edges.clear();
  for (auto& kv : off)
    if (kv.second != 0.0) out_edges.push_back({kv.first.first, kv.first.second, -kv.second});
}

void center_plain(vector<double>& x, bool anchored) {  // multigrid.rs:446-453
  if (anchored) return;
  double sum = neg_zero_sum();
  for (double v : x) sum += v;
  double mean = sum / (double)x.size();
  for (double& v : x) v -= mean;
}

// multigrid/line.rs:25-190: exact banded solves of lines (or small planes) of
// one level, used as block Gauss-Seidel on anisotropic grids.
struct BlockSmoother {
  vector<int> axes;
  size_t length = 0, bandwidth = 0;
  vector<size_t> nodes, outside_offsets, outside_columns;
  vector<double> outside_values, factor, scratch;

  int build(const Matrix& m, const size_t n[3], const vector<int>& group, const bool periodic[3]) {
    axes = group;
    size_t count = m.diagonal.size();
    length = 1;
    for (int a : axes) length *= n[a];
    nodes.clear();
    for (size_t i = 0; i < count; ++i) {
      size_t q[3] = {i % n[0], i / n[0] % n[1], i / (n[0] * n[1])};
      bool base = true;
      for (int a : axes) base &= q[a] == 0;
      if (!base) continue;
      for (size_t at = 0; at < length; ++at) {
        // a periodic cycle is folded (0, n-1, 1, n-2, ...) to keep the band narrow
        size_t remainder = at;
        for (int axis : axes) {
          size_t k = remainder % n[axis];
          remainder /= n[axis];
          q[axis] = !periodic[axis] ? k : (k % 2 == 0 ? k / 2 : n[axis] - 1 - k / 2);
        }
        nodes.push_back((q[2] * n[1] + q[1]) * n[0] + q[0]);
      }
    }
    vector<size_t> position(count, 0);
    for (size_t at = 0; at < nodes.size(); ++at) position[nodes[at]] = at;
    bandwidth = 0;
    for (size_t i = 0; i < count; ++i) {
      size_t pi = position[i];
      for (size_t at = m.offsets[i]; at < m.offsets[i + 1]; ++at) {
        size_t pj = position[m.columns[at]];
        if (pi / length == pj / length) bandwidth = std::max(bandwidth, pi > pj ? pi - pj : pj - pi);
      }
    }
    size_t stride = bandwidth + 1;
    outside_offsets.assign(1, 0);
    outside_columns.clear();
    outside_values.clear();
    for (size_t at = 0; at < nodes.size(); ++at) {
      size_t cell = nodes[at];
      for (size_t e = m.offsets[cell]; e < m.offsets[cell + 1]; ++e) {
        size_t col = m.columns[e];
        if (position[col] / length != at / length) {
          outside_columns.push_back(col);
          outside_values.push_back(m.values[e]);
        }
      }
      outside_offsets.push_back(outside_columns.size());
    }
    factor.assign(count * stride, 0.0);
    for (size_t at = 0; at < nodes.size(); ++at) {
      size_t cell = nodes[at];
      factor[at * stride] = m.diagonal[cell];
      for (size_t e = m.offsets[cell]; e < m.offsets[cell + 1]; ++e) {
        size_t col = position[m.columns[e]];
        if (col / length == at / length && col < at) factor[at * stride + at - col] += m.values[e];
      }
    }
    for (size_t line = 0; line < count / length; ++line) {
      size_t base = line * length * stride;
      for (size_t i = 0; i < length; ++i)
        for (size_t j = (i > bandwidth ? i - bandwidth : 0); j <= i; ++j) {
          size_t row = base + i * stride, other = base + j * stride;
          double value = factor[row + i - j];
          for (size_t k = (i > bandwidth ? i - bandwidth : 0); k < j; ++k) value -= factor[row + i - k] * factor[other + j - k];
          if (i == j) {
            if (!std::isfinite(value) || value <= 0.0)
              return fail(36, "CFD-E036: numerics.pressure_preconditioner: multigrid line Cholesky lost positive definiteness");
            factor[row] = std::sqrt(value);
          } else {
            factor[row + i - j] = value / factor[other];
          }
        }
    }
    for (size_t row = 0; row < count; ++row) factor[row * stride] = 1.0 / factor[row * stride];
    scratch.assign(length, 0.0);
    return 0;
  }
  // returns the blocks solved (line.rs:155-190)
  int64_t sweep(const vector<double>& rhs, vector<double>& x, bool backward) {
    size_t count = nodes.size() / length, stride = bandwidth + 1;
    for (size_t at = 0; at < count; ++at) {
      size_t line = backward ? count - 1 - at : at;
      size_t start = line * length, base = start * stride;
      for (size_t i = 0; i < length; ++i) {
        double value = rhs[nodes[start + i]];
        for (size_t e = outside_offsets[start + i]; e < outside_offsets[start + i + 1]; ++e)
          value -= outside_values[e] * x[outside_columns[e]];
        scratch[i] = value;
      }
      for (size_t i = 0; i < length; ++i) {
        size_t row = base + i * stride;
        double value = scratch[i];
        for (size_t j = (i > bandwidth ? i - bandwidth : 0); j < i; ++j) value -= factor[row + i - j] * scratch[j];
        scratch[i] = value * factor[row];
      }
      for (size_t i = length; i-- > 0;) {
        double value = scratch[i];
        for (size_t j = i + 1; j < std::min(length, i + bandwidth + 1); ++j)
          value -= factor[base + j * stride + j - i] * scratch[j];
        scratch[i] = value * factor[base + i * stride];
      }
      for (size_t i = 0; i < length; ++i) x[nodes[start + i]] = scratch[i];
    }
    return (int64_t)count;
  }
};

struct Level {  // multigrid.rs:435-521
  Matrix matrix;
  Transfer transfer;
  vector<double> rhs, x, residual, cholesky;
  bool anchored = false;
  vector<BlockSmoother> lines;
  Level(Matrix m, Transfer t, bool anch) : matrix(std::move(m)), transfer(std::move(t)), anchored(anch) {
    size_t n = matrix.diagonal.size();
    rhs.assign(n, 0.0);
    x.assign(n, 0.0);
    residual.assign(n, 0.0);
  }
  int factor() {
    size_t skip = anchored ? 0 : 1;
    size_t n = x.size() - skip;
    vector<double> l(n * n, 0.0);
    for (size_t row = skip; row < x.size(); ++row) {
      size_t i = row - skip;
      l[i * n + i] = matrix.diagonal[row];
      for (size_t at = matrix.offsets[row]; at < matrix.offsets[row + 1]; ++at) {
        size_t col = matrix.columns[at];
        if (col >= skip) l[i * n + col - skip] += matrix.values[at];
      }
    }
    for (size_t i = 0; i < n; ++i)
      for (size_t j = 0; j <= i; ++j) {
        double value = l[i * n + j];
        for (size_t k = 0; k < j; ++k) value -= l[i * n + k] * l[j * n + k];
        if (i == j) {
          if (!std::isfinite(value) || value <= 0.0)
            return fail(35, "CFD-E035: numerics.pressure_preconditioner: multigrid coarse Cholesky lost positive definiteness");
          l[i * n + j] = std::sqrt(value);
        } else {
          l[i * n + j] = value / l[j * n + j];
        }
      }
    cholesky = std::move(l);
    return 0;
  }
  void direct() {
    size_t skip = anchored ? 0 : 1;
    size_t n = x.size() - skip;
    std::fill(x.begin(), x.end(), 0.0);
    for (size_t i = 0; i < n; ++i) {
      double value = rhs[i + skip];
      for (size_t j = 0; j < i; ++j) value -= cholesky[i * n + j] * x[j + skip];
      x[i + skip] = value / cholesky[i * n + i];
    }
    for (size_t i = n; i-- > 0;) {
      double value = x[i + skip];
      for (size_t j = i + 1; j < n; ++j) value -= cholesky[j * n + i] * x[j + skip];
      x[i + skip] = value / cholesky[i * n + i];
    }
    center_plain(x, anchored);
  }
};

// masked.rs:16-70
struct Direct {
  vector<size_t> indices;
  vector<double> factor;
};
int direct_build(const Matrix& a, const vector<size_t>& component, bool anchored, Direct& d) {
  d.indices.assign(component.begin() + (anchored ? 0 : 1), component.end());
  size_t n = d.indices.size();
  std::map<size_t, size_t> local;
  for (size_t j = 0; j < n; ++j) local[d.indices[j]] = j;
  d.factor.assign(n * n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    size_t row = d.indices[i];
    d.factor[i * n + i] = a.diagonal[row];
    for (size_t at = a.offsets[row]; at < a.offsets[row + 1]; ++at) {
      auto it = local.find(a.columns[at]);
      if (it != local.end()) d.factor[i * n + it->second] += a.values[at];
    }
  }
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j <= i; ++j) {
      double v = d.factor[i * n + j];
      for (size_t k = 0; k < j; ++k) v -= d.factor[i * n + k] * d.factor[j * n + k];
      if (i == j) {
        if (!std::isfinite(v) || v <= 0.0)
          return fail(35, "CFD-E035: numerics.pressure_preconditioner: masked coarse Cholesky lost positive definiteness");
        d.factor[i * n + j] = std::sqrt(v);
      } else {
        d.factor[i * n + j] = v / d.factor[j * n + j];
      }
    }
  return 0;
}
void direct_solve(const Direct& d, const vector<double>& rhs, vector<double>& x) {
  size_t n = d.indices.size();
  for (size_t i = 0; i < n; ++i) {
    size_t row = d.indices[i];
    double v = rhs[row];
    for (size_t j = 0; j < i; ++j) v -= d.factor[i * n + j] * x[d.indices[j]];
    x[row] = v / d.factor[i * n + i];
  }
  for (size_t i = n; i-- > 0;) {
    size_t row = d.indices[i];
    double v = x[row];
    for (size_t j = i + 1; j < n; ++j) v -= d.factor[j * n + i] * x[d.indices[j]];
    x[row] = v / d.factor[i * n + i];
  }
}

struct MaskedLevel {  // masked.rs:72-81
  Matrix a;
  vector<size_t> parent;
  vector<vector<size_t>> closed;
  vector<double> rhs, x, residual;
  vector<Direct> direct;
  void sweep(bool backward) {  // masked.rs:158-172
    size_t n = x.size();
    for (size_t at = 0; at < n; ++at) {
      size_t i = backward ? n - 1 - at : at;
      if (a.diagonal[i] == 0.0) {
        x[i] = 0.0;
        continue;
      }
      double v = rhs[i];
      for (size_t k = a.offsets[i]; k < a.offsets[i + 1]; ++k) v -= a.values[k] * x[a.columns[k]];
      x[i] = v / a.diagonal[i];
    }
  }
};
void center_components(vector<double>& x, const vector<vector<size_t>>& closed) {  // masked.rs:82-89
  for (auto& cells : closed) {
    double sum = neg_zero_sum();
    for (size_t i : cells) sum += x[i];
    double mean = sum / (double)cells.size();
    for (size_t i : cells) x[i] -= mean;
  }
}

// MultigridWork (multigrid.rs:78-114), counted where the cycles increment it
struct Work {
  int64_t cycles = 0, smoothing = 0, residual = 0, coarse = 0;
  int64_t restriction = 0, prolongation = 0, block_smoothing = 0, block_solves = 0, block_visits = 0;
};

struct Hierarchy {
  bool masked = false;
  vector<Level> levels;
  vector<MaskedLevel> mlevels;
  vector<size_t> fluid_indices;
  vector<int64_t> level_cells, level_nonzeros;
  vector<std::array<int64_t, 3>> level_dimensions;  // unmasked only (multigrid.rs:562-563)
};

void cycle_plain(vector<Level>& levels, size_t at, Work& work) {  // multigrid.rs:667-726
  Level& level = levels[at];
  center_plain(level.rhs, level.anchored);
  if (at + 1 == levels.size()) {
    level.direct();
    work.coarse += 1;
    return;
  }
  std::fill(level.x.begin(), level.x.end(), 0.0);
  // multigrid.rs:676-688: one block sweep per side on a level with block smoothers
  const int sweeps = level.lines.empty() ? SWEEPS : 1;
  if (!level.lines.empty()) work.block_visits += (int64_t)level.x.size();
  for (int pass = 0; pass < sweeps; ++pass) {
    if (level.lines.empty()) {
      level.matrix.sweep(level.rhs, level.x, false);
    } else {
      work.block_solves += level.lines[(size_t)pass % level.lines.size()].sweep(level.rhs, level.x, false);
      work.block_smoothing += (int64_t)level.x.size();
    }
  }
  center_plain(level.x, level.anchored);
  level.matrix.product(level.x, level.residual);
  for (size_t i = 0; i < level.residual.size(); ++i) level.residual[i] = level.rhs[i] - level.residual[i];
  center_plain(level.residual, level.anchored);
  Level& next = levels[at + 1];
  std::fill(next.rhs.begin(), next.rhs.end(), 0.0);
  for (size_t i = 0; i < level.residual.size(); ++i) {
    double r = level.residual[i];
    for (size_t k = level.transfer.offsets[i]; k < level.transfer.offsets[i + 1]; ++k)
      next.rhs[level.transfer.columns[k]] += level.transfer.weights[k] * r;
  }
  cycle_plain(levels, at + 1, work);
  Level& lv = levels[at];
  Level& nx = levels[at + 1];
  for (size_t i = 0; i < lv.x.size(); ++i) {
    double s = neg_zero_sum();
    for (size_t k = lv.transfer.offsets[i]; k < lv.transfer.offsets[i + 1]; ++k)
      s += lv.transfer.weights[k] * nx.x[lv.transfer.columns[k]];
    lv.x[i] += s;
  }
  const int lsweeps = lv.lines.empty() ? SWEEPS : 1;
  for (int pass = 0; pass < lsweeps; ++pass) {
    if (lv.lines.empty()) {
      lv.matrix.sweep(lv.rhs, lv.x, true);
    } else {
      work.block_solves += lv.lines[(size_t)(lsweeps - 1 - pass) % lv.lines.size()].sweep(lv.rhs, lv.x, true);
      work.block_smoothing += (int64_t)lv.x.size();
    }
  }
  center_plain(lv.x, lv.anchored);
  int64_t cells = (int64_t)lv.x.size();
  work.smoothing += 2 * lsweeps * cells;
  work.residual += cells;
  work.restriction += cells;
  work.prolongation += cells;
}

void cycle_masked(vector<MaskedLevel>& levels, size_t at, Work& work) {  // masked.rs:304-342
  MaskedLevel& level = levels[at];
  center_components(level.rhs, level.closed);
  std::fill(level.x.begin(), level.x.end(), 0.0);
  if (at + 1 == levels.size()) {
    for (auto& d : level.direct) direct_solve(d, level.rhs, level.x);
    center_components(level.x, level.closed);
    work.coarse += 1;
    return;
  }
  for (int s = 0; s < SWEEPS; ++s) level.sweep(false);
  center_components(level.x, level.closed);
  level.a.product(level.x, level.residual);
  for (size_t i = 0; i < level.residual.size(); ++i) level.residual[i] = level.rhs[i] - level.residual[i];
  center_components(level.residual, level.closed);
  MaskedLevel& next = levels[at + 1];
  std::fill(next.rhs.begin(), next.rhs.end(), 0.0);
  for (size_t i = 0; i < level.residual.size(); ++i) next.rhs[level.parent[i]] += level.residual[i];
  cycle_masked(levels, at + 1, work);
  MaskedLevel& lv = levels[at];
  MaskedLevel& nx = levels[at + 1];
  for (size_t i = 0; i < lv.x.size(); ++i) lv.x[i] += nx.x[lv.parent[i]];
  for (int s = 0; s < SWEEPS; ++s) lv.sweep(true);
  center_components(lv.x, lv.closed);
  int64_t n = (int64_t)lv.x.size();
  work.smoothing += 2 * SWEEPS * n;
  work.residual += n;
  work.restriction += n;
  work.prolongation += n;
}

}  // namespace

// ---------------------------------------------------------------- Grid
struct orc_grid {
  size_t n[3];
  double dx[3];
  double volume;
  int boundary[3];
  int preconditioner;
  vector<char> fluid;
  vector<Face> faces;
  vector<std::array<size_t, 6>> cell_faces;  // [axis*2 + side]
  vector<size_t> component;
  vector<vector<size_t>> components;
  vector<char> anchored;
  size_t face_dims[3][3];
  size_t count() const { return fluid.size(); }
  orc_case flow;  // flow description for inlet values (pointers cleared)
  std::unique_ptr<Factor> factor;
  std::unique_ptr<Hierarchy> hierarchy;
  std::unique_ptr<Hierarchy> masked;  // masked.rs hierarchy of any grid (inspection)
  double weight(const Face& f) const { return f.area / (f.distance * volume); }
};

namespace {
// grid.rs:246-464
orc_grid* grid_build(const orc_case& c) {
  auto g = std::make_unique<orc_grid>();
  for (int a = 0; a < 3; ++a) {
    if (c.cells[a] < 1) {
      fail(4, "CFD-E004: grid.cells must be positive");
      return nullptr;
    }
    g->n[a] = (size_t)c.cells[a];
    g->boundary[a] = c.boundary[a];
  }
  g->preconditioner = c.preconditioner;
  g->flow = c;
  g->flow.blocks = nullptr;
  g->flow.n_blocks = 0;
  const size_t* n = g->n;
  for (int a = 0; a < 3; ++a) g->dx[a] = c.size_m[a] / (double)n[a];
  g->volume = 1.0 * g->dx[0] * g->dx[1] * g->dx[2];
  size_t count = n[0] * n[1] * n[2];
  g->fluid.assign(count, 1);
  g->cell_faces.assign(count, std::array<size_t, 6>{NONE, NONE, NONE, NONE, NONE, NONE});
  g->component.assign(count, NONE);
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d) g->face_dims[a][d] = n[d];
  for (int b = 0; b < c.n_blocks; ++b) {
    Raster r(c.blocks[b], n, g->dx);
    r.for_each(n, [&](size_t i) { g->fluid[i] = 0; });
    if (r.count() == 0) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "CFD-E028: geometry[%d] (%s) contains no cell center, so the grid does not represent it", b,
                    c.blocks[b].n_vertices > 0 ? "prism" : "box");
      fail(28, buf);
      return nullptr;
    }
  }
  bool any = false;
  for (char f : g->fluid) any |= f != 0;
  if (!any) {
    fail(30, "CFD-E030: geometry contains no fluid cells");
    return nullptr;
  }
  for (int a = 0; a < 3; ++a) {
    bool periodic = c.boundary[a] == ORC_PERIODIC;
    if (!periodic) g->face_dims[a][a] += 1;
    size_t dims[3] = {g->face_dims[a][0], g->face_dims[a][1], g->face_dims[a][2]};
    for (size_t k = 0; k < dims[2]; ++k)
      for (size_t j = 0; j < dims[1]; ++j)
        for (size_t i = 0; i < dims[0]; ++i) {
          size_t q[3] = {i, j, k};
          size_t t = q[a];
          size_t lo0 = NONE, hi0 = NONE;
          if (t > 0) {
            size_t p[3] = {q[0], q[1], q[2]};
            p[a] -= 1;
            lo0 = (p[2] * n[1] + p[1]) * n[0] + p[0];
          } else if (periodic) {
            size_t p[3] = {q[0], q[1], q[2]};
            p[a] = n[a] - 1;
            lo0 = (p[2] * n[1] + p[1]) * n[0] + p[0];
          }
          if (t < n[a]) hi0 = (q[2] * n[1] + q[1]) * n[0] + q[0];
          size_t lo = (lo0 != NONE && g->fluid[lo0]) ? lo0 : NONE;
          size_t hi = (hi0 != NONE && g->fluid[hi0]) ? hi0 : NONE;
          double center_z = ((double)q[2] + (a == 2 ? 0.0 : 0.5)) * g->dx[2];
          int kind;
          if (lo != NONE && hi != NONE)
            kind = ORC_INTERIOR;
          else if (lo == NONE && hi == NONE)
            kind = ORC_BLOCKED;
          else if (lo0 != NONE && hi0 != NONE)
            kind = ORC_WALLF;
          else {
            switch (c.boundary[a]) {
              case ORC_WALL:
                kind = ORC_WALLF;
                break;
              case ORC_SLIP:
                kind = ORC_SLIPF;
                break;
              case ORC_GROUND_OPEN_TOP:
                kind = t == 0 ? ORC_WALLF : ORC_SLIPF;
                break;
              case ORC_OPEN: {
                double normal = t == 0 ? -1.0 : 1.0;
                double v3[3];
                flow_vector(c, center_z, v3);
                double v = v3[a] * normal;
                kind = v < -1e-12 ? ORC_INLET : (v > 1e-12 ? ORC_OUTLET : ORC_SLIPF);
                break;
              }
              default:
                kind = ORC_BLOCKED;
            }
          }
          size_t f = g->faces.size();
          if (lo != NONE) g->cell_faces[lo][a * 2 + 1] = f;
          if (hi != NONE) g->cell_faces[hi][a * 2 + 0] = f;
          Face face;
          face.axis = a;
          face.coord[0] = i;
          face.coord[1] = j;
          face.coord[2] = k;
          face.center_z = center_z;
          face.lo = lo;
          face.hi = hi;
          face.kind = kind;
          face.area = g->volume / g->dx[a];
          face.distance = kind == ORC_INTERIOR ? g->dx[a] : 0.5 * g->dx[a];
          g->faces.push_back(face);
        }
  }
  for (size_t start = 0; start < count; ++start) {
    if (!g->fluid[start] || g->component[start] != NONE) continue;
    size_t id = g->components.size();
    vector<size_t> members;
    std::deque<size_t> queue{start};
    g->component[start] = id;
    bool anch = false;
    while (!queue.empty()) {
      size_t i = queue.front();
      queue.pop_front();
      members.push_back(i);
      for (int s = 0; s < 6; ++s) {
        const Face& f = g->faces[g->cell_faces[i][s]];
        anch |= f.kind == ORC_OUTLET;
        if (f.kind != ORC_INTERIOR) continue;
        size_t j = f.lo == i ? f.hi : f.lo;
        if (g->component[j] == NONE) {
          g->component[j] = id;
          queue.push_back(j);
        }
      }
    }
    g->components.push_back(std::move(members));
    g->anchored.push_back(anch ? 1 : 0);
  }
  return g.release();
}

// pressure.rs:118-222
std::unique_ptr<Factor> factor_build(const orc_grid& g) {
  size_t count = g.count();
  auto F = std::make_unique<Factor>();
  vector<size_t>& row_offsets = F->row_offsets;
  row_offsets.assign(count + 1, 0);
  for (auto& f : g.faces)
    if (f.kind == ORC_INTERIOR && f.lo != f.hi) row_offsets[std::max(f.lo, f.hi) + 1] += 1;
  for (size_t row = 0; row < count; ++row) row_offsets[row + 1] += row_offsets[row];
  size_t capacity = row_offsets[count];
  vector<size_t>& columns = F->columns;
  vector<double>& lower = F->lower;
  columns.assign(capacity, 0);
  lower.assign(capacity, 0.0);
  vector<size_t> counts(count, 0);
  vector<double>& diag = F->diagonal;
  diag.assign(count, 0.0);
  for (auto& f : g.faces) {
    double w = g.weight(f);
    if (f.kind == ORC_INTERIOR) {
      if (f.lo == f.hi) continue;
      size_t i = f.lo, j = f.hi;
      diag[i] += w;
      diag[j] += w;
      size_t row = i > j ? i : j, col = i > j ? j : i;
      size_t start = row_offsets[row], end = start + counts[row];
      size_t found = NONE;
      for (size_t k = start; k < end; ++k)
        if (columns[k] == col) {
          found = k;
          break;
        }
      if (found != NONE) {
        lower[found] -= w;
      } else {
        columns[end] = col;
        lower[end] = -w;
        counts[row] += 1;
      }
    } else if (f.kind == ORC_OUTLET) {
      diag[f.lo != NONE ? f.lo : f.hi] += w;
    }
  }
  size_t write = 0, read = 0;
  for (size_t row = 0; row < count; ++row) {
    size_t next = row_offsets[row + 1];
    size_t end = read + counts[row];
    for (size_t offset = read + 1; offset < end; ++offset) {
      size_t column = columns[offset];
      double value = lower[offset];
      size_t at = offset;
      while (at > read && columns[at - 1] > column) {
        columns[at] = columns[at - 1];
        lower[at] = lower[at - 1];
        at -= 1;
      }
      columns[at] = column;
      lower[at] = value;
    }
    row_offsets[row] = write;
    for (size_t offset = read; offset < end; ++offset) {
      columns[write] = columns[offset];
      lower[write] = lower[offset];
      write += 1;
    }
    read = next;
  }
  row_offsets[count] = write;
  columns.resize(write);
  lower.resize(write);
  double maxdiag = 1.0;
  for (double d : diag) maxdiag = std::fmax(maxdiag, d);
  for (size_t row = 0; row < count; ++row) {
    size_t start = row_offsets[row], end = row_offsets[row + 1];
    for (size_t offset = start; offset < end; ++offset) {
      size_t col = columns[offset];
      double dot = 0.0;
      size_t other = row_offsets[col], other_end = row_offsets[col + 1];
      for (size_t previous = start; previous < offset; ++previous) {
        size_t k = columns[previous];
        while (other < other_end && columns[other] < k) other += 1;
        if (other < other_end && columns[other] == k) dot += lower[previous] * lower[other];
      }
      lower[offset] = (lower[offset] - dot) / diag[col];
    }
    double squared = neg_zero_sum();
    for (size_t o = start; o < end; ++o) squared += lower[o] * lower[o];
    diag[row] = std::sqrt(std::fmax(diag[row] + 1e-6 * maxdiag - squared, 1e-12 * maxdiag));
  }
  return F;
}
// pressure.rs:223-241
void factor_apply(const Factor& F, const double* r, double* z, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    double value = r[i];
    for (size_t k = F.row_offsets[i]; k < F.row_offsets[i + 1]; ++k) value -= F.lower[k] * z[F.columns[k]];
    z[i] = value / F.diagonal[i];
  }
  for (size_t i = n; i-- > 0;) {
    double value = z[i] / F.diagonal[i];
    z[i] = value;
    for (size_t k = F.row_offsets[i]; k < F.row_offsets[i + 1]; ++k) z[F.columns[k]] -= F.lower[k] * value;
  }
}

// multigrid.rs:33-47: the axes whose face weight is within a factor 4 of the
// largest, when the weights spread by more than 25 over the axes of more than
// two cells; empty otherwise.
vector<int> line_axes(const orc_grid& g) {
  double weights[3];
  for (int a = 0; a < 3; ++a) weights[a] = 1.0 / (g.dx[a] * g.dx[a]);
  for (double w : weights)
    if (!std::isfinite(w) || w <= 0.0) return {};
  double minimum = INFINITY;
  for (int a = 0; a < 3; ++a)
    if (g.n[a] > 2) minimum = std::fmin(minimum, weights[a]);
  double maximum = 0.0;
  for (double w : weights) maximum = std::fmax(maximum, w);
  if (maximum / minimum <= 25.0) return {};
  vector<int> out;
  for (int a = 0; a < 3; ++a)
    if (weights[a] >= maximum / 4.0) out.push_back(a);
  return out;
}

// masked.rs:90-126
void masked_components(const Matrix& a, const vector<double>& anchors, vector<vector<size_t>>& result, vector<char>& anchored) {
  vector<char> seen(anchors.size(), 0);
  for (size_t start = 0; start < seen.size(); ++start) {
    if (seen[start]) continue;
    seen[start] = 1;
    std::deque<size_t> queue{start};
    vector<size_t> cells;
    bool fixed = false;
    while (!queue.empty()) {
      size_t i = queue.front();
      queue.pop_front();
      cells.push_back(i);
      fixed |= anchors[i] > 0.0;
      for (size_t k = a.offsets[i]; k < a.offsets[i + 1]; ++k) {
        size_t j = a.columns[k];
        if (!seen[j]) {
          seen[j] = 1;
          queue.push_back(j);
        }
      }
    }
    result.push_back(std::move(cells));
    anchored.push_back(fixed ? 1 : 0);
  }
}
// masked.rs:127-156
void masked_aggregate(const Matrix& a, const vector<std::array<size_t, 3>>& coords, vector<size_t>& parent,
                      vector<std::array<size_t, 3>>& coarse) {
  size_t n = coords.size();
  vector<std::array<size_t, 3>> bins(n);
  for (size_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) bins[i][d] = coords[i][d] / 2;
  parent.assign(n, NONE);
  coarse.clear();
  for (size_t start = 0; start < n; ++start) {
    if (parent[start] != NONE) continue;
    size_t id = coarse.size();
    coarse.push_back(bins[start]);
    parent[start] = id;
    std::deque<size_t> queue{start};
    while (!queue.empty()) {
      size_t i = queue.front();
      queue.pop_front();
      for (size_t k = a.offsets[i]; k < a.offsets[i + 1]; ++k) {
        size_t j = a.columns[k];
        if (parent[j] == NONE && bins[j] == bins[start]) {
          parent[j] = id;
          queue.push_back(j);
        }
      }
    }
  }
}

// masked.rs:181-291
int masked_build(const orc_grid& g, Hierarchy& h) {
  h.masked = true;
  vector<size_t> compact(g.count(), NONE);
  for (size_t i = 0; i < g.count(); ++i)
    if (g.fluid[i]) {
      compact[i] = h.fluid_indices.size();
      h.fluid_indices.push_back(i);
    }
  vector<std::array<size_t, 3>> coords;
  for (size_t i : h.fluid_indices) coords.push_back({i % g.n[0], i / g.n[0] % g.n[1], i / (g.n[0] * g.n[1])});
  vector<double> anchors(coords.size(), 0.0);
  vector<Edge> edges;
  for (auto& f : g.faces) {
    double w = g.weight(f);
    if (f.kind == ORC_INTERIOR && f.lo != f.hi)
      edges.push_back({compact[f.lo], compact[f.hi], w});
    else if (f.kind == ORC_OUTLET)
      anchors[compact[f.lo != NONE ? f.lo : f.hi]] += w;
  }
  for (;;) {
    Matrix a = Matrix::build(edges, anchors);
    vector<vector<size_t>> regions;
    vector<char> anch;
    masked_components(a, anchors, regions, anch);
    MaskedLevel level;
    for (size_t r = 0; r < regions.size(); ++r)
      if (!anch[r]) level.closed.push_back(regions[r]);
    size_t n = coords.size();
    level.a = std::move(a);
    level.rhs.assign(n, 0.0);
    level.x.assign(n, 0.0);
    level.residual.assign(n, 0.0);
    bool small = true;
    for (auto& r : regions) small &= r.size() <= COARSE_CELLS;
    if (small) {
      for (size_t r = 0; r < regions.size(); ++r) {
        Direct d;
        int e = direct_build(level.a, regions[r], anch[r] != 0, d);
        if (e) return e;
        level.direct.push_back(std::move(d));
      }
      h.mlevels.push_back(std::move(level));
      break;
    }
    vector<size_t> parent;
    vector<std::array<size_t, 3>> coarse;
    masked_aggregate(level.a, coords, parent, coarse);
    if (coarse.size() == n) {
      bool all_zero = true;
      for (auto& p : coords) {
        for (int d = 0; d < 3; ++d) p[d] /= 2;
        all_zero &= p[0] == 0 && p[1] == 0 && p[2] == 0;
      }
      if (all_zero) return fail(35, "CFD-E035: numerics.pressure_preconditioner: masked aggregation failed to reduce a connected component");
      continue;
    }
    std::map<std::pair<size_t, size_t>, double> sums;
    for (auto& e : edges) {
      size_t i = parent[e.i], j = parent[e.j];
      if (i != j) {
        auto key = std::make_pair(std::min(i, j), std::max(i, j));
        auto it = sums.find(key);
        if (it == sums.end()) it = sums.emplace(key, 0.0).first;
        it->second += e.w;
      }
    }
    edges.clear();
    for (auto& kv : sums) edges.push_back({kv.first.first, kv.first.second, kv.second});
    vector<double> ac(coarse.size(), 0.0);
    for (size_t i = 0; i < anchors.size(); ++i) ac[parent[i]] += anchors[i];
    anchors = std::move(ac);
    coords = std::move(coarse);
    level.parent = std::move(parent);
    h.mlevels.push_back(std::move(level));
  }
  for (auto& l : h.mlevels) {
    h.level_cells.push_back((int64_t)l.x.size());
    h.level_nonzeros.push_back((int64_t)(l.a.values.size() + l.x.size()));
  }
  return 0;
}

// multigrid.rs:530-645 (point smoothing only)
int hierarchy_build(const orc_grid& g, std::unique_ptr<Hierarchy>& out) {
  if (g.components.empty()) return fail(34, "CFD-E034: numerics.pressure_preconditioner: multigrid requires fluid cells");
  auto h = std::make_unique<Hierarchy>();
  bool any_solid = false;
  for (char f : g.fluid) any_solid |= !f;
  if (any_solid || g.components.size() != 1) {
    int e = masked_build(g, *h);
    if (e) return e;
    out = std::move(h);
    return 0;
  }
  const vector<int> axes = line_axes(g);
  size_t n[3] = {g.n[0], g.n[1], g.n[2]};
  vector<double> centers[3];
  double extent[3];
  bool periodic[3];
  for (int a = 0; a < 3; ++a) {
    for (size_t i = 0; i < n[a]; ++i) centers[a].push_back((double)i + 0.5);
    extent[a] = (double)n[a];
    periodic[a] = false;
    for (auto& f : g.faces)
      if (f.axis == a && f.coord[a] == 0 && f.kind == ORC_INTERIOR) {
        periodic[a] = true;
        break;
      }
  }
  bool anch = g.anchored[0] != 0;
  vector<double> anchors(g.count(), 0.0);
  vector<Edge> edges;
  for (auto& f : g.faces) {
    double w = g.weight(f);
    if (f.kind == ORC_INTERIOR)
      edges.push_back({f.lo, f.hi, w});
    else if (f.kind == ORC_OUTLET)
      anchors[f.lo != NONE ? f.lo : f.hi] += w;
  }
  for (;;) {
    h->level_dimensions.push_back({(int64_t)n[0], (int64_t)n[1], (int64_t)n[2]});
    Matrix matrix = Matrix::build(edges, anchors);
    size_t count = anchors.size();
    if (count <= COARSE_CELLS) {
      Level last(std::move(matrix), Transfer(), anch);
      int e = last.factor();
      if (e) return e;
      h->levels.push_back(std::move(last));
      break;
    }
    vector<double> coarse[3];
    Transfer transfer;
    vector<vector<int>> groups;
    if (axes.empty()) {
      transfer = transfer_build(n, centers, periodic, extent, coarse);
    } else {
      // multigrid.rs:582-611: keep the weak axes until the stiff ones have collapsed
      bool stiff_can_coarsen = false;
      for (int a : axes) stiff_can_coarsen |= n[a] > 2;
      bool sel[3];
      for (int a = 0; a < 3; ++a) {
        bool in_axes = std::find(axes.begin(), axes.end(), a) != axes.end();
        sel[a] = n[a] > 2 && (!stiff_can_coarsen || in_axes);
      }
      transfer = transfer_build(n, centers, periodic, extent, coarse, sel);
      size_t span = 1;
      for (int a : axes) span *= n[a];
      if (axes.size() > 1 && span <= 16) {
        groups.push_back(axes);
      } else {
        int best = axes[0];
        for (int a : axes)
          if (g.dx[a] < g.dx[best]) best = a;
        groups.push_back({best});
      }
    }
    Level level(std::move(matrix), Transfer(), anch);
    for (auto& grp : groups) {
      BlockSmoother b;
      int e = b.build(level.matrix, n, grp, periodic);
      if (e) return e;
      level.lines.push_back(std::move(b));
    }
    for (int a = 0; a < 3; ++a) n[a] = coarse[a].size();
    vector<Edge> ce;
    vector<double> ca;
    galerkin(transfer, edges, anchors, n[0] * n[1] * n[2], ce, ca);
    edges = std::move(ce);
    for (int a = 0; a < 3; ++a) centers[a] = coarse[a];
    level.transfer = std::move(transfer);
    h->levels.push_back(std::move(level));
    anchors = std::move(ca);
  }
  for (auto& l : h->levels) {
    h->level_cells.push_back((int64_t)l.x.size());
    h->level_nonzeros.push_back((int64_t)(l.matrix.values.size() + l.x.size()));
  }
  out = std::move(h);
  return 0;
}

void hierarchy_apply(Hierarchy& h, const double* rhs, double* z, size_t count, Work& work) {
  if (h.masked) {
    for (size_t j = 0; j < h.fluid_indices.size(); ++j) h.mlevels[0].rhs[j] = rhs[h.fluid_indices[j]];
    cycle_masked(h.mlevels, 0, work);
    std::fill(z, z + count, 0.0);
    for (size_t j = 0; j < h.fluid_indices.size(); ++j) z[h.fluid_indices[j]] = h.mlevels[0].x[j];
  } else {
    std::copy(rhs, rhs + count, h.levels[0].rhs.begin());
    cycle_plain(h.levels, 0, work);
    std::copy(h.levels[0].x.begin(), h.levels[0].x.end(), z);
  }
  work.cycles += 1;
}

struct OperatorEntry {
  bool interior;
  size_t lo, hi;
  double weight;
};

double dot(const vector<double>& x, const vector<double>& y) {
  double s = neg_zero_sum();
  for (size_t i = 0; i < x.size(); ++i) s += x[i] * y[i];
  return s;
}

}  // namespace

// ---------------------------------------------------------------- C ABI
extern "C" {

const char* orc_last_error(void) { return g_error.c_str(); }

orc_grid* orc_grid_build(const orc_case* c) { return grid_build(*c); }
void orc_grid_free(orc_grid* g) { delete g; }

void orc_grid_counts(const orc_grid* g, int64_t* counts) {
  size_t fluid = 0, anch = 0;
  for (char f : g->fluid) fluid += f ? 1 : 0;
  for (char a : g->anchored) anch += a ? 1 : 0;
  counts[0] = (int64_t)g->count();
  counts[1] = (int64_t)g->faces.size();
  counts[2] = (int64_t)fluid;
  counts[3] = (int64_t)g->components.size();
  counts[4] = (int64_t)anch;
  for (int a = 0; a < 3; ++a) counts[5 + a] = (int64_t)(g->face_dims[a][0] * g->face_dims[a][1] * g->face_dims[a][2]);
}

void orc_grid_faces(const orc_grid* g, int32_t* axis, int64_t* lo, int64_t* hi, int32_t* kind, double* weight) {
  for (size_t f = 0; f < g->faces.size(); ++f) {
    const Face& F = g->faces[f];
    if (axis) axis[f] = F.axis;
    if (lo) lo[f] = F.lo == NONE ? -1 : (int64_t)F.lo;
    if (hi) hi[f] = F.hi == NONE ? -1 : (int64_t)F.hi;
    if (kind) kind[f] = F.kind;
    if (weight) weight[f] = g->weight(F);
  }
}

void orc_grid_fluid(const orc_grid* g, uint8_t* fluid) {
  for (size_t i = 0; i < g->count(); ++i) fluid[i] = g->fluid[i] ? 1 : 0;
}

void orc_grid_components(const orc_grid* g, int64_t* component, uint8_t* anchored) {
  for (size_t i = 0; i < g->count(); ++i) component[i] = g->component[i] == NONE ? -1 : (int64_t)g->component[i];
  for (size_t c = 0; c < g->components.size(); ++c) anchored[c] = g->anchored[c] ? 1 : 0;
}

void orc_apply(const orc_grid* g, const double* x, double* y) {
  std::fill(y, y + g->count(), 0.0);
  for (auto& f : g->faces) {
    double w = g->weight(f);
    if (f.kind == ORC_INTERIOR) {
      double d = w * (x[f.lo] - x[f.hi]);
      y[f.lo] += d;
      y[f.hi] -= d;
    } else if (f.kind == ORC_OUTLET) {
      size_t i = f.lo != NONE ? f.lo : f.hi;
      y[i] += w * x[i];
    }
  }
}

}  // extern "C"

namespace {
// pressure.rs:254-272 Operator::build: the interior and outlet faces in face order
// (built once per solve, as the reference's solve_scaled does at 383-386)
vector<OperatorEntry> operator_build(const orc_grid* g) {
  vector<OperatorEntry> entries;
  entries.reserve(g->faces.size());
  for (auto& f : g->faces) {
    double w = g->weight(f);
    if (f.kind == ORC_INTERIOR)
      entries.push_back({true, f.lo, f.hi, w});
    else if (f.kind == ORC_OUTLET)
      entries.push_back({false, f.lo != NONE ? f.lo : f.hi, 0, w});
  }
  return entries;
}
// pressure.rs:273-285 Operator::apply
void operator_entries_apply(const vector<OperatorEntry>& entries, size_t count, const double* x, double* y) {
  std::fill(y, y + count, 0.0);
  for (auto& e : entries) {
    if (e.interior) {
      double d = e.weight * (x[e.lo] - x[e.hi]);
      y[e.lo] += d;
      y[e.hi] -= d;
    } else {
      y[e.lo] += e.weight * x[e.lo];
    }
  }
}
}  // namespace

extern "C" {

void orc_operator_apply(const orc_grid* g, const double* x, double* y) {
  operator_entries_apply(operator_build(g), g->count(), x, y);
}

void orc_center_nullspaces(const orc_grid* g, double* x) {
  for (size_t id = 0; id < g->components.size(); ++id) {
    if (g->anchored[id]) continue;
    const auto& m = g->components[id];
    double sum = neg_zero_sum();
    for (size_t i : m) sum += x[i];
    double mean = sum / (double)m.size();
    for (size_t i : m) x[i] -= mean;
  }
}

void orc_ic0_diagonal(orc_grid* g, double* diag) {
  if (!g->factor) g->factor = factor_build(*g);
  std::copy(g->factor->diagonal.begin(), g->factor->diagonal.end(), diag);
}

void orc_ic0_apply(orc_grid* g, const double* r, double* z) {
  if (!g->factor) g->factor = factor_build(*g);
  factor_apply(*g->factor, r, z, g->count());
}

int orc_mg_apply(orc_grid* g, const double* r, double* z) {
  if (!g->hierarchy) {
    int e = hierarchy_build(*g, g->hierarchy);
    if (e) return e;
  }
  Work w;
  hierarchy_apply(*g->hierarchy, r, z, g->count(), w);
  return 0;
}

int orc_mg_stats(orc_grid* g, int64_t* level_cells, int64_t* level_nonzeros, int cap, int* masked) {
  if (!g->hierarchy) {
    int e = hierarchy_build(*g, g->hierarchy);
    if (e) return -e;
  }
  const Hierarchy& h = *g->hierarchy;
  int n = (int)h.level_cells.size();
  for (int i = 0; i < n && i < cap; ++i) {
    level_cells[i] = h.level_cells[i];
    level_nonzeros[i] = h.level_nonzeros[i];
  }
  if (masked) *masked = h.masked ? 1 : 0;
  return n;
}

int orc_mg_shape(orc_grid* g, int64_t* dims, int64_t* blocks, int cap, int* n_blocks) {
  if (!g->hierarchy) {
    int e = hierarchy_build(*g, g->hierarchy);
    if (e) return -e;
  }
  const Hierarchy& h = *g->hierarchy;
  int n = (int)h.level_dimensions.size();
  for (int i = 0; i < n && i < cap; ++i)
    for (int a = 0; a < 3; ++a) dims[3 * i + a] = h.level_dimensions[(size_t)i][(size_t)a];
  // BlockSmootherStats (line.rs:200-209) without the storage bytes
  int k = 0;
  for (size_t l = 0; l < h.levels.size(); ++l)
    for (const BlockSmoother& b : h.levels[l].lines) {
      if (k >= cap) break;
      int64_t* o = blocks + 6 * k++;
      o[0] = (int64_t)l;
      o[1] = 0;
      for (int a : b.axes) o[1] |= int64_t(1) << a;
      o[2] = (int64_t)b.length;
      o[3] = (int64_t)(b.nodes.size() / b.length);
      o[4] = (int64_t)b.bandwidth;
      o[5] = (int64_t)b.axes.size();
    }
  *n_blocks = k;
  return n;
}

int orc_poly_contains(const double* xy, int32_t n, double x, double y) { return poly_contains(xy, n, x, y) ? 1 : 0; }

namespace {
int masked_of(orc_grid* g, const Hierarchy** out) {
  if (!g->masked) {
    auto h = std::make_unique<Hierarchy>();
    int e = masked_build(*g, *h);
    if (e) return e;
    g->masked = std::move(h);
  }
  *out = g->masked.get();
  return 0;
}
}  // namespace

int orc_masked_level_sizes(orc_grid* g, int level, int64_t* sizes) {
  const Hierarchy* h = nullptr;
  if (int e = masked_of(g, &h)) return -e;
  const int levels = (int)h->mlevels.size();
  if (level < 0 || level >= levels) return levels;
  const MaskedLevel& L = h->mlevels[(size_t)level];
  int64_t members = 0;
  for (const auto& c : L.closed) members += (int64_t)c.size();
  sizes[0] = (int64_t)L.x.size();
  sizes[1] = (int64_t)L.a.columns.size();
  sizes[2] = (int64_t)L.closed.size();
  sizes[3] = members;
  sizes[4] = L.parent.empty() ? 0 : 1;
  sizes[5] = (int64_t)L.direct.size();
  return levels;
}

int orc_masked_level(orc_grid* g, int level, int64_t* parent, int64_t* offsets, int64_t* columns, double* values,
                     double* diagonal, int64_t* closed_ptr, int64_t* closed_members, int64_t* fluid_indices) {
  const Hierarchy* h = nullptr;
  if (int e = masked_of(g, &h)) return -e;
  if (level < 0 || level >= (int)h->mlevels.size()) return -34;
  const MaskedLevel& L = h->mlevels[(size_t)level];
  const size_t n = L.x.size();
  for (size_t i = 0; i < L.parent.size(); ++i) parent[i] = (int64_t)L.parent[i];
  for (size_t i = 0; i <= n; ++i) offsets[i] = (int64_t)L.a.offsets[i];
  for (size_t k = 0; k < L.a.columns.size(); ++k) {
    columns[k] = (int64_t)L.a.columns[k];
    values[k] = L.a.values[k];
  }
  for (size_t i = 0; i < n; ++i) diagonal[i] = L.a.diagonal[i];
  int64_t at = 0;
  closed_ptr[0] = 0;
  for (size_t c = 0; c < L.closed.size(); ++c) {
    for (size_t i : L.closed[c]) closed_members[at++] = (int64_t)i;
    closed_ptr[c + 1] = at;
  }
  if (fluid_indices)
    for (size_t j = 0; j < h->fluid_indices.size(); ++j) fluid_indices[j] = (int64_t)h->fluid_indices[j];
  return 0;
}

int orc_solve_scaled(orc_grid* g, const double* b, const double* scale, double tol, int64_t max, double* x_out,
                     orc_info* info) {
  orc_info local;
  std::memset(&local, 0, sizeof local);
  size_t count = g->count();
  for (size_t id = 0; id < g->components.size(); ++id) {
    if (g->anchored[id]) continue;
    double sum = neg_zero_sum(), sc = neg_zero_sum();
    for (size_t i : g->components[id]) sum += b[i];
    for (size_t i : g->components[id]) sc += std::fmax(std::fabs(b[i]), scale ? scale[i] : 0.0);
    if (std::fabs(sum) > 1e-9 * sc) {
      if (info) *info = local;
      return fail(31, "CFD-E031: incompatible flux in a closed fluid component");
    }
  }
  bool mg = g->preconditioner == ORC_MULTIGRID;
  if (mg) {
    if (g->components.empty()) return fail(34, "CFD-E034: numerics.pressure_preconditioner: multigrid requires fluid cells");
    if (!g->hierarchy) {
      int e = hierarchy_build(*g, g->hierarchy);
      if (e) return e;
      local.hierarchy_builds = 1;
    }
  } else if (!g->factor) {
    g->factor = factor_build(*g);
  }
  Work work;
  auto precondition = [&](const vector<double>& r, vector<double>& z) {
    if (mg)
      hierarchy_apply(*g->hierarchy, r.data(), z.data(), count, work);
    else
      factor_apply(*g->factor, r.data(), z.data(), count);
    local.preconditioner_applications += 1;
  };
  vector<double> x(count, 0.0);
  vector<double> r(b, b + count);
  orc_center_nullspaces(g, r.data());
  vector<double> compatible_b = r;
  double norm = std::sqrt(dot(r, r));
  auto finish = [&](int code) {
    local.cycles = work.cycles;
    local.smoothing_cell_updates = work.smoothing;
    local.residual_cell_updates = work.residual;
    local.coarse_solves = work.coarse;
    local.restriction_cell_updates = work.restriction;
    local.prolongation_cell_updates = work.prolongation;
    local.block_smoothing_cell_updates = work.block_smoothing;
    local.block_solves = work.block_solves;
    local.block_level_cell_visits = work.block_visits;
    if (info) *info = local;
    return code;
  };
  if (norm < 1e-24) {
    std::copy(x.begin(), x.end(), x_out);
    return finish(0);
  }
  vector<double> z(count, 0.0);
  precondition(r, z);
  orc_center_nullspaces(g, z.data());
  vector<double> p = z;
  double rz = dot(r, z);
  vector<double> ap(count, 0.0);
  const vector<OperatorEntry> op = operator_build(g);
  for (int64_t it = 0; it < max; ++it) {
    operator_entries_apply(op, count, p.data(), ap.data());
    local.operator_applications += 1;
    double denom = dot(p, ap);
    if (denom <= 0.0 || !std::isfinite(denom)) return finish(fail(32, "CFD-E032: pressure CG lost positive definiteness"));
    double alpha = rz / denom;
    for (size_t i = 0; i < count; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * ap[i];
    }
    double rel = std::sqrt(dot(r, r)) / norm;
    if (rel <= tol) {
      orc_apply(g, x.data(), ap.data());
      local.operator_applications += 1;
      local.true_residual_checks += 1;
      double s = neg_zero_sum();
      for (size_t i = 0; i < count; ++i) {
        double d = ap[i] - compatible_b[i];
        s += d * d;
      }
      double true_rel = std::sqrt(s) / norm;
      if (true_rel <= tol * 10.0 + 1e-12) {
        orc_center_nullspaces(g, x.data());
        std::copy(x.begin(), x.end(), x_out);
        local.iterations = it + 1;
        local.relative_residual = true_rel;
        return finish(0);
      }
    }
    precondition(r, z);
    orc_center_nullspaces(g, z.data());
    double next = dot(r, z);
    double beta = next / rz;
    for (size_t i = 0; i < count; ++i) p[i] = z[i] + beta * p[i];
    rz = next;
  }
  local.iterations = max;
  char buf[96];
  std::snprintf(buf, sizeof buf, "CFD-E033: pressure iteration limit %lld", (long long)max);
  return finish(fail(33, buf));
}

void orc_projection_rhs(const orc_grid* g, const double* u, double dt, double* b, double* terms) {
  for (size_t i = 0; i < g->count(); ++i) {
    b[i] = -0.0;
    terms[i] = 0.0;
    if (!g->fluid[i]) continue;
    double divergence = 0.0, sc = 0.0;
    for (int a = 0; a < 3; ++a) {
      size_t lo = g->cell_faces[i][a * 2], hi = g->cell_faces[i][a * 2 + 1];
      divergence += (u[hi] - u[lo]) / g->dx[a];
      sc += (std::fabs(u[lo]) + std::fabs(u[hi])) / g->dx[a];
    }
    b[i] = -divergence / dt;
    terms[i] = sc / dt;
  }
}

int orc_project(orc_grid* g, double* u, double* p, double dt, double tol, int64_t max, orc_info* info) {
  size_t count = g->count();
  vector<double> b(count), terms(count), correction(count);
  orc_projection_rhs(g, u, dt, b.data(), terms.data());
  int e = orc_solve_scaled(g, b.data(), terms.data(), tol, max, correction.data(), info);
  if (e) return e;
  for (size_t f = 0; f < g->faces.size(); ++f) {
    const Face& F = g->faces[f];
    double d = 0.0;
    if (F.kind == ORC_INTERIOR)
      d = (correction[F.hi] - correction[F.lo]) / F.distance;
    else if (F.kind == ORC_OUTLET)
      d = F.lo != NONE ? -correction[F.lo] / F.distance : correction[F.hi] / F.distance;
    u[f] -= dt * d;
  }
  for (size_t i = 0; i < count; ++i) p[i] += correction[i];
  orc_center_nullspaces(g, p);
  return 0;
}

void orc_divergence(const orc_grid* g, const double* u, double* div) {
  for (size_t i = 0; i < g->count(); ++i) {
    if (!g->fluid[i]) {
      div[i] = 0.0;
      continue;
    }
    double s = neg_zero_sum();
    for (int a = 0; a < 3; ++a) s += (u[g->cell_faces[i][a * 2 + 1]] - u[g->cell_faces[i][a * 2]]) / g->dx[a];
    div[i] = s;
  }
}

void orc_gradient(const orc_grid* g, const double* p, double* grad) {
  for (size_t f = 0; f < g->faces.size(); ++f) {
    const Face& F = g->faces[f];
    if (F.kind == ORC_INTERIOR)
      grad[f] = (p[F.hi] - p[F.lo]) / F.distance;
    else if (F.kind == ORC_OUTLET)
      grad[f] = F.lo != NONE ? -p[F.lo] / F.distance : p[F.hi] / F.distance;
    else
      grad[f] = 0.0;
  }
}

void orc_enforce_velocity(const orc_grid* g, double* u) {
  for (size_t f = 0; f < g->faces.size(); ++f) {
    const Face& F = g->faces[f];
    if (F.kind == ORC_WALLF || F.kind == ORC_SLIPF || F.kind == ORC_BLOCKED) {
      u[f] = 0.0;
    } else if (F.kind == ORC_INLET) {
      double v3[3];
      flow_vector(g->flow, F.center_z, v3);
      u[f] = v3[F.axis];
    }
  }
}

void orc_flow_vector(const orc_case* c, double z, double* v3) { flow_vector(*c, z, v3); }

}  // extern "C"
