// Host set-up of the device grid: case -> per-cell coupling codes, fluid
// components and the IC(0) factor diagonal.
//
// The topology must be the reference's (grid.rs:246-464), so the mask
// rasterization (geometry.rs:113-193, case.rs:436-443) and the face
// classification follow the same rules; they are stored per cell (six side
// bits) instead of as a face list, which is all the device needs.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>

#include "internal.hpp"

namespace cfdgpu {
namespace {

constexpr int64_t NONE = -1;

// Polygon membership by even-odd crossings (geometry.rs:13-26).
bool in_polygon(const double* p, int n, double x, double y) {
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

double center(int64_t q, double dx) { return ((double)q + 0.5) * dx; }

// First index in [0, n) whose cell center is not below `bound` (geometry.rs:96-107,124-129).
int64_t first_center_at_or_above(int64_t n, double dx, double bound, bool nan_means_all) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (center(mid, dx) < bound || (nan_means_all && std::isnan(bound)))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

}  // namespace

bool block_ranges(const cfd_block& b, const int64_t n[3], const double dx[3], int64_t r0[3], int64_t r1[3]) {
  for (int a = 0; a < 3; ++a) {
    r0[a] = first_center_at_or_above(n[a], dx[a], b.min_m[a], true);
    r1[a] = std::max(first_center_at_or_above(n[a], dx[a], b.max_m[a], false), r0[a]);
    if (r0[a] == r1[a]) return false;
  }
  return true;
}

namespace {

// Marks the cells whose centers a block contains; returns how many (per block, overlaps counted).
int64_t rasterize(const cfd_block& b, const int64_t n[3], const double dx[3], std::vector<uint8_t>& fluid) {
  int64_t r0[3], r1[3];
  if (!block_ranges(b, n, dx, r0, r1)) return 0;
  // Every center of the z range lies within the z bounds: one test per column decides it.
  double z = center(r0[2], dx[2]);
  int64_t count = 0;
  for (int64_t j = r0[1]; j < r1[1]; ++j)
    for (int64_t i = r0[0]; i < r1[0]; ++i) {
      double p[3] = {center(i, dx[0]), center(j, dx[1]), z};
      bool inside = true;
      for (int a = 0; a < 3; ++a) inside = inside && p[a] >= b.min_m[a] && p[a] < b.max_m[a];
      if (inside && b.n_vertices > 0) inside = in_polygon(b.footprint_xy, b.n_vertices, p[0], p[1]);
      if (!inside) continue;
      count += r1[2] - r0[2];
      for (int64_t k = r0[2]; k < r1[2]; ++k) fluid[(size_t)((k * n[1] + j) * n[0] + i)] = 0;
    }
  return count;
}

// Flow::vector (case.rs:345-367) at height z: only its sign classifies open faces.
void flow_vector(const cfd_case& c, double z, double v[3]) {
  if (c.flow_kind == CFD_FLOW_NEUTRAL_RANS) {
    double z0 = c.roughness_m;
    double u = c.speed_m_s * std::log((std::max(z, 0.0) + z0) / z0) / std::log((c.reference_height_m + z0) / z0);
    double t = c.direction_deg_from * (M_PI / 180.0);
    v[0] = -u * std::sin(t);
    v[1] = -u * std::cos(t);
    v[2] = 0.0;
  } else {
    for (int a = 0; a < 3; ++a) v[a] = c.velocity_m_s[a];
  }
}

}  // namespace

void open_outlets(const cfd_case& c, const HostGrid& g, std::vector<uint8_t>& out) {
  const int64_t nz = g.n[2];
  out.assign((size_t)(6 * nz), 0);
  for (int a = 0; a < 3; ++a) {
    if (c.boundary[a] != CFD_OPEN) continue;
    for (int side = 0; side < 2; ++side)
      for (int64_t k = 0; k < nz; ++k) {
        // the face's height: the cell center, or the bottom or top face of a z boundary
        double face_z = a == 2 ? (double)(side == 0 ? 0 : nz) * g.dx[2] : ((double)k + 0.5) * g.dx[2];
        double v[3];
        flow_vector(c, face_z, v);
        double normal = side == 0 ? -1.0 : 1.0;
        out[(size_t)((2 * a + side) * nz + k)] = v[a] * normal > 1e-12 ? 1 : 0;
      }
  }
}

int grid_header(const cfd_case& c, HostGrid& g) {
  int64_t total = 1;
  for (int a = 0; a < 3; ++a) {
    int64_t minimum = c.boundary[a] == CFD_PERIODIC ? 1 : 2;
    if (c.cells[a] < minimum || c.cells[a] > 4096) {
      set_error("CFD-E004: grid.cells needs 2..4096 cells per axis (1..4096 on a periodic axis)");
      return CFD_EINVAL;
    }
    if (!(c.size_m[a] > 0.0) || !std::isfinite(c.size_m[a])) {
      set_error("CFD-E003: domain.size_m must be finite and positive");
      return CFD_EINVAL;
    }
    if (c.boundary[a] < CFD_PERIODIC || c.boundary[a] > CFD_GROUND_OPEN_TOP) {
      set_error("cfd_case.boundary: unknown boundary code");
      return CFD_EINVAL;
    }
    total *= c.cells[a];
  }
  if (total > 100000000) {
    set_error("CFD-E006: at most 1e8 cells");
    return CFD_EINVAL;
  }
  for (int a = 0; a < 3; ++a) {
    g.n[a] = c.cells[a];
    g.boundary[a] = c.boundary[a];
    g.periodic[a] = c.boundary[a] == CFD_PERIODIC;
    g.dx[a] = c.size_m[a] / (double)c.cells[a];
  }
  g.preconditioner = c.preconditioner;
  g.count = total;
  g.volume = g.dx[0] * g.dx[1] * g.dx[2];
  for (int a = 0; a < 3; ++a) {
    double area = g.volume / g.dx[a];
    g.w_int[a] = area / (g.dx[a] * g.volume);
    g.w_out[a] = area / ((0.5 * g.dx[a]) * g.volume);
  }
  for (int a = 0; a < 3; ++a) g.face_count[a] = total / g.n[a] * (g.n[a] + (g.periodic[a] ? 0 : 1));
  return 0;
}

int block_error(const cfd_case& c, int b) {
  char buf[200];
  std::snprintf(buf, sizeof buf,
                "CFD-E028: geometry[%d] (%s) contains no cell center, so the grid does not represent it; enlarge it or "
                "refine the grid",
                b, c.blocks[b].n_vertices > 0 ? "prism" : "box");
  set_error(buf);
  return CFD_EINVAL;
}

int count_fluid(HostGrid& g) {
  g.fluid_count = 0;
  for (uint8_t f : g.fluid) g.fluid_count += f;
  if (g.fluid_count == 0) {
    set_error("CFD-E030: geometry contains no fluid cells");
    return CFD_EINVAL;
  }
  return 0;
}

int build_host_grid(const cfd_case& c, HostGrid& g) {
  int rc = grid_header(c, g);
  if (rc) return rc;
  const int64_t total = g.count;
  const int64_t* n = g.n;
  g.fluid.assign((size_t)total, 1);
  for (int b = 0; b < c.n_blocks; ++b)
    if (rasterize(c.blocks[b], n, g.dx, g.fluid) == 0) return block_error(c, b);
  if ((rc = count_fluid(g))) return rc;

  // Face classification (grid.rs:341-429), per cell side. A side couples the cell
  // into the pressure operator when its face is Interior or Outlet. A one-cell
  // periodic axis joins a cell to itself: that face carries no pressure difference.
  g.codes.assign((size_t)total, 0);
  const int64_t stride[3] = {1, n[0], n[0] * n[1]};
  for (int64_t k = 0; k < n[2]; ++k)
    for (int64_t j = 0; j < n[1]; ++j)
      for (int64_t i = 0; i < n[0]; ++i) {
        int64_t c0 = (k * n[1] + j) * n[0] + i;
        if (!g.fluid[(size_t)c0]) continue;
        uint8_t code = kFluid;
        int64_t q[3] = {i, j, k};
        for (int a = 0; a < 3; ++a) {
          int64_t t = q[a];
          for (int side = 0; side < 2; ++side) {
            int64_t other = NONE;
            if (side == 0) {
              if (t > 0)
                other = c0 - stride[a];
              else if (g.periodic[a])
                other = c0 + (n[a] - 1) * stride[a];
            } else {
              if (t + 1 < n[a])
                other = c0 + stride[a];
              else if (g.periodic[a])
                other = c0 - (n[a] - 1) * stride[a];
            }
            bool coupled = false;
            if (other != NONE) {
              coupled = other != c0 && g.fluid[(size_t)other];  // Interior (or Wall against a solid)
            } else if (c.boundary[a] == CFD_OPEN) {
              // Domain face of an open axis: Outlet where the case flow leaves.
              double face_z = a == 2 ? (double)(side == 0 ? t : t + 1) * g.dx[2] : ((double)k + 0.5) * g.dx[2];
              double v[3];
              flow_vector(c, face_z, v);
              double normal = side == 0 ? -1.0 : 1.0;
              coupled = v[a] * normal > 1e-12;
            }
            if (coupled) code |= side_bit(a, side);
          }
        }
        g.codes[(size_t)c0] = code;
      }

  find_components(g);
  closed_lists(g);
  return 0;
}

// Fluid components by breadth-first search from the lowest unassigned cell
// (grid.rs:430-462): the same partition and numbering as the reference; a
// component with a coupled domain face (an outlet) is anchored.
void find_components(HostGrid& g) {
  const int64_t total = g.count;
  const int64_t* n = g.n;
  const int64_t stride[3] = {1, n[0], n[0] * n[1]};
  g.component.assign((size_t)total, -1);
  g.anchored.clear();
  std::vector<int64_t> queue;
  for (int64_t start = 0; start < total; ++start) {
    if (!g.fluid[(size_t)start] || g.component[(size_t)start] != -1) continue;
    int32_t id = (int32_t)g.anchored.size();
    bool anchored = false;
    queue.clear();
    queue.push_back(start);
    g.component[(size_t)start] = id;
    for (size_t head = 0; head < queue.size(); ++head) {
      int64_t c0 = queue[head];
      uint8_t code = g.codes[(size_t)c0];
      int64_t q[3] = {c0 % n[0], c0 / n[0] % n[1], c0 / (n[0] * n[1])};
      for (int a = 0; a < 3; ++a)
        for (int side = 0; side < 2; ++side) {
          if (!(code & side_bit(a, side))) continue;
          int64_t t = q[a];
          int64_t other;
          if (side == 0)
            other = t > 0 ? c0 - stride[a] : (g.periodic[a] ? c0 + (n[a] - 1) * stride[a] : NONE);
          else
            other = t + 1 < n[a] ? c0 + stride[a] : (g.periodic[a] ? c0 - (n[a] - 1) * stride[a] : NONE);
          if (other == NONE) {
            anchored = true;  // a coupled domain face is an outlet
            continue;
          }
          if (g.component[(size_t)other] == -1) {
            g.component[(size_t)other] = id;
            queue.push_back(other);
          }
        }
    }
    g.anchored.push_back(anchored ? 1 : 0);
  }
}

// The grid of a caller's built Grid (cfd_grid_desc): the mask, the periodic axes
// and the faces give the coupling codes (an Interior face couples its two cells
// unless they are one cell of a one-cell periodic axis; an Outlet face couples
// its cell to the outside), the weights come from the faces' own area and
// distance (pressure.rs:257), and the components from the codes as in
// build_host_grid. strict: weights that differ within an axis and face kind are
// CFD_EUNSUPPORTED; otherwise they leave that weight NaN (it then matches no grid).
int grid_from_desc(const cfd_grid_desc& d, HostGrid& g, bool strict) {
  char buf[256];
  int64_t total = 1;
  for (int a = 0; a < 3; ++a) {
    const int64_t minimum = d.periodic[a] ? 1 : 2;
    if (d.n[a] < minimum || d.n[a] > 4096) {
      set_error("CFD-E004: grid.cells needs 2..4096 cells per axis (1..4096 on a periodic axis)");
      return CFD_EINVAL;
    }
    if (!(d.dx[a] > 0.0) || !std::isfinite(d.dx[a])) {
      set_error("cfd_grid_desc: cell sizes must be finite and positive");
      return CFD_EINVAL;
    }
    total *= d.n[a];
  }
  if (total > 100000000) {
    set_error("CFD-E006: at most 1e8 cells");
    return CFD_EINVAL;
  }
  if (!(d.volume > 0.0) || !std::isfinite(d.volume) || !d.fluid || !d.faces) {
    set_error("cfd_grid_desc: a positive finite volume, the fluid mask and the faces are required");
    return CFD_EINVAL;
  }
  if (d.preconditioner != CFD_IC0 && d.preconditioner != CFD_MULTIGRID) {
    set_error("numerics.pressure_preconditioner must be ic0 or multigrid");
    return CFD_EINVAL;
  }
  int64_t fdim[3][3], faces = 0, first[3];
  for (int a = 0; a < 3; ++a) {
    g.n[a] = d.n[a];
    g.dx[a] = d.dx[a];
    g.periodic[a] = d.periodic[a] != 0;
    g.boundary[a] = g.periodic[a] ? CFD_PERIODIC : CFD_WALL;
    for (int b = 0; b < 3; ++b) fdim[a][b] = d.n[b] + (b == a && !g.periodic[a] ? 1 : 0);
    first[a] = faces;
    g.face_count[a] = fdim[a][0] * fdim[a][1] * fdim[a][2];
    faces += g.face_count[a];
  }
  if (d.face_count != faces) {
    std::snprintf(buf, sizeof buf, "cfd_grid_desc: %lld faces given, the grid has %lld", (long long)d.face_count,
                  (long long)faces);
    set_error(buf);
    return CFD_EINVAL;
  }
  g.count = total;
  g.volume = d.volume;
  g.preconditioner = d.preconditioner;
  g.fluid.assign(d.fluid, d.fluid + total);
  for (uint8_t& f : g.fluid) f = f ? 1 : 0;
  if (int rc = count_fluid(g)) return rc;
  const int64_t* n = g.n;
  const int64_t stride[3] = {1, n[0], n[0] * n[1]};
  g.codes.assign((size_t)total, 0);
  for (int64_t c0 = 0; c0 < total; ++c0)
    if (g.fluid[(size_t)c0]) g.codes[(size_t)c0] = kFluid;
  const uint64_t unset = ~0ULL;
  uint64_t wbits[2][3] = {{unset, unset, unset}, {unset, unset, unset}};
  bool mixed[2][3] = {{false, false, false}, {false, false, false}};
  int64_t wface[2][3] = {{NONE, NONE, NONE}, {NONE, NONE, NONE}};  // the face that set each weight
  for (int a = 0; a < 3; ++a) {
    for (int64_t e = 0; e < g.face_count[a]; ++e) {
      const int64_t f = first[a] + e;
      const cfd_face& F = d.faces[f];
      const int64_t q[3] = {e % fdim[a][0], e / fdim[a][0] % fdim[a][1], e / (fdim[a][0] * fdim[a][1])};
      const int64_t t = q[a];
      const int64_t base = q[0] * stride[0] + q[1] * stride[1] + q[2] * stride[2] - t * stride[a];
      const int64_t above = t < n[a] ? base + t * stride[a] : NONE;
      const int64_t below = t > 0 ? base + (t - 1) * stride[a] : (g.periodic[a] ? base + (n[a] - 1) * stride[a] : NONE);
      const int64_t lo = below != NONE && g.fluid[(size_t)below] ? below : NONE;
      const int64_t hi = above != NONE && g.fluid[(size_t)above] ? above : NONE;
      if (F.axis != a || F.lo != lo || F.hi != hi || F.kind < CFD_FACE_INTERIOR || F.kind > CFD_FACE_BLOCKED) {
        std::snprintf(buf, sizeof buf,
                      "cfd_grid_desc: face %lld is not the Grid's (axis %d, lo %lld, hi %lld expected; kind 0..5)",
                      (long long)f, a, (long long)lo, (long long)hi);
        set_error(buf);
        return CFD_EINVAL;
      }
      const bool domain = below == NONE || above == NONE;
      int role = -1;  // 0 interior weight, 1 outlet weight
      if (F.kind == CFD_FACE_INTERIOR) {
        if (lo == NONE || hi == NONE) {
          std::snprintf(buf, sizeof buf, "cfd_grid_desc: interior face %lld lacks a fluid cell on one side", (long long)f);
          set_error(buf);
          return CFD_EINVAL;
        }
        if (lo != hi) {
          g.codes[(size_t)lo] |= side_bit(a, 1);
          g.codes[(size_t)hi] |= side_bit(a, 0);
          role = 0;
        }
      } else if (F.kind == CFD_FACE_OUTLET) {
        if (!domain || (lo == NONE) == (hi == NONE)) {
          std::snprintf(buf, sizeof buf, "cfd_grid_desc: outlet face %lld is not a domain face of one fluid cell",
                        (long long)f);
          set_error(buf);
          return CFD_EINVAL;
        }
        if (hi != NONE)
          g.codes[(size_t)hi] |= side_bit(a, 0);
        else
          g.codes[(size_t)lo] |= side_bit(a, 1);
        role = 1;
      }
      if (role < 0) continue;
      const double w = F.area / (F.distance * g.volume);
      uint64_t bits;
      std::memcpy(&bits, &w, sizeof bits);
      if (!std::isfinite(w) || !(w > 0.0)) {
        std::snprintf(buf, sizeof buf, "cfd_grid_desc: face %lld has no positive finite pressure weight", (long long)f);
        set_error(buf);
        return CFD_EINVAL;
      }
      if (wbits[role][a] == unset) {
        wbits[role][a] = bits;
        wface[role][a] = f;
      } else if (wbits[role][a] != bits) {
        if (strict) {
          std::snprintf(buf, sizeof buf,
                        "cfd_grid_desc: faces %lld and %lld (axis %d) carry different %s pressure weights; the device "
                        "operator holds one interior and one outlet weight per axis",
                        (long long)wface[role][a], (long long)f, a, role ? "outlet" : "interior");
          set_error(buf);
          return CFD_EUNSUPPORTED;
        }
        mixed[role][a] = true;
      }
    }
  }
  for (int a = 0; a < 3; ++a) {
    // an axis without faces of a kind keeps the weight Grid::build gives them
    const double area = g.volume / g.dx[a];
    double w[2] = {area / (g.dx[a] * g.volume), area / ((0.5 * g.dx[a]) * g.volume)};
    for (int r = 0; r < 2; ++r) {
      if (wbits[r][a] != unset) std::memcpy(&w[r], &wbits[r][a], sizeof(double));
      if (mixed[r][a]) w[r] = std::nan("");
    }
    g.w_int[a] = w[0];
    g.w_out[a] = w[1];
  }
  find_components(g);
  if (d.component) {
    bool same = d.component_count == (int64_t)g.anchored.size();
    for (int64_t c0 = 0; c0 < total && same; ++c0) same = d.component[c0] == (int64_t)g.component[(size_t)c0];
    for (int64_t id = 0; id < d.component_count && same && d.anchored; ++id)
      same = (d.anchored[id] != 0) == (g.anchored[(size_t)id] != 0);
    if (!same) {
      set_error("cfd_grid_desc: the components or anchored flags differ from those its faces give (grid.rs:430-462)");
      return CFD_EINVAL;
    }
  }
  closed_lists(g);
  return 0;
}

void closed_lists(HostGrid& g) {
  const int64_t total = g.count;
  // Closed components: gauge lists, sorted by cell index.
  g.closed_of_comp.assign(g.anchored.size(), -1);
  int32_t n_closed = 0;
  for (size_t id = 0; id < g.anchored.size(); ++id)
    if (!g.anchored[id]) g.closed_of_comp[id] = n_closed++;
  g.closed_offsets.assign((size_t)n_closed + 1, 0);
  g.closed_index.assign((size_t)total, -1);
  for (int64_t c0 = 0; c0 < total; ++c0) {
    int32_t id = g.component[(size_t)c0];
    if (id < 0) continue;
    int32_t cl = g.closed_of_comp[(size_t)id];
    if (cl < 0) continue;
    g.closed_index[(size_t)c0] = cl;
    g.closed_offsets[(size_t)cl + 1] += 1;
    g.codes[(size_t)c0] |= kClosed;
  }
  for (int32_t cl = 0; cl < n_closed; ++cl) g.closed_offsets[(size_t)cl + 1] += g.closed_offsets[(size_t)cl];
  g.closed_members.assign((size_t)g.closed_offsets[(size_t)n_closed], 0);
  {
    std::vector<int64_t> fill(g.closed_offsets.begin(), g.closed_offsets.end() - 1);
    for (int64_t c0 = 0; c0 < total; ++c0) {
      int32_t cl = g.closed_index[(size_t)c0];
      if (cl >= 0) g.closed_members[(size_t)fill[(size_t)cl]++] = (int32_t)c0;
    }
  }
}

// Diagonal of the IC(0) factor of A + 1e-6*maxdiag*I (pressure.rs:118-222) in
// the reference's arithmetic order. The diagonal of A accumulates the face
// weights in face order (x-faces, then y-faces, then z-faces; within an axis
// the lower face first); the lower entries of a row are l = a / D[col] (no
// fill exists without a 3-cell periodic axis, so the reference's dot term is
// exactly zero), summed as squares in ascending column order.
void build_ic0_diagonal(HostGrid& g) {
  const int64_t* n = g.n;
  const int64_t total = g.count;
  const int64_t stride[3] = {1, n[0], n[0] * n[1]};
  g.ic0_fill = false;
  for (int a = 0; a < 3; ++a) g.ic0_fill |= g.periodic[a] && n[a] == 3;
  std::vector<double> raw((size_t)total, 0.0);
  double maxdiag = 1.0;
  for (int64_t c0 = 0; c0 < total; ++c0) {
    uint8_t code = g.codes[(size_t)c0];
    double d = 0.0;
    int64_t q[3] = {c0 % n[0], c0 / n[0] % n[1], c0 / (n[0] * n[1])};
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        if (!(code & side_bit(a, side))) continue;
        bool boundary = side == 0 ? q[a] == 0 : q[a] + 1 == n[a];
        d += (boundary && !g.periodic[a]) ? g.w_out[a] : g.w_int[a];
      }
    raw[(size_t)c0] = d;
    maxdiag = std::fmax(maxdiag, d);
  }
  g.ic0_diag.assign((size_t)total, 0.0);
  double* D = g.ic0_diag.data();
  for (int64_t c0 = 0; c0 < total; ++c0) {
    uint8_t code = g.codes[(size_t)c0];
    int64_t q[3] = {c0 % n[0], c0 / n[0] % n[1], c0 / (n[0] * n[1])};
    double squared = -0.0;
    // Lower neighbours in ascending index order: z wrap, z lo, y wrap, y lo, x wrap, x lo.
    for (int a = 2; a >= 0; --a) {
      if (!g.periodic[a] || n[a] == 1) {
        if (q[a] > 0 && (code & side_bit(a, 0))) {
          double l = -g.w_int[a] / D[(size_t)(c0 - stride[a])];
          squared += l * l;
        }
        continue;
      }
      if (n[a] == 2) {
        // Two faces join the same pair of cells; the reference merges them into one entry.
        if (q[a] == 1 && (code & side_bit(a, 0))) {
          double l = (-g.w_int[a] - g.w_int[a]) / D[(size_t)(c0 - stride[a])];
          squared += l * l;
        }
        continue;
      }
      if (q[a] == n[a] - 1 && (code & side_bit(a, 1))) {
        double l = -g.w_int[a] / D[(size_t)(c0 - (n[a] - 1) * stride[a])];
        squared += l * l;
      }
      if (q[a] > 0 && (code & side_bit(a, 0))) {
        double l = -g.w_int[a] / D[(size_t)(c0 - stride[a])];
        squared += l * l;
      }
    }
    D[(size_t)c0] = std::sqrt(std::fmax(raw[(size_t)c0] + 1e-6 * maxdiag - squared, 1e-12 * maxdiag));
  }
  if (g.ic0_fill) build_ic0_fill(g, raw, maxdiag);
}

// The factor of a grid with a periodic axis of exactly 3 cells, where two lower
// neighbours of a cell can be joined themselves (the three cells of the axis)
// and an entry takes the dot term of Factor::build (pressure.rs:197-209): the
// lower entries of each row (merged per column in face order, sorted by
// column), each l = (a - sum_k l_row,k l_col,k) / D[col] over the row's earlier
// columns k that the column's row also holds, then the diagonal as above. Also
// the entries the device sweeps gather, per cell in their slot order (per axis
// z, y, x: the wrap neighbour, then the neighbour on the axis): 6 forward
// (l of the cell's row at its lower neighbours) and 6 backward (l of its upper
// neighbours' rows at the cell).
void build_ic0_fill(HostGrid& g, const std::vector<double>& raw, double maxdiag) {
  const int64_t* n = g.n;
  const int64_t total = g.count;
  const int64_t stride[3] = {1, n[0], n[0] * n[1]};
  struct Entry {
    int64_t col;
    double l;
  };
  std::vector<std::vector<Entry>> rows((size_t)total);
  // lower neighbours in ascending index order (z wrap, z lo, y wrap, y lo, x wrap, x lo)
  auto lower_of = [&](int64_t c0, uint8_t code, const int64_t q[3], auto f) {
    for (int a = 2; a >= 0; --a) {
      if (!g.periodic[a] || n[a] == 1) {
        if (q[a] > 0 && (code & side_bit(a, 0))) f(c0 - stride[a], -g.w_int[a], a, 1);
        continue;
      }
      if (n[a] == 2) {
        if (q[a] == 1 && (code & side_bit(a, 0))) f(c0 - stride[a], -g.w_int[a] - g.w_int[a], a, 1);
        continue;
      }
      if (q[a] == n[a] - 1 && (code & side_bit(a, 1))) f(c0 - (n[a] - 1) * stride[a], -g.w_int[a], a, 0);
      if (q[a] > 0 && (code & side_bit(a, 0))) f(c0 - stride[a], -g.w_int[a], a, 1);
    }
  };
  double* D = g.ic0_diag.data();
  for (int64_t c0 = 0; c0 < total; ++c0) {
    const uint8_t code = g.codes[(size_t)c0];
    const int64_t q[3] = {c0 % n[0], c0 / n[0] % n[1], c0 / (n[0] * n[1])};
    auto& row = rows[(size_t)c0];
    lower_of(c0, code, q, [&](int64_t col, double a, int, int) { row.push_back(Entry{col, a}); });
    for (size_t e = 0; e < row.size(); ++e) {
      const auto& other = rows[(size_t)row[e].col];
      double dot = 0.0;
      size_t o = 0;
      for (size_t p = 0; p < e; ++p) {
        while (o < other.size() && other[o].col < row[p].col) ++o;
        if (o < other.size() && other[o].col == row[p].col) dot += row[p].l * other[o].l;
      }
      row[e].l = (row[e].l - dot) / D[(size_t)row[e].col];
    }
    double squared = -0.0;
    for (const Entry& e : row) squared += e.l * e.l;
    D[(size_t)c0] = std::sqrt(std::fmax(raw[(size_t)c0] + 1e-6 * maxdiag - squared, 1e-12 * maxdiag));
  }
  auto entry = [&](int64_t r, int64_t col) {
    for (const Entry& e : rows[(size_t)r])
      if (e.col == col) return e.l;
    return 0.0;
  };
  g.ic0_fill_coef.assign((size_t)total * 12, 0.0);
  for (int64_t c0 = 0; c0 < total; ++c0) {
    const uint8_t code = g.codes[(size_t)c0];
    const int64_t q[3] = {c0 % n[0], c0 / n[0] % n[1], c0 / (n[0] * n[1])};
    double* fwd = &g.ic0_fill_coef[(size_t)c0 * 12];
    double* bwd = fwd + 6;
    lower_of(c0, code, q, [&](int64_t col, double, int a, int slot) { fwd[2 * (2 - a) + slot] = entry(c0, col); });
    // upper neighbours: the wrap neighbour, then the neighbour on the axis
    for (int a = 2; a >= 0; --a) {
      const int m = 2 * (2 - a);
      if (!g.periodic[a] || n[a] == 1) {
        if (q[a] < n[a] - 1 && (code & side_bit(a, 1))) bwd[m + 1] = entry(c0 + stride[a], c0);
      } else if (n[a] == 2) {
        if (q[a] == 0 && (code & side_bit(a, 1))) bwd[m + 1] = entry(c0 + stride[a], c0);
      } else {
        if (q[a] == 0 && (code & side_bit(a, 0))) bwd[m] = entry(c0 + (n[a] - 1) * stride[a], c0);
        if (q[a] < n[a] - 1 && (code & side_bit(a, 1))) bwd[m + 1] = entry(c0 + stride[a], c0);
      }
    }
  }
}

}  // namespace cfdgpu
