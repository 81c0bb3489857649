/*
 * CPU oracle for the pressure-projection path: a serial FP64 restatement of the
 * reference's grid construction, pressure operator, PCG, IC(0) and multigrid.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline. The GPU product never links or calls it.
 *
 * The case description below has the same memory layout as `cfd_case` in
 * include/cfdgpu.h so that one ctypes structure feeds both; it is declared
 * again here so the oracle builds from its own sources.
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PERIODIC = 0, ORC_OPEN = 1, ORC_WALL = 2, ORC_SLIP = 3, ORC_GROUND_OPEN_TOP = 4 };
enum { ORC_FLOW_VECTOR = 0, ORC_FLOW_NEUTRAL_RANS = 1 };
enum { ORC_IC0 = 0, ORC_MULTIGRID = 1 };
/* FaceKind order of grid.rs:60-81 */
enum { ORC_INTERIOR = 0, ORC_WALLF = 1, ORC_SLIPF = 2, ORC_INLET = 3, ORC_OUTLET = 4, ORC_BLOCKED = 5 };

typedef struct {
  double min_m[3];
  double max_m[3];
  int32_t n_vertices; /* 0: axis-aligned box; else polygon prism footprint */
  int32_t reserved;
  const double* footprint_xy; /* 2*n_vertices values, x0 y0 x1 y1 ... */
} orc_block;

typedef struct {
  int64_t cells[3];
  double size_m[3];
  int32_t boundary[3];
  int32_t flow_kind;
  double velocity_m_s[3];
  double speed_m_s, reference_height_m, roughness_m, direction_deg_from;
  int32_t preconditioner;
  int32_t n_blocks;
  const orc_block* blocks;
} orc_case;

typedef struct {
  int64_t iterations;
  double relative_residual;
  int64_t operator_applications;
  int64_t preconditioner_applications;
  int64_t true_residual_checks;
  int64_t cycles;
  int64_t smoothing_cell_updates;
  int64_t residual_cell_updates;
  int64_t coarse_solves;
  int64_t hierarchy_builds;
  int64_t restriction_cell_updates;
  int64_t prolongation_cell_updates;
  int64_t block_smoothing_cell_updates;
  int64_t block_solves;
  int64_t block_level_cell_visits;
} orc_info;

typedef struct orc_grid orc_grid;

/* Grid::build (grid.rs:246-464). Returns NULL and sets the error text on failure. */
orc_grid* orc_grid_build(const orc_case* c);
void orc_grid_free(orc_grid* g);
const char* orc_last_error(void);

/* counts[0..6] = cells, faces, fluid cells, components, anchored components, face_dims product per axis (3) */
void orc_grid_counts(const orc_grid* g, int64_t* counts);
/* Per face: axis, lo, hi (-1 for none), kind. */
void orc_grid_faces(const orc_grid* g, int32_t* axis, int64_t* lo, int64_t* hi, int32_t* kind, double* weight);
void orc_grid_fluid(const orc_grid* g, uint8_t* fluid);
/* component id per cell (-1 solid), anchored flag per component */
void orc_grid_components(const orc_grid* g, int64_t* component, uint8_t* anchored);

/* pressure::apply (pressure.rs:293-312) and the compact Operator (246-286). */
void orc_apply(const orc_grid* g, const double* x, double* y);
void orc_operator_apply(const orc_grid* g, const double* x, double* y);
/* pressure::center_nullspaces (pressure.rs:316-325) */
void orc_center_nullspaces(const orc_grid* g, double* x);

/* pressure::solve_scaled (pressure.rs:342-436). scale may be NULL.
 * Returns 0 on success, else an error code (31..36) with the text in orc_last_error(). */
int orc_solve_scaled(orc_grid* g, const double* b, const double* scale, double tol, int64_t max, double* x,
                     orc_info* info);

/* projection_rhs (pressure.rs:440-459) and project (469-502) */
void orc_projection_rhs(const orc_grid* g, const double* u, double dt, double* b, double* terms);
int orc_project(orc_grid* g, double* u, double* p, double dt, double tol, int64_t max, orc_info* info);

/* Factor::build / Factor::apply (pressure.rs:118-241), exposed for kernel parity. */
void orc_ic0_apply(orc_grid* g, const double* r, double* z);
void orc_ic0_diagonal(orc_grid* g, double* diag);
/* Hierarchy::apply (multigrid.rs:656-665 / masked.rs:292-302): one V-cycle. Returns error code. */
int orc_mg_apply(orc_grid* g, const double* r, double* z);
/* MultigridStats (multigrid.rs:55-72): writes up to cap level sizes, returns the level count (<0 on error). */
int orc_mg_stats(orc_grid* g, int64_t* level_cells, int64_t* level_nonzeros, int cap, int* masked);
/* Level dimensions of the unmasked hierarchy (none for masked) and, 6 values per
 * block smoother: level, axes bitmask, unknowns per block, block count,
 * half-bandwidth, axis count (line.rs:200-209). Returns the dimension rows. */
int orc_mg_shape(orc_grid* g, int64_t* dims, int64_t* blocks, int cap, int* n_blocks);

/* geometry.rs:13-26 contains(): the even-odd crossing test of one point
 * against a footprint of n vertices (x0 y0 x1 y1 ...). */
int orc_poly_contains(const double* xy, int32_t n, double x, double y);
/* The masked aggregation hierarchy (masked.rs:181-291) of the grid, built on
 * first use for any grid (also one the reference would give the unmasked
 * hierarchy). sizes: unknowns, stored off-diagonal entries, closed components,
 * closed members, whether the level has parents, dense blocks. Returns the
 * level count, or minus the error code. */
int orc_masked_level_sizes(orc_grid* g, int level, int64_t* sizes);
/* One level's arrays: parent (unknowns, when the level has parents), CSR
 * offsets (unknowns + 1), columns and values, diagonal, closed component
 * offsets (closed + 1) and members in the level's order, and (may be NULL)
 * the fluid cell of each level-0 unknown. Returns 0 or minus the error code. */
int orc_masked_level(orc_grid* g, int level, int64_t* parent, int64_t* offsets, int64_t* columns, double* values,
                     double* diagonal, int64_t* closed_ptr, int64_t* closed_members, int64_t* fluid_indices);

/* Grid::divergence / gradient (grid.rs:612-665) */
void orc_divergence(const orc_grid* g, const double* u, double* div);
void orc_gradient(const orc_grid* g, const double* p, double* grad);
/* Grid::enforce_velocity (grid.rs:683-691) */
void orc_enforce_velocity(const orc_grid* g, double* u);
/* Flow::vector (case.rs:345-367) */
void orc_flow_vector(const orc_case* c, double z, double* v3);

#ifdef __cplusplus
}
#endif
#endif
