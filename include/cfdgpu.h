/*
 * cfdgpu — MI355X-native pressure-projection solver for CFD-craft's staggered
 * uniform Cartesian grids. C-ABI boundary: plain pointers and sizes, opaque
 * handles, no exceptions across the ABI, error text through cfd_last_error().
 *
 * The entry points replace the reference's pressure seam
 * (crates/cfd-craft/src/pressure.rs) and the grid topology it consumes
 * (crates/cfd-craft/src/grid.rs). Each declaration cites the function it
 * stands in for. Reference citations are path:line in the CFD-craft tree.
 *
 * Error codes returned by the solve entry points are the numbers of the
 * reference's coded errors (CFD-E031 -> 31, ...), CFD_CANCELLED for
 * "cancelled", CFD_WALL_TIME for "wall_time_limit" (solver.rs:118-139), or
 * CFD_EINVAL / CFD_EDEVICE for misuse and device failures. On any error
 * the outputs of solve are left untouched and project() leaves u and p
 * unchanged (pressure.rs:468).
 *
 * Threading: distinct grids may be used from distinct threads at once; one
 * grid is not re-entrant (the reference serializes solves on one grid through
 * its multigrid cache mutex, pressure.rs:71-77).
 */
#ifndef CFDGPU_H
#define CFDGPU_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ case */
/* Boundary of one axis pair (case.rs:186-245 Boundary). */
enum { CFD_PERIODIC = 0, CFD_OPEN = 1, CFD_WALL = 2, CFD_SLIP = 3, CFD_GROUND_OPEN_TOP = 4 };
/* Flow model, as far as it classifies open faces (case.rs:345-367 Flow::vector):
 * a constant vector (prescribed_uniform / laminar) or the neutral_rans log law. */
enum { CFD_FLOW_VECTOR = 0, CFD_FLOW_NEUTRAL_RANS = 1 };
/* numerics.pressure_preconditioner (case.rs:850-879). */
enum { CFD_IC0 = 0, CFD_MULTIGRID = 1 };

/* One geometry entry (case.rs:415-459 Block): an axis-aligned box, or a
 * polygon prism when n_vertices > 0. */
typedef struct {
  double min_m[3];
  double max_m[3];
  int32_t n_vertices;
  int32_t reserved;
  const double* footprint_xy; /* 2*n_vertices doubles: x0 y0 x1 y1 ... */
} cfd_block;

/* The keys of a case that fix the pressure operator (case.rs:88-130):
 * domain.size_m, grid.cells, boundaries, flow, geometry and
 * numerics.pressure_preconditioner. */
typedef struct {
  int64_t cells[3];
  double size_m[3];
  int32_t boundary[3];
  int32_t flow_kind;
  double velocity_m_s[3];
  double speed_m_s, reference_height_m, roughness_m, direction_deg_from;
  int32_t preconditioner;
  int32_t n_blocks;
  const cfd_block* blocks;
} cfd_case;

/* ------------------------------------------------------------------ results */
enum {
  CFD_OK = 0,
  CFD_E031 = 31, /* incompatible flux in a closed fluid component */
  CFD_E032 = 32, /* pressure CG lost positive definiteness */
  CFD_E033 = 33, /* pressure iteration limit */
  CFD_E034 = 34, /* invalid multigrid topology or weights */
  CFD_E035 = 35, /* coarse Cholesky lost positive definiteness */
  CFD_E036 = 36, /* line (block smoother) Cholesky lost positive definiteness */
  CFD_CANCELLED = 100,
  CFD_WALL_TIME = 101,
  CFD_EINVAL = 200,
  CFD_EUNSUPPORTED = 201,
  CFD_EDEVICE = 202,
  CFD_ENOMEM = 203
};

/* LinearInfo (pressure.rs:29-36) plus the Span work counters
 * (profiling.rs:112-135) and MultigridWork (multigrid.rs:78-114). */
typedef struct {
  int64_t iterations;
  double relative_residual; /* recomputed (true) residual / |b_c| */
  double recurrent_residual;
  int64_t operator_applications;
  int64_t preconditioner_applications;
  int64_t true_residual_checks;
  int64_t cycles;
  int64_t smoothing_cell_updates;
  int64_t residual_cell_updates;
  int64_t coarse_solves;
  int64_t hierarchy_builds;
  double seconds; /* device time of the solve, HIP events */
  /* MultigridWork (multigrid.rs:74-116) beyond the fields above */
  int64_t restriction_cell_updates;
  int64_t prolongation_cell_updates;
  int64_t block_smoothing_cell_updates;
  int64_t block_solves;
  int64_t block_level_cell_visits;
  int64_t hierarchy_levels;
  int64_t hierarchy_cells;
  int64_t hierarchy_nonzeros;
  int64_t hierarchy_bytes;
  /* host wall seconds of the hierarchy build that hierarchy_builds reports (the
   * hierarchy is built when the grid is created or a cycle is first selected;
   * the first solve after the build reports it, as the reference's lazy build
   * does, pressure.rs:71-93) */
  double hierarchy_build_seconds;
} cfd_linear_info;

/* BlockSmootherStats (multigrid/line.rs:9-22): the line or small-plane block
 * smoother of one anisotropic level of the unmasked hierarchy. */
typedef struct {
  int64_t level;
  int64_t n_axes;
  int64_t axes[3]; /* ascending; the first n_axes are used */
  int64_t block_unknowns;
  int64_t block_count;
  int64_t half_bandwidth;
  int64_t allocated_bytes; /* device storage of the blocks' order, factors and outside rows */
} cfd_block_smoother_stats;

/* MultigridStats (multigrid.rs:55-72). level_dims: cell dimensions per level of
 * the structured hierarchy (trilinear Galerkin), zero for the
 * aggregation hierarchy, whose levels are irregular (masked.rs:262). */
typedef struct {
  int32_t levels;
  int32_t reserved;
  int64_t level_cells[32];
  int64_t level_nonzeros[32];
  int64_t level_dims[32][3];
  int64_t allocated_bytes;
  int64_t n_block_smoothers;
  cfd_block_smoother_stats block_smoothers[32];
  /* cfd_case_multigrid_stats only (0 elsewhere): FNV-1a digest of the built
   * hierarchy's indices and values, equal for two builds that agree bit for bit */
  uint64_t digest;
} cfd_mg_stats;

/* Control::check as a callback (solver.rs:122-139): 0 continue, 1 cancelled,
 * 2 wall_time_limit. Called before the first PCG iteration and then at every
 * read-back of the solver state, at most 8 iterations apart (pressure.rs:387-389);
 * up to 16 further iterations may be queued on the device by then, and they end
 * without effect. */
typedef int (*cfd_poll_fn)(void* user);

/* Solver options beyond the reference's case keys (serving and precision). */
enum { CFD_PRECISION_FP64 = 0, CFD_PRECISION_MIXED = 1 };
/* Multigrid cycle (numerics.pressure_preconditioner = multigrid):
 *  CFD_CYCLE_RB         the reference's aggregation hierarchy (masked.rs:181-291)
 *                       with red-black Gauss-Seidel and a scaled coarse correction
 *                       (the throughput path, FP64 or mixed);
 *  CFD_CYCLE_REFERENCE  the reference's own cycle: its hierarchy and lexicographic
 *                       Gauss-Seidel (multigrid.rs:667-726, masked.rs:304-342), FP64;
 *  CFD_CYCLE_ZEBRA      the reference's hierarchy (with its line and small-plane
 *                       block smoothers and semicoarsening on anisotropic grids)
 *                       swept by colour classes instead of lexicographically: a
 *                       few launches per sweep; for anisotropic solid-free grids,
 *                       where point smoothers stall. FP64. */
enum { CFD_CYCLE_RB = 0, CFD_CYCLE_REFERENCE = 1, CFD_CYCLE_ZEBRA = 3 };  /* 2: not a cycle */
typedef struct {
  int32_t precision;   /* FP64 (plain) or FP32 multigrid/IC(0)-apply inside FP64 PCG */
  int32_t warm_start;  /* 1: start PCG from the grid's resident previous solution */
  int32_t poll_every;  /* PCG iterations between host polls (multiple of 8; default 8) */
  int32_t cycle;       /* CFD_CYCLE_* (multigrid only) */
} cfd_options;

typedef struct cfd_grid cfd_grid;

/* ------------------------------------------------------------------ grid */
/* Grid::build (grid.rs:246-464): masks the geometry, classifies every face,
 * finds the fluid components and uploads the topology (1 byte per cell) to
 * the current HIP device. Builds the IC(0) factor diagonal or the multigrid
 * hierarchy for the selected preconditioner (pressure.rs:55-97). Errors:
 * CFD-E028/E030 as CFD_EINVAL with the reference text. */
int cfd_grid_create(const cfd_case* c, cfd_grid** out);
void cfd_grid_destroy(cfd_grid* g);
/* The multigrid cache's Signature::matches (multigrid.rs:158-222, checked before
 * every multigrid solve at pressure.rs:69-92): *same = 1 when the case gives the
 * grid's operator exactly (cell counts, spacing bits, fluid mask, anchored
 * components, every interior and outlet face with its weight bits, periodic
 * axes, preconditioner), else 0. A caller that keeps a context per grid and
 * lets the grid change calls this before a solve and recreates the context on
 * 0; the next solve then reports hierarchy_builds = 1, as the reference's does.
 * Builds the case's topology on the host (O(cells)); errors of Grid::build are
 * returned as by cfd_grid_create. */
int cfd_grid_matches(const cfd_grid* g, const cfd_case* c, int* same);

/* FaceKind (grid.rs:60-81). */
enum { CFD_FACE_INTERIOR = 0, CFD_FACE_WALL = 1, CFD_FACE_SLIP = 2, CFD_FACE_INLET = 3, CFD_FACE_OUTLET = 4,
       CFD_FACE_BLOCKED = 5 };
/* One entry of Grid::faces (grid.rs:85-108): normal axis, kind, the fluid cells
 * below and above (-1 for None; across a periodic boundary the wrapped cell),
 * area (m^2) and gradient distance (m). A caller may have edited area or
 * distance after Grid::build (pressure.rs:244). */
typedef struct {
  int32_t axis;
  int32_t kind;
  int64_t lo, hi;
  double area;
  double distance;
} cfd_face;
/* The built Grid as the pressure path reads it (grid.rs:114-141): cell counts,
 * cell sizes and volume, periodic axes (face_dims[a][a] == n[a]), the fluid
 * mask, every face in Grid::faces order (x-faces, y-faces, z-faces, each x
 * fastest over face_dims), and optionally Grid::component (-1 for solid
 * cells) with Grid::anchored, which are then checked against the faces. */
typedef struct {
  int64_t n[3];
  double dx[3];
  double volume;
  int32_t periodic[3];
  const uint8_t* fluid;
  int64_t face_count;
  const cfd_face* faces;
  const int64_t* component;  /* optional (NULL) */
  int64_t component_count;
  const uint8_t* anchored;   /* optional (NULL), component_count entries */
  int32_t preconditioner;    /* CFD_IC0 or CFD_MULTIGRID */
} cfd_grid_desc;
/* The grid context from a Grid the caller already built, in place of
 * cfd_grid_create's own Grid::build (the pressure path's view of the Rust
 * Grid: Operator::build reads the faces, pressure.rs:255-279). The faces'
 * order, kinds and cells are checked against the mask and the periodic axes
 * (CFD_EINVAL naming the first face that differs). The device operator holds
 * one interior and one outlet weight area / (distance * volume) per axis: a
 * grid whose interior (or outlet) faces of one axis carry different weight
 * bits, as after a caller's edit of one face, is refused with
 * CFD_EUNSUPPORTED naming the face; Wall, Slip, Inlet and Blocked faces carry
 * no pressure weight and may be edited freely. */
int cfd_grid_create_desc(const cfd_grid_desc* d, cfd_grid** out);
/* Signature::matches (multigrid.rs:158-222) against a descriptor: *same = 1
 * when the descriptor gives the grid's operator exactly (cell counts, spacing
 * and volume bits, fluid mask, anchoring, every interior and outlet face with
 * its weight bits, periodic axes, preconditioner), else 0 (also when its
 * weights vary within an axis, which cfd_grid_create_desc then refuses). */
int cfd_grid_matches_desc(const cfd_grid* g, const cfd_grid_desc* d, int* same);
/* counts: cells, faces, fluid cells, components, anchored components, x/y/z face counts. */
int cfd_grid_counts(const cfd_grid* g, int64_t counts[8]);
/* Per cell coupling code: bit 2a = lower face on axis a couples (interior or
 * outlet), bit 2a+1 = upper face, bit 6 = fluid, bit 7 = in a closed component. */
int cfd_grid_codes(const cfd_grid* g, uint8_t* codes);
/* Grid.component / Grid.anchored (grid.rs:129-138). */
int cfd_grid_components(const cfd_grid* g, int64_t* component, uint8_t* anchored);
int cfd_grid_set_options(cfd_grid* g, const cfd_options* o);
/* The IC(0) factor diagonal of pressure.rs:191-214 (host copy, for parity checks). */
int cfd_ic0_diagonal(const cfd_grid* g, double* diag);
/* multigrid_stats (pressure.rs:100-106) of the device hierarchy. */
int cfd_multigrid_stats(const cfd_grid* g, cfd_mg_stats* s);

/* Host-only inspection of Grid::build (grid.rs:246-464) without a device: the
 * counts of cfd_grid_counts, and optionally the per-cell coupling codes, the
 * component of each cell, the anchored flag of each component (at most
 * anchored_cap) and the IC(0) factor diagonal (pressure.rs:191-214). Any
 * output pointer may be NULL. */
int cfd_case_topology(const cfd_case* c, int64_t counts[8], uint8_t* codes, int64_t* component, uint8_t* anchored,
                      int64_t anchored_cap, double* ic0_diagonal);

/* The same topology built on the current device (Grid::build, grid.rs:246-464):
 * mask rasterization, face codes and union-find components in HIP kernels, the
 * result copied back. cfd_grid_create builds its grids this way unless
 * CFD_GPU_SETUP=0. */
int cfd_case_topology_device(const cfd_case* c, int64_t counts[8], uint8_t* codes, int64_t* component,
                             uint8_t* anchored, int64_t anchored_cap);

/* Host-only MultigridStats (multigrid.rs:55-72) of the hierarchy that `cycle`
 * builds for a case: unknowns and stored entries per level. CFD_CYCLE_REFERENCE:
 * the reference's own hierarchy (masked aggregation, masked.rs:181-291, or
 * trilinear Galerkin, multigrid.rs:530-645); CFD_CYCLE_RB: the masked
 * aggregation down to the dense coarsest level. */
int cfd_case_multigrid_stats(const cfd_case* c, int32_t cycle, cfd_mg_stats* s);

/* ------------------------------------------------------------------ host-buffer seam */
/* pressure::apply (pressure.rs:293-312): y = A x. */
int cfd_apply(cfd_grid* g, const double* x, double* y);
/* pressure::center_nullspaces (pressure.rs:316-325). */
int cfd_center_nullspaces(cfd_grid* g, double* x);
/* pressure::solve_scaled (pressure.rs:342-436); scale may be NULL (= solve, 329-337). */
int cfd_solve_scaled(cfd_grid* g, const double* b, const double* roundoff_scale, double tol, int64_t max_iterations,
                     cfd_poll_fn poll, void* poll_user, double* x, cfd_linear_info* info);
/* pressure::project (pressure.rs:469-502): u has one value per face in the
 * reference's face order (x-faces, y-faces, z-faces, each x fastest). */
int cfd_project(cfd_grid* g, double* u, double* p, double dt, double tol, int64_t max_iterations, cfd_poll_fn poll,
                void* poll_user, cfd_linear_info* info);
/* The same projection with no device allocation, for callers that cannot
 * allocate (code under HIP graph capture elsewhere in the process, tight
 * serving loops, harnesses that forbid hipMalloc). No counterpart in the
 * reference. cfd_project stages u and p in device buffers it allocates per
 * call; this entry point stages them in a static device array of 6 Mi doubles
 * (48 MiB) that the library holds on every GPU it runs on until the process
 * ends, and launches the PCG iterations kernel by kernel (no graph is
 * instantiated). A grid whose faces plus cells exceed 6 Mi values (about 1.4 M
 * cells) is refused with CFD_EINVAL before any device work. All calls on one
 * GPU share that GPU's array: calls that overlap on one GPU (two threads, two
 * grids) race and must be serialized by the caller. Same contract as
 * cfd_project: u and p are left unchanged on any error. */
int cfd_project_static(cfd_grid* g, double* u, double* p, double dt, double tol, int64_t max_iterations,
                       cfd_poll_fn poll, void* poll_user, cfd_linear_info* info);
/* projection_rhs (pressure.rs:440-459). */
int cfd_projection_rhs(cfd_grid* g, const double* u, double dt, double* b, double* terms);
/* One preconditioner application z = M r: Factor::apply (pressure.rs:223-241)
 * or one V-cycle (multigrid.rs:656-665). */
int cfd_precondition(cfd_grid* g, const double* r, double* z);

/* ------------------------------------------------------------------ device-resident seam */
/* The same operations on device pointers (HBM), ordered on the grid's stream. */
int cfd_apply_device(cfd_grid* g, const double* d_x, double* d_y);
int cfd_solve_scaled_device(cfd_grid* g, const double* d_b, const double* d_scale, double tol, int64_t max_iterations,
                            cfd_poll_fn poll, void* poll_user, double* d_x, cfd_linear_info* info);
int cfd_precondition_device(cfd_grid* g, const double* d_r, double* d_z);
int cfd_project_device(cfd_grid* g, double* d_u, double* d_p, double dt, double tol, int64_t max_iterations,
                       cfd_poll_fn poll, void* poll_user, cfd_linear_info* info);

/* Measurement hook (the GPU counterpart of the reference's phase spans,
 * profiling.rs:217-281): average device duration of one launch of a hot kernel
 * on the grid's stream, from HIP events around `reps` back-to-back launches over
 * the grid's own work vectors, and the kernel's algorithmic HBM bytes per
 * launch (each array crossing HBM once). kernel: 0 fused p update + operator +
 * p.Ap of a first iteration (p = z); 1 fine-level red-black half-sweep of the
 * multigrid smoother; 2 x/r update with |r|^2; 3 one IC(0) application (both
 * sweeps); 4 one V-cycle; 5 the operator kernel of a later iteration
 * (p = z + beta p_old). */
int cfd_profile_kernel(cfd_grid* g, int kernel, int reps, double* avg_seconds, double* bytes_per_launch);

/* Kernel classes of cfd_profile_solve. */
enum {
  CFD_PROBE_FINE_HALF = 1,        /* fine red-black half-sweep */
  CFD_PROBE_COARSE_HALF = 2,      /* coarse half-sweep of one colour class */
  CFD_PROBE_OPERATOR = 3,         /* p update, A p and p . Ap */
  CFD_PROBE_AXPY = 4,             /* x and r update with |r|^2 */
  CFD_PROBE_FINE_RESTRICT = 5,    /* fine residual and restriction */
  CFD_PROBE_FINE_RZ = 6,          /* last fine half-sweep with z and r . z */
  CFD_PROBE_FINE_FIRST = 7,       /* first fine half-sweep from zero */
  CFD_PROBE_COARSE_RESTRICT = 8,  /* coarse restriction */
  CFD_PROBE_COARSEST = 9,         /* dense coarsest solve */
  CFD_PROBE_FINE_PROLONG = 10,    /* coarse correction of the fine red cells */
  CFD_PROBE_IC0 = 11,             /* one IC(0) application */
  CFD_PROBE_TRUE_RESIDUAL = 12    /* recomputed residual and convergence decision */
};
typedef struct {
  int64_t launches;  /* launches of the class that did the solve's work */
  double seconds;    /* their summed device duration (a HIP event pair around each) */
  double bytes;      /* their summed algorithmic HBM bytes (0 where not counted) */
  int64_t dropped;   /* launches queued past convergence (they end at once), not counted */
  int32_t truncated; /* 1 when the event pool ran out (later launches not counted) */
  int32_t reserved;
} cfd_kernel_probe;
/* Measurement hook (the GPU counterpart of the reference's phase spans,
 * profiling.rs:217-281): one solve_scaled on device buffers with the iterations
 * launched kernel by kernel and every launch of one kernel class bracketed by a
 * pair of HIP events on the grid's stream; the class's launch count, summed
 * durations and algorithmic bytes. */
int cfd_profile_solve(cfd_grid* g, const double* d_b, const double* d_scale, double tol, int64_t max_iterations,
                      double* d_x, int kernel, cfd_kernel_probe* out, cfd_linear_info* info);

/* ------------------------------------------------------------------ serving */
/* Continuous batching of independent pressure solves (wind-direction campaigns,
 * optimisation jobs): no counterpart in the reference, whose concurrent solves
 * are separate processes or threads (validation/urban_campaign.py:197-290).
 * A server admits grids against an HBM budget by their measured footprint and
 * runs up to max_active requests at once, each on its grid's own stream; a
 * request that arrives while others are on the GPU starts as soon as a slot and
 * its grid are free. Results do not depend on what else runs (fixed-shape
 * reductions). One server is driven by one host thread. */
typedef struct cfd_server cfd_server;
/* hbm_budget_bytes 0: 90 % of the free device memory. */
int cfd_server_create(int max_active, size_t hbm_budget_bytes, cfd_server** out);
/* Batching: requests on red-black multigrid grids without closed components
 * (nx a multiple of 4) run in lanes, each a stream running a batch of up to 8
 * solves per precision with one launch per stage of a PCG iteration for the
 * whole batch; lanes run concurrently. Each solve's result is bitwise the one
 * its grid gives alone. lanes 0: every request on its grid's own stream; the
 * default is 3. Set before the first request. Grids admitted while lanes > 0
 * take two fine cells per thread (the batched layout; the same values). */
int cfd_server_set_batching(cfd_server* s, int lanes);
void cfd_server_destroy(cfd_server* s);
/* cfd_grid_create under the budget; CFD_ENOMEM (nothing allocated) when it does not fit. */
int cfd_server_admit(cfd_server* s, const cfd_case* c, cfd_grid** out);
/* Queued admission: the case is copied and its grid built by one of the
 * server's builder threads (two; CFD_SERVE_BUILDERS 1..4) as soon as the budget
 * has room for it beside the admitted grids and the builds under way; builds
 * start in ticket order (a grid that does not fit waits for cfd_server_release
 * of others, and the admissions behind it wait too) and may end out of order.
 * Grid set-ups so overlap the running solves. Room is judged by an estimate (the cells
 * times the largest footprint per cell measured so far, 256 B a cell before
 * the first grid) and checked against the measured footprint after the build
 * (a grid over the estimate that does not fit is freed and waits again with its
 * measured size). cfd_server_admit does not wait for queued admissions. */
int cfd_server_admit_async(cfd_server* s, const cfd_case* c, uint64_t* ticket);
/* done = 0 (out NULL) while the admission waits or builds; done = 1 with the
 * grid in out once admitted, or with the build's error code (CFD_ENOMEM when
 * the grid alone exceeds the budget). The ticket is forgotten once done. */
int cfd_server_admitted(cfd_server* s, uint64_t ticket, int* done, cfd_grid** out);
int cfd_server_release(cfd_server* s, cfd_grid* g);
int cfd_server_usage(const cfd_server* s, size_t* used, size_t* budget);
size_t cfd_grid_footprint(const cfd_grid* g);
/* Queues a solve of grid g (device pointers); precision CFD_PRECISION_FP64 or MIXED. */
int cfd_server_submit(cfd_server* s, cfd_grid* g, const double* d_b, const double* d_scale, double tol,
                      int64_t max_iterations, int precision, double* d_x, uint64_t* ticket);
/* Advances every running request without blocking; returns the number running or queued
 * (a negative CFD error code when the device failed). */
int cfd_server_progress(cfd_server* s);
int cfd_server_wait(cfd_server* s, uint64_t ticket);
/* done = 1 when the request has finished and its result can be taken (no wait). */
int cfd_server_finished(cfd_server* s, uint64_t ticket, int* done);
int cfd_server_drain(cfd_server* s);
/* Result of a finished request (then forgotten): the solve's code, info, latency
 * (submit to finish) and queueing time (submit to start), in seconds. */
int cfd_server_result(cfd_server* s, uint64_t ticket, cfd_linear_info* info, double* latency_s, double* queued_s);

/* Workload helper (no counterpart in the reference's pressure path, whose
 * campaign scripts form their right-hand sides on the host): d_dst = d_src +
 * rel * rms(d_src) * xi on the fluid cells of g (d_src elsewhere), rms over the
 * fluid cells and xi a standard normal field fixed by seed (SplitMix64 mixes of
 * the seed and the cell index, Box-Muller), the chained right-hand sides of a
 * wind-direction campaign formed in HBM. Synchronous; not while a solve on g
 * runs (it shares the grid's reduction buffer). d_src == d_dst is allowed. */
int cfd_grid_perturb(cfd_grid* g, const double* d_src, double* d_dst, double rel, uint64_t seed);

/* Device memory helpers for callers without their own allocator. */
int cfd_device_alloc(size_t bytes, void** d_ptr);
int cfd_device_free(void* d_ptr);
int cfd_copy_to_device(void* d_dst, const void* h_src, size_t bytes);
int cfd_copy_to_host(void* h_dst, const void* d_src, size_t bytes);
int cfd_device_synchronize(void);

const char* cfd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
