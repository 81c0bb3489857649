// Plain-old-data structures shared by the kernels and the host launchers.
#pragma once
#include <cstdint>

namespace cfdgpu {

// Geometry of a structured level as the kernels see it.
struct Geo {
  int n[3];
  int per[3];  // periodic flags
  long long N;
  long long sy, sz;  // strides of y and z (x is 1)
  double wi[3];      // interior face weight per axis (fine level)
  double wo[3];      // outlet face weight per axis (fine level)
};

// Reduction grids: a fixed number of blocks writes one partial each.
constexpr int kReduceBlock = 1024;
constexpr int kReduceGrid = 512;
// Regions of a grid's partials buffer, one per reduction site of an iteration.
constexpr int P_PAP = 0;                // p . Ap
constexpr int P_RR = kReduceGrid;       // |r|^2
constexpr int P_TR = 2 * kReduceGrid;   // |A x - b_c|^2
constexpr int P_RZ = 3 * kReduceGrid;   // r . z
constexpr int P_CS = 4 * kReduceGrid;   // component sums (2 per block or chunk)

// Device-resident PCG scalars (FP64) and control words.
enum : int {
  ST_RUNNING = 0,
  ST_CONVERGED = 1,
  ST_ZERO = 2,
  ST_E031 = 31,
  ST_E032 = 32,
  ST_E033 = 33,
};
struct PcgState {
  double rz;        // r . z of the current iteration
  double denom;     // p . Ap
  double alpha, beta;
  double norm;      // |b_c|
  double rel;       // recurrent |r| / norm
  double true_rel;  // recomputed |A x - b_c| / norm
  double tol;
  double accept;    // 10 tol + 1e-12
  double spare[7];
  long long it;     // completed iterations
  long long max_it;
  long long true_checks;
  int status;
  int check;  // recurrent residual crossed tol: recompute the true residual
  int first;  // next p update is p = z_c
  int spare_i;
  unsigned int ticket;
  unsigned int pad[3];
};

// The reference's face list (grid.rs:341-429) as three blocks of face_dims each.
struct FaceGeo {
  long long off[3];
  int dims[3][3];
  double dx[3];
  long long total;
};

// One coarse level of the aggregation hierarchy on the device: a CSR matrix over
// aggregates. In the colour-sorted numbering the rows of colour c are the range
// [cptr[c], cptr[c+1]) and form an independent set.
constexpr int kMaxColours = 16;
template <class T>
struct AggDevLevel {
  int n;
  int ncol;
  int nnz;  // stored off-diagonal entries (the CSR form)
  int cptr[kMaxColours + 1];
  const int* ptr;  // off-diagonal entries of each row
  const int* col;
  const T* val;
  const T* diag;  // 0: a node without couplings (its value stays 0)
  T* b;
  T* x;
  const int* parent;     // node of the next level (null at the coarsest)
  const int* child_ptr;  // finer nodes of each node
  const int* child;
  // the same rows padded to K <= kEllWidth entries (null when a row is longer,
  // the level has 2^24 nodes or more, or more than 256 distinct off-diagonal
  // values): entry k of row i at [k * n + i] holds the column in its low 24 bits
  // and the index of its value in etab in the high 8; padding points at the row
  // itself with the value 0
  int K;
  const unsigned* ecol;
  const T* etab;
};
constexpr int kEllWidth = 8;

// Colour-separated copy of the fine level for the red-black cycle: the cells of
// colour c (i + j + k = c mod 2) at [c * half + (i >> 1) + hx * (j + ny * k)],
// hx = (nx + 1) / 2, half = hx * ny * nz. A half-sweep then reads only the
// other colour's iterate and its own right-hand side and codes.
template <class T>
struct FineRB {
  long long half;
  int hx;
  const uint8_t* codes;
  T* b;
  T* x;
  int two_cells;     // half-sweeps take two cells per thread (nx a multiple of 4; chosen at set-up): 1 in
                     // tiles of 128 x 4 entries, 2 consecutive threads on consecutive pairs of a colour
  const int* rpar;   // per red cell (red layout): node of the first coarse level, -1 without one
  const int* bpar;   // per 2x2x2 block of cells (x fastest): the common node of the first coarse
                     // level of its cells in the grid (-1: all solid), kMixedParent when they differ
  unsigned long long mhx = 0, mny = 0;  // ceil(2^40 / hx), ceil(2^40 / ny): row and layer of a linear entry
};
constexpr int kMixedParent = -2147483647 - 1;

// Coarsest level: one dense block per connected component. Blocks of at most
// 64 nodes take one wavefront each; larger ones ("big", listed separately) one
// wavefront per row.
constexpr int kSmallBlock = 64;
struct AggCoarsest {
  int nblocks;
  const int* bptr;        // members of block k: rows[bptr[k] .. bptr[k+1])
  const int* rows;
  const long long* goff;  // block k: G[goff[k] ..] row-major m x m
  const double* G;
  int nbig;
  int big_rows;           // rows of all big blocks
  const int* big_block;   // block index of big block b
  const int* big_first;   // first row (over all big blocks) of big block b, nbig + 1 entries
  int nsmall;
  const int* small_block;  // block indices of the small blocks
};

}  // namespace cfdgpu
