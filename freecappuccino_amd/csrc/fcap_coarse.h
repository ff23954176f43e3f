// Aggregation coarse space of the flexible-CG preconditioner (fcap_coarse.hip).
//
// The cells are split into aggregates of `agg` consecutive positions of a cell
// ordering (the network's Morton ordering where it exists: each aggregate is
// then a compact brick of the mesh).  With P the n x nc indicator matrix of the
// aggregates, the coarse operator is A_c = P^T A P (nc x nc, symmetric positive
// definite when A is), factored by dense Cholesky and inverted once per
// assembled matrix; one application of the correction is
//     z <- z + P A_c^{-1} P^T (r - A z).
#pragma once

#include "fcap_internal.h"
#include "fcap_kernels.h"

namespace fcap {

// Cells per aggregate by default, and the most aggregates a system gets: a
// larger mesh gets larger aggregates, so the dense inverse stays at most
// kCoarseMaxAggregates^2 floats.
constexpr int kCoarseCells = 4096;
constexpr int kCoarseMaxAggregates = 4096;

// exclusive prefix sum of n ints, and a radix sort of (64-bit key, int) pairs, on the
// stream (fcap_sort.hip)
void exclusive_scan_i32(hipStream_t s, const int *in, int *out, int n, DevBuf &tmp);
void sort_pairs_u64(hipStream_t s, const unsigned long long *key_in, unsigned long long *key_out, const int *val_in,
                    int *val_out, int n, DevBuf &tmp);

struct CoarseSpace {
    int n = 0, agg = 0, nc = 0;
    long long version = -1;   // assembly the operator was built for
    bool ok = false;          // factored: the correction is applied
    int fbits = 40;           // fixed point of afix: units of 2^-fbits amax
    const int *perm = nullptr, *inv = nullptr;   // position -> cell, cell -> position (null: identity)
    DevBuf afix;    // P^T A P in units of 2^-fbits amax, nc x nc int64, column-major
    DevBuf ac;      // P^T A P, then its inverse (lower triangle), FP64 column-major
    DevBuf ainv;    // the inverse, FP32 row-major (both triangles)
    DevBuf amax;    // largest |a_ij| (FP64 bits as uint64)
    DevBuf info;    // factorisation status (int)
    DevBuf rc, yc;  // P^T (r - A z) and A_c^{-1} P^T (r - A z), FP32
    // Each aggregate's cells in ascending cell order (slots I agg .. (I + 1) agg - 1;
    // null with cell order): the restriction reads r and z in runs of consecutive cells.
    DevBuf cells;
    const int *cell_of = nullptr;
    // P^T A as the restriction reads it, by slot q of aggregate I (cell j = cells[q]):
    // colsum[q] = sum over rows i of I of a_ij, and row j's entries whose column lies in
    // another aggregate, (column cell, a) in xcol/xval[xoff[q] .. xoff[q + 1]) (FP32,
    // from the FP32 matrix)
    DevBuf colsum, xoff, xcol, xval, xcnt, scan_tmp;
    long long nx = -1;               // outside entries (-1: not counted yet)
    const int *cross_ptr = nullptr;  // the CSR row pointer they were counted for
};

// Cells per aggregate for n cells and a requested size (a multiple of 256).
int coarse_aggregate_cells(int n, int requested);

// Build and factor A_c for the FP64 CSR matrix (0-based ptr/ja, n rows) on the
// context's stream.  cs.ok is false afterwards when A_c is not positive definite
// (the correction is then skipped).  perm/inv may be null (cell order).
// max_row: the most entries of a row (bounds the fixed-point sums); a32: the
// FP32 copy of the matrix the preconditioner applies; trans: the CSR position
// of each entry's transpose.
void coarse_setup(fc_context *ctx, CoarseSpace &cs, int n, long long nnz, const int *ptr, const int *ja,
                  const double *a, const float *a32, const int *trans, const int *perm, const int *inv, int agg,
                  int max_row, long long version);

// P^T A P as built (FP64, column-major nc x nc) into host memory.
void coarse_matrix_host(fc_context *ctx, const CoarseSpace &cs, double *out);

// z = zf + P A_c^{-1} P^T (r - A zf) with the FP32 matrix, P^T (r - A zf)
// formed as P^T r - (P^T A) zf without a pass over the matrix; z may alias zf.
// Kernels return at once when sc->done is set.
void coarse_apply(hipStream_t s, CoarseSpace &cs, const float *r, const float *zf, float *z, const Scalars *sc);

// Release the context's dense-solver handle (created on first use).
void coarse_release(fc_context *ctx);

}  // namespace fcap
