// Learned preconditioner: message-passing network over the cell graph (fcap_gnn.hip).
#pragma once

#include "fcap_coarse.h"
#include "fcap_internal.h"
#include "fcap_kernels.h"

namespace fcap {

constexpr int kGnnH = 32;                                  // hidden width
constexpr int kGnnL = 4;                                   // message-passing layers
constexpr int kGnnLayerStride = 2 * kGnnH * kGnnH + kGnnH + 3;   // Ws, Wm, b, gate(3)
constexpr int kGnnOffEncW = 0;
constexpr int kGnnOffEncB = kGnnH;
constexpr int kGnnOffLayer0 = 2 * kGnnH;
constexpr int kGnnOffDecW = kGnnOffLayer0 + kGnnL * kGnnLayerStride;
constexpr int kGnnOffDecB = kGnnOffDecW + kGnnH;
constexpr int kGnnParams = kGnnOffDecB + 1;

// Per-system graph data the network reads: packed BF16 edge features per CSR
// entry, FP32 diagonal, transposed-entry index, FP32 matrix copy.
struct GnnGraph {
    int n = 0, nnz = 0;
    const int *ptr = nullptr, *ja = nullptr, *diagpos = nullptr;
    DevBuf ef, dg, trans, a32;
    // The off-diagonal entries again as an ELL over the cells: slot k of cell i
    // at [k * n + i], in CSR order; padding slots hold the cell itself and zero
    // features and lie past ecnt[i].  Lanes of a wave own consecutive cells, so
    // their neighbour and feature loads are coalesced.
    int ew = 0;                      // ELL width: the largest off-diagonal count
    DevBuf ecnt, enbr, eef, eeft;    // count; neighbour j; features of (i,j); of (j,i)
    // Processing order of the 64-cell blocks of the layer kernels: the blocks
    // sorted by the Morton code of their centroid, so that the blocks in flight
    // on one XCD cover a compact region and the rows their neighbours gather
    // are still in that XCD's L2 (in index order a plane of a structured mesh
    // separates two neighbours).
    int nblk = 0;
    DevBuf border;
    // Cluster layout of the inference layers (k_gnn_layer_cl).  The network's
    // cells are renumbered: positions sorted by the Morton code of the cells'
    // per-axis centroid ranks, so each run of kGnnClCells consecutive positions
    // is a compact cluster (an 8 x 8 x 4 brick of a structured hex box, graded
    // or not).  A workgroup stages its cluster's rows and its halo -- the
    // distinct outside neighbours, at most kGnnClHalo -- in LDS and gathers every
    // message from there.  cperm: position -> cell (n_pad = ncl * kGnnClCells
    // entries, -1 past n); cinv: cell -> position; per cluster the halo count and
    // kGnnClHalo halo positions; per slot pair (2q, 2q + 1) of a position the two
    // LDS row indices as 16-bit halves at [q * n_pad + pos] (0xffff: padding
    // slot); the edge features of the slots at [k * n_pad + pos].  Built once per
    // structure (cl_n); the features again at every assembly.
    int cl_n = -1, ncl = 0;
    bool cl_ok = false;              // every halo fits
    int cl_halo_max = 0;             // largest halo found
    DevBuf cperm, cinv, hcnt, halo, lidx, efc;
    // gate pairs of every layer for one set of weights (k_cl_gates), tagged with
    // the weights' version; -1 after the features change
    DevBuf cgate;
    long long gate_wver = -1;
    // coarse correction of flexible CG: aggregates of consecutive positions of the
    // cluster layout's order (cell order without one), rebuilt per assembly
    CoarseSpace coarse;
    DevBuf zscratch;   // the Jacobi steps' second vector
    // structure the per-structure parts were built for (Structure::key, numCells)
    uint64_t skey = 0;
    int n_struct = -1;
    bool ready = false;
};

constexpr int kGnnClCells = 256;   // cells per cluster (threads of the layer's workgroup)
constexpr int kGnnClHalo = 384;    // halo rows a cluster may stage

// (key, value) pairs sorted by key on the stream (radix sort; stable)
void sort_pairs_u64(hipStream_t s, const unsigned long long *key_in, unsigned long long *key_out, const int *val_in,
                    int *val_out, int n, DevBuf &tmp);

// Build the graph data of an assembled system (device arrays of the system).
void gnn_prepare_graph(fc_context *ctx, GnnGraph &g, const Structure &st, const int *kface, const int *owner0,
                       const int *ij, const int *ji, int nif, const double *arx, const double *ary,
                       const double *arz, const double *xc, const double *yc, const double *zc, const double *a);

// Preconditioner hook used by flexible CG: z = P(r) on the context's stream;
// kernels return at once when sc->done is set.  light: the application leaves
// the network out (the Jacobi guess r / a_ii takes its place).
struct PrecondHook {
    void *obj = nullptr;
    void (*apply)(void *obj, hipStream_t s, const float *r, float *z, const Scalars *sc, bool light) = nullptr;
};

}  // namespace fcap
