// Device kernels shared by the solver drivers: SpMV family, fused reductions,
// level-scheduled triangular sweeps and pivot recurrences, vector updates.
//
// Floating-point order: every per-row sum runs over the row's entries in CSR
// order starting from the value the reference starts from, products rounded
// before the sum (compiled with -ffp-contract=off), so FP64 SpMV, residual,
// sweeps and pivots are bit-identical to the sequential loops of
// src/iccg.f90 / src/bicgstab.f90.  Reductions (dot products, L1/L2 norms) use
// a fixed two-stage order: per-block tree sums, then one block summing the
// block partials; results are deterministic but not bit-identical to a
// sequential sum.
#pragma once

#include <hip/hip_runtime.h>

namespace fcap {

constexpr int kWave = 64;
constexpr int kSpmvWaves = 4;                    // waves per SpMV block
constexpr int kSpmvRows = kWave * kSpmvWaves;    // rows per SpMV block
constexpr int kSpmvCap = 512;                    // LDS entries staged per wave (8 per row on average)
constexpr int kRedBlock = 256;                   // block size of the reduction kernels
constexpr int kMaxRedGrid = 2048;                // grid cap of grid-stride reductions
constexpr int kFcgM = 16;                        // most directions truncated flexible CG can keep
constexpr double kFcgInnerEta = 1e-2;            // default inner tolerance of flexible CG
constexpr int kXcds = 8;                         // XCDs of an MI355X (workgroups are dealt round-robin)
constexpr int kFinBlock = 1024;                  // block size of the finalize kernels
constexpr int kPartRows = 4096;                  // rows per part of the part-wise triangular sweeps

// Device-side solver state shared by the kernels of one solve.  Kernels of an
// iteration return immediately once `done` is set, so iterations can be
// enqueued in batches and the host reads `done` once per batch.
struct Scalars {
    double res0, resl, rsm, resmax, small;
    double s0, sk, bet, alf, pkapk;            // PCG
    double om, gam, beto, ukreso, vr, vv;      // BiCGStab
    double bnorm2, rnorm2, rel;                // L2 quantities (mixed precision / L2 stop)
    double inner_tol2, inner_r0;               // inner-solve stopping threshold
    double fcg_pr;                             // flexible CG: p.r of this step
    double fcg_pqh[kFcgM], fcg_bet[kFcgM];     // flexible CG: p.q and beta per history slot
    double stag_ref, stag_min;                 // flexible CG: best ||r||^2 before / within the window
    double stag_factor;                        // flexible CG: a window's best must beat stag_factor x stag_ref
    int stag_every;                            // flexible CG: stagnation window (0: no check)
    int fcg_m;                                 // flexible CG: directions kept (1..kFcgM)
    int iters, done, status, nsw;
    int stop_rule;     // 0: L1 rsm < resmax; 1: L2 rel < resmax; 2: rnorm2 <= inner_tol2
    int early_exit;    // 1: stop before iterating when res0 < 1e-13 (src/iccg.f90:67-71)
};

enum : int { kStatusRunning = -100 };

// --------------------------------------------------------------- reductions

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

// Block sum (fixed order), result valid in thread 0.
template <int BLOCK>
__device__ inline double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < BLOCK / kWave; ++i) r += sh[i];
    }
    __syncthreads();
    return r;
}

}  // namespace fcap
