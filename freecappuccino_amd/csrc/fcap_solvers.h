// Krylov solver drivers (fcap_solvers.hip).
#pragma once

#include "fcap_internal.h"
#include "fcap_kernels.h"
#include "fcap_linalg.h"

namespace fcap {

struct Workspace {
    // FP64 vectors
    DevBuf r, z, p, q, x, reso, uk, vk, d, tmp;
    // FP32 vectors and matrix copies
    DevBuf r32, z32, p32, q32, e32, a32, d32;
    // reduction partials, summed scalars, device solver state, error flag
    DevBuf pa, pb, sums, sc, err;
    // dependency-driven sweeps: their ticket state and the forward sweep's result
    DevBuf sf, sf_tmp;
    // flexible CG: the last m directions p and their images A p (FP32, m x n each)
    DevBuf ph, qh;
    // the current solve's factor in level-packed order (PackedFactor), FP64 and FP32
    DevBuf pk_af, pk_df, pk_ab, pk_db, pk_af32, pk_df32, pk_ab32, pk_db32;
    int n = 0, nnz = 0;
    void ensure(int n_, int nnz_);
};

struct SolveArgs {
    const Structure *st = nullptr;
    const double *a = nullptr;    // device, nnz
    const double *su = nullptr;   // device, n
    double *fi = nullptr;         // device, n (initial guess in, solution out)
    OcView oc;
    double sor = 1e-8;
    int nsw = 1000;
    int variant = FC_VARIANT_SRC;
    int stop_rule = FC_STOP_L1;
    // relative L2 reduction of each inner FP32 solve (1e-4 was measured slower for
    // FP32 IC(0)-PCG at 192^3: 878 against 698 iterations, its FP32 recursive
    // residual drifting from the true one below ~1e-4)
    double inner_eta = 1e-3;
    int max_outer = 40;
};

// FP64 solvers with the reference's semantics.  On return ws.r holds the
// final (recursively updated) residual vector.
fc_status solve_pcg64(fc_context *ctx, const SolveArgs &args, bool dic, fc_solve_info *info);
fc_status solve_bicgstab64(fc_context *ctx, const SolveArgs &args, fc_solve_info *info);
// FP32 inner solves (DIC-PCG or DILU-BiCGStab) inside FP64 iterative refinement.
fc_status solve_mixed(fc_context *ctx, const SolveArgs &args, bool bicgstab, fc_solve_info *info);

// Allocation-free Jacobi-PCG (fc_dpcg_noalloc): the device addresses of its
// file-scope scratch are taken once per device (noalloc_init, with that device
// current); the solve then uses caller-owned device arrays and that scratch only.
void noalloc_init(int device);
size_t noalloc_bytes();
int noalloc_max_cells();
fc_status solve_dpcg_noalloc(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
                             const double *a, const double *su, double *fi, double *res, double sor, int nsw,
                             fc_solve_info *info);
void cast_d2f(hipStream_t s, long long n, const double *x, float *y);
// Truncated flexible CG (each direction A-orthogonalised against the previous
// kFcgM) in FP32 with preconditioner `P` inside FP64 iterative refinement; on
// stagnation or a non-finite iterate the solve restarts from the last refined
// iterate with FP64 DIC-PCG (status FC_FALLBACK_USED).
struct PrecondHook;
fc_status solve_fcg(fc_context *ctx, const SolveArgs &args, const PrecondHook &P, fc_solve_info *info);
// ||su - A fi||_2 / ||su||_2 in FP64 (O-C terms included).
double true_rel_l2(fc_context *ctx, const SolveArgs &args);

}  // namespace fcap
