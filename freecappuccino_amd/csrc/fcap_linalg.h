// Launchers of the linear-algebra kernels (fcap_linalg.hip).
#pragma once

#include "fcap_internal.h"
#include "fcap_kernels.h"

namespace fcap {

// O-C couplings of a solve (device pointers, 0-based cells), per-cell lists in cut order.
struct OcView {
    int noc = 0;
    const int *ijl = nullptr, *ijr = nullptr, *co_ptr = nullptr, *co_idx = nullptr;
    const double *al = nullptr, *ar = nullptr;
};

int red_grid(int n);
int spmv_grid(int n);
// SpMV block order for the whole process: 0 natural (default), 1 XCD-aware

// mode 0: y = A x; 1: y = b - A x with L1 (pa) and L2^2 (pb) partials and an
// optional FP32 copy y32; 2: y = A x with partial y.v (pa); 3: as 2 plus y.y (pb)
void spmv_launch64(hipStream_t s, int mode, int n, const int *ptr, const int *ja, const double *a, const double *x,
                   const double *b, const double *v, double *y, float *y32, double *pa, double *pb,
                   const Scalars *sc);
void spmv_launch32(hipStream_t s, int mode, int n, const int *ptr, const int *ja, const float *a, const float *x,
                   const float *b, const float *v, float *y, double *pa, double *pb, const Scalars *sc);
// y = x + omega (b - A x) / d (a damped-Jacobi step; y must not alias x)
void jacobi_step32(hipStream_t s, int n, const int *ptr, const int *ja, const float *a, const float *x, const float *b,
                   const float *d, float omega, float *y, const Scalars *sc);
// y = b - A x, no partial sums (y must not alias x)
void residual32(hipStream_t s, int n, const int *ptr, const int *ja, const float *a, const float *x, const float *b,
                float *y, const Scalars *sc);
void oc_launch64(hipStream_t s, int n, const OcView &oc, const double *x, double *y, int sign, const Scalars *sc);

// State of the dependency-driven sweeps: the next block of 64 packed rows to
// hand out, the number of waves that have left the kernel (the last one resets
// both for the next launch), a sticky abort flag raised when a wait exceeds
// its time limit, that limit in microseconds (0: the built-in one) and the
// rate of the device's wall clock in kHz.
struct SyncFree {
    unsigned ticket, fin;
    int abort, wait_us;
    unsigned rate_khz, pad;
};
// zeroed state with the context's wait limit, on the stream (not for graph capture)
void syncfree_reset(fc_context *ctx, hipStream_t s, SyncFree *sf);

// How the triangular sweeps run: one launch per level (default without sf), or
// one launch per sweep in which every row waits only for the rows it reads (sf;
// needs the packed factor; `waves` persistent one-wave workgroups, `tmp` holds
// the forward result).
struct SweepCfg {
    SyncFree *sf = nullptr;
    int waves = 0;
    void *tmp = nullptr;
    bool parts = false;   // part-wise schedule (Structure::rows_pf ...; factor packed in part order)
};
// part-wise sweeps (FC_SWEEP_PARTS): the structure has them, and at least half of each
// triangle's entries lie in their row's own part (else the dependency-driven schedule)
bool part_sweeps(const Structure &st);
// dependency-driven sweeps when the structure is packed at a width they handle
bool syncfree_sweeps(const Structure &st);

// The row-wise dependency-driven schedule pays where levels are narrow (deep chains).
// Where the average level holds at least kWideLevelRows rows, one launch per level
// keeps the GPU as busy and saves the polling: 96^3 permuted Kuhn tets (16 levels of
// 330K rows) take 0.66 ms per DILU sweep pair with level launches against 2.0 ms
// row-wise (with 256 to 4,096 waves).
constexpr long long kWideLevelRows = 65536;
inline bool wide_levels(const Structure &st) {
    const long long nlev = std::max(1, std::max(st.nlev_f, st.nlev_b));
    return (long long)st.n / nlev >= kWideLevelRows;
}

// The schedule the context's sweep mode gives on a structure: part-wise (parts),
// row-wise dependency-driven (dep only) or one launch per level (neither).  The
// default (FC_SWEEP_PARTS) falls back to the row-wise schedule, or to level launches
// where levels are wide; FC_SWEEP_DEPENDENT keeps the row-wise one.
struct SweepChoice {
    bool dep = false, parts = false;
};
inline SweepChoice choose_sweeps(const fc_context *ctx, const Structure &st) {
    SweepChoice c;
    const int mode = ctx->sweep_mode;
    if ((mode != FC_SWEEP_DEPENDENT && mode != FC_SWEEP_PARTS) || !syncfree_sweeps(st)) return c;
    if (mode == FC_SWEEP_PARTS && part_sweeps(st)) {
        c.dep = c.parts = true;
        return c;
    }
    c.dep = mode == FC_SWEEP_DEPENDENT || !wide_levels(st);
    return c;
}

// The factor's values in level-packed order (Structure::kpos_f / kpos_b): strict
// lower / upper entries as an ELL over the packed rows, and the pivots per
// packed row.  Filled once per solve by pack_factor*; the sweeps then read
// every operand of a row with coalesced loads and only gather w.
template <class T>
struct PackedFactor {
    const T *af = nullptr, *df = nullptr, *ab = nullptr, *db = nullptr;
};
bool packed_sweeps(const Structure &st);   // both triangles packed
void pack_factor64(hipStream_t s, const Structure &st, const double *a, const double *d, double *af, double *df,
                   double *ab, double *db);
void pack_factor32(hipStream_t s, const Structure &st, const float *a, const float *d, float *af, float *df,
                   float *ab, float *db);
// the same values in the part-wise order (Structure::rows_pf / rows_pb)
void pack_factor_parts64(hipStream_t s, const Structure &st, const double *a, const double *d, double *af,
                         double *df, double *ab, double *db);
void pack_factor_parts32(hipStream_t s, const Structure &st, const float *a, const float *d, float *af, float *df,
                         float *ab, float *db);

// DIC (src/iccg.f90:77-83) / DILU (src/bicgstab.f90:68-79) pivots: one launch per
// level, or with cfg->sf one dependency-driven launch (bit-identical).
void pivots_launch(hipStream_t s, const Structure &st, const double *a, double *d, bool dilu, int *err,
                   const SweepCfg *cfg = nullptr);
// z = M^-1 rhs with the sweeps of src/iccg.f90:94-111: level-packed kernels when
// `pf` is given (one dependency-driven launch per sweep with cfg->sf), one
// launch per level over the CSR rows otherwise.
void lu_launch64(hipStream_t s, const Structure &st, const double *a, const double *d, double small,
                 const double *rhs, double *w, const Scalars *sc, const SweepCfg *cfg = nullptr,
                 const PackedFactor<double> *pf = nullptr);
void lu_launch32(hipStream_t s, const Structure &st, const float *a, const float *d, float small, const float *rhs,
                 float *w, const Scalars *sc, const SweepCfg *cfg = nullptr, const PackedFactor<float> *pf = nullptr);

void dot_launch64(hipStream_t s, int n, const double *x, const double *y, double *pa);
void norms_launch64(hipStream_t s, int n, const double *x, double *pa, double *pb);
void sum_parts_launch(hipStream_t s, int nparts, const double *pa, const double *pb, double *out);

}  // namespace fcap
