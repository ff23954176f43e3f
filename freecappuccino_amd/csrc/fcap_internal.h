// Internal declarations of libfcap: context, device buffers, error plumbing,
// kernel launchers shared between translation units.
#pragma once

#include <algorithm>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fcap.h"

namespace fcap {

// Constants written as default-REAL literals in the reference and widened to
// real(dp): small = 1e-20 (src/modules_allocatable.f90:27), tol = 1e-13
// (src/iccg.f90:37), s0 = 1.e20 (src/iccg.f90:85), tiny = 1e-30
// (src/mesh_geometry_and_topology.f90:58).
constexpr double kSmallSrc = (double)1e-20f;
constexpr double kSmallTests = (double)1e-30f;
constexpr double kTol = (double)1e-13f;
constexpr double kS0 = (double)1e20f;
constexpr double kTiny = (double)1e-30f;

struct Error : std::runtime_error {
    fc_status code;
    Error(fc_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define FC_HIP_CHECK(call)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::fcap::Error(FC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define FC_REQUIRE(cond, code, msg)                                                        \
    do {                                                                                   \
        if (!(cond)) throw ::fcap::Error(code, msg);                                       \
    } while (0)

// Owning device allocation.
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p_ = o.p_; bytes_ = o.bytes_; o.p_ = nullptr; o.bytes_ = 0; }
        return *this;
    }
    void alloc(size_t bytes) {
        release();
        if (bytes == 0) bytes = 16;
        hipError_t e = hipMalloc(&p_, bytes);
        if (e != hipSuccess) {
            p_ = nullptr;
            throw Error(FC_ERR_ALLOC, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
        }
        bytes_ = bytes;
    }
    // Grow-only reallocation (contents not preserved).
    void ensure(size_t bytes) { if (bytes > bytes_ || !p_) alloc(bytes); }
    void release() { if (p_) { (void)hipFree(p_); p_ = nullptr; bytes_ = 0; } }
    template <class T> T *as() const { return static_cast<T *>(p_); }
    void *get() const { return p_; }
    size_t bytes() const { return bytes_; }

  private:
    void *p_ = nullptr;
    size_t bytes_ = 0;
};

// Upload helpers (synchronous w.r.t. the stream: callers own the host arrays).
void h2d(void *dst, const void *src, size_t bytes, hipStream_t s);
void d2h(void *dst, const void *src, size_t bytes, hipStream_t s);

// Device CSR structure with the analysis needed by the solvers.
struct Structure {
    int n = 0, nnz = 0;
    uint64_t key = 0;           // checksum of (n, nnz, ioffset, ja)
    DevBuf ioffset, ja, diag;   // 0-based on the device
    // level schedules of the forward (strict lower) and backward (strict upper) DAGs
    int nlev_f = 0, nlev_b = 0;
    DevBuf rows_f, rows_b;      // rows in level order
    std::vector<int> lvl_f, lvl_b;   // host copy of the level pointers (nlev+1)
    DevBuf lvl_f_d, lvl_b_d;         // device copy of the level pointers
    int max_level_f = 0, max_level_b = 0; // largest level size
    DevBuf tpos;                // DILU: position of (j,i) in row j for each strict-lower (i,j), -1 if none
    bool has_tpos = false;
    bool levels_ready = false;
    // Level-packed sweeps: per direction, an ELL of the strict-lower (forward) or
    // strict-upper (backward) entries of the rows in level order; slot k of
    // packed row p sits at [k * n + p], entries in CSR order, padding 0.
    int ell_f = 0, ell_b = 0;        // ELL widths (0: not packed)
    DevBuf cnt_f, cnt_b;             // entries per packed row
    DevBuf kpos_f, kpos_b;           // CSR position of each slot
    DevBuf col_f, col_b;             // column of each slot
    // Part-wise sweeps: the rows split into parts of kPartRows consecutive indices
    // (in which the forward sweep reads only lower and the backward sweep only
    // higher indices); each part's rows in the order of their levels, packed like
    // the level order above (same widths).  int_frac: the fraction of the
    // triangle's entries whose column lies in the row's own part.  meta: per
    // packed row, its entry count | its step within its 64-row window << 8 | the
    // window's number of steps << 16 (the static schedule of the part sweeps).
    bool parts = false;
    DevBuf rows_pf, cnt_pf, kpos_pf, col_pf, meta_pf;
    DevBuf rows_pb, cnt_pb, kpos_pb, col_pb, meta_pb;
    double int_frac_f = 0.0, int_frac_b = 0.0;
};

struct Workspace;   // solver vectors, partial sums, device scalars

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    void create() { if (!a) { FC_HIP_CHECK(hipEventCreate(&a)); FC_HIP_CHECK(hipEventCreate(&b)); } }
    void destroy() { if (a) { (void)hipEventDestroy(a); (void)hipEventDestroy(b); a = b = nullptr; } }
    void start(hipStream_t s) { create(); FC_HIP_CHECK(hipEventRecord(a, s)); }
    double stop_ms(hipStream_t s) {
        FC_HIP_CHECK(hipEventRecord(b, s));
        FC_HIP_CHECK(hipEventSynchronize(b));
        float ms = 0.f;
        FC_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
};

}  // namespace fcap

struct fc_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string last_error;
    int variant = FC_VARIANT_SRC;
    int report = 0;
    std::string report_name = "p";
    // structure cache of the host-pointer solvers
    std::unique_ptr<fcap::Structure> cache;
    std::unique_ptr<fcap::Workspace> ws;
    fcap::Timer timer;
    int num_cus = 256;
    // replay solver iterations as HIP graphs (fc_set_graphs; off: launch kernels directly)
    int use_graphs = 1;
    // sums of the FP64 solvers in index order, as the reference forms them
    // (fc_set_reduction_order FC_REDUCE_SEQUENTIAL), instead of a fixed tree
    int reduce_seq = 0;
    // triangular sweeps: FC_SWEEP_LEVELS, FC_SWEEP_DEPENDENT or FC_SWEEP_PARTS (fc_set_sweep_mode)
    int sweep_mode = 3;
    int part_waves = 0;   // persistent waves of the dependency-driven sweeps (0: automatic)
    // wall-clock limit of one wait of the dependency-driven sweeps, microseconds
    // (fc_set_sweep_wait_limit; 0: the built-in 2 s)
    int sf_wait_us = 0;
    int wall_clock_khz = 100000;   // rate of the device's wall clock (wall_clock64())
    // flexible CG stagnation test (fc_set_fcg_stagnation): after the first window of
    // fcg_stag_window iterations, every window's best ||r||^2 must fall below
    // fcg_stag_factor x the best before it (window 0: no test)
    int fcg_stag_window = 64;
    double fcg_stag_factor = 0.81;
    // flexible CG directions kept, rounded up to 4, 8, 12 or 16 (fc_set_fcg_window;
    // 0: 4 with the coarse correction, 12 without)
    int fcg_m = 0;
    // inner tolerance of the FP32 solves in FP64 refinement (fc_set_inner_tolerance;
    // 0: the method's default)
    double inner_eta = 0.0;
    int gnn_smooth = 3;       // damped-Jacobi steps after the network in flexible CG (fc_set_fcg_smoothing)
    float gnn_omega = 0.8f;   // their damping
    int gnn_layout = 0;       // FC_NET_CLUSTERS or FC_NET_CELL_ORDER (fc_set_network_layout)
    // cells per aggregate of flexible CG's coarse correction (fc_set_fcg_coarse; 0: off)
    int fcg_coarse = 4096;
    // flexible CG iterations per network application (fc_set_fcg_network_period;
    // 0: 2 with the coarse correction, 1 without)
    int fcg_net_period = 0;
    void *dense_handle = nullptr;   // rocBLAS handle of the coarse factorisation (created on first use)
    // pinned host flag block read once per iteration batch
    int *host_flags = nullptr;
};

namespace fcap {

// persistent waves of the part-wise sweeps: the context's setting, else three per
// two CUs (192^3 box, FP64 IC(0)-PCG with 4,096-row parts: 256 waves 1.03 s, 384
// 0.93 s, 512 0.95 s, 768 1.03 s; with 2,048-row parts 512 waves were best at
// 1.15 s, with 1,024-row parts 1.71 s, with 8,192-row parts 1.61 s: bigger parts
// cross fewer part boundaries until too few of them run at once)
inline int part_wave_count(const fc_context *ctx) {
    return ctx->part_waves > 0 ? ctx->part_waves : 3 * ctx->num_cus / 2;
}

// persistent waves of the row-wise dependency-driven sweeps and pivots: one 64-row
// block per wave for the rows of an average level, at least one wave per CU and at
// most sixteen.  A deep chain (192^3 box in natural order: 574 levels of 12K rows)
// then runs with one wave per CU -- more waves only add polling -- and a shallow one
// (96^3 permuted Kuhn tets: 16 levels of 330K rows) with enough waves to keep its
// gathers in flight.
inline int syncfree_wave_count(const fc_context *ctx, const Structure &st) {
    if (ctx->part_waves > 0) return ctx->part_waves;
    const long long nlev = std::max(1, std::max(st.nlev_f, st.nlev_b));
    const long long blocks = ((long long)st.n / nlev + 63) / 64;
    return (int)std::min<long long>(16LL * ctx->num_cus, std::max<long long>(ctx->num_cus, blocks));
}

// ---------------------------------------------------------------- launchers
// Structure analysis (fcap_mesh.hip)
void build_structure_from_host_csr(fc_context *ctx, Structure &st, int n, int nnz,
                                   const int *ioffset1, const int *ja1, const int *diag1);
void build_structure_from_faces_dev(fc_context *ctx, Structure &st, int n, int nif,
                                    const int *owner0_d, const int *neigh0_d, int *ij_slot_d,
                                    int *ji_slot_d);
void analyse_levels(fc_context *ctx, Structure &st);
inline void ensure_levels(fc_context *ctx, Structure &st) {
    if (!st.levels_ready) { analyse_levels(ctx, st); st.levels_ready = true; }
}
void to_one_based(fc_context *ctx, int *d, long long n);
void sub_one_dev(fc_context *ctx, const int *in, int *out, long long n);
void fill_int(fc_context *ctx, int *d, long long n, int v);
void build_tpos(fc_context *ctx, Structure &st);
uint64_t csr_checksum(int n, int nnz, const int *ioffset, const int *ja);

// Geometry (fcap_mesh.hip); all device pointers, 0-based node/cell ids.
struct GeometryDev {
    DevBuf arx, ary, arz, xf, yf, zf, vol, xc, yc, zc, facint;
};
struct MeshDev {
    int nnodes = 0, nfaces = 0, nif = 0, n = 0;
    DevBuf xyz, face_ptr, face_nodes, owner, neighbour;   // 0-based ids
    DevBuf cf_ptr, cf_face;   // cell -> faces in ascending face order
    GeometryDev g;
};
void upload_mesh(fc_context *ctx, MeshDev &m, int nnodes, const double *xyz, int nfaces, int nif,
                 const int *face_ptr, const int *face_nodes1, const int *owner1,
                 const int *neighbour1, int ncells);
void compute_geometry(fc_context *ctx, MeshDev &m);

// Assembly (fcap_assembly.hip)
struct WallDev {
    int nwal = 0, iwall_faces_start = 0, iwall_start = 0;
    DevBuf cw_ptr, cw_idx;   // cell -> wall indices (0..nwal-1) in wall order
};
void build_wall_lists(fc_context *ctx, WallDev &w, const int *owner0_d, int n, int nwal,
                      int iwall_faces_start, int iwall_start);
struct OcDev {
    int noc = 0, ioc_faces_start = 0;
    DevBuf ijl, ijr, foc, al, ar;    // 0-based cells
    DevBuf co_ptr, co_idx;           // cell -> O-C indices, encoded (i<<1 | side)
};
void build_oc_lists(fc_context *ctx, OcDev &o, int n);
// Fields of the pressure-correction assembly (device, 0-based cells; gradients
// 3 per cell, component fastest, as dUdxi(3, numCells))
struct RhieChowDev {
    const double *xc, *yc, *zc, *vol, *den, *apu, *apv, *apw, *u, *v, *w, *p;
    const double *dudx, *dvdx, *dwdx, *dpdx;
};
// Scratch of grad_gauss_dev and calcp_nonorth_dev (previous iterate, per-face and per-cut
// values); a caller that keeps one across calls saves the allocations
struct GradScratch {
    DevBuf old, gf, gc;
};
// calcp's non-orthogonal corrector source (su overwritten, flmass and fmoc += fmcor)
void calcp_nonorth_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0, const int *neigh0,
                       const int *kface, const int *kface_order, const double *xf, const double *yf,
                       const double *zf, const double *arx, const double *ary, const double *arz, const double *facint,
                       const double *xc, const double *yc, const double *zc, const double *vol, const double *den,
                       const double *apu, const double *dpdx, const OcDev &oc, double *su, double *flmass,
                       double *fmoc, GradScratch *scratch = nullptr);
// adjustMassFlow: inlet sources, outlet fluxes scaled to flomas (0-based faces and values)
void adjust_mass_flow_dev(fc_context *ctx, int n, const int *owner0_all, const double *arx, const double *ary,
                          const double *arz, int ninl, int inlet_face0, const double *fmi, int nout,
                          int outlet_face0, int outlet_val0, double flomas, const double *den, double *u, double *v,
                          double *w, double *su, double *fmo);
// calcp's flux, velocity and pressure corrections (pref0: 0-based reference cell)
void calcp_correct_dev(fc_context *ctx, int n, int nif, const int *owner0, const int *neigh0, const int *ij_slot,
                       const double *a, const double *pp, const double *dpdx, const double *vol, const double *apu,
                       const double *apv, const double *apw, double urf, int pref0, int noc, const int *ijl,
                       const int *ijr, const double *ar, double *fmoc, double *flmass, double *u, double *v,
                       double *w, double *p);
// Boundary groups of the Gauss gradient in the reference's order (inlet, outlet,
// symmetry, wall, pressure outlet): item q of the sequence over all groups is
// face fstart[g] + (q - qstart[g]) with its value at vstart[g] + (q - qstart[g])
// (0-based); qstart[5] is the number of items.
struct GradGroups {
    int qstart[6];
    int fstart[5], vstart[5];
};
// per-cell lists of boundary items, each in sequence order
void build_boundary_lists(fc_context *ctx, const GradGroups &G, const int *owner0_all, int n, DevBuf &cb_ptr,
                          DevBuf &cb_q);
// (x, y, z component arrays of n) -> out(3, n), component fastest
void interleave3_dev(fc_context *ctx, int n, const double *gx, const double *gy, const double *gz, double *out);
// grad_gauss: nigrad iterations of the Gauss gradient of u (cells, then boundary values)
void grad_gauss_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0, const int *neigh0,
                    const int *kface, const int *kface_order, const double *xf, const double *yf,
                    const double *zf, const double *arx, const double *ary, const double *arz, const double *facint,
                    const double *xc, const double *yc, const double *zc, const double *vol, const double *u,
                    const GradGroups &G, const int *cb_ptr, const int *cb_q, const OcDev &oc, const int *ocface,
                    int nigrad, double *dudx, double *dudy, double *dudz,
                    GradScratch *scratch = nullptr);
// one velocity component's diagonal, source and 1/a_P (calcuvw), a's diagonal slots updated
void momentum_rows_dev(fc_context *ctx, const Structure &st, double *a, const double *sp, const double *src,
                       const double *phi, double urf, double *su, double *ap);
// calcp's face loops (a and su overwritten, flmass of the inner faces, O-C al/ar/fmoc)
void calcp_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0, const int *neigh0,
               const int *ij_slot, const int *ji_slot, const int *kface, const int *kface_order, const double *xf,
               const double *yf, const double *zf, const double *arx, const double *ary, const double *arz,
               const double *facint, const RhieChowDev &F, const OcDev &oc, double *a, double *su, double *flmass,
               double *fmoc, bool piso = false);
// correctBoundaryConditionsVelocity: outlet values and fluxes scaled to flomas (fac_scratch:
// one device double), symmetry values projected (0-based faces and values)
void velocity_bc_dev(fc_context *ctx, const int *owner0_all, const double *arx, const double *ary, const double *arz,
                     int nout, int outlet_face0, int outlet_val0, int nsym, int sym_face0, int sym_val0, double flomas,
                     const double *den, double *u, double *v, double *w, double *fmo, double *fac_scratch);
// bpres at one stage on p (numTotal); dpdx (3, numCells) is read from stage 2 on
void bpres_dev(fc_context *ctx, const GradGroups &G, int noc, int oc_face0, int oc_val0, const int *owner0_all,
               const double *xc, const double *yc, const double *zc, const double *xf, const double *yf,
               const double *zf, const double *dpdx, int stage, double *p);
// PISO's reference cell: row pref0 of a zeroed, diagonal 1, su(pref0) = p(pref0)
void pin_reference_dev(fc_context *ctx, const Structure &st, int pref0, const double *p, double *a, double *su);
// range/consistency check of caller-supplied owner/neighbour and slot maps (device, 0-based)
void validate_faces(fc_context *ctx, const Structure &st, int nif, const int *own, const int *nei, const int *ij,
                    const int *ji);
void laplacian_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0,
                   const int *neigh0, const int *ij_slot, const int *ji_slot,
                   const double *arx, const double *ary, const double *arz,
                   const double *facint, const double *xc, const double *yc, const double *zc,
                   const double *xf, const double *yf, const double *zf, const double *mu,
                   const double *phi, const WallDev &w, const OcDev &oc, const int *kface_order,
                   double *a, double *su);
void convdiff_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0,
                  const int *neigh0, const int *ij_slot, const int *ji_slot,
                  const double *arx, const double *ary, const double *arz, const double *facint,
                  const double *xc, const double *yc, const double *zc, const double *xf,
                  const double *yf, const double *zf, const double *vis, const double *flux,
                  const WallDev &w, double urf, double *a);
void face_flux_dev(fc_context *ctx, int nif, const int *owner0, const int *neigh0,
                   const double *facint, const double *arx, const double *ary, const double *arz,
                   const double *u, const double *v, const double *w, double *flux);
void rhs_divergence_dev(fc_context *ctx, const Structure &st, const int *kface, const int *kface_order,
                        const double *flux, double *su);
// face id of each off-diagonal CSR entry (-1 on the diagonal), and whether every
// row lists its faces in ascending face order when read in column order
void build_kface(fc_context *ctx, const Structure &st, int nif, const int *ij_slot,
                 const int *ji_slot, DevBuf &kface, DevBuf &kface_order, bool &needs_order);

// Linear algebra (fcap_linalg.hip)
void spmv64(fc_context *ctx, const Structure &st, const double *a, const double *x, double *y);
void spmv32(fc_context *ctx, const Structure &st, const float *a, const float *x, float *y);
void pivots(fc_context *ctx, const Structure &st, const double *a, double *d, bool dilu);
void lu_apply64(fc_context *ctx, const Structure &st, const double *a, const double *d,
                double small, const double *r, double *z);

}  // namespace fcap
