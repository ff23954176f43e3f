// SpMV, reductions, level-scheduled incomplete factorisations and sweeps.
//
// SpMV: one wavefront per block of 64 consecutive rows.  The wave stages the
// block's slice of a/ja (contiguous in CSR) into LDS with coalesced loads,
// then each lane walks its own row from LDS in CSR order, gathering x from
// global memory.  The per-row order of operations is the reference's
// (src/iccg.f90:129-134), so the FP64 result is bit-identical.
//
// Triangular sweeps and pivot recurrences are level-scheduled: the rows of a
// level of the lower (forward) or upper (backward) dependency DAG have no
// mutual dependencies; each level is one launch over its rows, each row
// computed in the reference's order (src/iccg.f90:77-111, src/bicgstab.f90:68-79).

#include <algorithm>

#include "fcap_internal.h"
#include "fcap_kernels.h"
#include "fcap_linalg.h"

namespace fcap {

namespace {

inline int grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    return (int)std::max<long long>(1, std::min<long long>(g, 1 << 30));
}

}  // namespace

int red_grid(int n) { return std::min(kMaxRedGrid, grid_for(n, kRedBlock)); }
int spmv_grid(int n) { return grid_for(n, kSpmvRows); }
// launch grid of a kernel with an XCD-aware block order: `blocks` rounded up to a multiple of kXcds
static int xcd_grid(int blocks) { return (blocks + kXcds - 1) / kXcds * kXcds; }

// SpMV block order: workgroup b takes row block b.  (An XCD-aware order, each
// XCD sweeping one contiguous eighth of the rows, measured slower at 192^3:
// 4.86 against 5.10 TB/s.)
constexpr int g_spmv_xcd = 0;

// ------------------------------------------------------------------- SpMV

template <class T, class AT, int MODE>
__global__ __launch_bounds__(kSpmvRows) void k_spmv(int n, const int *__restrict__ ptr, const int *__restrict__ ja,
                                                    const AT *__restrict__ a, const T *__restrict__ x,
                                                    const T *__restrict__ b, const T *__restrict__ v,
                                                    T *__restrict__ y, float *__restrict__ y32,
                                                    double *__restrict__ partA, double *__restrict__ partB,
                                                    const Scalars *__restrict__ sc, int xcd, float omega) {
    // MODE 0: y = A x; 1: y = b - A x with |y| and y^2 partials; 2: y = A x with y.v
    // partials; 3: y = A x with y.v and y^2 partials; 4: y = x + omega (b - A x) / v;
    // 5: y = b - A x
    constexpr bool kRes = MODE == 1 || MODE == 4 || MODE == 5;
    __shared__ AT sa[kSpmvWaves][kSpmvCap];
    __shared__ int sj[kSpmvWaves][kSpmvCap];
    __shared__ double red[kSpmvWaves];
    if (sc && sc->done) return;
    // xcd = 1, XCD-aware order: workgroups go to the 8 XCDs round-robin, so
    // workgroup b takes row block (b % 8) * per + b / 8 and each XCD sweeps one
    // contiguous eighth of the rows (the grid is then a multiple of 8).
    const int nblk = (int)(((long long)n + kSpmvRows - 1) / kSpmvRows);
    const int per = gridDim.x / kXcds;
    const int blk = xcd ? (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds) : (int)blockIdx.x;
    if (blk >= nblk) return;
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const long long row0 = (long long)blk * kSpmvRows + (long long)w * kWave;
    const long long rowl = row0 + lane;
    const bool valid = rowl < n;
    const int row = (int)rowl;
    int nrows = (int)std::min<long long>(std::max<long long>(n - row0, 0), kWave);
    int s0 = 0, cnt = 0;
    if (nrows > 0) {
        s0 = ptr[row0];
        cnt = ptr[row0 + nrows] - s0;
    }
    const bool staged = nrows > 0 && cnt <= kSpmvCap;
    if (staged) {
        // all loads of the slice are issued before the first LDS store, so a
        // wave keeps up to 2*kSpmvCap/64 loads in flight; a and ja are read
        // once, so they are streamed (non-temporal) and leave the caches to x
        constexpr int kChunks = kSpmvCap / kWave;
        AT va[kChunks];
        int vj[kChunks];
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int j = c * kWave + lane;
            if (j < cnt) {
                va[c] = __builtin_nontemporal_load(a + s0 + j);
                vj[c] = __builtin_nontemporal_load(ja + s0 + j);
            }
        }
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int j = c * kWave + lane;
            if (j < cnt) {
                sa[w][j] = va[c];
                sj[w][j] = vj[c];
            }
        }
    }
    __syncthreads();
    double pa = 0.0, pb = 0.0;
    if (valid) {
        const int rs = ptr[row], re = ptr[row + 1];
        T acc = kRes ? b[row] : T(0);
        if (staged) {
            // gathers issued four at a time; the sum still runs in CSR order
            int k = rs - s0;
            const int ke = re - s0;
            for (; k + 4 <= ke; k += 4) {
                const T x0 = x[sj[w][k]], x1 = x[sj[w][k + 1]], x2 = x[sj[w][k + 2]], x3 = x[sj[w][k + 3]];
                const T p0 = (T)sa[w][k] * x0, p1 = (T)sa[w][k + 1] * x1;
                const T p2 = (T)sa[w][k + 2] * x2, p3 = (T)sa[w][k + 3] * x3;
                if (kRes) { acc = acc - p0; acc = acc - p1; acc = acc - p2; acc = acc - p3; }
                else { acc = acc + p0; acc = acc + p1; acc = acc + p2; acc = acc + p3; }
            }
            for (; k < ke; ++k) {
                T prod = (T)sa[w][k] * x[sj[w][k]];
                acc = kRes ? acc - prod : acc + prod;
            }
        } else {
            for (int k = rs; k < re; ++k) {
                T prod = (T)a[k] * x[ja[k]];
                acc = kRes ? acc - prod : acc + prod;
            }
        }
        if (MODE == 4)
            y[row] = x[row] + (T)omega * (acc / v[row]);
        else
            y[row] = acc;
        if (MODE == 1) {
            if (y32) y32[row] = (float)acc;
            pa = fabs((double)acc);
            pb = (double)acc * (double)acc;
        } else if (MODE == 2) {
            pa = (double)acc * (double)v[row];
        } else if (MODE == 3) {
            pa = (double)acc * (double)v[row];
            pb = (double)acc * (double)acc;
        }
    }
    if (MODE != 0 && MODE != 4 && MODE != 5) {
        pa = wave_sum(pa);
        if (lane == 0) red[w] = pa;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int i = 0; i < kSpmvWaves; ++i) s += red[i];
            partA[blk] = s;
        }
        if (MODE == 1 || MODE == 3) {
            __syncthreads();
            pb = wave_sum(pb);
            if (lane == 0) red[w] = pb;
            __syncthreads();
            if (threadIdx.x == 0) {
                double s = 0.0;
                for (int i = 0; i < kSpmvWaves; ++i) s += red[i];
                partB[blk] = s;
            }
        }
    }
}

// O-C couplings in cut order, per cell (src/iccg.f90:59-62,137-140):
// sign -1 for the residual form, +1 for the product form.
template <class T>
__global__ void k_oc_apply(int n, const int *__restrict__ co_ptr, const int *__restrict__ co_idx,
                           const int *__restrict__ ijl, const int *__restrict__ ijr, const double *__restrict__ al,
                           const double *__restrict__ ar, const T *__restrict__ x, T *__restrict__ y, int sign,
                           const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = co_ptr[i], e = co_ptr[i + 1];
    if (s == e) return;
    T acc = y[i];
    for (int q = s; q < e; ++q) {
        int code = co_idx[q], o = code >> 1;
        T prod = (code & 1) ? (T)al[o] * x[ijl[o]] : (T)ar[o] * x[ijr[o]];
        acc = sign < 0 ? acc - prod : acc + prod;
    }
    y[i] = acc;
}

void spmv_launch64(hipStream_t s, int mode, int n, const int *ptr, const int *ja, const double *a, const double *x,
                   const double *b, const double *v, double *y, float *y32, double *pa, double *pb,
                   const Scalars *sc) {
    const int xo = g_spmv_xcd;
    const int g = xo ? xcd_grid(spmv_grid(n)) : spmv_grid(n);
    switch (mode) {
    case 0: k_spmv<double, double, 0><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, v, y, y32, pa, pb, sc, xo, 0.f); break;
    case 1: k_spmv<double, double, 1><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, v, y, y32, pa, pb, sc, xo, 0.f); break;
    case 2: k_spmv<double, double, 2><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, v, y, y32, pa, pb, sc, xo, 0.f); break;
    default: k_spmv<double, double, 3><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, v, y, y32, pa, pb, sc, xo, 0.f); break;
    }
}

void spmv_launch32(hipStream_t s, int mode, int n, const int *ptr, const int *ja, const float *a, const float *x,
                   const float *b, const float *v, float *y, double *pa, double *pb, const Scalars *sc) {
    const int xo = g_spmv_xcd;
    const int g = xo ? xcd_grid(spmv_grid(n)) : spmv_grid(n);
    switch (mode) {
    case 0: k_spmv<float, float, 0><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, v, y, nullptr, pa, pb, sc, xo, 0.f); break;
    case 1: k_spmv<float, float, 1><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, v, y, nullptr, pa, pb, sc, xo, 0.f); break;
    case 2: k_spmv<float, float, 2><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, v, y, nullptr, pa, pb, sc, xo, 0.f); break;
    default: k_spmv<float, float, 3><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, v, y, nullptr, pa, pb, sc, xo, 0.f); break;
    }
}

void jacobi_step32(hipStream_t s, int n, const int *ptr, const int *ja, const float *a, const float *x, const float *b,
                   const float *d, float omega, float *y, const Scalars *sc) {
    const int xo = g_spmv_xcd;
    const int g = xo ? xcd_grid(spmv_grid(n)) : spmv_grid(n);
    k_spmv<float, float, 4><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, d, y, nullptr, nullptr, nullptr, sc, xo,
                                                    omega);
}

void residual32(hipStream_t s, int n, const int *ptr, const int *ja, const float *a, const float *x, const float *b,
                float *y, const Scalars *sc) {
    const int xo = g_spmv_xcd;
    const int g = xo ? xcd_grid(spmv_grid(n)) : spmv_grid(n);
    k_spmv<float, float, 5><<<g, kSpmvRows, 0, s>>>(n, ptr, ja, a, x, b, nullptr, y, nullptr, nullptr, nullptr, sc,
                                                    xo, 0.f);
}

void oc_launch64(hipStream_t s, int n, const OcView &oc, const double *x, double *y, int sign, const Scalars *sc) {
    if (!oc.noc) return;
    k_oc_apply<double><<<grid_for(n, 256), 256, 0, s>>>(n, oc.co_ptr, oc.co_idx, oc.ijl, oc.ijr, oc.al, oc.ar, x, y,
                                                        sign, sc);
}

// ------------------------------------------------------- pivots and sweeps

__global__ void k_dic_level(const int *__restrict__ rows, int beg, int cnt, const int *__restrict__ ptr,
                            const int *__restrict__ ja, const int *__restrict__ diag, const double *__restrict__ a,
                            double *__restrict__ d, int *__restrict__ err) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    int i = rows[beg + t];
    double di = a[diag[i]];
    for (int k = ptr[i]; k < diag[i]; ++k) {
        double ak = a[k];
        di = di - (ak * ak) * d[ja[k]];
    }
    if (!(di != 0.0) || !isfinite(di)) atomicOr(err, 1);
    d[i] = 1.0 / di;
}

__global__ void k_dilu_level(const int *__restrict__ rows, int beg, int cnt, const int *__restrict__ ptr,
                             const int *__restrict__ ja, const int *__restrict__ diag, const int *__restrict__ tpos,
                             const double *__restrict__ a, double *__restrict__ d, int *__restrict__ err) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    int i = rows[beg + t];
    double di = a[diag[i]];
    for (int k = ptr[i]; k < diag[i]; ++k) di = di - (a[k] * d[ja[k]]) * a[tpos[k]];
    if (!(di != 0.0) || !isfinite(di)) atomicOr(err, 1);
    d[i] = 1.0 / di;
}

// forward sweep of (D+L) w = rhs: w_i = (rhs_i - sum_{k<diag} a_k w_j) d_i
template <class T>
__global__ void k_fwd_level(const int *__restrict__ rows, int beg, int cnt, const int *__restrict__ ptr,
                            const int *__restrict__ ja, const int *__restrict__ diag, const T *__restrict__ a,
                            const T *__restrict__ d, const T *__restrict__ rhs, T *__restrict__ w,
                            const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    int i = rows[beg + t];
    T acc = rhs[i];
    for (int k = ptr[i]; k < diag[i]; ++k) acc = acc - a[k] * w[ja[k]];
    w[i] = acc * d[i];
}

// backward sweep in place: z_i = (w_i/(d_i+small) - sum_{k>diag} a_k z_j) d_i
template <class T>
__global__ void k_bwd_level(const int *__restrict__ rows, int beg, int cnt, const int *__restrict__ ptr,
                            const int *__restrict__ ja, const int *__restrict__ diag, const T *__restrict__ a,
                            const T *__restrict__ d, T small, T *__restrict__ w, const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    int i = rows[beg + t];
    T di = d[i];
    T acc = w[i] / (di + small);
    for (int k = diag[i] + 1; k < ptr[i + 1]; ++k) acc = acc - a[k] * w[ja[k]];
    w[i] = acc * di;
}

template <int W>
void pivots_syncfree_w(hipStream_t s, const Structure &st, const double *a, double *d, bool dilu, int *err,
                       const SweepCfg &cfg);

void pivots_launch(hipStream_t s, const Structure &st, const double *a, double *d, bool dilu, int *err,
                   const SweepCfg *cfg) {
    if (cfg && cfg->sf && syncfree_sweeps(st)) {
        FC_REQUIRE(!dilu || st.tpos.get(), FC_ERR_ARGUMENT, "DILU pivots need the transpose positions");
        switch (st.ell_f) {
        case 3: pivots_syncfree_w<3>(s, st, a, d, dilu, err, *cfg); break;
        case 4: pivots_syncfree_w<4>(s, st, a, d, dilu, err, *cfg); break;
        default: pivots_syncfree_w<8>(s, st, a, d, dilu, err, *cfg); break;
        }
        FC_HIP_CHECK(hipGetLastError());
        return;
    }
    const int *rows = st.rows_f.as<int>();
    for (int l = 0; l < st.nlev_f; ++l) {
        int beg = st.lvl_f[l], cnt = st.lvl_f[l + 1] - beg;
        if (dilu)
            k_dilu_level<<<grid_for(cnt, 256), 256, 0, s>>>(rows, beg, cnt, st.ioffset.as<int>(), st.ja.as<int>(),
                                                             st.diag.as<int>(), st.tpos.as<int>(), a, d, err);
        else
            k_dic_level<<<grid_for(cnt, 256), 256, 0, s>>>(rows, beg, cnt, st.ioffset.as<int>(), st.ja.as<int>(),
                                                            st.diag.as<int>(), a, d, err);
    }
    FC_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------- level-packed sweeps

// packed row p of a level: operands at [k * n + p] (coalesced across the
// level's threads); all of a row's slots are loaded, then all gathers of w
// issued, then the sum runs in CSR order over the row's own entries only, so
// the arithmetic is that of k_fwd_level / k_bwd_level.  Padding slots hold
// column 0 and value 0 and are never added.
template <class T, int W>
__global__ void k_fwd_packed(int beg, int cnt, int n, const int *__restrict__ rows, const int *__restrict__ cntp,
                             const int *__restrict__ col, const T *__restrict__ av, const T *__restrict__ dv,
                             const T *__restrict__ rhs, T *__restrict__ w, int ell, const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int p = beg + t;
    const int i = rows[p];
    const int c = cntp[p];
    T acc = rhs[i];
    if (W > 0) {
        int cj[W > 0 ? W : 1];
        T ak[W > 0 ? W : 1], wk[W > 0 ? W : 1];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            cj[k] = col[(size_t)k * n + p];
            ak[k] = av[(size_t)k * n + p];
        }
#pragma unroll
        for (int k = 0; k < W; ++k) wk[k] = w[cj[k]];
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k < c) acc = acc - ak[k] * wk[k];
    } else {
        for (int k = 0; k < c; ++k) acc = acc - av[(size_t)k * n + p] * w[col[(size_t)k * n + p]];
    }
    (void)ell;
    w[i] = acc * dv[p];
}

template <class T, int W>
__global__ void k_bwd_packed(int beg, int cnt, int n, const int *__restrict__ rows, const int *__restrict__ cntp,
                             const int *__restrict__ col, const T *__restrict__ av, const T *__restrict__ dv, T small,
                             T *__restrict__ w, const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int p = beg + t;
    const int i = rows[p];
    const int c = cntp[p];
    const T di = dv[p];
    if (W > 0) {
        int cj[W > 0 ? W : 1];
        T ak[W > 0 ? W : 1], wk[W > 0 ? W : 1];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            cj[k] = col[(size_t)k * n + p];
            ak[k] = av[(size_t)k * n + p];
        }
#pragma unroll
        for (int k = 0; k < W; ++k) wk[k] = w[cj[k]];
        T acc = w[i] / (di + small);
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k < c) acc = acc - ak[k] * wk[k];
        w[i] = acc * di;
    } else {
        T acc = w[i] / (di + small);
        for (int k = 0; k < c; ++k) acc = acc - av[(size_t)k * n + p] * w[col[(size_t)k * n + p]];
        w[i] = acc * di;
    }
}

// values of the factor into packed order (one pass per solve)
template <class T>
__global__ void k_pack_factor(int n, int ell, const int *__restrict__ rows, const int *__restrict__ cntp,
                              const int *__restrict__ kpos, const T *__restrict__ a, const T *__restrict__ d,
                              T *__restrict__ av, T *__restrict__ dv) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int c = cntp[p];
    for (int k = 0; k < ell; ++k) {
        const size_t o = (size_t)k * n + p;
        av[o] = k < c ? a[kpos[o]] : T(0);
    }
    dv[p] = d[rows[p]];
}

// Each triangle's ELL is stored at exactly the width its kernel reads (3, 4 or
// 8 slots per row, or its widest row for the generic W = 0 kernel).
template <class T, int W>
void fwd_packed_w(hipStream_t s, const Structure &st, const PackedFactor<T> &pf, const T *rhs, T *w,
                  const Scalars *sc) {
    const int *rf = st.rows_f.as<int>();
    for (int l = 0; l < st.nlev_f; ++l) {
        const int beg = st.lvl_f[l], cnt = st.lvl_f[l + 1] - beg;
        k_fwd_packed<T, W><<<grid_for(cnt, 256), 256, 0, s>>>(beg, cnt, st.n, rf, st.cnt_f.as<int>(),
                                                              st.col_f.as<int>(), pf.af, pf.df, rhs, w, st.ell_f, sc);
    }
}

template <class T, int W>
void bwd_packed_w(hipStream_t s, const Structure &st, const PackedFactor<T> &pf, T small, T *w, const Scalars *sc) {
    const int *rb = st.rows_b.as<int>();
    for (int l = 0; l < st.nlev_b; ++l) {
        const int beg = st.lvl_b[l], cnt = st.lvl_b[l + 1] - beg;
        k_bwd_packed<T, W><<<grid_for(cnt, 256), 256, 0, s>>>(beg, cnt, st.n, rb, st.cnt_b.as<int>(),
                                                              st.col_b.as<int>(), pf.ab, pf.db, small, w, sc);
    }
}

template <class T>
void sweeps_packed(hipStream_t s, const Structure &st, const PackedFactor<T> &pf, T small, const T *rhs, T *w,
                   const Scalars *sc) {
    switch (st.ell_f) {
    case 3: fwd_packed_w<T, 3>(s, st, pf, rhs, w, sc); break;
    case 4: fwd_packed_w<T, 4>(s, st, pf, rhs, w, sc); break;
    case 8: fwd_packed_w<T, 8>(s, st, pf, rhs, w, sc); break;
    default: fwd_packed_w<T, 0>(s, st, pf, rhs, w, sc); break;
    }
    switch (st.ell_b) {
    case 3: bwd_packed_w<T, 3>(s, st, pf, small, w, sc); break;
    case 4: bwd_packed_w<T, 4>(s, st, pf, small, w, sc); break;
    case 8: bwd_packed_w<T, 8>(s, st, pf, small, w, sc); break;
    default: bwd_packed_w<T, 0>(s, st, pf, small, w, sc); break;
    }
}

template <class T>
void pack_factor_t(hipStream_t s, const Structure &st, const T *a, const T *d, T *af, T *df, T *ab, T *db) {
    FC_REQUIRE(packed_sweeps(st), FC_ERR_ARGUMENT, "structure has no level-packed sweeps");
    const int n = st.n;
    k_pack_factor<T><<<grid_for(n, 256), 256, 0, s>>>(n, st.ell_f, st.rows_f.as<int>(), st.cnt_f.as<int>(),
                                                      st.kpos_f.as<int>(), a, d, af, df);
    k_pack_factor<T><<<grid_for(n, 256), 256, 0, s>>>(n, st.ell_b, st.rows_b.as<int>(), st.cnt_b.as<int>(),
                                                      st.kpos_b.as<int>(), a, d, ab, db);
    FC_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------- dependency-driven sweeps
// One launch per sweep.  Persistent one-wave workgroups take blocks of 64
// consecutive packed rows (level order) from a ticket counter, so every row a
// block reads lies in a block handed out earlier, to a wave that is running or
// has finished: the waits always end.  The output starts filled with a
// signalling-NaN pattern that arithmetic never produces; a row polls the
// outputs of the rows it reads until none holds the pattern, computes exactly
// as k_fwd_packed / k_bwd_packed (same operands, same order), and stores its
// value with one coherent store, which is also its "ready" signal.  A wave
// whose wait exceeds the time limit, or that sees the abort flag, raises the
// flag and leaves; the host then reports the solve as failed.

// Waves in flight poll the coherent level of the memory system; the fewer there
// are, the faster the chain of levels advances (192^3 box: 256 waves 1.87 s,
// 512 1.96 s, 2048 2.55 s, 4096 3.61 s per FP64 IC(0)-PCG solve).  A wait is
// bounded in wall-clock time (wall_clock64(), a constant-rate counter), not in
// polls, so the bound does not shrink with a deeper chain or a faster clock:
// it only ends waits that cannot end otherwise.
constexpr unsigned long long kSfWaitUs = 2000000ull;   // 2 s

// the wait limit in wall-clock ticks (sf->wait_us microseconds, 0: kSfWaitUs)
__device__ inline unsigned long long sf_wait_ticks(const SyncFree *sf, unsigned long long rate_khz) {
    const unsigned long long us = sf->wait_us > 0 ? (unsigned long long)sf->wait_us : kSfWaitUs;
    return us * rate_khz / 1000ull + 1ull;
}

template <class T> struct SfBits;
template <> struct SfBits<double> {
    using U = unsigned long long;
    static constexpr U kMark = 0x7FF4DEADBEEF0001ull;   // signalling NaN
};
template <> struct SfBits<float> {
    using U = unsigned;
    static constexpr U kMark = 0x7FA0BEEFu;               // signalling NaN
};

template <class T>
__device__ inline typename SfBits<T>::U sf_load_bits(const T *p) {
    return __hip_atomic_load(reinterpret_cast<const typename SfBits<T>::U *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ inline void sf_store(T *p, T v) {
    using U = typename SfBits<T>::U;
    U b;
    __builtin_memcpy(&b, &v, sizeof(U));
    __hip_atomic_store(reinterpret_cast<U *>(p), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__global__ void k_sf_mark(int n, T *__restrict__ w, const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using U = typename SfBits<T>::U;
    reinterpret_cast<U *>(w)[i] = SfBits<T>::kMark;
}

// BWD == false: out[i] = (rhs[i] - sum a w) d over the strict lower triangle;
// BWD == true:  out[i] = (fwd[i] / (d + small) - sum a w) d over the strict upper one.
template <class T, int W, bool BWD>
__global__ __launch_bounds__(64) void k_sweep_syncfree(int n, const int *__restrict__ rows,
                                                        const int *__restrict__ cntp, const int *__restrict__ col,
                                                        const T *__restrict__ av, const T *__restrict__ dv,
                                                        const T *__restrict__ in, T small, T *out, SyncFree *sf,
                                                        const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    using U = typename SfBits<T>::U;
    const int lane = threadIdx.x;
    const unsigned nblk = (unsigned)((n + 63) / 64);
    const unsigned long long limit = sf_wait_ticks(sf, sf->rate_khz);
    for (;;) {
        unsigned b = 0;
        if (lane == 0) b = __hip_atomic_fetch_add(&sf->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b = __shfl(b, 0);
        if (b >= nblk) break;
        const int p = (int)b * 64 + lane;
        bool done = p >= n;
        int i = 0, c = 0;
        int cj[W];
        T ak[W], wk[W];
        T acc = T(0), di = T(0);
        unsigned ready = 0;
        if (!done) {
            i = rows[p];
            c = cntp[p];
            di = dv[p];
#pragma unroll
            for (int k = 0; k < W; ++k) {
                cj[k] = col[(size_t)k * n + p];
                ak[k] = av[(size_t)k * n + p];
                wk[k] = T(0);
            }
            acc = BWD ? in[i] / (di + small) : in[i];
        }
        int spins = 0;
        const unsigned long long t0 = wall_clock64();
        for (;;) {
            if (!done) {
                bool all = true;
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    if (k < c && !((ready >> k) & 1u)) {
                        const U bits = sf_load_bits(out + cj[k]);
                        if (bits != SfBits<T>::kMark) {
                            T v;
                            __builtin_memcpy(&v, &bits, sizeof(U));
                            wk[k] = v;
                            ready |= 1u << k;
                        } else {
                            all = false;
                        }
                    }
                }
                if (all) {
                    T r = acc;
#pragma unroll
                    for (int k = 0; k < W; ++k)
                        if (k < c) r = r - ak[k] * wk[k];
                    sf_store(out + i, r * di);
                    done = true;
                }
            }
            if (__all(done)) break;
            ++spins;
            bool give_up = false;
            if ((spins & 63) == 0) {
                give_up = wall_clock64() - t0 > limit ||
                          __hip_atomic_load(&sf->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
            }
            if (__any(give_up)) {
                if (lane == 0) __hip_atomic_store(&sf->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (lane == 0) {
        const unsigned f = __hip_atomic_fetch_add(&sf->fin, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (f == gridDim.x - 1) {
            __hip_atomic_store(&sf->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&sf->fin, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------- part-wise sweeps
// One launch per sweep.  The rows are split into parts of kPartRows consecutive
// indices: the forward sweep reads only rows of lower index and the backward one
// only rows of higher index, so a part depends only on the parts before it (after
// it, backward) in index order, and tickets hand the parts out in that order: every
// part a part reads has been handed to a wave that is running or has finished.
// A persistent wave computes its part 64 packed rows (a window) at a time, the
// rows in the order of their levels.  The part's values live in the wave's LDS.
// A window first waits until every input it reads outside the part is in global
// memory (polled like k_sweep_syncfree's: the signalling-NaN mark means "not yet"),
// then computes its rows in the static schedule of Structure::meta_*: at step s the
// rows of step s, whose inputs inside the part are in LDS by then (earlier steps or
// earlier windows).  Only this wave touches its LDS and a wave's LDS operations
// complete in order, so the steps need no barrier.  A row's value is computed
// exactly as in k_sweep_syncfree and published with one coherent store at the end
// of its window.
//
// The loads run two windows ahead: a window's operands are loaded two windows
// before it, its right-hand side and the first poll of its outside inputs one
// window before; three register sets rotate through the unrolled window loop, and
// every lane issues the same loads and stores on every path (the padding lanes of
// a part's last window repeat its last row: same loads, same value, same store),
// so the waits on the vector-memory counter only cover loads issued a window ago.
// Waits are bounded in wall-clock time as in k_sweep_syncfree.
template <class T, int W>
struct PartRow {
    int i, meta;
    int cj[W];
    T ak[W];
    T di;
};
template <class T, int W>
struct PartWin {
    PartRow<T, W> o;
    T rhs;
    typename SfBits<T>::U pb[W];
};

template <class T, int W>
__device__ __forceinline__ void part_row_load(PartRow<T, W> &o, int q, int last, int n, const int *__restrict__ rows,
                                              const int *__restrict__ meta, const int *__restrict__ col,
                                              const T *__restrict__ av, const T *__restrict__ dv) {
    q = min(q, last);
    o.i = rows[q];
    o.meta = meta[q];
    o.di = dv[q];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        o.cj[k] = col[(size_t)k * n + q];
        o.ak[k] = av[(size_t)k * n + q];
    }
}

template <bool BWD>
__device__ __forceinline__ bool part_inside(int j, int r0, int r1) {
    return BWD ? j < r1 : j >= r0;
}

// the right-hand side and the first poll of the row's inputs outside the part:
// every slot issues a load, of the row's own output where the slot is inside the
// part or padding, so that every lane issues the same loads (the address is
// blended arithmetically: a select between two operand fields would keep the
// register sets in scratch memory)
template <class T, int W, bool BWD>
__device__ __forceinline__ void part_inputs(PartWin<T, W> &S, int r0, int r1, const T *__restrict__ in,
                                            const T *out) {
    S.rhs = in[S.o.i];
    const int c = S.o.meta & 255;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int j = S.o.cj[k];
        const int own = -(int)(k >= c || part_inside<BWD>(j, r0, r1));   // all ones or zero
        S.pb[k] = sf_load_bits(out + (j + ((S.o.i - j) & own)));
    }
}

// one window: S0 its operands and inputs, S1 the next window (operands loaded, its
// inputs issued here), N2 the window after (operands issued here); false: gave up
template <class T, int W, bool BWD>
__device__ __forceinline__ bool part_window(PartWin<T, W> &S0, PartWin<T, W> &S1, PartRow<T, W> &N2, int q0, int r0,
                                            int r1, int n, typename SfBits<T>::U *lv, const int *__restrict__ rows,
                                            const int *__restrict__ meta, const int *__restrict__ col,
                                            const T *__restrict__ av, const T *__restrict__ dv,
                                            const T *__restrict__ in, T small, T *out, SyncFree *sf,
                                            unsigned long long limit) {
    using U = typename SfBits<T>::U;
    const int lane = threadIdx.x;
    const int c = S0.o.meta & 255, step = (S0.o.meta >> 8) & 255;
    const int nst = __shfl((S0.o.meta >> 16) & 255, 0);
    T wk[W];
    unsigned need = 0;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        wk[k] = __builtin_bit_cast(T, S0.pb[k]);
        if (k < c && !part_inside<BWD>(S0.o.cj[k], r0, r1) && S0.pb[k] == SfBits<T>::kMark) need |= 1u << k;
    }
    if (__any(need != 0)) {
        const unsigned long long t0 = wall_clock64();
        do {
            __builtin_amdgcn_s_sleep(1);
            U r[W];
#pragma unroll
            for (int k = 0; k < W; ++k) r[k] = ((need >> k) & 1u) ? sf_load_bits(out + S0.o.cj[k]) : SfBits<T>::kMark;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                if (((need >> k) & 1u) && r[k] != SfBits<T>::kMark) {
                    wk[k] = __builtin_bit_cast(T, r[k]);
                    need &= ~(1u << k);
                }
            }
            const bool give_up = wall_clock64() - t0 > limit ||
                                 __hip_atomic_load(&sf->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
            if (__any(give_up)) {
                if (lane == 0) __hip_atomic_store(&sf->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
        } while (__any(need != 0));
    }
    const int last = r1 - 1;
    __builtin_amdgcn_sched_barrier(0);
    part_inputs<T, W, BWD>(S1, r0, r1, in, out);
    __builtin_amdgcn_sched_barrier(0);
    part_row_load<T, W>(N2, q0 + 128 + lane, last, n, rows, meta, col, av, dv);
    __builtin_amdgcn_sched_barrier(0);
    const T acc = BWD ? S0.rhs / (S0.o.di + small) : S0.rhs;
    T v = T(0);
    for (int st = 0; st < nst; ++st) {
        if (step == st) {
            U b[W];
#pragma unroll
            for (int k = 0; k < W; ++k) {
                const bool ins = k < c && part_inside<BWD>(S0.o.cj[k], r0, r1);
                b[k] = lv[ins ? S0.o.cj[k] - r0 : 0];
            }
            T r = acc;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                T x = wk[k];
                if (k < c && part_inside<BWD>(S0.o.cj[k], r0, r1)) x = __builtin_bit_cast(T, b[k]);
                if (k < c) r = r - S0.o.ak[k] * x;
            }
            v = r * S0.o.di;
            lv[S0.o.i - r0] = __builtin_bit_cast(U, v);
        }
    }
    sf_store(out + S0.o.i, v);
    return true;
}

template <class T, int W, bool BWD>
__global__ __launch_bounds__(64) void k_sweep_parts(int n, const int *__restrict__ rows, const int *__restrict__ meta,
                                                     const int *__restrict__ col, const T *__restrict__ av,
                                                     const T *__restrict__ dv, const T *__restrict__ in, T small,
                                                     T *out, SyncFree *sf, const Scalars *__restrict__ sc) {
    using U = typename SfBits<T>::U;
    __shared__ U lv[kPartRows];
    if (sc && sc->done) return;
    const int lane = threadIdx.x;
    const unsigned nparts = (unsigned)((n + kPartRows - 1) / kPartRows);
    const unsigned long long limit = sf_wait_ticks(sf, sf->rate_khz);
    for (;;) {
        unsigned t = 0;
        if (lane == 0) t = __hip_atomic_fetch_add(&sf->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        t = __shfl(t, 0);
        if (t >= nparts) break;
        const int part = BWD ? (int)(nparts - 1u - t) : (int)t;
        const int r0 = part * kPartRows, r1 = min(n, r0 + kPartRows);
        // the same issue order as inside the loop (inputs of A before the operands of
        // B), so that the loop head's waits on its two entries nearly agree
        PartWin<T, W> A, B, C;
        part_row_load<T, W>(A.o, r0 + lane, r1 - 1, n, rows, meta, col, av, dv);
        __builtin_amdgcn_sched_barrier(0);
        part_inputs<T, W, BWD>(A, r0, r1, in, out);
        __builtin_amdgcn_sched_barrier(0);
        part_row_load<T, W>(B.o, r0 + 64 + lane, r1 - 1, n, rows, meta, col, av, dv);
        __builtin_amdgcn_sched_barrier(0);
        bool ok = true;
        for (int q0 = r0;;) {
            if (!(ok = part_window<T, W, BWD>(A, B, C.o, q0, r0, r1, n, lv, rows, meta, col, av, dv, in, small, out,
                                               sf, limit)))
                break;
            if ((q0 += 64) >= r1) break;
            if (!(ok = part_window<T, W, BWD>(B, C, A.o, q0, r0, r1, n, lv, rows, meta, col, av, dv, in, small, out,
                                               sf, limit)))
                break;
            if ((q0 += 64) >= r1) break;
            if (!(ok = part_window<T, W, BWD>(C, A, B.o, q0, r0, r1, n, lv, rows, meta, col, av, dv, in, small, out,
                                               sf, limit)))
                break;
            if ((q0 += 64) >= r1) break;
        }
        if (!ok) break;
    }
    if (lane == 0) {
        const unsigned f = __hip_atomic_fetch_add(&sf->fin, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (f == gridDim.x - 1) {
            __hip_atomic_store(&sf->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&sf->fin, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// The pivot recurrences of k_dic_level / k_dilu_level with the same schedule:
// d_i = 1 / (a_ii - sum over the strict lower row, in CSR order, of a_k^2 d_j
// (DIC) or (a_k d_j) a_kT (DILU)); W bounds the strict lower row (the packed
// width of the forward triangle).
template <int W, bool DILU>
__global__ __launch_bounds__(64) void k_pivots_syncfree(int n, const int *__restrict__ rows, const int *__restrict__ ptr,
                                                        const int *__restrict__ ja, const int *__restrict__ diag,
                                                        const int *__restrict__ tpos, const double *__restrict__ a,
                                                        double *d, int *__restrict__ err, SyncFree *sf) {
    using U = SfBits<double>::U;
    const int lane = threadIdx.x;
    const unsigned nblk = (unsigned)((n + 63) / 64);
    const unsigned long long limit = sf_wait_ticks(sf, sf->rate_khz);
    for (;;) {
        unsigned b = 0;
        if (lane == 0) b = __hip_atomic_fetch_add(&sf->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b = __shfl(b, 0);
        if (b >= nblk) break;
        const int p = (int)b * 64 + lane;
        bool done = p >= n;
        int i = 0, c = 0;
        int cj[W];
        double a1[W], a2[W], dk[W];
        double di = 0.0;
        unsigned ready = 0;
        if (!done) {
            i = rows[p];
            const int k0 = ptr[i], kd = diag[i];
            c = kd - k0;
            di = a[kd];
#pragma unroll
            for (int k = 0; k < W; ++k) {
                const bool in = k < c;
                cj[k] = in ? ja[k0 + k] : 0;
                a1[k] = in ? a[k0 + k] : 0.0;
                a2[k] = in && DILU ? a[tpos[k0 + k]] : 0.0;
                dk[k] = 0.0;
            }
            if (c > W) {   // not a structure these sweeps were set up for: report, never wait
                atomicOr(err, 2);
                sf_store(d + i, __builtin_nan(""));
                done = true;
            }
        }
        int spins = 0;
        const unsigned long long t0 = wall_clock64();
        for (;;) {
            if (!done) {
                bool all = true;
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    if (k < c && !((ready >> k) & 1u)) {
                        const U bits = sf_load_bits(d + cj[k]);
                        if (bits != SfBits<double>::kMark) {
                            __builtin_memcpy(&dk[k], &bits, sizeof(U));
                            ready |= 1u << k;
                        } else {
                            all = false;
                        }
                    }
                }
                if (all) {
                    double v = di;
#pragma unroll
                    for (int k = 0; k < W; ++k) {
                        if (k < c) v = DILU ? v - (a1[k] * dk[k]) * a2[k] : v - (a1[k] * a1[k]) * dk[k];
                    }
                    if (!(v != 0.0) || !isfinite(v)) atomicOr(err, 1);
                    sf_store(d + i, 1.0 / v);
                    done = true;
                }
            }
            if (__all(done)) break;
            ++spins;
            bool give_up = false;
            if ((spins & 63) == 0) {
                give_up = wall_clock64() - t0 > limit ||
                          __hip_atomic_load(&sf->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
            }
            if (__any(give_up)) {
                if (lane == 0) __hip_atomic_store(&sf->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (lane == 0) {
        const unsigned f = __hip_atomic_fetch_add(&sf->fin, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (f == gridDim.x - 1) {
            __hip_atomic_store(&sf->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&sf->fin, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int W>
void pivots_syncfree_w(hipStream_t s, const Structure &st, const double *a, double *d, bool dilu, int *err,
                       const SweepCfg &cfg) {
    const int n = st.n;
    const int waves = std::max(1, std::min(cfg.waves, (n + 63) / 64));
    k_sf_mark<double><<<grid_for(n, 256), 256, 0, s>>>(n, d, nullptr);
    if (dilu)
        k_pivots_syncfree<W, true><<<waves, 64, 0, s>>>(n, st.rows_f.as<int>(), st.ioffset.as<int>(),
                                                         st.ja.as<int>(), st.diag.as<int>(), st.tpos.as<int>(), a, d,
                                                         err, cfg.sf);
    else
        k_pivots_syncfree<W, false><<<waves, 64, 0, s>>>(n, st.rows_f.as<int>(), st.ioffset.as<int>(),
                                                          st.ja.as<int>(), st.diag.as<int>(), nullptr, a, d, err,
                                                          cfg.sf);
}

void syncfree_reset(fc_context *ctx, hipStream_t s, SyncFree *sf) {
    SyncFree h{};
    h.wait_us = ctx->sf_wait_us;
    h.rate_khz = (unsigned)ctx->wall_clock_khz;
    FC_HIP_CHECK(hipMemcpyAsync(sf, &h, sizeof(SyncFree), hipMemcpyHostToDevice, s));
    FC_HIP_CHECK(hipStreamSynchronize(s));   // h lives on this stack frame
}

bool syncfree_sweeps(const Structure &st) {
    auto ok = [](int w) { return w == 3 || w == 4 || w == 8; };
    return packed_sweeps(st) && ok(st.ell_f) && ok(st.ell_b);
}

bool part_sweeps(const Structure &st) {
    return syncfree_sweeps(st) && st.parts && st.int_frac_f >= 0.5 && st.int_frac_b >= 0.5;
}

template <class T, int WF, int WB>
void lu_parts_w(hipStream_t s, const Structure &st, const PackedFactor<T> &pf, T small, const T *rhs, T *w,
                const Scalars *sc, const SweepCfg &cfg) {
    const int n = st.n;
    T *tmp = static_cast<T *>(cfg.tmp);
    const int nparts = (n + kPartRows - 1) / kPartRows;
    const int waves = std::max(1, std::min(cfg.waves, nparts));
    k_sf_mark<T><<<grid_for(n, 256), 256, 0, s>>>(n, tmp, sc);
    k_sweep_parts<T, WF, false><<<waves, 64, 0, s>>>(n, st.rows_pf.as<int>(), st.meta_pf.as<int>(), st.col_pf.as<int>(),
                                                       pf.af, pf.df, rhs, small, tmp, cfg.sf, sc);
    k_sf_mark<T><<<grid_for(n, 256), 256, 0, s>>>(n, w, sc);
    k_sweep_parts<T, WB, true><<<waves, 64, 0, s>>>(n, st.rows_pb.as<int>(), st.meta_pb.as<int>(), st.col_pb.as<int>(),
                                                      pf.ab, pf.db, tmp, small, w, cfg.sf, sc);
}

template <class T, int WF>
void lu_parts_b(hipStream_t s, const Structure &st, const PackedFactor<T> &pf, T small, const T *rhs, T *w,
                const Scalars *sc, const SweepCfg &cfg) {
    switch (st.ell_b) {
    case 3: lu_parts_w<T, WF, 3>(s, st, pf, small, rhs, w, sc, cfg); break;
    case 4: lu_parts_w<T, WF, 4>(s, st, pf, small, rhs, w, sc, cfg); break;
    default: lu_parts_w<T, WF, 8>(s, st, pf, small, rhs, w, sc, cfg); break;
    }
}

template <class T>
void lu_parts(hipStream_t s, const Structure &st, const PackedFactor<T> &pf, T small, const T *rhs, T *w,
              const Scalars *sc, const SweepCfg &cfg) {
    FC_REQUIRE(part_sweeps(st) && cfg.tmp && cfg.sf, FC_ERR_ARGUMENT, "part-wise sweeps not set up");
    switch (st.ell_f) {
    case 3: lu_parts_b<T, 3>(s, st, pf, small, rhs, w, sc, cfg); break;
    case 4: lu_parts_b<T, 4>(s, st, pf, small, rhs, w, sc, cfg); break;
    default: lu_parts_b<T, 8>(s, st, pf, small, rhs, w, sc, cfg); break;
    }
    FC_HIP_CHECK(hipGetLastError());
}

template <class T, int WF, int WB>
void lu_syncfree_w(hipStream_t s, const Structure &st, const PackedFactor<T> &pf, T small, const T *rhs, T *w,
                   const Scalars *sc, const SweepCfg &cfg) {
    const int n = st.n;
    T *tmp = static_cast<T *>(cfg.tmp);
    const int waves = std::max(1, std::min(cfg.waves, (n + 63) / 64));
    k_sf_mark<T><<<grid_for(n, 256), 256, 0, s>>>(n, tmp, sc);
    k_sweep_syncfree<T, WF, false><<<waves, 64, 0, s>>>(n, st.rows_f.as<int>(), st.cnt_f.as<int>(),
                                                          st.col_f.as<int>(), pf.af, pf.df, rhs, small, tmp, cfg.sf,
                                                          sc);
    k_sf_mark<T><<<grid_for(n, 256), 256, 0, s>>>(n, w, sc);
    k_sweep_syncfree<T, WB, true><<<waves, 64, 0, s>>>(n, st.rows_b.as<int>(), st.cnt_b.as<int>(),
                                                         st.col_b.as<int>(), pf.ab, pf.db, tmp, small, w, cfg.sf, sc);
}

template <class T, int WF>
void lu_syncfree_b(hipStream_t s, const Structure &st, const PackedFactor<T> &pf, T small, const T *rhs, T *w,
                   const Scalars *sc, const SweepCfg &cfg) {
    switch (st.ell_b) {
    case 3: lu_syncfree_w<T, WF, 3>(s, st, pf, small, rhs, w, sc, cfg); break;
    case 4: lu_syncfree_w<T, WF, 4>(s, st, pf, small, rhs, w, sc, cfg); break;
    default: lu_syncfree_w<T, WF, 8>(s, st, pf, small, rhs, w, sc, cfg); break;
    }
}

template <class T>
void lu_syncfree(hipStream_t s, const Structure &st, const PackedFactor<T> &pf, T small, const T *rhs, T *w,
                 const Scalars *sc, const SweepCfg &cfg) {
    FC_REQUIRE(syncfree_sweeps(st) && cfg.tmp && cfg.sf, FC_ERR_ARGUMENT, "dependency-driven sweeps not set up");
    switch (st.ell_f) {
    case 3: lu_syncfree_b<T, 3>(s, st, pf, small, rhs, w, sc, cfg); break;
    case 4: lu_syncfree_b<T, 4>(s, st, pf, small, rhs, w, sc, cfg); break;
    default: lu_syncfree_b<T, 8>(s, st, pf, small, rhs, w, sc, cfg); break;
    }
    FC_HIP_CHECK(hipGetLastError());
}

template <class T>
void lu_launch_t(hipStream_t s, const Structure &st, const T *a, const T *d, T small, const T *rhs, T *w,
                 const Scalars *sc, const SweepCfg *cfg, const PackedFactor<T> *pf) {
    if (cfg && cfg->sf && cfg->parts && pf && pf->af) {
        lu_parts<T>(s, st, *pf, small, rhs, w, sc, *cfg);
        return;
    }
    if (cfg && cfg->sf && pf && pf->af) {
        lu_syncfree<T>(s, st, *pf, small, rhs, w, sc, *cfg);
        return;
    }
    if (pf && pf->af) {
        sweeps_packed<T>(s, st, *pf, small, rhs, w, sc);
        return;
    }
    const int *ptr = st.ioffset.as<int>(), *ja = st.ja.as<int>(), *dg = st.diag.as<int>();
    const int *rf = st.rows_f.as<int>(), *rb = st.rows_b.as<int>();
    for (int l = 0; l < st.nlev_f; ++l) {
        int beg = st.lvl_f[l], cnt = st.lvl_f[l + 1] - beg;
        k_fwd_level<T><<<grid_for(cnt, 256), 256, 0, s>>>(rf, beg, cnt, ptr, ja, dg, a, d, rhs, w, sc);
    }
    for (int l = 0; l < st.nlev_b; ++l) {
        int beg = st.lvl_b[l], cnt = st.lvl_b[l + 1] - beg;
        k_bwd_level<T><<<grid_for(cnt, 256), 256, 0, s>>>(rb, beg, cnt, ptr, ja, dg, a, d, small, w, sc);
    }
}

void lu_launch64(hipStream_t s, const Structure &st, const double *a, const double *d, double small,
                 const double *rhs, double *w, const Scalars *sc, const SweepCfg *cfg,
                 const PackedFactor<double> *pf) {
    lu_launch_t<double>(s, st, a, d, small, rhs, w, sc, cfg, pf);
}

void lu_launch32(hipStream_t s, const Structure &st, const float *a, const float *d, float small, const float *rhs,
                 float *w, const Scalars *sc, const SweepCfg *cfg, const PackedFactor<float> *pf) {
    lu_launch_t<float>(s, st, a, d, small, rhs, w, sc, cfg, pf);
}

bool packed_sweeps(const Structure &st) { return st.levels_ready && st.ell_f > 0 && st.ell_b > 0; }

void pack_factor64(hipStream_t s, const Structure &st, const double *a, const double *d, double *af, double *df,
                   double *ab, double *db) {
    pack_factor_t<double>(s, st, a, d, af, df, ab, db);
}

void pack_factor32(hipStream_t s, const Structure &st, const float *a, const float *d, float *af, float *df,
                   float *ab, float *db) {
    pack_factor_t<float>(s, st, a, d, af, df, ab, db);
}

template <class T>
void pack_factor_parts_t(hipStream_t s, const Structure &st, const T *a, const T *d, T *af, T *df, T *ab, T *db) {
    FC_REQUIRE(st.parts && packed_sweeps(st), FC_ERR_ARGUMENT, "structure has no part-wise sweeps");
    const int n = st.n;
    k_pack_factor<T><<<grid_for(n, 256), 256, 0, s>>>(n, st.ell_f, st.rows_pf.as<int>(), st.cnt_pf.as<int>(),
                                                      st.kpos_pf.as<int>(), a, d, af, df);
    k_pack_factor<T><<<grid_for(n, 256), 256, 0, s>>>(n, st.ell_b, st.rows_pb.as<int>(), st.cnt_pb.as<int>(),
                                                      st.kpos_pb.as<int>(), a, d, ab, db);
    FC_HIP_CHECK(hipGetLastError());
}

void pack_factor_parts64(hipStream_t s, const Structure &st, const double *a, const double *d, double *af,
                         double *df, double *ab, double *db) {
    pack_factor_parts_t<double>(s, st, a, d, af, df, ab, db);
}

void pack_factor_parts32(hipStream_t s, const Structure &st, const float *a, const float *d, float *af, float *df,
                         float *ab, float *db) {
    pack_factor_parts_t<float>(s, st, a, d, af, df, ab, db);
}

// ----------------------------------------------------------- reductions

__global__ void k_sum_parts(int nparts, const double *__restrict__ pa, const double *__restrict__ pb,
                            double *__restrict__ out) {
    __shared__ double sh[kFinBlock / kWave];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kFinBlock) {
        a += pa[i];
        if (pb) b += pb[i];
    }
    a = block_sum<kFinBlock>(a, sh);
    b = block_sum<kFinBlock>(b, sh);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

template <class TA, class TB, int OP>
__global__ __launch_bounds__(kRedBlock) void k_reduce(long long n, const TA *__restrict__ x, const TB *__restrict__ y,
                                                      double *__restrict__ pa, double *__restrict__ pb) {
    __shared__ double sh[kRedBlock / kWave];
    double a = 0.0, b = 0.0;
    for (long long i = blockIdx.x * (long long)kRedBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kRedBlock) {
        double xv = (double)x[i];
        if (OP == 0) a += xv * (double)y[i];        // dot
        else if (OP == 1) { a += fabs(xv); b += xv * xv; }   // L1 and L2^2
    }
    a = block_sum<kRedBlock>(a, sh);
    if (OP == 1) b = block_sum<kRedBlock>(b, sh);
    if (threadIdx.x == 0) {
        pa[blockIdx.x] = a;
        if (OP == 1) pb[blockIdx.x] = b;
    }
}

void dot_launch64(hipStream_t s, int n, const double *x, const double *y, double *pa) {
    k_reduce<double, double, 0><<<red_grid(n), kRedBlock, 0, s>>>(n, x, y, pa, nullptr);
}

void norms_launch64(hipStream_t s, int n, const double *x, double *pa, double *pb) {
    k_reduce<double, double, 1><<<red_grid(n), kRedBlock, 0, s>>>(n, x, x, pa, pb);
}

void sum_parts_launch(hipStream_t s, int nparts, const double *pa, const double *pb, double *out) {
    k_sum_parts<<<1, kFinBlock, 0, s>>>(nparts, pa, pb, out);
}

}  // namespace fcap
