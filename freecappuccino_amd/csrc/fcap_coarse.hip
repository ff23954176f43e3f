// Aggregation coarse space of the flexible-CG preconditioner: P^T A P assembled
// in integer fixed point (exact sums, so the result does not depend on the order
// of the atomics), factored and inverted by rocSOLVER once per assembled matrix,
// and applied as restriction -> dense FP32 product -> prolongation.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "fcap_coarse.h"
#include "fcap_kernels.h"

namespace fcap {
namespace {

constexpr int kCoBlock = 256;   // threads per block; aggregates are whole multiples of it

inline int grid_for(long long n, int block) {
    return (int)std::max<long long>(1, std::min<long long>((n + block - 1) / block, 1 << 20));
}

inline int ceil_log2(long long v) {
    int b = 0;
    while ((1ll << b) < v) ++b;
    return b;
}

#define FC_BLAS_CHECK(call)                                                                                 \
    do {                                                                                                    \
        const rocblas_status st_ = (call);                                                                  \
        if (st_ != rocblas_status_success)                                                                  \
            throw Error(FC_ERR_HIP, std::string("rocSOLVER/rocBLAS call failed: ") + #call + " status " +   \
                                        std::to_string((int)st_));                                          \
    } while (0)

// largest |a_ij|, as the bits of a non-negative double (their order is the values' order)
__global__ __launch_bounds__(kCoBlock) void k_co_absmax(long long nnz, const double *__restrict__ a,
                                                       unsigned long long *__restrict__ amax) {
    double m = 0.0;
    for (long long k = (long long)blockIdx.x * kCoBlock + threadIdx.x; k < nnz; k += (long long)gridDim.x * kCoBlock)
        m = fmax(m, fabs(a[k]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(amax, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_co_zero(long long cnt, long long *__restrict__ x) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += (long long)gridDim.x * blockDim.x)
        x[k] = 0;
}

__device__ inline double fix_scale(const unsigned long long *amax, int fbits) {
    const double m = __longlong_as_double((long long)*amax);
    return m > 0.0 ? ldexp(1.0, fbits) / m : 1.0;
}

// A_c[I][J] += a_ij for every entry of the rows at positions p (block = 256
// positions of one aggregate).  Entries inside the row's own aggregate are summed
// per block and added once; the others go straight to their (I, J) element.
// Column-major: element (I, J) at I + J nc.
__global__ __launch_bounds__(kCoBlock) void k_co_accum(int n, int agg, int nc, const int *__restrict__ ptr,
                                                      const int *__restrict__ ja, const double *__restrict__ a,
                                                      const int *__restrict__ perm, const int *__restrict__ inv,
                                                      const unsigned long long *__restrict__ amax, int fbits,
                                                      long long *__restrict__ afix) {
    __shared__ long long part[kCoBlock / kWave];
    const double scale = fix_scale(amax, fbits);
    const int base = blockIdx.x * kCoBlock;
    const int I = base / agg;
    const int p = base + (int)threadIdx.x;
    long long own = 0;
    if (p < n) {
        const int i = perm ? perm[p] : p;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            const int j = ja[k];
            const int J = (inv ? inv[j] : j) / agg;
            const long long v = llrint(a[k] * scale);
            if (J == I)
                own += v;
            else
                atomicAdd(reinterpret_cast<unsigned long long *>(afix + (size_t)J * nc + I), (unsigned long long)v);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) own += __shfl_down(own, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = own;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < kCoBlock / kWave; ++w) t += part[w];
        atomicAdd(reinterpret_cast<unsigned long long *>(afix + (size_t)I * nc + I), (unsigned long long)t);
    }
}

__global__ void k_co_to_f64(long long cnt, const long long *__restrict__ afix, const unsigned long long *__restrict__ amax,
                            int fbits, double *__restrict__ ac) {
    const double inv_scale = 1.0 / fix_scale(amax, fbits);
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += (long long)gridDim.x * blockDim.x)
        ac[k] = (double)afix[k] * inv_scale;
}

// FP32 row-major copy of the symmetric inverse from its lower triangle (column-major)
__global__ void k_co_inv32(int nc, const double *__restrict__ ac, float *__restrict__ ainv) {
    const long long cnt = (long long)nc * nc;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(k / nc), c = (int)(k % nc);
        const int hi = max(r, c), lo = min(r, c);
        ainv[k] = (float)ac[hi + (size_t)lo * nc];
    }
}

// (aggregate, cell) keys of the positions, to list each aggregate's cells in cell order
__global__ void k_co_cell_keys(int n, int agg, const int *__restrict__ perm, unsigned long long *__restrict__ key,
                               int *__restrict__ val) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int j = perm[p];
        key[p] = ((unsigned long long)(p / agg) << 32) | (unsigned)j;
        val[p] = j;
    }
}

// Per slot q (cell j = cells[q] of aggregate I = q / agg): the number of entries of
// row j whose column lies in another aggregate.
__global__ void k_co_cross_count(int n, int agg, const int *__restrict__ ptr, const int *__restrict__ ja,
                                 const int *__restrict__ cells, const int *__restrict__ inv, int *__restrict__ cnt) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p <= n; p += gridDim.x * blockDim.x) {
        int c = 0;
        if (p < n) {
            const int j = cells ? cells[p] : p, I = p / agg;
            for (int k = ptr[j]; k < ptr[j + 1]; ++k) c += ((inv ? inv[ja[k]] : ja[k]) / agg != I);
        }
        cnt[p] = c;
    }
}

// colsum[p] = sum over the rows i of I of a_ij (j = cell at p: the transposes of
// row j's entries inside I), and row j's entries outside I in CSR order.
__global__ void k_co_cross_fill(int n, int agg, const int *__restrict__ ptr, const int *__restrict__ ja,
                                const float *__restrict__ a32, const int *__restrict__ trans,
                                const int *__restrict__ cells, const int *__restrict__ inv,
                                const int *__restrict__ xoff, float *__restrict__ colsum, int *__restrict__ xcol,
                                float *__restrict__ xval) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int j = cells ? cells[p] : p, I = p / agg;
        int o = xoff[p];
        float s = 0.f;
        for (int k = ptr[j]; k < ptr[j + 1]; ++k) {
            const int c = ja[k];
            if ((inv ? inv[c] : c) / agg == I) {
                s += a32[trans ? trans[k] : k];
            } else {
                xcol[o] = c;
                xval[o] = a32[k];
                ++o;
            }
        }
        colsum[p] = s;
    }
}

// rc[I] = sum over I of (r - A z) = sum_{j in I} (r_j - colsum_j z_j) - sum of the
// aggregate's outside entries a z_c (one block per aggregate, a fixed order of sums)
__global__ __launch_bounds__(kCoBlock) void k_co_restrict(int n, int agg, const int *__restrict__ cells,
                                                         const float *__restrict__ r, const float *__restrict__ z,
                                                         const float *__restrict__ colsum,
                                                         const int *__restrict__ xoff, const int *__restrict__ xcol,
                                                         const float *__restrict__ xval, float *__restrict__ rc,
                                                         const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    __shared__ float part[kCoBlock / kWave];
    const int I = blockIdx.x;
    const int beg = I * agg, end = min(n, beg + agg);
    // four positions (and four outside entries) per trip, their loads issued together
    float v = 0.f;
    int p = beg + (int)threadIdx.x;
    for (; p + 3 * kCoBlock < end; p += 4 * kCoBlock) {
        int j[4];
        float c[4], rv[4], zv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            j[u] = cells ? cells[p + u * kCoBlock] : p + u * kCoBlock;
            c[u] = colsum[p + u * kCoBlock];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            rv[u] = r[j[u]];
            zv[u] = z[j[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) v += rv[u] - c[u] * zv[u];
    }
    for (; p < end; p += kCoBlock) {
        const int j = cells ? cells[p] : p;
        v += r[j] - colsum[p] * z[j];
    }
    const int e1 = xoff[end];
    int e = xoff[beg] + (int)threadIdx.x;
    for (; e + 3 * kCoBlock < e1; e += 4 * kCoBlock) {
        int col[4];
        float a[4], zv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            col[u] = xcol[e + u * kCoBlock];
            a[u] = xval[e + u * kCoBlock];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) zv[u] = z[col[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) v -= a[u] * zv[u];
    }
    for (; e < e1; e += kCoBlock) v -= xval[e] * z[xcol[e]];
    v = wave_sum(v);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int w = 0; w < kCoBlock / kWave; ++w) s += part[w];
        rc[I] = s;
    }
}

// yc = A_c^{-1} rc: one wave per row
__global__ __launch_bounds__(kCoBlock) void k_co_gemv(int nc, const float *__restrict__ ainv, const float *__restrict__ rc,
                                                     float *__restrict__ yc, const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    const int row = blockIdx.x * (kCoBlock / kWave) + (int)(threadIdx.x / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    if (row >= nc) return;
    const float *ar = ainv + (size_t)row * nc;
    float v = 0.f;
    for (int c = lane; c < nc; c += kWave) v = fmaf(ar[c], rc[c], v);
    v = wave_sum(v);
    if (lane == 0) yc[row] = v;
}

// z = zf + yc[aggregate of the cell]
__global__ void k_co_prolong(int n, int agg, const int *__restrict__ inv, const float *__restrict__ zf,
                             const float *__restrict__ yc, float *__restrict__ z, const Scalars *__restrict__ sc) {
    if (sc && sc->done) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        z[i] = zf[i] + yc[(inv ? inv[i] : i) / agg];
}

rocblas_handle dense_handle(fc_context *ctx) {
    if (!ctx->dense_handle) {
        rocblas_handle h = nullptr;
        FC_BLAS_CHECK(rocblas_create_handle(&h));
        ctx->dense_handle = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(ctx->dense_handle);
    FC_BLAS_CHECK(rocblas_set_stream(h, ctx->stream));
    return h;
}

}  // namespace

int coarse_aggregate_cells(int n, int requested) {
    const long long need = ((long long)n + kCoarseMaxAggregates - 1) / kCoarseMaxAggregates;
    const long long grown = (need + kCoBlock - 1) / kCoBlock * kCoBlock;
    return (int)std::max<long long>(requested, grown);
}

void coarse_setup(fc_context *ctx, CoarseSpace &cs, int n, long long nnz, const int *ptr, const int *ja,
                  const double *a, const float *a32, const int *trans, const int *perm, const int *inv, int agg,
                  int max_row, long long version) {
    FC_REQUIRE(agg >= kCoBlock && agg % kCoBlock == 0, FC_ERR_ARGUMENT, "aggregate size must be a multiple of 256");
    const hipStream_t s = ctx->stream;
    // the outside entries' positions depend on the structure and the aggregates only
    const bool same_cross = cs.nx >= 0 && cs.n == n && cs.agg == agg && cs.perm == perm && cs.inv == inv &&
                            cs.cross_ptr == ptr;
    cs.ok = false;
    cs.n = n;
    cs.agg = agg;
    cs.nc = (int)(((long long)n + agg - 1) / agg);
    cs.perm = perm;
    cs.inv = inv;
    cs.version = version;
    const int nc = cs.nc;
    const long long cnt = (long long)nc * nc;
    // fixed point: |a_ij| <= 2^fbits, at most agg * max_row terms per element, all below 2^62
    cs.fbits = std::min(40, 61 - ceil_log2(agg) - ceil_log2(std::max(max_row, 1)));
    cs.afix.ensure(sizeof(long long) * (size_t)cnt);
    cs.ac.ensure(sizeof(double) * (size_t)cnt);
    cs.ainv.ensure(sizeof(float) * (size_t)cnt);
    cs.amax.ensure(sizeof(unsigned long long));
    cs.info.ensure(sizeof(int) * 2);
    cs.rc.ensure(sizeof(float) * (size_t)nc);
    cs.yc.ensure(sizeof(float) * (size_t)nc);
    const int nnz_rows = n;
    k_co_zero<<<1, 1, 0, s>>>(1, reinterpret_cast<long long *>(cs.amax.get()));
    k_co_absmax<<<std::min(grid_for(nnz, kCoBlock), 4096), kCoBlock, 0, s>>>(nnz, a,
                                                                            cs.amax.as<unsigned long long>());
    k_co_zero<<<std::min(grid_for(cnt, 256), 8192), 256, 0, s>>>(cnt, cs.afix.as<long long>());
    k_co_accum<<<grid_for(nnz_rows, kCoBlock), kCoBlock, 0, s>>>(n, agg, nc, ptr, ja, a, perm, inv,
                                                                cs.amax.as<unsigned long long>(), cs.fbits,
                                                                cs.afix.as<long long>());
    k_co_to_f64<<<std::min(grid_for(cnt, 256), 8192), 256, 0, s>>>(cnt, cs.afix.as<long long>(),
                                                                   cs.amax.as<unsigned long long>(), cs.fbits,
                                                                   cs.ac.as<double>());
    FC_HIP_CHECK(hipGetLastError());
    rocblas_handle h = dense_handle(ctx);
    int *info = cs.info.as<int>();
    // both enqueued before one status read; dpotri of a failed factor only reports a
    // status of its own, and then neither result is used
    FC_BLAS_CHECK(rocsolver_dpotrf(h, rocblas_fill_lower, nc, cs.ac.as<double>(), nc, info));
    FC_BLAS_CHECK(rocsolver_dpotri(h, rocblas_fill_lower, nc, cs.ac.as<double>(), nc, info + 1));
    int hinfo[2] = {0, 0};
    FC_HIP_CHECK(hipMemcpyAsync(hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost, s));
    FC_HIP_CHECK(hipStreamSynchronize(s));
    if (hinfo[0] != 0 || hinfo[1] != 0) return;   // not positive definite: no correction
    k_co_inv32<<<std::min(grid_for(cnt, 256), 8192), 256, 0, s>>>(nc, cs.ac.as<double>(), cs.ainv.as<float>());
    // the restriction's operator P^T A
    cs.xoff.ensure(sizeof(int) * ((size_t)n + 1));
    cs.colsum.ensure(sizeof(float) * (size_t)n);
    const int gb = std::min(grid_for((long long)n + 1, 256), 8192);
    if (!same_cross) {
        cs.cell_of = nullptr;
        if (perm) {
            DevBuf ka(sizeof(unsigned long long) * (size_t)n), kb(sizeof(unsigned long long) * (size_t)n),
                va(sizeof(int) * (size_t)n);
            cs.cells.ensure(sizeof(int) * (size_t)n);
            k_co_cell_keys<<<gb, 256, 0, s>>>(n, agg, perm, ka.as<unsigned long long>(), va.as<int>());
            sort_pairs_u64(s, ka.as<unsigned long long>(), kb.as<unsigned long long>(), va.as<int>(),
                           cs.cells.as<int>(), n, cs.scan_tmp);
            cs.cell_of = cs.cells.as<int>();
        }
        cs.xcnt.ensure(sizeof(int) * ((size_t)n + 1));
        k_co_cross_count<<<gb, 256, 0, s>>>(n, agg, ptr, ja, cs.cell_of, inv, cs.xcnt.as<int>());
        exclusive_scan_i32(s, cs.xcnt.as<int>(), cs.xoff.as<int>(), n + 1, cs.scan_tmp);
        int last = 0;
        FC_HIP_CHECK(hipMemcpyAsync(&last, cs.xoff.as<int>() + n, sizeof(int), hipMemcpyDeviceToHost, s));
        FC_HIP_CHECK(hipStreamSynchronize(s));
        cs.nx = last;
        cs.cross_ptr = ptr;
    }
    const int total = (int)cs.nx;
    cs.xcol.ensure(sizeof(int) * (size_t)std::max(total, 1));
    cs.xval.ensure(sizeof(float) * (size_t)std::max(total, 1));
    k_co_cross_fill<<<gb, 256, 0, s>>>(n, agg, ptr, ja, a32, trans, cs.cell_of, inv, cs.xoff.as<int>(),
                                       cs.colsum.as<float>(), cs.xcol.as<int>(), cs.xval.as<float>());
    FC_HIP_CHECK(hipGetLastError());
    cs.ok = true;
}

void coarse_matrix_host(fc_context *ctx, const CoarseSpace &cs, double *out) {
    const hipStream_t s = ctx->stream;
    const size_t cnt = (size_t)cs.nc * cs.nc;
    std::vector<long long> fix(cnt);
    unsigned long long bits = 0;
    FC_HIP_CHECK(hipMemcpyAsync(fix.data(), cs.afix.get(), sizeof(long long) * cnt, hipMemcpyDeviceToHost, s));
    FC_HIP_CHECK(hipMemcpyAsync(&bits, cs.amax.get(), sizeof bits, hipMemcpyDeviceToHost, s));
    FC_HIP_CHECK(hipStreamSynchronize(s));
    double m = 0.0;
    std::memcpy(&m, &bits, sizeof m);
    const double inv_scale = m > 0.0 ? m / std::ldexp(1.0, cs.fbits) : 1.0;
    for (size_t k = 0; k < cnt; ++k) out[k] = (double)fix[k] * inv_scale;
}

void coarse_apply(hipStream_t s, CoarseSpace &cs, const float *r, const float *zf, float *z, const Scalars *sc) {
    const int n = cs.n;
    k_co_restrict<<<cs.nc, kCoBlock, 0, s>>>(n, cs.agg, cs.cell_of, r, zf, cs.colsum.as<float>(), cs.xoff.as<int>(),
                                             cs.xcol.as<int>(), cs.xval.as<float>(), cs.rc.as<float>(), sc);
    k_co_gemv<<<grid_for(cs.nc, kCoBlock / kWave), kCoBlock, 0, s>>>(cs.nc, cs.ainv.as<float>(), cs.rc.as<float>(),
                                                                     cs.yc.as<float>(), sc);
    k_co_prolong<<<std::min(grid_for(n, 256), 8192), 256, 0, s>>>(n, cs.agg, cs.inv, zf, cs.yc.as<float>(), z, sc);
}

void coarse_release(fc_context *ctx) {
    if (ctx && ctx->dense_handle) {
        (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->dense_handle));
        ctx->dense_handle = nullptr;
    }
}

}  // namespace fcap
