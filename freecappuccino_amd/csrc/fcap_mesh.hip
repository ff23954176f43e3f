// Mesh geometry, CSR structure and level analysis on the GPU.
//
// Geometry follows src/mesh_geometry_and_topology.f90:859-1081 operation by
// operation (this file is compiled with -ffp-contract=off).  The reference
// accumulates each cell's volume and centroid over its faces in face order;
// here one thread per cell walks its faces in ascending face order, so every
// accumulation happens in the same order and the results are bit-identical.
//
// The CSR structure (src/sparse_matrix.f90:42-172) is the sorted set of
// (row, col) pairs {(i,i)} U {(o,n)} U {(n,o)}; it is built by a counting
// pass, a per-row insertion sort and per-face searches for the slot maps
// (csr_to_k, src/utils.f90:76-128).

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "fcap_internal.h"
#include "fcap_kernels.h"

namespace fcap {

namespace {

constexpr int kBlock = 256;

inline int grid_for(long long n, int block = kBlock) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (int)std::min<long long>(g, 1 << 30);
}

__global__ void k_sub_one(const int *__restrict__ in, int *__restrict__ out, long long n) {
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] - 1;
}

__global__ void k_add_one(int *__restrict__ x, long long n) {
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] + 1;
}

__global__ void k_fill_int(int *__restrict__ x, long long n, int v) {
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// error flag bits written by validation kernels
enum : int { kErrRange = 1, kErrDuplicate = 2, kErrNoDiag = 4, kErrUnsorted = 8, kErrAsym = 16,
             kErrNoTranspose = 32 };

void exclusive_scan(fc_context *ctx, const int *in, int *out, int n) {
    size_t tmp = 0;
    FC_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, ctx->stream));
    DevBuf t(tmp);
    FC_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(t.get(), tmp, in, out, n, ctx->stream));
}

int read_flag(fc_context *ctx, const int *dflag) {
    int h = 0;
    d2h(&h, dflag, sizeof(int), ctx->stream);
    return h;
}

// ----------------------------------------------------------------- CSR build

__global__ void k_row_counts(const int *__restrict__ own, const int *__restrict__ nei, int nif, int n,
                             int *__restrict__ cnt, int *__restrict__ err) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int o = own[f], nb = nei[f];
    if (o < 0 || o >= n || nb < 0 || nb >= n || o == nb) {
        atomicOr(err, kErrRange);
        return;
    }
    atomicAdd(&cnt[o], 1);
    atomicAdd(&cnt[nb], 1);
}

__global__ void k_init_counts(int *__restrict__ cnt, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = 1;      // the diagonal
    if (i == n) cnt[i] = 0;
}

__global__ void k_place_diag(const int *__restrict__ ptr, int *__restrict__ cursor, int *__restrict__ ja, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ja[ptr[i]] = i;
    cursor[i] = ptr[i] + 1;
}

__global__ void k_place_faces(const int *__restrict__ own, const int *__restrict__ nei, int nif,
                              int *__restrict__ cursor, int *__restrict__ ja) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int o = own[f], nb = nei[f];
    ja[atomicAdd(&cursor[o], 1)] = nb;
    ja[atomicAdd(&cursor[nb], 1)] = o;
}

// Sort each row's columns (rows are short), locate the diagonal, reject duplicates.
__global__ void k_sort_rows(const int *__restrict__ ptr, int *__restrict__ ja, int *__restrict__ diag, int n,
                            int *__restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = ptr[i], e = ptr[i + 1];
    for (int k = s + 1; k < e; ++k) {
        int v = ja[k];
        int m = k - 1;
        while (m >= s && ja[m] > v) {
            ja[m + 1] = ja[m];
            --m;
        }
        ja[m + 1] = v;
    }
    int dpos = -1;
    for (int k = s; k < e; ++k) {
        if (k > s && ja[k] == ja[k - 1]) atomicOr(err, kErrDuplicate);
        if (ja[k] == i && dpos < 0) dpos = k;
    }
    diag[i] = dpos;
}

// csr_to_k (src/utils.f90:76-128): first k in row i with ja(k) == j, else -1 (0 in 1-based).
__device__ inline int csr_to_k(const int *__restrict__ ptr, const int *__restrict__ ja, int i, int j) {
    for (int k = ptr[i]; k < ptr[i + 1]; ++k)
        if (ja[k] == j) return k;
    return -1;
}

__global__ void k_slot_maps(const int *__restrict__ own, const int *__restrict__ nei, int nif,
                            const int *__restrict__ ptr, const int *__restrict__ ja, int *__restrict__ ij,
                            int *__restrict__ ji) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int o = own[f], nb = nei[f];
    ij[f] = csr_to_k(ptr, ja, o, nb);
    ji[f] = csr_to_k(ptr, ja, nb, o);
}

// ------------------------------------------------- validation of a host CSR

__global__ void k_validate_csr(const int *__restrict__ ptr, const int *__restrict__ ja, const int *__restrict__ diag,
                               int n, int nnz, int *__restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = ptr[i], e = ptr[i + 1];
    if (s < 0 || e > nnz || e < s) { atomicOr(err, kErrRange); return; }
    int d = diag[i];
    if (d < s || d >= e || ja[d] != i) { atomicOr(err, kErrNoDiag); return; }
    for (int k = s; k < e; ++k) {
        int j = ja[k];
        if (j < 0 || j >= n) { atomicOr(err, kErrRange); return; }
        if (k > s && ja[k - 1] >= j) atomicOr(err, kErrUnsorted);
        if (j != i) {
            // structural symmetry: (j, i) exists (binary search in the sorted row j)
            int lo = ptr[j], hi = ptr[j + 1] - 1;
            bool found = false;
            while (lo <= hi) {
                int mid = (lo + hi) >> 1;
                int c = ja[mid];
                if (c == i) { found = true; break; }
                if (c < i) lo = mid + 1; else hi = mid - 1;
            }
            if (!found) atomicOr(err, kErrAsym);
        }
    }
}

// ------------------------------------------------------------ level analysis
// Kahn's algorithm over the strict-lower DAG (forward sweep) or the
// strict-upper DAG (backward sweep).  With a structurally symmetric pattern,
// the dependents of row j in the forward DAG are the columns of its strict
// upper part, and vice versa.

__global__ void k_indeg(const int *__restrict__ ptr, const int *__restrict__ diag, int n, int upper,
                        int *__restrict__ indeg, int *__restrict__ level, int *__restrict__ frontier,
                        int *__restrict__ count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = upper ? (ptr[i + 1] - 1 - diag[i]) : (diag[i] - ptr[i]);
    indeg[i] = c;
    if (c == 0) {
        level[i] = 0;
        frontier[atomicAdd(count, 1)] = i;
    }
}

__global__ void k_kahn_step(const int *__restrict__ ptr, const int *__restrict__ ja, const int *__restrict__ diag,
                            int upper, const int *__restrict__ cur, int ncur, int lev,
                            int *__restrict__ indeg, int *__restrict__ level, int *__restrict__ next,
                            int *__restrict__ count) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncur) return;
    int j = cur[t];
    int s, e;
    if (upper) { s = ptr[j]; e = diag[j]; }            // backward DAG: rows i < j with U(i,j)
    else { s = diag[j] + 1; e = ptr[j + 1]; }          // forward DAG: rows i > j with L(i,j)
    for (int k = s; k < e; ++k) {
        int i = ja[k];
        if (atomicSub(&indeg[i], 1) == 1) {
            level[i] = lev + 1;
            next[atomicAdd(count, 1)] = i;
        }
    }
}

__global__ void k_iota(int *__restrict__ x, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = i;
}

void levels_one(fc_context *ctx, Structure &st, bool upper, DevBuf &rows_out, std::vector<int> &lvl_ptr,
                int &nlev, int &maxlev) {
    const int n = st.n;
    hipStream_t s = ctx->stream;
    DevBuf indeg(sizeof(int) * n), level(sizeof(int) * n), fa(sizeof(int) * n), fb(sizeof(int) * n),
        cnt(sizeof(int) * 2);
    FC_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(int) * 2, s));
    k_indeg<<<grid_for(n), kBlock, 0, s>>>(st.ioffset.as<int>(), st.diag.as<int>(), n, upper ? 1 : 0,
                                          indeg.as<int>(), level.as<int>(), fa.as<int>(), cnt.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    int ncur = read_flag(ctx, cnt.as<int>());
    std::vector<int> sizes;
    int *cur = fa.as<int>(), *nxt = fb.as<int>();
    int lev = 0, done = 0;
    while (ncur > 0) {
        sizes.push_back(ncur);
        done += ncur;
        FC_HIP_CHECK(hipMemsetAsync(cnt.as<int>() + 1, 0, sizeof(int), s));
        k_kahn_step<<<grid_for(ncur), kBlock, 0, s>>>(st.ioffset.as<int>(), st.ja.as<int>(), st.diag.as<int>(),
                                                     upper ? 1 : 0, cur, ncur, lev, indeg.as<int>(),
                                                     level.as<int>(), nxt, cnt.as<int>() + 1);
        FC_HIP_CHECK(hipGetLastError());
        ncur = read_flag(ctx, cnt.as<int>() + 1);
        std::swap(cur, nxt);
        ++lev;
    }
    FC_REQUIRE(done == n, FC_ERR_STRUCTURE, "level analysis: dependency graph is not acyclic");
    nlev = (int)sizes.size();
    lvl_ptr.assign(nlev + 1, 0);
    maxlev = 0;
    for (int l = 0; l < nlev; ++l) {
        lvl_ptr[l + 1] = lvl_ptr[l] + sizes[l];
        maxlev = std::max(maxlev, sizes[l]);
    }
    // rows ordered by (level, row): stable radix sort of the levels with row ids as values
    DevBuf ids(sizeof(int) * n), keys_out(sizeof(int) * n);
    rows_out.alloc(sizeof(int) * n);
    k_iota<<<grid_for(n), kBlock, 0, s>>>(ids.as<int>(), n);
    FC_HIP_CHECK(hipGetLastError());
    int bits = 1;
    while ((1 << bits) <= nlev && bits < 31) ++bits;
    size_t tmp = 0;
    FC_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, level.as<int>(), keys_out.as<int>(),
                                                    ids.as<int>(), rows_out.as<int>(), n, 0, bits, s));
    DevBuf t(tmp);
    FC_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(t.get(), tmp, level.as<int>(), keys_out.as<int>(),
                                                    ids.as<int>(), rows_out.as<int>(), n, 0, bits, s));
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

__global__ void k_tpos(const int *__restrict__ ptr, const int *__restrict__ ja, const int *__restrict__ diag, int n,
                       int *__restrict__ tpos, int *__restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = ptr[i]; k < diag[i]; ++k) {
        int j = ja[k];
        // src/bicgstab.f90:71-75: scan row j from its diagonal for column i
        int l = diag[j], e = ptr[j + 1];
        for (; l < e; ++l)
            if (ja[l] == i) break;
        if (l == e) {
            atomicOr(err, kErrNoTranspose);
            tpos[k] = -1;
        } else {
            tpos[k] = l;
        }
    }
}

// ------------------------------------------------------------------ geometry

__global__ void k_cell_face_count(const int *__restrict__ own, const int *__restrict__ nei, int nfaces, int nif,
                                  int n, int *__restrict__ cnt, int *__restrict__ err) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfaces) return;
    int o = own[f];
    if (o < 0 || o >= n) { atomicOr(err, kErrRange); return; }
    atomicAdd(&cnt[o], 1);
    if (f < nif) {
        int nb = nei[f];
        if (nb < 0 || nb >= n) { atomicOr(err, kErrRange); return; }
        atomicAdd(&cnt[nb], 1);
    }
}

__global__ void k_cell_face_fill(const int *__restrict__ own, const int *__restrict__ nei, int nfaces, int nif,
                                 int *__restrict__ cursor, int *__restrict__ cf) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfaces) return;
    cf[atomicAdd(&cursor[own[f]], 1)] = 2 * f;          // even: the cell owns the face
    if (f < nif) cf[atomicAdd(&cursor[nei[f]], 1)] = 2 * f + 1;   // odd: neighbour side
}

__global__ void k_copy_int(const int *__restrict__ in, int *__restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

__global__ void k_sort_segments(const int *__restrict__ ptr, int *__restrict__ v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = ptr[i], e = ptr[i + 1];
    for (int k = s + 1; k < e; ++k) {
        int x = v[k];
        int m = k - 1;
        while (m >= s && v[m] > x) { v[m + 1] = v[m]; --m; }
        v[m + 1] = x;
    }
}

__global__ void k_face_node_check(const int *__restrict__ fptr, const int *__restrict__ fn, int nfaces,
                                  int nnodes, int *__restrict__ err) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfaces) return;
    int s = fptr[f], e = fptr[f + 1];
    if (e - s < 3) { atomicOr(err, kErrRange); return; }
    for (int k = s; k < e; ++k)
        if (fn[k] < 0 || fn[k] >= nnodes) { atomicOr(err, kErrRange); return; }
}

// sweep 1, face part: area vector and face centre (:876-948)
__global__ void k_face_area_centre(const double *__restrict__ xyz, const int *__restrict__ fptr,
                                   const int *__restrict__ fn, int nfaces, double *__restrict__ arx,
                                   double *__restrict__ ary, double *__restrict__ arz, double *__restrict__ xf,
                                   double *__restrict__ yf, double *__restrict__ zf) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfaces) return;
    const double half = 0.5;
    int s = fptr[f], nn = fptr[f + 1] - s;
    const int *nd = fn + s;
    double x1 = xyz[3 * nd[0]], y1 = xyz[3 * nd[0] + 1], z1 = xyz[3 * nd[0] + 2];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int i = 0; i < nn - 2; ++i) {
        const double *p2 = xyz + 3 * nd[i + 1];
        const double *p3 = xyz + 3 * nd[i + 2];
        double px = p2[0] - x1, py = p2[1] - y1, pz = p2[2] - z1;
        double qx = p3[0] - x1, qy = p3[1] - y1, qz = p3[2] - z1;
        double nx = py * qz - pz * qy;
        double ny = pz * qx - px * qz;
        double nz = px * qy - py * qx;
        ax = ax + half * nx;
        ay = ay + half * ny;
        az = az + half * nz;
    }
    arx[f] = ax;
    ary[f] = ay;
    arz[f] = az;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < nn; ++i) {
        sx = sx + xyz[3 * nd[i]];
        sy = sy + xyz[3 * nd[i] + 1];
        sz = sz + xyz[3 * nd[i] + 2];
    }
    xf[f] = sx / (double)nn;
    yf[f] = sy / (double)nn;
    zf[f] = sz / (double)nn;
}

// sweep 1, cell part: volume accumulated over the cell's faces in face order (:917-921)
__global__ void k_cell_volume(const double *__restrict__ xyz, const int *__restrict__ fptr,
                              const int *__restrict__ fn, const int *__restrict__ cfp, const int *__restrict__ cf,
                              int n, double *__restrict__ vol) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double one_third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    double v = 0.0;
    for (int q = cfp[c]; q < cfp[c + 1]; ++q) {
        int code = cf[q];
        int f = code >> 1;
        bool nb_side = code & 1;
        int s = fptr[f], nn = fptr[f + 1] - s;
        const int *nd = fn + s;
        double x1 = xyz[3 * nd[0]], y1 = xyz[3 * nd[0] + 1], z1 = xyz[3 * nd[0] + 2];
        for (int i = 0; i < nn - 2; ++i) {
            const double *p2 = xyz + 3 * nd[i + 1];
            const double *p3 = xyz + 3 * nd[i + 2];
            double px = p2[0] - x1, py = p2[1] - y1, pz = p2[2] - z1;
            double qx = p3[0] - x1, qy = p3[1] - y1, qz = p3[2] - z1;
            double nx = py * qz - pz * qy;
            double ny = pz * qx - px * qz;
            double nz = px * qy - py * qx;
            double cx = one_third * ((p3[0] + p2[0]) + x1);
            double cy = one_third * ((p3[1] + p2[1]) + y1);
            double cz = one_third * ((p3[2] + p2[2]) + z1);
            double part;
            if (!nb_side) part = sixth * ((cx * nx + cy * ny) + cz * nz);
            else part = sixth * ((cx * (-nx) + cy * (-ny)) + cz * (-nz));
            v = v + part;
        }
    }
    vol[c] = v;
}

// centroid_component_part_polymesh (:160-183), evaluated left to right
__device__ inline double centroid_part(double ax, double bx, double cx, double nx, double vol) {
    double t = 1.0 / (2.0 * vol);
    t = t * 1.0;
    t = t / 24.0;
    t = t * nx;
    double ab = ax + bx, bc = bx + cx, ca = cx + ax;
    double s = ab * ab + bc * bc;
    s = s + ca * ca;
    return t * s;
}

// sweep 2: cell centres (:974-1013)
__global__ void k_cell_centre(const double *__restrict__ xyz, const int *__restrict__ fptr,
                              const int *__restrict__ fn, const int *__restrict__ cfp, const int *__restrict__ cf,
                              int n, const double *__restrict__ vol, double *__restrict__ xc,
                              double *__restrict__ yc, double *__restrict__ zc) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double vo = vol[c];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int q = cfp[c]; q < cfp[c + 1]; ++q) {
        int code = cf[q];
        int f = code >> 1;
        bool nb_side = code & 1;
        int s = fptr[f], nn = fptr[f + 1] - s;
        const int *nd = fn + s;
        double x1 = xyz[3 * nd[0]], y1 = xyz[3 * nd[0] + 1], z1 = xyz[3 * nd[0] + 2];
        for (int i = 0; i < nn - 2; ++i) {
            const double *p2 = xyz + 3 * nd[i + 1];
            const double *p3 = xyz + 3 * nd[i + 2];
            double px = p2[0] - x1, py = p2[1] - y1, pz = p2[2] - z1;
            double qx = p3[0] - x1, qy = p3[1] - y1, qz = p3[2] - z1;
            double nx = py * qz - pz * qy;
            double ny = pz * qx - px * qz;
            double nz = px * qy - py * qx;
            if (!nb_side) {
                sx = sx + centroid_part(x1, p2[0], p3[0], nx, vo);
                sy = sy + centroid_part(y1, p2[1], p3[1], ny, vo);
                sz = sz + centroid_part(z1, p2[2], p3[2], nz, vo);
            } else {
                sx = sx + centroid_part(x1, p3[0], p2[0], -nx, vo);
                sy = sy + centroid_part(y1, p3[1], p2[1], -ny, vo);
                sz = sz + centroid_part(z1, p3[2], p2[2], -nz, vo);
            }
        }
    }
    xc[c] = sx;
    yc[c] = sy;
    zc[c] = sz;
}

// find_intersection_point (:201-274): the symbolic expression evaluated left to right
__device__ inline void intersection_point(double x1, double y1, double z1, double x2, double y2, double z2,
                                          double x3, double y3, double z3, double x4, double y4, double z4,
                                          double x5, double y5, double z5, double &xj, double &yj,
                                          double &zj) {
    double num = x2 * (y3 * z4 - y4 * z3);
    num = num - x1 * (y3 * z4 - y4 * z3);
    num = num - x3 * (y2 * z4 - y4 * z2);
    num = num + x1 * (y2 * z4 - y4 * z2);
    num = num + x3 * (y1 * z4 - y4 * z1);
    num = num - x2 * (y1 * z4 - y4 * z1);
    num = num + x4 * (y2 * z3 - y3 * z2);
    num = num - x1 * (y2 * z3 - y3 * z2);
    num = num - x4 * (y1 * z3 - y3 * z1);
    num = num + x2 * (y1 * z3 - y3 * z1);
    num = num + x4 * (y1 * z2 - y2 * z1);
    num = num - x3 * (y1 * z2 - y2 * z1);
    double dz = z5 - z4, dy = y5 - y4, dx = x5 - x4;
    double den = x2 * (y3 * dz - dy * z3);
    den = den - x1 * (y3 * dz - dy * z3);
    den = den - x3 * (y2 * dz - dy * z2);
    den = den + x1 * (y2 * dz - dy * z2);
    den = den + x3 * (y1 * dz - dy * z1);
    den = den - x2 * (y1 * dz - dy * z1);
    den = den + dx * (y2 * z3 - y3 * z2);
    den = den - dx * (y1 * z3 - y3 * z1);
    den = den + dx * (y1 * z2 - y2 * z1);
    den = den + kTiny;
    double t = -(num / den);
    xj = x4 + dx * t;
    yj = y4 + dy * t;
    zj = z4 + dz * t;
}

// sweep 3: interpolation factor (:1040-1081)
__global__ void k_facint(const double *__restrict__ xyz, const int *__restrict__ fptr, const int *__restrict__ fn,
                         const int *__restrict__ own, const int *__restrict__ nei, int nif,
                         const double *__restrict__ xc, const double *__restrict__ yc,
                         const double *__restrict__ zc, double *__restrict__ facint) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int p = own[f], q = nei[f];
    const int *nd = fn + fptr[f];
    double xp = xc[p], yp = yc[p], zp = zc[p];
    double xq = xc[q], yq = yc[q], zq = zc[q];
    double xpn = xq - xp, ypn = yq - yp, zpn = zq - zp;
    double dpn = sqrt((xpn * xpn + ypn * ypn) + zpn * zpn);
    const double *a = xyz + 3 * nd[0], *b = xyz + 3 * nd[1], *c = xyz + 3 * nd[2];
    double xj, yj, zj;
    intersection_point(a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], xp, yp, zp, xq, yq, zq, xj, yj, zj);
    xpn = xj - xp;
    ypn = yj - yp;
    zpn = zj - zp;
    double djn = sqrt((xpn * xpn + ypn * ypn) + zpn * zpn);
    facint[f] = djn / dpn;
}

// widest strict-lower (upper = false) or strict-upper row
__global__ void k_tri_width(const int *__restrict__ ptr, const int *__restrict__ diag, int n, bool upper,
                            int *__restrict__ wmax) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = upper ? ptr[i + 1] - diag[i] - 1 : diag[i] - ptr[i];
    atomicMax(wmax, c);
}

// ELL of the rows in level order: slot k of packed row p at [k * n + p]
__global__ void k_tri_ell(const int *__restrict__ rows, const int *__restrict__ ptr, const int *__restrict__ diag,
                          const int *__restrict__ ja, int n, int ell, bool upper, int *__restrict__ cnt,
                          int *__restrict__ kpos, int *__restrict__ col) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = rows[p];
    const int s = upper ? diag[i] + 1 : ptr[i];
    const int c = upper ? ptr[i + 1] - s : diag[i] - s;
    cnt[p] = c;
    for (int k = 0; k < ell; ++k) {
        const size_t o = (size_t)k * n + p;
        kpos[o] = k < c ? s + k : 0;
        col[o] = k < c ? ja[s + k] : 0;
    }
}

}  // namespace

// ---------------------------------------------------------------- host side

// Copies between a caller's array and the library's device buffers.  The direction
// is taken from the pointers (unified addressing), so a caller's array may also be
// device memory: the copy then stays on the device.
void h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    FC_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

void d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    FC_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

uint64_t csr_checksum(int n, int nnz, const int *ioffset, const int *ja) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint64_t v) { h ^= v; h *= 1099511628211ull; };
    mix((uint64_t)n);
    mix((uint64_t)nnz);
    for (int i = 0; i <= n; ++i) mix((uint64_t)(uint32_t)ioffset[i]);
    for (int k = 0; k < nnz; ++k) mix((uint64_t)(uint32_t)ja[k]);
    return h;
}

static void check_err(fc_context *ctx, const DevBuf &err, const char *what) {
    int e = read_flag(ctx, err.as<int>());
    if (e == 0) return;
    std::string m = std::string(what) + ":";
    if (e & kErrRange) m += " index out of range;";
    if (e & kErrDuplicate) m += " duplicate owner/neighbour pair;";
    if (e & kErrNoDiag) m += " diag does not point at the diagonal;";
    if (e & kErrUnsorted) m += " row columns not strictly ascending;";
    if (e & kErrAsym) m += " pattern not structurally symmetric;";
    if (e & kErrNoTranspose) m += " transposed entry missing;";
    throw Error(FC_ERR_STRUCTURE, m);
}

// Level-packed ELL of one triangle (rows in level order); width 0 (not packed)
// when a row is wider than kMaxEll, which would make the padding dominate.
static void pack_triangle(fc_context *ctx, Structure &st, bool upper) {
    constexpr int kMaxEll = 16;
    hipStream_t s = ctx->stream;
    const int n = st.n;
    DevBuf wmax(sizeof(int));
    FC_HIP_CHECK(hipMemsetAsync(wmax.get(), 0, sizeof(int), s));
    k_tri_width<<<grid_for(n), kBlock, 0, s>>>(st.ioffset.as<int>(), st.diag.as<int>(), n, upper, wmax.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    const int widest = read_flag(ctx, wmax.as<int>());
    // stored width = the width the sweep kernel is compiled for (3, 4, 8), or the
    // widest row for the generic kernel; slots past a row's count are padding
    const int ell = widest <= 3 ? 3 : widest <= 4 ? 4 : widest <= 8 ? 8 : widest;
    int &width = upper ? st.ell_b : st.ell_f;
    DevBuf &cnt = upper ? st.cnt_b : st.cnt_f;
    DevBuf &kpos = upper ? st.kpos_b : st.kpos_f;
    DevBuf &col = upper ? st.col_b : st.col_f;
    width = 0;
    if (widest < 1 || widest > kMaxEll) {
        cnt.release();
        kpos.release();
        col.release();
        return;
    }
    cnt.alloc(sizeof(int) * (size_t)n);
    kpos.alloc(sizeof(int) * (size_t)ell * n);
    col.alloc(sizeof(int) * (size_t)ell * n);
    k_tri_ell<<<grid_for(n), kBlock, 0, s>>>((upper ? st.rows_b : st.rows_f).as<int>(), st.ioffset.as<int>(),
                                             st.diag.as<int>(), st.ja.as<int>(), n, ell, upper, cnt.as<int>(),
                                             kpos.as<int>(), col.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    FC_HIP_CHECK(hipStreamSynchronize(s));
    width = ell;
}

// The part-wise order of one triangle: rows in level order, stably grouped by part
// (index / kPartRows), so that each part's rows keep their level order; packed
// with the level-order ELL's width.  Also counts the entries whose column lies in
// the row's own part, and gives each row its step in the static schedule of its
// window (the 64 consecutive packed rows of its part it is computed with): one
// more than the latest step of the rows it reads inside that window, 0 if it reads
// none there.  A window's rows only read rows of lower level, which come earlier in
// the part's order, so a step only reads values of earlier steps or earlier windows.
static void pack_parts(fc_context *ctx, Structure &st, bool upper) {
    hipStream_t s = ctx->stream;
    const int n = st.n, ell = upper ? st.ell_b : st.ell_f;
    DevBuf &rows_p = upper ? st.rows_pb : st.rows_pf;
    DevBuf &cnt = upper ? st.cnt_pb : st.cnt_pf;
    DevBuf &kpos = upper ? st.kpos_pb : st.kpos_pf;
    DevBuf &col = upper ? st.col_pb : st.col_pf;
    DevBuf &meta = upper ? st.meta_pb : st.meta_pf;
    std::vector<int> lev((size_t)n), order((size_t)n);
    d2h(lev.data(), (upper ? st.rows_b : st.rows_f).get(), sizeof(int) * (size_t)n, s);
    const int nparts = (n + kPartRows - 1) / kPartRows;
    std::vector<int> start((size_t)nparts + 1, 0);
    for (int p = 0; p < n; ++p) start[lev[p] / kPartRows + 1] += 1;
    for (int t = 0; t < nparts; ++t) start[t + 1] += start[t];
    for (int p = 0; p < n; ++p) order[start[lev[p] / kPartRows]++] = lev[p];
    rows_p.alloc(sizeof(int) * (size_t)n);
    h2d(rows_p.get(), order.data(), sizeof(int) * (size_t)n, s);
    cnt.alloc(sizeof(int) * (size_t)n);
    kpos.alloc(sizeof(int) * (size_t)ell * n);
    col.alloc(sizeof(int) * (size_t)ell * n);
    k_tri_ell<<<grid_for(n), kBlock, 0, s>>>(rows_p.as<int>(), st.ioffset.as<int>(), st.diag.as<int>(),
                                             st.ja.as<int>(), n, ell, upper, cnt.as<int>(), kpos.as<int>(),
                                             col.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    // entries inside their row's part
    std::vector<int> c((size_t)n), cl((size_t)ell * n);
    d2h(c.data(), cnt.get(), sizeof(int) * (size_t)n, s);
    d2h(cl.data(), col.get(), sizeof(int) * (size_t)ell * n, s);
    long long inside = 0, total = 0;
    for (int p = 0; p < n; ++p) {
        const int part = order[p] / kPartRows;
        for (int k = 0; k < c[p]; ++k) {
            inside += cl[(size_t)k * n + p] / kPartRows == part;
            total += 1;
        }
    }
    (upper ? st.int_frac_b : st.int_frac_f) = total ? (double)inside / (double)total : 0.0;
    // static schedule: meta = count | step << 8 | steps of the window << 16
    std::vector<int> pos((size_t)n), step((size_t)n, 0), m((size_t)n);
    for (int p = 0; p < n; ++p) pos[order[p]] = p;
    for (int w0 = 0; w0 < n; w0 += 64) {
        const int part = w0 / kPartRows;
        const int w1 = std::min({n, w0 + 64, (part + 1) * kPartRows});
        int top = 0;
        for (int p = w0; p < w1; ++p) {
            int sp = 0;
            for (int k = 0; k < c[p]; ++k) {
                const int j = cl[(size_t)k * n + p];
                if (j / kPartRows != part) continue;
                const int pj = pos[j];
                FC_REQUIRE(pj < p, FC_ERR_STRUCTURE, "part order: an input after its row");
                if (pj >= w0) sp = std::max(sp, step[pj] + 1);
            }
            step[p] = sp;
            top = std::max(top, sp);
        }
        for (int p = w0; p < w1; ++p) m[p] = c[p] | step[p] << 8 | (top + 1) << 16;
    }
    meta.alloc(sizeof(int) * (size_t)n);
    h2d(meta.get(), m.data(), sizeof(int) * (size_t)n, s);
}

void analyse_levels(fc_context *ctx, Structure &st) {
    levels_one(ctx, st, false, st.rows_f, st.lvl_f, st.nlev_f, st.max_level_f);
    levels_one(ctx, st, true, st.rows_b, st.lvl_b, st.nlev_b, st.max_level_b);
    st.lvl_f_d.alloc(sizeof(int) * st.lvl_f.size());
    h2d(st.lvl_f_d.get(), st.lvl_f.data(), sizeof(int) * st.lvl_f.size(), ctx->stream);
    st.lvl_b_d.alloc(sizeof(int) * st.lvl_b.size());
    h2d(st.lvl_b_d.get(), st.lvl_b.data(), sizeof(int) * st.lvl_b.size(), ctx->stream);
    pack_triangle(ctx, st, false);
    pack_triangle(ctx, st, true);
    st.parts = false;
    if (st.ell_f > 0 && st.ell_b > 0) {
        pack_parts(ctx, st, false);
        pack_parts(ctx, st, true);
        st.parts = true;
    }
}

void build_tpos(fc_context *ctx, Structure &st) {
    if (st.has_tpos) return;
    DevBuf err(sizeof(int));
    FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), ctx->stream));
    st.tpos.alloc(sizeof(int) * (size_t)st.nnz);
    k_tpos<<<grid_for(st.n), kBlock, 0, ctx->stream>>>(st.ioffset.as<int>(), st.ja.as<int>(), st.diag.as<int>(),
                                                       st.n, st.tpos.as<int>(), err.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    check_err(ctx, err, "DILU transposed positions");
    st.has_tpos = true;
}

void build_structure_from_host_csr(fc_context *ctx, Structure &st, int n, int nnz, const int *ioffset1,
                                   const int *ja1, const int *diag1) {
    FC_REQUIRE(n > 0 && nnz >= n, FC_ERR_ARGUMENT, "n/nnz");
    FC_REQUIRE(ioffset1[0] == 1 && ioffset1[n] == nnz + 1, FC_ERR_STRUCTURE, "ioffset must run from 1 to nnz+1");
    hipStream_t s = ctx->stream;
    st.n = n;
    st.nnz = nnz;
    st.ioffset.alloc(sizeof(int) * (n + 1));
    st.ja.alloc(sizeof(int) * (size_t)nnz);
    st.diag.alloc(sizeof(int) * n);
    DevBuf tmp(sizeof(int) * (size_t)std::max(nnz, n + 1));
    h2d(tmp.get(), ioffset1, sizeof(int) * (n + 1), s);
    k_sub_one<<<grid_for(n + 1), kBlock, 0, s>>>(tmp.as<int>(), st.ioffset.as<int>(), n + 1);
    FC_HIP_CHECK(hipStreamSynchronize(s));
    h2d(tmp.get(), ja1, sizeof(int) * (size_t)nnz, s);
    k_sub_one<<<grid_for(nnz), kBlock, 0, s>>>(tmp.as<int>(), st.ja.as<int>(), nnz);
    FC_HIP_CHECK(hipStreamSynchronize(s));
    h2d(tmp.get(), diag1, sizeof(int) * n, s);
    k_sub_one<<<grid_for(n), kBlock, 0, s>>>(tmp.as<int>(), st.diag.as<int>(), n);
    FC_HIP_CHECK(hipGetLastError());
    DevBuf err(sizeof(int));
    FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), s));
    k_validate_csr<<<grid_for(n), kBlock, 0, s>>>(st.ioffset.as<int>(), st.ja.as<int>(), st.diag.as<int>(), n, nnz,
                                                  err.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    check_err(ctx, err, "CSR");
    st.levels_ready = false;
    st.has_tpos = false;
}

void build_structure_from_faces_dev(fc_context *ctx, Structure &st, int n, int nif, const int *own0,
                                    const int *nei0, int *ij_slot, int *ji_slot) {
    hipStream_t s = ctx->stream;
    const int nnz = 2 * nif + n;
    st.n = n;
    st.nnz = nnz;
    st.ioffset.alloc(sizeof(int) * (n + 1));
    st.ja.alloc(sizeof(int) * (size_t)nnz);
    st.diag.alloc(sizeof(int) * n);
    DevBuf cnt(sizeof(int) * (n + 1)), cursor(sizeof(int) * (n + 1)), err(sizeof(int));
    FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), s));
    k_init_counts<<<grid_for(n + 1), kBlock, 0, s>>>(cnt.as<int>(), n);
    k_row_counts<<<grid_for(nif), kBlock, 0, s>>>(own0, nei0, nif, n, cnt.as<int>(), err.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    check_err(ctx, err, "owner/neighbour");
    exclusive_scan(ctx, cnt.as<int>(), st.ioffset.as<int>(), n + 1);
    k_place_diag<<<grid_for(n), kBlock, 0, s>>>(st.ioffset.as<int>(), cursor.as<int>(), st.ja.as<int>(), n);
    k_place_faces<<<grid_for(nif), kBlock, 0, s>>>(own0, nei0, nif, cursor.as<int>(), st.ja.as<int>());
    k_sort_rows<<<grid_for(n), kBlock, 0, s>>>(st.ioffset.as<int>(), st.ja.as<int>(), st.diag.as<int>(), n,
                                              err.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    check_err(ctx, err, "CSR build");
    k_slot_maps<<<grid_for(nif), kBlock, 0, s>>>(own0, nei0, nif, st.ioffset.as<int>(), st.ja.as<int>(), ij_slot,
                                                ji_slot);
    FC_HIP_CHECK(hipGetLastError());
    st.levels_ready = false;
    st.has_tpos = false;
}

void upload_mesh(fc_context *ctx, MeshDev &m, int nnodes, const double *xyz, int nfaces, int nif,
                 const int *face_ptr, const int *face_nodes1, const int *owner1, const int *neighbour1,
                 int ncells) {
    FC_REQUIRE(nnodes > 0 && nfaces > 0 && nif >= 0 && nif <= nfaces && ncells > 0, FC_ERR_ARGUMENT,
               "mesh sizes");
    FC_REQUIRE(face_ptr[0] == 0, FC_ERR_ARGUMENT, "face_ptr[0] must be 0");
    hipStream_t s = ctx->stream;
    m.nnodes = nnodes;
    m.nfaces = nfaces;
    m.nif = nif;
    m.n = ncells;
    const int ntot_nodes = face_ptr[nfaces];
    m.xyz.alloc(sizeof(double) * 3 * (size_t)nnodes);
    h2d(m.xyz.get(), xyz, sizeof(double) * 3 * (size_t)nnodes, s);
    m.face_ptr.alloc(sizeof(int) * (nfaces + 1));
    h2d(m.face_ptr.get(), face_ptr, sizeof(int) * (nfaces + 1), s);
    DevBuf tmp(sizeof(int) * (size_t)std::max(ntot_nodes, nfaces));
    m.face_nodes.alloc(sizeof(int) * (size_t)ntot_nodes);
    h2d(tmp.get(), face_nodes1, sizeof(int) * (size_t)ntot_nodes, s);
    k_sub_one<<<grid_for(ntot_nodes), kBlock, 0, s>>>(tmp.as<int>(), m.face_nodes.as<int>(), ntot_nodes);
    FC_HIP_CHECK(hipStreamSynchronize(s));
    m.owner.alloc(sizeof(int) * nfaces);
    h2d(tmp.get(), owner1, sizeof(int) * nfaces, s);
    k_sub_one<<<grid_for(nfaces), kBlock, 0, s>>>(tmp.as<int>(), m.owner.as<int>(), nfaces);
    FC_HIP_CHECK(hipStreamSynchronize(s));
    m.neighbour.alloc(sizeof(int) * std::max(nif, 1));
    if (nif) {
        h2d(tmp.get(), neighbour1, sizeof(int) * nif, s);
        k_sub_one<<<grid_for(nif), kBlock, 0, s>>>(tmp.as<int>(), m.neighbour.as<int>(), nif);
    }
    FC_HIP_CHECK(hipGetLastError());
    DevBuf err(sizeof(int));
    FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), s));
    k_face_node_check<<<grid_for(nfaces), kBlock, 0, s>>>(m.face_ptr.as<int>(), m.face_nodes.as<int>(), nfaces,
                                                          nnodes, err.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    check_err(ctx, err, "faces");
    // cell -> faces, ascending face order
    DevBuf cnt(sizeof(int) * (ncells + 1)), cursor(sizeof(int) * (ncells + 1));
    FC_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(int) * (ncells + 1), s));
    k_cell_face_count<<<grid_for(nfaces), kBlock, 0, s>>>(m.owner.as<int>(), m.neighbour.as<int>(), nfaces, nif,
                                                          ncells, cnt.as<int>(), err.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    check_err(ctx, err, "owner/neighbour");
    m.cf_ptr.alloc(sizeof(int) * (ncells + 1));
    exclusive_scan(ctx, cnt.as<int>(), m.cf_ptr.as<int>(), ncells + 1);
    int total = 0;
    d2h(&total, m.cf_ptr.as<int>() + ncells, sizeof(int), s);
    m.cf_face.alloc(sizeof(int) * (size_t)std::max(total, 1));
    k_copy_int<<<grid_for(ncells), kBlock, 0, s>>>(m.cf_ptr.as<int>(), cursor.as<int>(), ncells);
    k_cell_face_fill<<<grid_for(nfaces), kBlock, 0, s>>>(m.owner.as<int>(), m.neighbour.as<int>(), nfaces, nif,
                                                         cursor.as<int>(), m.cf_face.as<int>());
    k_sort_segments<<<grid_for(ncells), kBlock, 0, s>>>(m.cf_ptr.as<int>(), m.cf_face.as<int>(), ncells);
    FC_HIP_CHECK(hipGetLastError());
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

void compute_geometry(fc_context *ctx, MeshDev &m) {
    hipStream_t s = ctx->stream;
    GeometryDev &g = m.g;
    const size_t F = sizeof(double) * m.nfaces, C = sizeof(double) * m.n;
    g.arx.ensure(F); g.ary.ensure(F); g.arz.ensure(F);
    g.xf.ensure(F); g.yf.ensure(F); g.zf.ensure(F);
    g.vol.ensure(C); g.xc.ensure(C); g.yc.ensure(C); g.zc.ensure(C);
    g.facint.ensure(sizeof(double) * std::max(m.nif, 1));
    k_face_area_centre<<<grid_for(m.nfaces), kBlock, 0, s>>>(m.xyz.as<double>(), m.face_ptr.as<int>(),
                                                             m.face_nodes.as<int>(), m.nfaces, g.arx.as<double>(),
                                                             g.ary.as<double>(), g.arz.as<double>(),
                                                             g.xf.as<double>(), g.yf.as<double>(), g.zf.as<double>());
    k_cell_volume<<<grid_for(m.n), kBlock, 0, s>>>(m.xyz.as<double>(), m.face_ptr.as<int>(), m.face_nodes.as<int>(),
                                                   m.cf_ptr.as<int>(), m.cf_face.as<int>(), m.n, g.vol.as<double>());
    k_cell_centre<<<grid_for(m.n), kBlock, 0, s>>>(m.xyz.as<double>(), m.face_ptr.as<int>(), m.face_nodes.as<int>(),
                                                   m.cf_ptr.as<int>(), m.cf_face.as<int>(), m.n, g.vol.as<double>(),
                                                   g.xc.as<double>(), g.yc.as<double>(), g.zc.as<double>());
    if (m.nif)
        k_facint<<<grid_for(m.nif), kBlock, 0, s>>>(m.xyz.as<double>(), m.face_ptr.as<int>(), m.face_nodes.as<int>(),
                                                    m.owner.as<int>(), m.neighbour.as<int>(), m.nif,
                                                    g.xc.as<double>(), g.yc.as<double>(), g.zc.as<double>(),
                                                    g.facint.as<double>());
    FC_HIP_CHECK(hipGetLastError());
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

// exported for the C ABI layer: 0-based device ints -> 1-based, in place
void to_one_based(fc_context *ctx, int *d, long long n) {
    if (n <= 0) return;
    k_add_one<<<grid_for(n), kBlock, 0, ctx->stream>>>(d, n);
    FC_HIP_CHECK(hipGetLastError());
}

void fill_int(fc_context *ctx, int *d, long long n, int v) {
    if (n <= 0) return;
    k_fill_int<<<grid_for(n), kBlock, 0, ctx->stream>>>(d, n, v);
    FC_HIP_CHECK(hipGetLastError());
}

void sub_one_dev(fc_context *ctx, const int *in, int *out, long long n) {
    if (n <= 0) return;
    k_sub_one<<<grid_for(n), kBlock, 0, ctx->stream>>>(in, out, n);
    FC_HIP_CHECK(hipGetLastError());
}

void scan_dev(fc_context *ctx, const int *in, int *out, int n) { exclusive_scan(ctx, in, out, n); }

void sort_segments_dev(fc_context *ctx, const int *ptr, int *v, int nseg) {
    k_sort_segments<<<grid_for(nseg), kBlock, 0, ctx->stream>>>(ptr, v, nseg);
    FC_HIP_CHECK(hipGetLastError());
}

}  // namespace fcap
