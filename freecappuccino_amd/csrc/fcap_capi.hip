// C ABI of libfcap (declared in include/fcap.h).
//
// Every entry point converts exceptions to an fc_status and records the
// message in the context; nothing aborts.  Host-pointer entry points upload
// their inputs, compute on the GPU and copy the outputs back.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "fcap_coarse.h"
#include "fcap_internal.h"
#include "fcap_solvers.h"
#include "fcap_system.h"

using namespace fcap;

namespace {

template <class F>
fc_status guarded(fc_context *ctx, F &&f) {
    if (!ctx) return FC_ERR_ARGUMENT;
    try {
        FC_HIP_CHECK(hipSetDevice(ctx->device));
        return f();
    } catch (const Error &e) {
        ctx->last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        return FC_ERR_ALLOC;
    }
}

// 1-based host ints -> 0-based device ints
void upload_idx(fc_context *ctx, DevBuf &dst, const int *src1, long long count) {
    dst.alloc(sizeof(int) * (size_t)std::max<long long>(count, 1));
    if (count <= 0) return;
    DevBuf tmp(sizeof(int) * (size_t)count);
    h2d(tmp.get(), src1, sizeof(int) * (size_t)count, ctx->stream);
    sub_one_dev(ctx, tmp.as<int>(), dst.as<int>(), count);
    FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void upload_dbl(fc_context *ctx, DevBuf &dst, const double *src, long long count) {
    dst.alloc(sizeof(double) * (size_t)std::max<long long>(count, 1));
    if (count > 0 && src) h2d(dst.get(), src, sizeof(double) * (size_t)count, ctx->stream);
}

// 0-based device ints -> 1-based host ints
void download_idx(fc_context *ctx, int *dst1, const DevBuf &src, long long count) {
    if (count <= 0) return;
    DevBuf tmp(sizeof(int) * (size_t)count);
    FC_HIP_CHECK(hipMemcpyAsync(tmp.get(), src.get(), sizeof(int) * (size_t)count, hipMemcpyDeviceToDevice,
                                ctx->stream));
    to_one_based(ctx, tmp.as<int>(), count);
    d2h(dst1, tmp.get(), sizeof(int) * (size_t)count, ctx->stream);
}

void download_dbl(fc_context *ctx, double *dst, const void *src, long long count) {
    if (count > 0 && dst) d2h(dst, src, sizeof(double) * (size_t)count, ctx->stream);
}

Workspace &workspace(fc_context *ctx, int n, int nnz) {
    if (!ctx->ws) ctx->ws.reset(new Workspace());
    ctx->ws->ensure(n, nnz);
    return *ctx->ws;
}

// Cached device structure of a host CSR (n, nnz, ioffset, ja, diag checksum).
Structure &cached_structure(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja, const int *diag) {
    FC_REQUIRE(n > 0 && nnz > 0 && ioffset && ja && diag, FC_ERR_ARGUMENT, "CSR arguments");
    uint64_t key = csr_checksum(n, nnz, ioffset, ja);
    for (int i = 0; i < n; ++i) key = (key ^ (uint64_t)(uint32_t)diag[i]) * 1099511628211ull;
    if (!ctx->cache || ctx->cache->key != key || ctx->cache->n != n || ctx->cache->nnz != nnz) {
        ctx->cache.reset(new Structure());
        build_structure_from_host_csr(ctx, *ctx->cache, n, nnz, ioffset, ja, diag);
        ctx->cache->key = key;
    }
    return *ctx->cache;
}

void fortran_e10_3(char *buf, size_t len, double v) {
    // 1PE10.3: d.dddE+xx; three-digit exponents drop the 'E' (d.ddd+xxx)
    char tmp[64];
    std::snprintf(tmp, sizeof tmp, "%.3E", v);
    const char *e = std::strchr(tmp, 'E');
    if (e && std::strlen(e + 2) >= 3) {
        std::string m(tmp, e - tmp);
        std::snprintf(buf, len, "%10s", (m + std::string(e + 1)).c_str());
    } else {
        std::snprintf(buf, len, "%10s", tmp);
    }
}

void report_line(fc_context *ctx, const char *head, const fc_solve_info &info) {
    if (!ctx->report) return;
    char r0[32], rl[32];
    fortran_e10_3(r0, sizeof r0, info.res0);
    fortran_e10_3(rl, sizeof rl, info.status == FC_EARLY_EXIT ? info.res0 : info.resl);
    std::printf("%sSolving for %s, Initial residual = %s, Final residual = %s, No Iterations %d\n", head,
                ctx->report_name.c_str(), r0, rl, info.iters);
    std::fflush(stdout);
}

struct OcHost {
    DevBuf ijl, ijr, al, ar, co_ptr, co_idx;
};

OcView make_oc(fc_context *ctx, OcHost &h, int n, int noc, const int *ijl, const int *ijr, const double *al,
               const double *ar) {
    OcView v;
    if (noc <= 0) return v;
    FC_REQUIRE(ijl && ijr && al && ar, FC_ERR_ARGUMENT, "O-C arrays");
    OcDev od;
    od.noc = noc;
    upload_idx(ctx, od.ijl, ijl, noc);
    upload_idx(ctx, od.ijr, ijr, noc);
    build_oc_lists(ctx, od, n);
    h.ijl = std::move(od.ijl);
    h.ijr = std::move(od.ijr);
    h.co_ptr = std::move(od.co_ptr);
    h.co_idx = std::move(od.co_idx);
    upload_dbl(ctx, h.al, al, noc);
    upload_dbl(ctx, h.ar, ar, noc);
    v.noc = noc;
    v.ijl = h.ijl.as<int>();
    v.ijr = h.ijr.as<int>();
    v.co_ptr = h.co_ptr.as<int>();
    v.co_idx = h.co_idx.as<int>();
    v.al = h.al.as<double>();
    v.ar = h.ar.as<double>();
    return v;
}

fc_status host_solve(fc_context *ctx, int kind, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
                     const double *a, const double *su, double *fi, double *res, int noc, const int *ijl,
                     const int *ijr, const double *al, const double *ar, double sor, int nsw,
                     fc_solve_info *info) {
    FC_REQUIRE(a && su && fi, FC_ERR_ARGUMENT, "a/su/fi");
    Structure &st = cached_structure(ctx, n, nnz, ioffset, ja, diag);
    if (kind != 2) ensure_levels(ctx, st);
    Workspace &ws = workspace(ctx, n, nnz);
    DevBuf da, dsu;
    upload_dbl(ctx, da, a, nnz);
    upload_dbl(ctx, dsu, su, n);
    h2d(ws.x.get(), fi, sizeof(double) * n, ctx->stream);
    OcHost och;
    SolveArgs args;
    args.st = &st;
    args.a = da.as<double>();
    args.su = dsu.as<double>();
    args.fi = ws.x.as<double>();
    args.oc = make_oc(ctx, och, n, ctx->variant == FC_VARIANT_TESTS ? 0 : noc, ijl, ijr, al, ar);
    args.sor = sor;
    args.nsw = nsw;
    args.variant = ctx->variant;
    fc_solve_info local;
    std::memset(&local, 0, sizeof local);
    fc_status s;
    if (kind == 1) s = solve_pcg64(ctx, args, true, &local);
    else if (kind == 2) s = solve_pcg64(ctx, args, false, &local);
    else s = solve_bicgstab64(ctx, args, &local);
    local.rel_l2 = true_rel_l2(ctx, args);
    d2h(fi, ws.x.get(), sizeof(double) * n, ctx->stream);
    if (res) d2h(res, ws.r.get(), sizeof(double) * n, ctx->stream);
    const char *head = kind == 1 ? "  PCG(IC0):  " : (kind == 2 ? "PCG(Jacobi):  " : "  BiCGStab(ILU(0)):  ");
    if (local.status == FC_EARLY_EXIT && kind == 2) head = "  PCG(Jacobi):  ";
    report_line(ctx, head, local);
    if (info) *info = local;
    return s;
}

}  // namespace

// =========================================================== context

extern "C" fc_status fc_init(fc_context **out, int device) {
    if (!out) return FC_ERR_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return FC_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return FC_ERR_ARGUMENT;
    fc_context *ctx = new fc_context();
    ctx->device = device;
    fc_status s = guarded(ctx, [&]() {
        FC_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->own_stream = true;
        FC_HIP_CHECK(hipHostMalloc((void **)&ctx->host_flags, sizeof(int) * 16, hipHostMallocDefault));
        hipDeviceProp_t prop;
        FC_HIP_CHECK(hipGetDeviceProperties(&prop, device));
        ctx->num_cus = prop.multiProcessorCount;
        int khz = 0;
        FC_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
        if (khz > 0) ctx->wall_clock_khz = khz;
        noalloc_init(device);   // this device's address of the allocation-free scratch
        return FC_OK;
    });
    if (s != FC_OK) {
        fc_finalize(ctx);
        return s;
    }
    *out = ctx;
    return FC_OK;
}

extern "C" fc_status fc_finalize(fc_context *ctx) {
    if (!ctx) return FC_ERR_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->cache.reset();
    ctx->ws.reset();
    ctx->timer.destroy();
    coarse_release(ctx);
    if (ctx->host_flags) (void)hipHostFree(ctx->host_flags);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FC_OK;
}

extern "C" fc_status fc_set_stream(fc_context *ctx, void *stream) {
    return guarded(ctx, [&]() {
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (ctx->own_stream) FC_HIP_CHECK(hipStreamDestroy(ctx->stream));
        if (stream) {
            ctx->stream = (hipStream_t)stream;
            ctx->own_stream = false;
        } else {
            FC_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            ctx->own_stream = true;
        }
        return FC_OK;
    });
}

extern "C" const char *fc_last_error(const fc_context *ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

extern "C" fc_status fc_set_report(fc_context *ctx, int on, const char *name) {
    if (!ctx) return FC_ERR_ARGUMENT;
    ctx->report = on;
    if (name) ctx->report_name = name;
    return FC_OK;
}

extern "C" fc_status fc_set_variant(fc_context *ctx, int variant) {
    if (!ctx || (variant != FC_VARIANT_SRC && variant != FC_VARIANT_TESTS)) return FC_ERR_ARGUMENT;
    ctx->variant = variant;
    return FC_OK;
}

extern "C" fc_status fc_set_inner_tolerance(fc_context *ctx, double eta) {
    if (!ctx || !(eta >= 0.0 && eta < 1.0)) return FC_ERR_ARGUMENT;
    ctx->inner_eta = eta;
    return FC_OK;
}

extern "C" fc_status fc_set_fcg_window(fc_context *ctx, int m) {
    if (!ctx || m < 0 || m > kFcgM) return FC_ERR_ARGUMENT;
    ctx->fcg_m = m;
    return FC_OK;
}

extern "C" fc_status fc_set_fcg_stagnation(fc_context *ctx, int window, double factor) {
    if (!ctx || window < 0 || !(factor >= 0.0 && factor <= 1.0)) return FC_ERR_ARGUMENT;
    ctx->fcg_stag_window = window;
    ctx->fcg_stag_factor = factor;
    return FC_OK;
}

extern "C" fc_status fc_set_fcg_smoothing(fc_context *ctx, int steps, double omega) {
    if (!ctx || steps < 0 || steps > 8 || !(omega > 0.0 && omega < 2.0)) return FC_ERR_ARGUMENT;
    ctx->gnn_smooth = steps;
    ctx->gnn_omega = (float)omega;
    return FC_OK;
}

extern "C" fc_status fc_set_fcg_coarse(fc_context *ctx, int cells_per_aggregate) {
    if (!ctx || cells_per_aggregate < 0 || cells_per_aggregate > (1 << 24) || cells_per_aggregate % 256 != 0)
        return FC_ERR_ARGUMENT;
    ctx->fcg_coarse = cells_per_aggregate;
    return FC_OK;
}

extern "C" fc_status fc_set_fcg_network_period(fc_context *ctx, int period) {
    if (!ctx || period < 0 || period > 4) return FC_ERR_ARGUMENT;
    ctx->fcg_net_period = period;
    return FC_OK;
}

extern "C" fc_status fc_set_network_layout(fc_context *ctx, int layout) {
    if (!ctx || (layout != FC_NET_CLUSTERS && layout != FC_NET_CELL_ORDER)) return FC_ERR_ARGUMENT;
    ctx->gnn_layout = layout;
    return FC_OK;
}

extern "C" fc_status fc_set_sweep_mode(fc_context *ctx, int mode) {
    if (!ctx || (mode != FC_SWEEP_LEVELS && mode != FC_SWEEP_DEPENDENT && mode != FC_SWEEP_PARTS))
        return FC_ERR_ARGUMENT;
    ctx->sweep_mode = mode;
    return FC_OK;
}

extern "C" fc_status fc_set_sweep_waves(fc_context *ctx, int waves) {
    if (!ctx || waves < 0 || waves > 65536) return FC_ERR_ARGUMENT;
    ctx->part_waves = waves;
    return FC_OK;
}

extern "C" fc_status fc_set_sweep_wait_limit(fc_context *ctx, int microseconds) {
    if (!ctx || microseconds < 0) return FC_ERR_ARGUMENT;
    ctx->sf_wait_us = microseconds;
    return FC_OK;
}

extern "C" fc_status fc_set_reduction_order(fc_context *ctx, int order) {
    if (!ctx || (order != FC_REDUCE_TREE && order != FC_REDUCE_SEQUENTIAL)) return FC_ERR_ARGUMENT;
    ctx->reduce_seq = order == FC_REDUCE_SEQUENTIAL;
    return FC_OK;
}

extern "C" fc_status fc_set_graphs(fc_context *ctx, int on) {
    if (!ctx) return FC_ERR_ARGUMENT;
    ctx->use_graphs = on ? 1 : 0;
    return FC_OK;
}

// ======================================================= mesh and CSR

extern "C" fc_status fc_mesh_geometry(fc_context *ctx, int numNodes, const double *xyz, int numFaces,
                                      int numInnerFaces, const int *face_ptr, const int *face_nodes, const int *owner,
                                      const int *neighbour, int numCells, double *arx, double *ary, double *arz,
                                      double *xf, double *yf, double *zf, double *vol, double *xc, double *yc,
                                      double *zc, double *facint) {
    return guarded(ctx, [&]() {
        FC_REQUIRE(xyz && face_ptr && face_nodes && owner && (neighbour || numInnerFaces == 0), FC_ERR_ARGUMENT,
                   "mesh arrays");
        MeshDev m;
        upload_mesh(ctx, m, numNodes, xyz, numFaces, numInnerFaces, face_ptr, face_nodes, owner, neighbour, numCells);
        compute_geometry(ctx, m);
        download_dbl(ctx, arx, m.g.arx.get(), numFaces);
        download_dbl(ctx, ary, m.g.ary.get(), numFaces);
        download_dbl(ctx, arz, m.g.arz.get(), numFaces);
        download_dbl(ctx, xf, m.g.xf.get(), numFaces);
        download_dbl(ctx, yf, m.g.yf.get(), numFaces);
        download_dbl(ctx, zf, m.g.zf.get(), numFaces);
        download_dbl(ctx, vol, m.g.vol.get(), numCells);
        download_dbl(ctx, xc, m.g.xc.get(), numCells);
        download_dbl(ctx, yc, m.g.yc.get(), numCells);
        download_dbl(ctx, zc, m.g.zc.get(), numCells);
        download_dbl(ctx, facint, m.g.facint.get(), numInnerFaces);
        return FC_OK;
    });
}

extern "C" fc_status fc_csr_build(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                                  const int *neighbour, int *ioffset, int *ja, int *diag, int *ij_slot,
                                  int *ji_slot) {
    return guarded(ctx, [&]() {
        FC_REQUIRE(numCells > 0 && numInnerFaces >= 0, FC_ERR_ARGUMENT, "sizes");
        FC_REQUIRE(ioffset && ja && diag, FC_ERR_ARGUMENT, "output arrays");
        DevBuf own, nei, ij, ji;
        upload_idx(ctx, own, owner, numInnerFaces);
        upload_idx(ctx, nei, neighbour, numInnerFaces);
        ij.alloc(sizeof(int) * std::max(numInnerFaces, 1));
        ji.alloc(sizeof(int) * std::max(numInnerFaces, 1));
        Structure st;
        build_structure_from_faces_dev(ctx, st, numCells, numInnerFaces, own.as<int>(), nei.as<int>(), ij.as<int>(),
                                       ji.as<int>());
        download_idx(ctx, ioffset, st.ioffset, numCells + 1);
        download_idx(ctx, ja, st.ja, st.nnz);
        download_idx(ctx, diag, st.diag, numCells);
        if (ij_slot) download_idx(ctx, ij_slot, ij, numInnerFaces);
        if (ji_slot) download_idx(ctx, ji_slot, ji, numInnerFaces);
        return FC_OK;
    });
}

// ============================================================ assembly

namespace {

struct FaceSet {
    DevBuf own, nei, ij, ji, kface, kord;
    bool needs_order = false;
};

// Upload owner/neighbour and slot maps of the inner faces and derive kface.
void upload_faces(fc_context *ctx, FaceSet &fs, const Structure &st, int nif, const int *owner, const int *neighbour,
                  const int *ij_slot, const int *ji_slot) {
    upload_idx(ctx, fs.own, owner, nif);
    upload_idx(ctx, fs.nei, neighbour, nif);
    upload_idx(ctx, fs.ij, ij_slot, nif);
    upload_idx(ctx, fs.ji, ji_slot, nif);
    validate_faces(ctx, st, nif, fs.own.as<int>(), fs.nei.as<int>(), fs.ij.as<int>(), fs.ji.as<int>());
    build_kface(ctx, st, nif, fs.ij.as<int>(), fs.ji.as<int>(), fs.kface, fs.kord, fs.needs_order);
}

}  // namespace

extern "C" fc_status fc_laplacian(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                                  const int *neighbour, const double *arx, const double *ary, const double *arz,
                                  const double *facint, const double *xc, const double *yc, const double *zc,
                                  const double *xf, const double *yf, const double *zf, const double *mu,
                                  const double *phi, int numTotal, int nwal, int iWallFacesStart, int iWallStart,
                                  int noc, int iOCFacesStart, const int *ijl, const int *ijr, const double *foc,
                                  const int *ioffset, const int *ja, const int *diag, const int *ij_slot,
                                  const int *ji_slot, int nnz, double *a, double *su, double *al, double *ar) {
    return guarded(ctx, [&]() {
        const int n = numCells, nif = numInnerFaces;
        FC_REQUIRE(n > 0 && nif >= 0 && nnz == 2 * nif + n, FC_ERR_ARGUMENT, "nnz must be 2*numInnerFaces+numCells");
        FC_REQUIRE(a && su && mu && phi, FC_ERR_ARGUMENT, "a/su/mu/phi");
        const int nfaces_needed = std::max({nif, nwal > 0 ? iWallFacesStart + nwal : 0,
                                            noc > 0 ? iOCFacesStart + noc : 0});
        FC_REQUIRE(numTotal >= n && (nwal <= 0 || iWallStart + nwal <= numTotal), FC_ERR_ARGUMENT, "numTotal");
        Structure &st = cached_structure(ctx, n, nnz, ioffset, ja, diag);
        FaceSet fs;
        upload_faces(ctx, fs, st, nif, owner, neighbour, ij_slot, ji_slot);
        // geometry: face arrays cover every face the loops touch
        DevBuf dx[6], dc[3], dfi, dmu, dphi, dsu, da;
        const double *fa[6] = {arx, ary, arz, xf, yf, zf};
        for (int k = 0; k < 6; ++k) upload_dbl(ctx, dx[k], fa[k], nfaces_needed);
        const double *ca[3] = {xc, yc, zc};
        for (int k = 0; k < 3; ++k) upload_dbl(ctx, dc[k], ca[k], n);
        upload_dbl(ctx, dfi, facint, nif);
        upload_dbl(ctx, dmu, mu, n);
        upload_dbl(ctx, dphi, phi, numTotal);
        upload_dbl(ctx, dsu, su, n);
        da.alloc(sizeof(double) * nnz);
        FC_HIP_CHECK(hipMemsetAsync(da.get(), 0, sizeof(double) * nnz, ctx->stream));   // a = 0 (fvm_laplacian.f90:33)
        WallDev w;
        if (nwal > 0) {
            DevBuf owner_all;
            upload_idx(ctx, owner_all, owner, iWallFacesStart + nwal);
            build_wall_lists(ctx, w, owner_all.as<int>(), n, nwal, iWallFacesStart, iWallStart);
        }
        OcDev oc;
        oc.noc = std::max(noc, 0);
        oc.ioc_faces_start = iOCFacesStart;
        if (noc > 0) {
            FC_REQUIRE(ijl && ijr && foc && al && ar, FC_ERR_ARGUMENT, "O-C arrays");
            upload_idx(ctx, oc.ijl, ijl, noc);
            upload_idx(ctx, oc.ijr, ijr, noc);
            upload_dbl(ctx, oc.foc, foc, noc);
            oc.al.alloc(sizeof(double) * noc);
            oc.ar.alloc(sizeof(double) * noc);
            build_oc_lists(ctx, oc, n);
        }
        laplacian_dev(ctx, st, nif, fs.own.as<int>(), fs.nei.as<int>(), fs.ij.as<int>(), fs.ji.as<int>(),
                      dx[0].as<double>(), dx[1].as<double>(), dx[2].as<double>(), dfi.as<double>(),
                      dc[0].as<double>(), dc[1].as<double>(), dc[2].as<double>(), dx[3].as<double>(),
                      dx[4].as<double>(), dx[5].as<double>(), dmu.as<double>(), dphi.as<double>(), w, oc,
                      fs.needs_order ? fs.kord.as<int>() : nullptr, da.as<double>(), dsu.as<double>());
        download_dbl(ctx, a, da.get(), nnz);
        download_dbl(ctx, su, dsu.get(), n);
        if (noc > 0) {
            download_dbl(ctx, al, oc.al.get(), noc);
            download_dbl(ctx, ar, oc.ar.get(), noc);
        }
        return FC_OK;
    });
}

extern "C" fc_status fc_convdiff(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                                 const int *neighbour, const double *arx, const double *ary, const double *arz,
                                 const double *facint, const double *xc, const double *yc, const double *zc,
                                 const double *xf, const double *yf, const double *zf, const double *vis,
                                 const double *flux, int nwal, int iWallFacesStart, const int *ioffset,
                                 const int *ja, const int *diag, const int *ij_slot, const int *ji_slot, int nnz,
                                 double urf, double *a) {
    return guarded(ctx, [&]() {
        const int n = numCells, nif = numInnerFaces;
        FC_REQUIRE(n > 0 && nif >= 0 && nnz == 2 * nif + n && urf > 0.0, FC_ERR_ARGUMENT, "sizes/urf");
        FC_REQUIRE(a && vis && flux, FC_ERR_ARGUMENT, "a/vis/flux");
        Structure &st = cached_structure(ctx, n, nnz, ioffset, ja, diag);
        FaceSet fs;
        upload_faces(ctx, fs, st, nif, owner, neighbour, ij_slot, ji_slot);
        const int nfaces_needed = std::max(nif, nwal > 0 ? iWallFacesStart + nwal : 0);
        DevBuf dx[6], dc[3], dfi, dvis, dflux, da;
        const double *fa[6] = {arx, ary, arz, xf, yf, zf};
        for (int k = 0; k < 6; ++k) upload_dbl(ctx, dx[k], fa[k], nfaces_needed);
        const double *ca[3] = {xc, yc, zc};
        for (int k = 0; k < 3; ++k) upload_dbl(ctx, dc[k], ca[k], n);
        upload_dbl(ctx, dfi, facint, nif);
        upload_dbl(ctx, dvis, vis, n);
        upload_dbl(ctx, dflux, flux, nif);
        da.alloc(sizeof(double) * nnz);
        FC_HIP_CHECK(hipMemsetAsync(da.get(), 0, sizeof(double) * nnz, ctx->stream));   // a = 0 (fvm_laplacian.f90:33)
        WallDev w;
        if (nwal > 0) {
            DevBuf owner_all;
            upload_idx(ctx, owner_all, owner, iWallFacesStart + nwal);
            build_wall_lists(ctx, w, owner_all.as<int>(), n, nwal, iWallFacesStart, 0);
        }
        convdiff_dev(ctx, st, nif, fs.own.as<int>(), fs.nei.as<int>(), fs.ij.as<int>(), fs.ji.as<int>(),
                     dx[0].as<double>(), dx[1].as<double>(), dx[2].as<double>(), dfi.as<double>(), dc[0].as<double>(),
                     dc[1].as<double>(), dc[2].as<double>(), dx[3].as<double>(), dx[4].as<double>(),
                     dx[5].as<double>(), dvis.as<double>(), dflux.as<double>(), w, urf, da.as<double>());
        download_dbl(ctx, a, da.get(), nnz);
        return FC_OK;
    });
}

// The face loops of calcp (facefluxmass) or PISO (facefluxmass_piso, which reads
// neither apv/apw nor p and dPdxi: those may be null) on host arrays.
static fc_status calcp_assemble_host(bool piso, fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                                     const int *neighbour, const double *xc, const double *yc, const double *zc,
                                     const double *vol, const double *xf, const double *yf, const double *zf,
                                     const double *arx, const double *ary, const double *arz, const double *facint,
                                     const double *den, const double *apu, const double *apv, const double *apw,
                                     const double *u, const double *v, const double *w, const double *p,
                                     const double *dUdxi, const double *dVdxi, const double *dWdxi,
                                     const double *dPdxi, int noc, int iOCFacesStart, const int *ijl,
                                     const int *ijr, const double *foc, const int *ioffset, const int *ja,
                                     const int *diag, const int *ij_slot, const int *ji_slot, int nnz, double *a,
                                     double *su, double *flmass, double *al, double *ar, double *fmoc) {
    return guarded(ctx, [&]() {
        const int n = numCells, nif = numInnerFaces;
        FC_REQUIRE(n > 0 && nif >= 0 && nnz == 2 * nif + n, FC_ERR_ARGUMENT, "nnz must be 2*numInnerFaces+numCells");
        FC_REQUIRE(a && su && (flmass || nif == 0), FC_ERR_ARGUMENT, "a/su/flmass");
        const double *cf[12] = {xc, yc, zc, vol, den, apu, apv, apw, u, v, w, p};
        for (int k = 0; k < 12; ++k)
            FC_REQUIRE(cf[k] != nullptr || (piso && (k == 6 || k == 7 || k == 11)), FC_ERR_ARGUMENT, "cell fields");
        const double *gf[4] = {dUdxi, dVdxi, dWdxi, dPdxi};
        for (int k = 0; k < 4; ++k)
            FC_REQUIRE(gf[k] != nullptr || (piso && k == 3), FC_ERR_ARGUMENT, "gradients");
        const int nfaces_needed = std::max(nif, noc > 0 ? iOCFacesStart + noc : 0);
        Structure &st = cached_structure(ctx, n, nnz, ioffset, ja, diag);
        FaceSet fs;
        upload_faces(ctx, fs, st, nif, owner, neighbour, ij_slot, ji_slot);
        DevBuf dx[6], dc[12], dg[4], dfi, dsu, da, dfl, dfm;
        const double *fa[6] = {arx, ary, arz, xf, yf, zf};
        for (int k = 0; k < 6; ++k) upload_dbl(ctx, dx[k], fa[k], nfaces_needed);
        for (int k = 0; k < 12; ++k) upload_dbl(ctx, dc[k], cf[k], cf[k] ? n : 0);
        for (int k = 0; k < 4; ++k) upload_dbl(ctx, dg[k], gf[k], gf[k] ? 3 * (size_t)n : 0);
        upload_dbl(ctx, dfi, facint, nif);
        dsu.alloc(sizeof(double) * n);
        da.alloc(sizeof(double) * nnz);
        dfl.alloc(sizeof(double) * std::max(nif, 1));
        FC_HIP_CHECK(hipMemsetAsync(da.get(), 0, sizeof(double) * nnz, ctx->stream));   // a = 0 (calcp:34)
        OcDev oc;
        oc.noc = std::max(noc, 0);
        oc.ioc_faces_start = iOCFacesStart;
        if (noc > 0) {
            FC_REQUIRE(ijl && ijr && foc && al && ar && fmoc, FC_ERR_ARGUMENT, "O-C arrays");
            upload_idx(ctx, oc.ijl, ijl, noc);
            upload_idx(ctx, oc.ijr, ijr, noc);
            upload_dbl(ctx, oc.foc, foc, noc);
            oc.al.alloc(sizeof(double) * noc);
            oc.ar.alloc(sizeof(double) * noc);
            dfm.alloc(sizeof(double) * noc);
            build_oc_lists(ctx, oc, n);
        }
        RhieChowDev F;
        F.xc = dc[0].as<double>();
        F.yc = dc[1].as<double>();
        F.zc = dc[2].as<double>();
        F.vol = dc[3].as<double>();
        F.den = dc[4].as<double>();
        F.apu = dc[5].as<double>();
        F.apv = dc[6].as<double>();
        F.apw = dc[7].as<double>();
        F.u = dc[8].as<double>();
        F.v = dc[9].as<double>();
        F.w = dc[10].as<double>();
        F.p = dc[11].as<double>();
        F.dudx = dg[0].as<double>();
        F.dvdx = dg[1].as<double>();
        F.dwdx = dg[2].as<double>();
        F.dpdx = dg[3].as<double>();
        calcp_dev(ctx, st, nif, fs.own.as<int>(), fs.nei.as<int>(), fs.ij.as<int>(), fs.ji.as<int>(),
                  fs.kface.as<int>(), fs.needs_order ? fs.kord.as<int>() : nullptr, dx[3].as<double>(),
                  dx[4].as<double>(), dx[5].as<double>(), dx[0].as<double>(), dx[1].as<double>(), dx[2].as<double>(),
                  dfi.as<double>(), F, oc, da.as<double>(), dsu.as<double>(), dfl.as<double>(),
                  noc > 0 ? dfm.as<double>() : nullptr, piso);
        download_dbl(ctx, a, da.get(), nnz);
        download_dbl(ctx, su, dsu.get(), n);
        if (nif) download_dbl(ctx, flmass, dfl.get(), nif);
        if (noc > 0) {
            download_dbl(ctx, al, oc.al.get(), noc);
            download_dbl(ctx, ar, oc.ar.get(), noc);
            download_dbl(ctx, fmoc, dfm.get(), noc);
        }
        return FC_OK;
    });
}

extern "C" fc_status fc_calcp_assemble(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                                       const int *neighbour, const double *xc, const double *yc, const double *zc,
                                       const double *vol, const double *xf, const double *yf, const double *zf,
                                       const double *arx, const double *ary, const double *arz,
                                       const double *facint, const double *den, const double *apu,
                                       const double *apv, const double *apw, const double *u, const double *v,
                                       const double *w, const double *p, const double *dUdxi, const double *dVdxi,
                                       const double *dWdxi, const double *dPdxi, int noc, int iOCFacesStart,
                                       const int *ijl, const int *ijr, const double *foc, const int *ioffset,
                                       const int *ja, const int *diag, const int *ij_slot, const int *ji_slot,
                                       int nnz, double *a, double *su, double *flmass, double *al, double *ar,
                                       double *fmoc) {
    return calcp_assemble_host(false, ctx, numCells, numInnerFaces, owner, neighbour, xc, yc, zc, vol, xf, yf, zf,
                               arx, ary, arz, facint, den, apu, apv, apw, u, v, w, p, dUdxi, dVdxi, dWdxi, dPdxi,
                               noc, iOCFacesStart, ijl, ijr, foc, ioffset, ja, diag, ij_slot, ji_slot, nnz, a, su,
                               flmass, al, ar, fmoc);
}

extern "C" fc_status fc_piso_assemble(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                                      const int *neighbour, const double *xc, const double *yc, const double *zc,
                                      const double *vol, const double *xf, const double *yf, const double *zf,
                                      const double *arx, const double *ary, const double *arz, const double *facint,
                                      const double *den, const double *apu, const double *u, const double *v,
                                      const double *w, const double *dUdxi, const double *dVdxi,
                                      const double *dWdxi, int noc, int iOCFacesStart, const int *ijl,
                                      const int *ijr, const double *foc, const int *ioffset, const int *ja,
                                      const int *diag, const int *ij_slot, const int *ji_slot, int nnz, double *a,
                                      double *su, double *flmass, double *al, double *ar, double *fmoc) {
    return calcp_assemble_host(true, ctx, numCells, numInnerFaces, owner, neighbour, xc, yc, zc, vol, xf, yf, zf,
                               arx, ary, arz, facint, den, apu, nullptr, nullptr, u, v, w, nullptr, dUdxi, dVdxi,
                               dWdxi, nullptr, noc, iOCFacesStart, ijl, ijr, foc, ioffset, ja, diag, ij_slot,
                               ji_slot, nnz, a, su, flmass, al, ar, fmoc);
}

extern "C" fc_status fc_calcp_nonorth(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                                     const int *neighbour, const double *xc, const double *yc, const double *zc,
                                     const double *vol, const double *xf, const double *yf, const double *zf,
                                     const double *arx, const double *ary, const double *arz, const double *facint,
                                     const double *den, const double *apu, const double *dPdxi, int noc,
                                     int iOCFacesStart, const int *ijl, const int *ijr, const double *foc,
                                     const int *ioffset, const int *ja, const int *diag, const int *ij_slot,
                                     const int *ji_slot, int nnz, double *su, double *flmass, double *fmoc) {
    return guarded(ctx, [&]() {
        const int n = numCells, nif = numInnerFaces;
        FC_REQUIRE(n > 0 && nif >= 0 && nnz == 2 * nif + n, FC_ERR_ARGUMENT, "nnz must be 2*numInnerFaces+numCells");
        FC_REQUIRE(su && (flmass || nif == 0), FC_ERR_ARGUMENT, "su/flmass");
        const double *cf[6] = {xc, yc, zc, vol, den, apu};
        for (const double *c : cf) FC_REQUIRE(c != nullptr, FC_ERR_ARGUMENT, "cell fields");
        FC_REQUIRE(dPdxi != nullptr, FC_ERR_ARGUMENT, "dPdxi");
        const int nfaces_needed = std::max(nif, noc > 0 ? iOCFacesStart + noc : 0);
        Structure &st = cached_structure(ctx, n, nnz, ioffset, ja, diag);
        FaceSet fs;
        upload_faces(ctx, fs, st, nif, owner, neighbour, ij_slot, ji_slot);
        DevBuf dx[6], dc[6], dg, dfi, dsu(sizeof(double) * n), dfl, dfm;
        const double *fa[6] = {arx, ary, arz, xf, yf, zf};
        for (int k = 0; k < 6; ++k) upload_dbl(ctx, dx[k], fa[k], nfaces_needed);
        for (int k = 0; k < 6; ++k) upload_dbl(ctx, dc[k], cf[k], n);
        upload_dbl(ctx, dg, dPdxi, 3 * (size_t)n);
        upload_dbl(ctx, dfi, facint, nif);
        upload_dbl(ctx, dfl, flmass, nif);
        OcDev oc;
        oc.noc = std::max(noc, 0);
        oc.ioc_faces_start = iOCFacesStart;
        if (noc > 0) {
            FC_REQUIRE(ijl && ijr && foc && fmoc, FC_ERR_ARGUMENT, "O-C arrays");
            upload_idx(ctx, oc.ijl, ijl, noc);
            upload_idx(ctx, oc.ijr, ijr, noc);
            upload_dbl(ctx, oc.foc, foc, noc);
            upload_dbl(ctx, dfm, fmoc, noc);
            build_oc_lists(ctx, oc, n);
        }
        calcp_nonorth_dev(ctx, st, nif, fs.own.as<int>(), fs.nei.as<int>(), fs.kface.as<int>(),
                          fs.needs_order ? fs.kord.as<int>() : nullptr, dx[3].as<double>(), dx[4].as<double>(),
                          dx[5].as<double>(), dx[0].as<double>(), dx[1].as<double>(), dx[2].as<double>(),
                          dfi.as<double>(), dc[0].as<double>(), dc[1].as<double>(), dc[2].as<double>(),
                          dc[3].as<double>(), dc[4].as<double>(), dc[5].as<double>(), dg.as<double>(), oc,
                          dsu.as<double>(), dfl.as<double>(), noc > 0 ? dfm.as<double>() : nullptr);
        download_dbl(ctx, su, dsu.get(), n);
        if (nif) download_dbl(ctx, flmass, dfl.get(), nif);
        if (noc > 0) download_dbl(ctx, fmoc, dfm.get(), noc);
        return FC_OK;
    });
}

extern "C" fc_status fc_adjust_mass_flow(fc_context *ctx, int numCells, int numTotal, int numFaces,
                                        const int *owner, const double *arx, const double *ary, const double *arz,
                                        int ninl, int iInletFacesStart, const double *fmi, int nout,
                                        int iOutletFacesStart, int iOutletStart, double flomas, const double *den,
                                        double *u, double *v, double *w, double *su, double *fmo) {
    return guarded(ctx, [&]() {
        const int n = numCells, nf = numFaces;
        FC_REQUIRE(n > 0 && numTotal >= n && nf >= 0 && ninl >= 0 && nout >= 0, FC_ERR_ARGUMENT, "sizes");
        FC_REQUIRE(ninl == 0 || (iInletFacesStart >= 0 && iInletFacesStart + ninl <= nf), FC_ERR_ARGUMENT,
                   "inlet faces outside the faces");
        FC_REQUIRE(nout == 0 || (iOutletFacesStart >= 0 && iOutletFacesStart + nout <= nf), FC_ERR_ARGUMENT,
                   "outlet faces outside the faces");
        FC_REQUIRE(nout == 0 || (iOutletStart >= n && iOutletStart + nout <= numTotal), FC_ERR_ARGUMENT,
                   "outlet values outside numCells+1..numTotal");
        FC_REQUIRE(owner && arx && ary && arz && den && u && v && w && su && (fmi || ninl == 0) && (fmo || nout == 0),
                   FC_ERR_ARGUMENT, "arguments");
        DevBuf own, dx[3], dden, duvw[3], dsu, dfmi, dfmo(sizeof(double) * (size_t)std::max(nout, 1));
        upload_idx(ctx, own, owner, nf);
        const double *fa[3] = {arx, ary, arz};
        for (int k = 0; k < 3; ++k) upload_dbl(ctx, dx[k], fa[k], nf);
        upload_dbl(ctx, dden, den, n);
        double *vel[3] = {u, v, w};
        for (int k = 0; k < 3; ++k) upload_dbl(ctx, duvw[k], vel[k], numTotal);
        upload_dbl(ctx, dsu, su, n);
        upload_dbl(ctx, dfmi, fmi, ninl);
        adjust_mass_flow_dev(ctx, n, own.as<int>(), dx[0].as<double>(), dx[1].as<double>(), dx[2].as<double>(), ninl,
                             iInletFacesStart, dfmi.as<double>(), nout, iOutletFacesStart, iOutletStart,
                             flomas, dden.as<double>(), duvw[0].as<double>(), duvw[1].as<double>(),
                             duvw[2].as<double>(), dsu.as<double>(), dfmo.as<double>());
        download_dbl(ctx, su, dsu.get(), n);
        for (int k = 0; k < 3; ++k) download_dbl(ctx, vel[k], duvw[k].get(), numTotal);
        download_dbl(ctx, fmo, dfmo.get(), nout);
        return FC_OK;
    });
}

// The boundary groups of bnd (count, faces before the group; inlet, outlet, symmetry,
// wall, pressure outlet) with their values after the cells, in the reference's order.
static GradGroups groups_of(const int *bnd, int n, int nif, int nf, int ntotal) {
    GradGroups G;
    int q = 0, v = n;
    for (int g = 0; g < 5; ++g) {
        const int cnt = bnd[2 * g], fs = bnd[2 * g + 1];
        FC_REQUIRE(cnt >= 0 && (cnt == 0 || (fs >= nif && fs + cnt <= nf)), FC_ERR_ARGUMENT,
                   "boundary group outside the boundary faces");
        G.qstart[g] = q;
        G.fstart[g] = fs;
        G.vstart[g] = v;
        q += cnt;
        v += cnt;
    }
    G.qstart[5] = q;
    FC_REQUIRE(v <= ntotal, FC_ERR_ARGUMENT, "boundary values beyond numTotal");
    return G;
}

extern "C" fc_status fc_bpres(fc_context *ctx, int numCells, int numTotal, int numFaces, int numInnerFaces,
                              const int *owner, const double *xc, const double *yc, const double *zc,
                              const double *xf, const double *yf, const double *zf, const int *bnd, int noc,
                              int iOCFacesStart, const double *dPdxi, int istage, double *p) {
    return guarded(ctx, [&]() {
        const int n = numCells, nf = numFaces;
        FC_REQUIRE(n > 0 && numInnerFaces >= 0 && nf >= numInnerFaces && numTotal >= n && noc >= 0, FC_ERR_ARGUMENT,
                   "sizes");
        FC_REQUIRE(owner && xc && yc && zc && xf && yf && zf && bnd && p && istage >= 1 && (istage == 1 || dPdxi),
                   FC_ERR_ARGUMENT, "arguments");
        const GradGroups G = groups_of(bnd, n, numInnerFaces, nf, numTotal);
        const int oc_val0 = n + G.qstart[5];
        FC_REQUIRE(noc == 0 || (iOCFacesStart >= 0 && iOCFacesStart + noc <= nf && oc_val0 + noc <= numTotal &&
                                noc <= nf),
                   FC_ERR_ARGUMENT, "O-C faces or values out of range");
        DevBuf own, dc[3], dx[3], dg, dp;
        upload_idx(ctx, own, owner, nf);
        const double *ca[3] = {xc, yc, zc};
        for (int k = 0; k < 3; ++k) upload_dbl(ctx, dc[k], ca[k], n);
        const double *fa[3] = {xf, yf, zf};
        for (int k = 0; k < 3; ++k) upload_dbl(ctx, dx[k], fa[k], nf);
        upload_dbl(ctx, dg, istage >= 2 ? dPdxi : nullptr, istage >= 2 ? 3 * (long long)n : 0);
        upload_dbl(ctx, dp, p, numTotal);
        bpres_dev(ctx, G, noc, iOCFacesStart, oc_val0, own.as<int>(), dc[0].as<double>(), dc[1].as<double>(),
                  dc[2].as<double>(), dx[0].as<double>(), dx[1].as<double>(), dx[2].as<double>(), dg.as<double>(),
                  istage, dp.as<double>());
        download_dbl(ctx, p, dp.get(), numTotal);
        return FC_OK;
    });
}

extern "C" fc_status fc_correct_velocity_bc(fc_context *ctx, int numCells, int numTotal, int numFaces,
                                            int numInnerFaces, const int *owner, const double *arx,
                                            const double *ary, const double *arz, const int *bnd, double flomas,
                                            const double *den, double *u, double *v, double *w, double *fmo) {
    return guarded(ctx, [&]() {
        const int n = numCells, nf = numFaces;
        FC_REQUIRE(n > 0 && numInnerFaces >= 0 && nf >= numInnerFaces && numTotal >= n, FC_ERR_ARGUMENT, "sizes");
        FC_REQUIRE(owner && arx && ary && arz && bnd && den && u && v && w, FC_ERR_ARGUMENT, "arguments");
        const GradGroups G = groups_of(bnd, n, numInnerFaces, nf, numTotal);
        const int nout = bnd[2], nsym = bnd[4];
        FC_REQUIRE(fmo || nout == 0, FC_ERR_ARGUMENT, "fmo");
        DevBuf own, dx[3], dden, duvw[3], dfmo(sizeof(double) * (size_t)std::max(nout, 1)), fac(sizeof(double));
        upload_idx(ctx, own, owner, nf);
        const double *fa[3] = {arx, ary, arz};
        for (int k = 0; k < 3; ++k) upload_dbl(ctx, dx[k], fa[k], nf);
        upload_dbl(ctx, dden, den, n);
        double *vel[3] = {u, v, w};
        for (int k = 0; k < 3; ++k) upload_dbl(ctx, duvw[k], vel[k], numTotal);
        velocity_bc_dev(ctx, own.as<int>(), dx[0].as<double>(), dx[1].as<double>(), dx[2].as<double>(), nout,
                        G.fstart[1], G.vstart[1], nsym, G.fstart[2], G.vstart[2], flomas, dden.as<double>(),
                        duvw[0].as<double>(), duvw[1].as<double>(), duvw[2].as<double>(), dfmo.as<double>(),
                        fac.as<double>());
        for (int k = 0; k < 3; ++k) download_dbl(ctx, vel[k], duvw[k].get(), numTotal);
        download_dbl(ctx, fmo, dfmo.get(), nout);
        return FC_OK;
    });
}

extern "C" fc_status fc_calcp_correct(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                                     const int *neighbour, const int *ioffset, const int *ja, const int *diag,
                                     const int *ij_slot, const int *ji_slot, int nnz, const double *a,
                                     const double *pp, const double *dPdxi, const double *vol, const double *apu,
                                     const double *apv, const double *apw, double urf, int pRefCell, int noc,
                                     const int *ijl, const int *ijr, const double *ar, double *fmoc,
                                     double *flmass, double *u, double *v, double *w, double *p) {
    return guarded(ctx, [&]() {
        const int n = numCells, nif = numInnerFaces;
        FC_REQUIRE(n > 0 && nif >= 0 && nnz == 2 * nif + n, FC_ERR_ARGUMENT, "nnz must be 2*numInnerFaces+numCells");
        FC_REQUIRE(pRefCell >= 1 && pRefCell <= n, FC_ERR_ARGUMENT, "pRefCell outside 1..numCells");
        FC_REQUIRE(a && pp && dPdxi && vol && apu && apv && apw && u && v && w && p && (flmass || nif == 0),
                   FC_ERR_ARGUMENT, "arguments");
        Structure &st = cached_structure(ctx, n, nnz, ioffset, ja, diag);
        FaceSet fs;
        upload_faces(ctx, fs, st, nif, owner, neighbour, ij_slot, ji_slot);
        DevBuf da, dpp, dg, dvol, dap[3], duvwp[4], dfl, dar, dfm;
        upload_dbl(ctx, da, a, nnz);
        upload_dbl(ctx, dpp, pp, n);
        upload_dbl(ctx, dg, dPdxi, 3 * (size_t)n);
        upload_dbl(ctx, dvol, vol, n);
        const double *aps[3] = {apu, apv, apw};
        for (int k = 0; k < 3; ++k) upload_dbl(ctx, dap[k], aps[k], n);
        double *flds[4] = {u, v, w, p};
        for (int k = 0; k < 4; ++k) upload_dbl(ctx, duvwp[k], flds[k], n);
        upload_dbl(ctx, dfl, flmass, nif);
        OcDev oc;
        oc.noc = std::max(noc, 0);
        if (noc > 0) {
            FC_REQUIRE(ijl && ijr && ar && fmoc, FC_ERR_ARGUMENT, "O-C arrays");
            upload_idx(ctx, oc.ijl, ijl, noc);
            upload_idx(ctx, oc.ijr, ijr, noc);
            build_oc_lists(ctx, oc, n);   // range check of ijl/ijr
            upload_dbl(ctx, dar, ar, noc);
            upload_dbl(ctx, dfm, fmoc, noc);
        }
        calcp_correct_dev(ctx, n, nif, fs.own.as<int>(), fs.nei.as<int>(), fs.ij.as<int>(), da.as<double>(),
                          dpp.as<double>(), dg.as<double>(), dvol.as<double>(), dap[0].as<double>(),
                          dap[1].as<double>(), dap[2].as<double>(), urf, pRefCell - 1, oc.noc,
                          noc > 0 ? oc.ijl.as<int>() : nullptr, noc > 0 ? oc.ijr.as<int>() : nullptr,
                          noc > 0 ? dar.as<double>() : nullptr, noc > 0 ? dfm.as<double>() : nullptr,
                          dfl.as<double>(), duvwp[0].as<double>(), duvwp[1].as<double>(), duvwp[2].as<double>(),
                          duvwp[3].as<double>());
        if (nif) download_dbl(ctx, flmass, dfl.get(), nif);
        if (noc > 0) download_dbl(ctx, fmoc, dfm.get(), noc);
        for (int k = 0; k < 4; ++k) download_dbl(ctx, flds[k], duvwp[k].get(), n);
        return FC_OK;
    });
}

extern "C" fc_status fc_grad_gauss(fc_context *ctx, int numCells, int numFaces, int numInnerFaces, const int *owner,
                                   const int *neighbour, const double *arx, const double *ary, const double *arz,
                                   const double *facint, const double *xc, const double *yc, const double *zc,
                                   const double *xf, const double *yf, const double *zf, const double *vol,
                                   const double *u, int numTotal, const int *bnd, int noc, const int *ijlFace,
                                   const int *ijl, const int *ijr, const double *foc, int nigrad, double *dudx,
                                   double *dudy, double *dudz) {
    return guarded(ctx, [&]() {
        const int n = numCells, nif = numInnerFaces, nf = numFaces;
        FC_REQUIRE(n > 0 && nif >= 0 && nf >= nif && numTotal >= n, FC_ERR_ARGUMENT, "sizes");
        FC_REQUIRE(owner && neighbour && u && vol && dudx && dudy && dudz && bnd && nigrad >= 1, FC_ERR_ARGUMENT,
                   "arguments");
        GradGroups G;
        int q = 0, v = n;
        for (int g = 0; g < 5; ++g) {
            const int cnt = bnd[2 * g], fs = bnd[2 * g + 1];
            FC_REQUIRE(cnt >= 0 && (cnt == 0 || (fs >= nif && fs + cnt <= nf)), FC_ERR_ARGUMENT,
                       "boundary group outside the boundary faces");
            G.qstart[g] = q;
            G.fstart[g] = fs;
            G.vstart[g] = v;
            q += cnt;
            v += cnt;
        }
        G.qstart[5] = q;
        FC_REQUIRE(v <= numTotal, FC_ERR_ARGUMENT, "numTotal smaller than cells + boundary values");
        DevBuf own, nei, ij, ji, kface, kord, own_all;
        upload_idx(ctx, own, owner, nif);
        upload_idx(ctx, nei, neighbour, nif);
        upload_idx(ctx, own_all, owner, nf);
        ij.alloc(sizeof(int) * std::max(nif, 1));
        ji.alloc(sizeof(int) * std::max(nif, 1));
        Structure st;
        build_structure_from_faces_dev(ctx, st, n, nif, own.as<int>(), nei.as<int>(), ij.as<int>(), ji.as<int>());
        bool needs = false;
        build_kface(ctx, st, nif, ij.as<int>(), ji.as<int>(), kface, kord, needs);
        DevBuf dx[6], dc[4], dfi, du, dg(sizeof(double) * 3 * (size_t)n), cb_ptr, cb_q, ocf;
        const double *fa[6] = {arx, ary, arz, xf, yf, zf};
        for (int k = 0; k < 6; ++k) upload_dbl(ctx, dx[k], fa[k], nf);
        const double *ca[4] = {xc, yc, zc, vol};
        for (int k = 0; k < 4; ++k) upload_dbl(ctx, dc[k], ca[k], n);
        upload_dbl(ctx, dfi, facint, nif);
        upload_dbl(ctx, du, u, numTotal);
        build_boundary_lists(ctx, G, own_all.as<int>(), n, cb_ptr, cb_q);
        OcDev oc;
        oc.noc = std::max(noc, 0);
        if (noc > 0) {
            FC_REQUIRE(ijlFace && ijl && ijr && foc, FC_ERR_ARGUMENT, "O-C arrays");
            upload_idx(ctx, oc.ijl, ijl, noc);
            upload_idx(ctx, oc.ijr, ijr, noc);
            upload_idx(ctx, ocf, ijlFace, noc);
            upload_dbl(ctx, oc.foc, foc, noc);
            build_oc_lists(ctx, oc, n);
            std::vector<int> hf(ijlFace, ijlFace + noc);
            for (int f : hf) FC_REQUIRE(f >= 1 && f <= nf, FC_ERR_ARGUMENT, "O-C face number out of range");
        }
        double *g = dg.as<double>();
        grad_gauss_dev(ctx, st, nif, own.as<int>(), nei.as<int>(), kface.as<int>(),
                       needs ? kord.as<int>() : nullptr, dx[3].as<double>(), dx[4].as<double>(), dx[5].as<double>(),
                       dx[0].as<double>(), dx[1].as<double>(), dx[2].as<double>(), dfi.as<double>(),
                       dc[0].as<double>(), dc[1].as<double>(), dc[2].as<double>(), dc[3].as<double>(),
                       du.as<double>(), G, cb_ptr.get() ? cb_ptr.as<int>() : nullptr,
                       cb_q.get() ? cb_q.as<int>() : nullptr, oc, noc > 0 ? ocf.as<int>() : nullptr, nigrad, g,
                       g + n, g + 2 * (size_t)n);
        download_dbl(ctx, dudx, g, n);
        download_dbl(ctx, dudy, g + n, n);
        download_dbl(ctx, dudz, g + 2 * (size_t)n, n);
        return FC_OK;
    });
}

extern "C" fc_status fc_face_flux_central(fc_context *ctx, int numInnerFaces, const int *owner, const int *neighbour,
                                          const double *facint, const double *arx, const double *ary,
                                          const double *arz, const double *u, const double *v, const double *w,
                                          int numCells, double *flux) {
    return guarded(ctx, [&]() {
        const int nif = numInnerFaces;
        FC_REQUIRE(nif >= 0 && numCells > 0 && flux, FC_ERR_ARGUMENT, "sizes");
        DevBuf own, nei, dfi, dax, day, daz, du, dv, dw, dfl;
        upload_idx(ctx, own, owner, nif);
        upload_idx(ctx, nei, neighbour, nif);
        upload_dbl(ctx, dfi, facint, nif);
        upload_dbl(ctx, dax, arx, nif);
        upload_dbl(ctx, day, ary, nif);
        upload_dbl(ctx, daz, arz, nif);
        upload_dbl(ctx, du, u, numCells);
        upload_dbl(ctx, dv, v, numCells);
        upload_dbl(ctx, dw, w, numCells);
        dfl.alloc(sizeof(double) * std::max(nif, 1));
        face_flux_dev(ctx, nif, own.as<int>(), nei.as<int>(), dfi.as<double>(), dax.as<double>(), day.as<double>(),
                      daz.as<double>(), du.as<double>(), dv.as<double>(), dw.as<double>(), dfl.as<double>());
        download_dbl(ctx, flux, dfl.get(), nif);
        return FC_OK;
    });
}

extern "C" fc_status fc_rhs_divergence(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                                       const int *neighbour, const double *flux, double *su) {
    return guarded(ctx, [&]() {
        const int n = numCells, nif = numInnerFaces;
        FC_REQUIRE(n > 0 && nif >= 0 && su && flux, FC_ERR_ARGUMENT, "sizes");
        DevBuf own, nei, ij, ji, dfl, dsu, kface, kord;
        upload_idx(ctx, own, owner, nif);
        upload_idx(ctx, nei, neighbour, nif);
        ij.alloc(sizeof(int) * std::max(nif, 1));
        ji.alloc(sizeof(int) * std::max(nif, 1));
        Structure st;
        build_structure_from_faces_dev(ctx, st, n, nif, own.as<int>(), nei.as<int>(), ij.as<int>(), ji.as<int>());
        bool needs = false;
        build_kface(ctx, st, nif, ij.as<int>(), ji.as<int>(), kface, kord, needs);
        upload_dbl(ctx, dfl, flux, nif);
        upload_dbl(ctx, dsu, su, n);
        rhs_divergence_dev(ctx, st, kface.as<int>(), needs ? kord.as<int>() : nullptr,
                           dfl.as<double>(), dsu.as<double>());
        download_dbl(ctx, su, dsu.get(), n);
        return FC_OK;
    });
}

// ============================================================= kernels

extern "C" fc_status fc_spmv(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja, const double *a,
                             const double *x, double *y) {
    return guarded(ctx, [&]() {
        FC_REQUIRE(n > 0 && nnz > 0 && ioffset && ja && a && x && y, FC_ERR_ARGUMENT, "spmv arguments");
        FC_REQUIRE(ioffset[0] == 1 && ioffset[n] == nnz + 1, FC_ERR_STRUCTURE, "ioffset");
        DevBuf dp, dj, da, dx, dy;
        upload_idx(ctx, dp, ioffset, n + 1);
        upload_idx(ctx, dj, ja, nnz);
        upload_dbl(ctx, da, a, nnz);
        upload_dbl(ctx, dx, x, n);
        dy.alloc(sizeof(double) * n);
        spmv_launch64(ctx->stream, 0, n, dp.as<int>(), dj.as<int>(), da.as<double>(), dx.as<double>(), nullptr,
                      nullptr, dy.as<double>(), nullptr, nullptr, nullptr, nullptr);
        FC_HIP_CHECK(hipGetLastError());
        download_dbl(ctx, y, dy.get(), n);
        return FC_OK;
    });
}

static fc_status pivots_common(fc_context *ctx, bool dilu, int n, int nnz, const int *ioffset, const int *ja,
                               const int *diag, const double *a, double *d) {
    return guarded(ctx, [&]() {
        FC_REQUIRE(a && d, FC_ERR_ARGUMENT, "a/d");
        Structure &st = cached_structure(ctx, n, nnz, ioffset, ja, diag);
        ensure_levels(ctx, st);
        if (dilu) build_tpos(ctx, st);
        DevBuf da, dd, err(sizeof(int));
        upload_dbl(ctx, da, a, nnz);
        dd.alloc(sizeof(double) * n);
        FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), ctx->stream));
        SweepCfg cfg;
        DevBuf sfs;
        if (choose_sweeps(ctx, st).dep) {
            sfs.alloc(sizeof(SyncFree));
            syncfree_reset(ctx, ctx->stream, sfs.as<SyncFree>());
            cfg.sf = sfs.as<SyncFree>();
            cfg.waves = syncfree_wave_count(ctx, st);
        }
        pivots_launch(ctx->stream, st, da.as<double>(), dd.as<double>(), dilu, err.as<int>(), &cfg);
        if (cfg.sf) {
            SyncFree hs;
            d2h(&hs, sfs.get(), sizeof(SyncFree), ctx->stream);
            FC_REQUIRE(!hs.abort, FC_ERR_HIP, "dependency-driven pivots: a wait exceeded its time limit");
        }
        int h = 0;
        d2h(&h, err.get(), sizeof(int), ctx->stream);
        download_dbl(ctx, d, dd.get(), n);
        return h ? FC_ERR_ZERO_PIVOT : FC_OK;
    });
}

extern "C" fc_status fc_dic_pivots(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                                   const int *diag, const double *a, double *d) {
    return pivots_common(ctx, false, n, nnz, ioffset, ja, diag, a, d);
}

extern "C" fc_status fc_dilu_pivots(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                                    const int *diag, const double *a, double *d) {
    return pivots_common(ctx, true, n, nnz, ioffset, ja, diag, a, d);
}

extern "C" fc_status fc_lu_apply(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
                                 const double *a, const double *d, const double *r, double *z) {
    return guarded(ctx, [&]() {
        FC_REQUIRE(a && d && r && z, FC_ERR_ARGUMENT, "a/d/r/z");
        Structure &st = cached_structure(ctx, n, nnz, ioffset, ja, diag);
        ensure_levels(ctx, st);
        DevBuf da, dd, dr, dz;
        upload_dbl(ctx, da, a, nnz);
        upload_dbl(ctx, dd, d, n);
        upload_dbl(ctx, dr, r, n);
        dz.alloc(sizeof(double) * n);
        const double small = ctx->variant == FC_VARIANT_TESTS ? kSmallTests : kSmallSrc;
        // the solvers' sweeps: level-packed operands when the structure has them
        DevBuf af, df, ab, db;
        PackedFactor<double> pf;
        const SweepChoice ch = choose_sweeps(ctx, st);
        const bool dep = ch.dep, parts = ch.parts;
        if (packed_sweeps(st)) {
            af.alloc(sizeof(double) * (size_t)st.ell_f * n);
            ab.alloc(sizeof(double) * (size_t)st.ell_b * n);
            df.alloc(sizeof(double) * n);
            db.alloc(sizeof(double) * n);
            if (parts)
                pack_factor_parts64(ctx->stream, st, da.as<double>(), dd.as<double>(), af.as<double>(),
                                    df.as<double>(), ab.as<double>(), db.as<double>());
            else
                pack_factor64(ctx->stream, st, da.as<double>(), dd.as<double>(), af.as<double>(), df.as<double>(),
                              ab.as<double>(), db.as<double>());
            pf.af = af.as<double>();
            pf.df = df.as<double>();
            pf.ab = ab.as<double>();
            pf.db = db.as<double>();
        }
        SweepCfg cfg;
        DevBuf sfs, sft;
        if (dep && pf.af) {
            sfs.alloc(sizeof(SyncFree));
            sft.alloc(sizeof(double) * n);
            syncfree_reset(ctx, ctx->stream, sfs.as<SyncFree>());
            cfg.sf = sfs.as<SyncFree>();
            cfg.tmp = sft.get();
            cfg.waves = parts ? part_wave_count(ctx) : syncfree_wave_count(ctx, st);
            cfg.parts = parts;
        }
        lu_launch64(ctx->stream, st, da.as<double>(), dd.as<double>(), small, dr.as<double>(), dz.as<double>(),
                    nullptr, &cfg, pf.af ? &pf : nullptr);
        FC_HIP_CHECK(hipGetLastError());
        if (cfg.sf) {
            SyncFree h;
            d2h(&h, sfs.get(), sizeof(SyncFree), ctx->stream);
            FC_REQUIRE(!h.abort, FC_ERR_HIP, "dependency-driven sweep: a wait exceeded its time limit");
        }
        download_dbl(ctx, z, dz.get(), n);
        return FC_OK;
    });
}

// ============================================================= solvers

extern "C" fc_status fc_iccg(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
                             const double *a, const double *su, double *fi, double *res, int noc, const int *ijl,
                             const int *ijr, const double *al, const double *ar, double sor, int nsw,
                             fc_solve_info *info) {
    return guarded(ctx, [&]() {
        return host_solve(ctx, 1, n, nnz, ioffset, ja, diag, a, su, fi, res, noc, ijl, ijr, al, ar, sor, nsw, info);
    });
}

extern "C" fc_status fc_dpcg(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
                             const double *a, const double *su, double *fi, double *res, int noc, const int *ijl,
                             const int *ijr, const double *al, const double *ar, double sor, int nsw,
                             fc_solve_info *info) {
    return guarded(ctx, [&]() {
        return host_solve(ctx, 2, n, nnz, ioffset, ja, diag, a, su, fi, res, noc, ijl, ijr, al, ar, sor, nsw, info);
    });
}

extern "C" fc_status fc_bicgstab(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                                 const int *diag, const double *a, const double *su, double *fi, double *res,
                                 int noc, const int *ijl, const int *ijr, const double *al, const double *ar,
                                 double sor, int nsw, fc_solve_info *info) {
    return guarded(ctx, [&]() {
        return host_solve(ctx, 3, n, nnz, ioffset, ja, diag, a, su, fi, res, noc, ijl, ijr, al, ar, sor, nsw, info);
    });
}

// ====================================================== resident system

extern "C" fc_status fc_sys_create(fc_context *ctx, fc_system **out, int numNodes, const double *xyz, int numFaces,
                                   int numInnerFaces, const int *face_ptr, const int *face_nodes, const int *owner,
                                   const int *neighbour, int numCells, int nwal, int iWallFacesStart,
                                   int iWallStart) {
    if (!out) return FC_ERR_ARGUMENT;
    *out = nullptr;
    fc_system *sys = new fc_system();
    sys->ctx = ctx;
    fc_status s = guarded(ctx, [&]() {
        FC_REQUIRE(nwal >= 0 && (nwal == 0 || iWallFacesStart + nwal <= numFaces), FC_ERR_ARGUMENT, "wall range");
        upload_mesh(ctx, sys->mesh, numNodes, xyz, numFaces, numInnerFaces, face_ptr, face_nodes, owner, neighbour,
                    numCells);
        compute_geometry(ctx, sys->mesh);
        const int nif = numInnerFaces, n = numCells;
        sys->ij.alloc(sizeof(int) * std::max(nif, 1));
        sys->ji.alloc(sizeof(int) * std::max(nif, 1));
        build_structure_from_faces_dev(ctx, sys->st, n, nif, sys->mesh.owner.as<int>(), sys->mesh.neighbour.as<int>(),
                                       sys->ij.as<int>(), sys->ji.as<int>());
        ensure_levels(ctx, sys->st);
        build_kface(ctx, sys->st, nif, sys->ij.as<int>(), sys->ji.as<int>(), sys->kface, sys->kord,
                    sys->needs_order);
        build_wall_lists(ctx, sys->walls, sys->mesh.owner.as<int>(), n, nwal, iWallFacesStart, iWallStart);
        sys->ntotal = n + (numFaces - nif);
        sys->a.alloc(sizeof(double) * sys->st.nnz);
        sys->su.alloc(sizeof(double) * n);
        sys->fi.alloc(sizeof(double) * n);
        FC_HIP_CHECK(hipMemsetAsync(sys->fi.get(), 0, sizeof(double) * n, ctx->stream));
        FC_HIP_CHECK(hipMemsetAsync(sys->su.get(), 0, sizeof(double) * n, ctx->stream));
        workspace(ctx, n, sys->st.nnz);
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return FC_OK;
    });
    if (s != FC_OK) {
        delete sys;
        return s;
    }
    *out = sys;
    return FC_OK;
}

extern "C" fc_status fc_sys_destroy(fc_system *sys) {
    if (!sys) return FC_ERR_ARGUMENT;
    if (sys->ctx) (void)hipStreamSynchronize(sys->ctx->stream);
    delete sys;
    return FC_OK;
}

static void sys_assemble_laplacian(fc_system *sys) {
    fc_context *ctx = sys->ctx;
    const int n = sys->st.n;
    GeometryDev &g = sys->mesh.g;
    if (sys->su0_zero) FC_HIP_CHECK(hipMemsetAsync(sys->su.get(), 0, sizeof(double) * n, ctx->stream));
    else FC_HIP_CHECK(hipMemcpyAsync(sys->su.get(), sys->su0.get(), sizeof(double) * n, hipMemcpyDeviceToDevice,
                                     ctx->stream));
    laplacian_dev(ctx, sys->st, sys->mesh.nif, sys->mesh.owner.as<int>(), sys->mesh.neighbour.as<int>(),
                  sys->ij.as<int>(), sys->ji.as<int>(), g.arx.as<double>(), g.ary.as<double>(), g.arz.as<double>(),
                  g.facint.as<double>(), g.xc.as<double>(), g.yc.as<double>(), g.zc.as<double>(), g.xf.as<double>(),
                  g.yf.as<double>(), g.zf.as<double>(), sys->mu.as<double>(), sys->phi.as<double>(), sys->walls,
                  sys->oc, sys->needs_order ? sys->kord.as<int>() : nullptr, sys->a.as<double>(),
                  sys->su.as<double>());
}

extern "C" fc_status fc_sys_assemble_laplacian(fc_system *sys, const double *mu, const double *phi,
                                               const double *su0) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        FC_REQUIRE(mu && phi, FC_ERR_ARGUMENT, "mu/phi");
        const int n = sys->st.n;
        upload_dbl(ctx, sys->mu, mu, n);
        upload_dbl(ctx, sys->phi, phi, sys->ntotal);
        sys->su0_zero = su0 == nullptr;
        if (su0) upload_dbl(ctx, sys->su0, su0, n);
        ctx->timer.start(ctx->stream);
        sys_assemble_laplacian(sys);
        sys->last_ms[0] = ctx->timer.stop_ms(ctx->stream);
        sys->assembly_version += 1;
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_assemble_laplacian_dev(fc_system *sys) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        FC_REQUIRE(sys->mu.get() && sys->phi.get(), FC_ERR_ARGUMENT, "no coefficients uploaded yet");
        ctx->timer.start(ctx->stream);
        sys_assemble_laplacian(sys);
        sys->last_ms[0] = ctx->timer.stop_ms(ctx->stream);
        sys->assembly_version += 1;
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_assemble_convdiff(fc_system *sys, const double *vis, const double *flux, double urf,
                                              const double *su0) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        FC_REQUIRE(vis && flux && urf > 0.0, FC_ERR_ARGUMENT, "vis/flux/urf");
        const int n = sys->st.n, nif = sys->mesh.nif;
        upload_dbl(ctx, sys->vis, vis, n);
        upload_dbl(ctx, sys->flux, flux, nif);
        GeometryDev &g = sys->mesh.g;
        ctx->timer.start(ctx->stream);
        convdiff_dev(ctx, sys->st, nif, sys->mesh.owner.as<int>(), sys->mesh.neighbour.as<int>(), sys->ij.as<int>(),
                     sys->ji.as<int>(), g.arx.as<double>(), g.ary.as<double>(), g.arz.as<double>(),
                     g.facint.as<double>(), g.xc.as<double>(), g.yc.as<double>(), g.zc.as<double>(),
                     g.xf.as<double>(), g.yf.as<double>(), g.zf.as<double>(), sys->vis.as<double>(),
                     sys->flux.as<double>(), sys->walls, urf, sys->a.as<double>());
        sys->last_ms[0] = ctx->timer.stop_ms(ctx->stream);
        sys->assembly_version += 1;
        if (su0) h2d(sys->su.get(), su0, sizeof(double) * n, ctx->stream);
        else FC_HIP_CHECK(hipMemsetAsync(sys->su.get(), 0, sizeof(double) * n, ctx->stream));
        return FC_OK;
    });
}

// True when p is device memory of the current device (used in place); host memory of
// any kind (pageable, pinned, unregistered) is staged through a copy instead.
static bool device_resident(const void *p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    int dev = -1;
    FC_HIP_CHECK(hipGetDevice(&dev));
    return at.type == hipMemoryTypeDevice && at.device == dev;
}

// A field of count doubles as a device pointer: p itself when it is device memory,
// else p copied into stage (which keeps its allocation across calls of one size).
static const double *field_on_device(fc_context *ctx, DevBuf &stage, const double *p, size_t count) {
    if (device_resident(p)) return p;
    stage.ensure(sizeof(double) * std::max<size_t>(count, 1));
    h2d(stage.get(), p, sizeof(double) * count, ctx->stream);
    return stage.as<double>();
}

// calcp's (or, piso, PISO's) face loops on the resident system; PISO reads neither
// apv/apw nor p and dPdxi, which may then be null
static fc_status sys_calcp_assemble(bool piso, fc_system *sys, const double *den, const double *apu,
                                    const double *apv, const double *apw, const double *u, const double *v,
                                    const double *w, const double *p, const double *dUdxi, const double *dVdxi,
                                    const double *dWdxi, const double *dPdxi, double *flmass) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        const int n = sys->st.n, nif = sys->mesh.nif;
        FC_REQUIRE(flmass || nif == 0, FC_ERR_ARGUMENT, "flmass");
        FC_REQUIRE(sys->oc.noc == 0, FC_ERR_ARGUMENT, "resident systems carry no O-C faces");
        const double *cf[8] = {den, apu, apv, apw, u, v, w, p};
        for (int k = 0; k < 8; ++k)
            FC_REQUIRE(cf[k] != nullptr || (piso && (k == 2 || k == 3 || k == 7)), FC_ERR_ARGUMENT, "cell fields");
        const double *gf[4] = {dUdxi, dVdxi, dWdxi, dPdxi};
        for (int k = 0; k < 4; ++k)
            FC_REQUIRE(gf[k] != nullptr || (piso && k == 3), FC_ERR_ARGUMENT, "gradients");
        const double *dcf[8], *dgf[4];
        for (int k = 0; k < 8; ++k) dcf[k] = cf[k] ? field_on_device(ctx, sys->cp_stage[k], cf[k], n) : nullptr;
        for (int k = 0; k < 4; ++k)
            dgf[k] = gf[k] ? field_on_device(ctx, sys->cp_stage[8 + k], gf[k], 3 * (size_t)n) : nullptr;
        const bool fl_dev = nif > 0 && device_resident(flmass);
        double *dfl = flmass;
        if (!fl_dev) {
            sys->cp_flmass.ensure(sizeof(double) * std::max(nif, 1));
            dfl = sys->cp_flmass.as<double>();
        }
        GeometryDev &g = sys->mesh.g;
        RhieChowDev F;
        F.xc = g.xc.as<double>();
        F.yc = g.yc.as<double>();
        F.zc = g.zc.as<double>();
        F.vol = g.vol.as<double>();
        F.den = dcf[0];
        F.apu = dcf[1];
        F.apv = dcf[2];
        F.apw = dcf[3];
        F.u = dcf[4];
        F.v = dcf[5];
        F.w = dcf[6];
        F.p = dcf[7];
        F.dudx = dgf[0];
        F.dvdx = dgf[1];
        F.dwdx = dgf[2];
        F.dpdx = dgf[3];
        ctx->timer.start(ctx->stream);
        FC_HIP_CHECK(hipMemsetAsync(sys->a.get(), 0, sizeof(double) * sys->st.nnz, ctx->stream));   // calcp:34
        calcp_dev(ctx, sys->st, nif, sys->mesh.owner.as<int>(), sys->mesh.neighbour.as<int>(), sys->ij.as<int>(),
                  sys->ji.as<int>(), sys->kface.as<int>(), sys->needs_order ? sys->kord.as<int>() : nullptr,
                  g.xf.as<double>(), g.yf.as<double>(), g.zf.as<double>(), g.arx.as<double>(), g.ary.as<double>(),
                  g.arz.as<double>(), g.facint.as<double>(), F, sys->oc, sys->a.as<double>(), sys->su.as<double>(),
                  dfl, nullptr, piso);
        // calcp: pp = 0 before the pressure-correction solve (calcp:112): the solution vector is
        // the initial guess of the next fc_sys_solve, and the L1 rule's res0 depends on it.  PISO
        // solves for p itself and starts from whatever it holds (its pp = 0 is commented out,
        // src/PISO_multiple_correction.f90:198-199), so the solution vector is left alone.
        if (!piso) FC_HIP_CHECK(hipMemsetAsync(sys->fi.get(), 0, sizeof(double) * n, ctx->stream));
        sys->last_ms[0] = ctx->timer.stop_ms(ctx->stream);
        sys->assembly_version += 1;
        if (nif && !fl_dev) download_dbl(ctx, flmass, dfl, nif);
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_calcp_assemble(fc_system *sys, const double *den, const double *apu, const double *apv,
                                           const double *apw, const double *u, const double *v, const double *w,
                                           const double *p, const double *dUdxi, const double *dVdxi,
                                           const double *dWdxi, const double *dPdxi, double *flmass) {
    return sys_calcp_assemble(false, sys, den, apu, apv, apw, u, v, w, p, dUdxi, dVdxi, dWdxi, dPdxi, flmass);
}

extern "C" fc_status fc_sys_piso_assemble(fc_system *sys, const double *den, const double *apu, const double *u,
                                          const double *v, const double *w, const double *dUdxi, const double *dVdxi,
                                          const double *dWdxi, double *flmass) {
    return sys_calcp_assemble(true, sys, den, apu, nullptr, nullptr, u, v, w, nullptr, dUdxi, dVdxi, dWdxi, nullptr,
                              flmass);
}

extern "C" fc_status fc_sys_pin_reference(fc_system *sys, int pRefCell, const double *p) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        const int n = sys->st.n;
        FC_REQUIRE(pRefCell >= 1 && pRefCell <= n, FC_ERR_ARGUMENT, "pRefCell outside 1..numCells");
        FC_REQUIRE(p != nullptr, FC_ERR_ARGUMENT, "p");
        const double *dp = field_on_device(ctx, sys->cp_stage[0], p, n);
        pin_reference_dev(ctx, sys->st, pRefCell - 1, dp, sys->a.as<double>(), sys->su.as<double>());
        sys->assembly_version += 1;
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return FC_OK;
    });
}

// An in/out array of count doubles as a device pointer: p itself when it is device memory, else
// p copied into stage; *copy_back tells the caller to copy stage back into p afterwards.
static double *inout_on_device(fc_context *ctx, DevBuf &stage, double *p, size_t count, bool *copy_back) {
    *copy_back = !device_resident(p);
    if (!*copy_back) return p;
    return const_cast<double *>(field_on_device(ctx, stage, p, count));
}

extern "C" fc_status fc_sys_calcp_correct(fc_system *sys, const double *pp, const double *dPdxi, const double *apu,
                                          const double *apv, const double *apw, double urf, int pRefCell,
                                          double *flmass, double *u, double *v, double *w, double *p) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        const int n = sys->st.n, nif = sys->mesh.nif;
        FC_REQUIRE(pRefCell >= 1 && pRefCell <= n, FC_ERR_ARGUMENT, "pRefCell outside 1..numCells");
        FC_REQUIRE(dPdxi && apu && apv && apw && u && v && w && p && (flmass || nif == 0), FC_ERR_ARGUMENT,
                   "arguments");
        FC_REQUIRE(sys->oc.noc == 0, FC_ERR_ARGUMENT, "resident systems carry no O-C faces");
        const double *dpp = pp ? field_on_device(ctx, sys->cp_stage[0], pp, n) : sys->fi.as<double>();
        const double *dg = field_on_device(ctx, sys->cp_stage[1], dPdxi, 3 * (size_t)n);
        const double *dap[3] = {field_on_device(ctx, sys->cp_stage[2], apu, n),
                                field_on_device(ctx, sys->cp_stage[3], apv, n),
                                field_on_device(ctx, sys->cp_stage[4], apw, n)};
        double *io[5] = {u, v, w, p, flmass};
        const size_t cnt[5] = {(size_t)n, (size_t)n, (size_t)n, (size_t)n, (size_t)nif};
        double *dio[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        bool back[5] = {false, false, false, false, false};
        for (int k = 0; k < 5; ++k) {
            if (cnt[k] == 0) continue;
            dio[k] = inout_on_device(ctx, k < 4 ? sys->cp_stage[5 + k] : sys->cp_flmass, io[k], cnt[k], &back[k]);
        }
        ctx->timer.start(ctx->stream);
        calcp_correct_dev(ctx, n, nif, sys->mesh.owner.as<int>(), sys->mesh.neighbour.as<int>(), sys->ij.as<int>(),
                          sys->a.as<double>(), dpp, dg, sys->mesh.g.vol.as<double>(), dap[0], dap[1], dap[2], urf,
                          pRefCell - 1, 0, nullptr, nullptr, nullptr, nullptr, dio[4], dio[0], dio[1], dio[2],
                          dio[3]);
        sys->last_ms[0] = ctx->timer.stop_ms(ctx->stream);
        for (int k = 0; k < 5; ++k)
            if (back[k]) download_dbl(ctx, io[k], dio[k], cnt[k]);
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_calcp_nonorth(fc_system *sys, const double *den, const double *apu,
                                          const double *dPdxi, double *flmass) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        const int n = sys->st.n, nif = sys->mesh.nif;
        FC_REQUIRE(den && apu && dPdxi && (flmass || nif == 0), FC_ERR_ARGUMENT, "arguments");
        FC_REQUIRE(sys->oc.noc == 0, FC_ERR_ARGUMENT, "resident systems carry no O-C faces");
        const double *dden = field_on_device(ctx, sys->cp_stage[0], den, n);
        const double *dapu = field_on_device(ctx, sys->cp_stage[1], apu, n);
        const double *dg = field_on_device(ctx, sys->cp_stage[2], dPdxi, 3 * (size_t)n);
        bool back = false;
        double *dfl = nif ? inout_on_device(ctx, sys->cp_flmass, flmass, nif, &back) : nullptr;
        GeometryDev &g = sys->mesh.g;
        OcDev none;
        ctx->timer.start(ctx->stream);
        calcp_nonorth_dev(ctx, sys->st, nif, sys->mesh.owner.as<int>(), sys->mesh.neighbour.as<int>(),
                          sys->kface.as<int>(), sys->needs_order ? sys->kord.as<int>() : nullptr, g.xf.as<double>(),
                          g.yf.as<double>(), g.zf.as<double>(), g.arx.as<double>(), g.ary.as<double>(),
                          g.arz.as<double>(), g.facint.as<double>(), g.xc.as<double>(), g.yc.as<double>(),
                          g.zc.as<double>(), g.vol.as<double>(), dden, dapu, dg, none, sys->su.as<double>(), dfl,
                          nullptr, &sys->grad_scratch);
        // the corrector's next pass starts again from pp = 0 (calcp:112)
        FC_HIP_CHECK(hipMemsetAsync(sys->fi.get(), 0, sizeof(double) * n, ctx->stream));
        sys->last_ms[0] = ctx->timer.stop_ms(ctx->stream);
        if (back) download_dbl(ctx, flmass, dfl, nif);
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return FC_OK;
    });
}


// Gauss gradient of du (device, the system's numTotal values) into out (device, (3, numCells))
static void sys_grad(fc_system *sys, const int *bnd, const double *du, int nigrad, double *out) {
    fc_context *ctx = sys->ctx;
    FC_REQUIRE(sys->oc.noc == 0, FC_ERR_ARGUMENT, "resident systems carry no O-C faces");
    const int n = sys->st.n, nif = sys->mesh.nif;
    const GradGroups G = groups_of(bnd, n, nif, sys->mesh.nfaces, sys->ntotal);
    if (!std::equal(bnd, bnd + 10, sys->grad_bnd)) {
        std::fill(sys->grad_bnd, sys->grad_bnd + 10, -1);
        build_boundary_lists(ctx, G, sys->mesh.owner.as<int>(), n, sys->grad_cb_ptr, sys->grad_cb_q);
        std::copy(bnd, bnd + 10, sys->grad_bnd);
    }
    sys->grad_xyz.ensure(sizeof(double) * 3 * (size_t)n);
    double *g = sys->grad_xyz.as<double>();
    GeometryDev &geo = sys->mesh.g;
    OcDev none;
    grad_gauss_dev(ctx, sys->st, nif, sys->mesh.owner.as<int>(), sys->mesh.neighbour.as<int>(), sys->kface.as<int>(),
                   sys->needs_order ? sys->kord.as<int>() : nullptr, geo.xf.as<double>(), geo.yf.as<double>(),
                   geo.zf.as<double>(), geo.arx.as<double>(), geo.ary.as<double>(), geo.arz.as<double>(),
                   geo.facint.as<double>(), geo.xc.as<double>(), geo.yc.as<double>(), geo.zc.as<double>(),
                   geo.vol.as<double>(), du, G, sys->grad_cb_ptr.get() ? sys->grad_cb_ptr.as<int>() : nullptr,
                   sys->grad_cb_q.get() ? sys->grad_cb_q.as<int>() : nullptr, none, nullptr, nigrad, g, g + n,
                   g + 2 * (size_t)n, &sys->grad_scratch);
    interleave3_dev(ctx, n, g, g + n, g + 2 * (size_t)n, out);
}

extern "C" fc_status fc_sys_grad_gauss(fc_system *sys, const int *bnd, const double *u, int nigrad,
                                       double *dudxi) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        FC_REQUIRE(bnd && u && dudxi && nigrad >= 1, FC_ERR_ARGUMENT, "arguments");
        const int n = sys->st.n;
        const double *du = field_on_device(ctx, sys->grad_u, u, sys->ntotal);
        const bool out_dev = device_resident(dudxi);
        double *out = dudxi;
        if (!out_dev) {
            sys->grad_out.ensure(sizeof(double) * 3 * (size_t)n);
            out = sys->grad_out.as<double>();
        }
        ctx->timer.start(ctx->stream);
        sys_grad(sys, bnd, du, nigrad, out);
        sys->last_ms[0] = ctx->timer.stop_ms(ctx->stream);
        if (!out_dev) download_dbl(ctx, dudxi, out, 3 * (size_t)n);
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_bpres(fc_system *sys, const int *bnd, int istage, const double *dPdxi, double *p) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        FC_REQUIRE(bnd && p && istage >= 1 && (istage == 1 || dPdxi), FC_ERR_ARGUMENT, "arguments");
        FC_REQUIRE(sys->oc.noc == 0, FC_ERR_ARGUMENT, "resident systems carry no O-C faces");
        const int n = sys->st.n;
        const GradGroups G = groups_of(bnd, n, sys->mesh.nif, sys->mesh.nfaces, sys->ntotal);
        bool back = false;
        double *dp = inout_on_device(ctx, sys->grad_u, p, sys->ntotal, &back);
        const double *dg = istage >= 2 ? field_on_device(ctx, sys->cp_stage[0], dPdxi, 3 * (size_t)n) : nullptr;
        GeometryDev &geo = sys->mesh.g;
        bpres_dev(ctx, G, 0, 0, 0, sys->mesh.owner.as<int>(), geo.xc.as<double>(), geo.yc.as<double>(),
                  geo.zc.as<double>(), geo.xf.as<double>(), geo.yf.as<double>(), geo.zf.as<double>(), dg, istage, dp);
        if (back) download_dbl(ctx, p, dp, sys->ntotal);
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_grad_pp(fc_system *sys, const int *bnd, int nipgrad, int nigrad, double *dPdxi) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        FC_REQUIRE(bnd && dPdxi && nipgrad >= 1 && nigrad >= 1, FC_ERR_ARGUMENT, "arguments");
        FC_REQUIRE(sys->oc.noc == 0, FC_ERR_ARGUMENT, "resident systems carry no O-C faces");
        const int n = sys->st.n;
        const GradGroups G = groups_of(bnd, n, sys->mesh.nif, sys->mesh.nfaces, sys->ntotal);
        // pp over numTotal: the solution's cells, boundary values from bpres
        sys->pp_total.ensure(sizeof(double) * (size_t)sys->ntotal);
        double *pp = sys->pp_total.as<double>();
        FC_HIP_CHECK(hipMemsetAsync(pp, 0, sizeof(double) * (size_t)sys->ntotal, ctx->stream));
        FC_HIP_CHECK(hipMemcpyAsync(pp, sys->fi.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
        const bool out_dev = device_resident(dPdxi);
        double *out = dPdxi;
        if (!out_dev) {
            sys->grad_out.ensure(sizeof(double) * 3 * (size_t)n);
            out = sys->grad_out.as<double>();
        }
        GeometryDev &geo = sys->mesh.g;
        ctx->timer.start(ctx->stream);
        for (int istage = 1; istage <= nipgrad; ++istage) {   // calcp:131-136
            bpres_dev(ctx, G, 0, 0, 0, sys->mesh.owner.as<int>(), geo.xc.as<double>(), geo.yc.as<double>(),
                      geo.zc.as<double>(), geo.xf.as<double>(), geo.yf.as<double>(), geo.zf.as<double>(), out, istage,
                      pp);
            sys_grad(sys, bnd, pp, nigrad, out);
        }
        sys->last_ms[0] = ctx->timer.stop_ms(ctx->stream);
        if (!out_dev) download_dbl(ctx, dPdxi, out, 3 * (size_t)n);
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_correct_velocity_bc(fc_system *sys, const int *bnd, double flomas, const double *den,
                                                double *u, double *v, double *w, double *fmo) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        FC_REQUIRE(bnd && den && u && v && w, FC_ERR_ARGUMENT, "arguments");
        const int n = sys->st.n;
        const GradGroups G = groups_of(bnd, n, sys->mesh.nif, sys->mesh.nfaces, sys->ntotal);
        const int nout = bnd[2], nsym = bnd[4];
        FC_REQUIRE(fmo || nout == 0, FC_ERR_ARGUMENT, "fmo");
        const double *dden = field_on_device(ctx, sys->cp_stage[0], den, n);
        double *io[3] = {u, v, w};
        double *dio[3];
        bool back[3];
        for (int k = 0; k < 3; ++k) dio[k] = inout_on_device(ctx, sys->cp_stage[1 + k], io[k], sys->ntotal, &back[k]);
        bool fback = false;
        double *dfmo = nout ? inout_on_device(ctx, sys->cp_stage[4], fmo, nout, &fback) : nullptr;
        sys->vbc_fac.ensure(sizeof(double));
        GeometryDev &geo = sys->mesh.g;
        velocity_bc_dev(ctx, sys->mesh.owner.as<int>(), geo.arx.as<double>(), geo.ary.as<double>(), geo.arz.as<double>(),
                        nout, G.fstart[1], G.vstart[1], nsym, G.fstart[2], G.vstart[2], flomas, dden, dio[0], dio[1],
                        dio[2], dfmo, sys->vbc_fac.as<double>());
        for (int k = 0; k < 3; ++k)
            if (back[k]) download_dbl(ctx, io[k], dio[k], sys->ntotal);
        if (fback) download_dbl(ctx, fmo, dfmo, nout);
        FC_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_solve_momentum(fc_system *sys, const double *const *sp, const double *const *src,
                                           const double *urf, double *const *phi, double *const *ap, int method,
                                           double sor, int nsw, fc_solve_info *info) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        FC_REQUIRE(sp && src && urf && phi && nsw > 0 && sor > 0.0, FC_ERR_ARGUMENT, "arguments");
        FC_REQUIRE(method == FC_METHOD_BICGSTAB || method == FC_METHOD_BICGSTAB_MIXED, FC_ERR_ARGUMENT,
                   "method must be FC_METHOD_BICGSTAB or FC_METHOD_BICGSTAB_MIXED");
        const int n = sys->st.n;
        for (int c = 0; c < 3; ++c)
            FC_REQUIRE(sp[c] && src[c] && phi[c] && urf[c] > 0.0, FC_ERR_ARGUMENT, "component arrays / urf");
        DevBuf dsp, dsrc, dap;
        dap.alloc(sizeof(double) * n);
        // calcuvw solves u, v and w whatever each solve's outcome (src/calcuvw.f90:441,498,556):
        // a component that runs out of nsw does not stop the next one; errors do.
        fc_status worst = FC_OK;
        for (int c = 0; c < 3; ++c) {
            upload_dbl(ctx, dsp, sp[c], n);
            upload_dbl(ctx, dsrc, src[c], n);
            h2d(sys->fi.get(), phi[c], sizeof(double) * n, ctx->stream);
            momentum_rows_dev(ctx, sys->st, sys->a.as<double>(), dsp.as<double>(), dsrc.as<double>(),
                              sys->fi.as<double>(), urf[c], sys->su.as<double>(), dap.as<double>());
            sys->assembly_version += 1;
            SolveArgs args;
            args.st = &sys->st;
            args.a = sys->a.as<double>();
            args.su = sys->su.as<double>();
            args.fi = sys->fi.as<double>();
            args.sor = sor;
            args.nsw = nsw;
            args.variant = FC_VARIANT_SRC;
            args.stop_rule = FC_STOP_L1;
            if (ctx->inner_eta > 0.0 && ctx->inner_eta < 1.0) args.inner_eta = ctx->inner_eta;
            fc_solve_info local;
            std::memset(&local, 0, sizeof local);
            const fc_status s = method == FC_METHOD_BICGSTAB ? solve_bicgstab64(ctx, args, &local)
                                                             : solve_mixed(ctx, args, true, &local);
            local.rel_l2 = true_rel_l2(ctx, args);
            if (info) info[c] = local;
            d2h(phi[c], sys->fi.get(), sizeof(double) * n, ctx->stream);
            if (ap && ap[c]) d2h(ap[c], dap.get(), sizeof(double) * n, ctx->stream);
            if (s < 0) return s;
            if (s == FC_NOT_CONVERGED) worst = FC_NOT_CONVERGED;
        }
        return worst;
    });
}

extern "C" fc_status fc_sys_set_rhs(fc_system *sys, const double *su) {
    if (!sys || !su) return FC_ERR_ARGUMENT;
    return guarded(sys->ctx, [&]() {
        h2d(sys->su.get(), su, sizeof(double) * sys->st.n, sys->ctx->stream);
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_set_solution(fc_system *sys, const double *fi) {
    if (!sys || !fi) return FC_ERR_ARGUMENT;
    return guarded(sys->ctx, [&]() {
        h2d(sys->fi.get(), fi, sizeof(double) * sys->st.n, sys->ctx->stream);
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_get_solution(fc_system *sys, double *fi) {
    if (!sys || !fi) return FC_ERR_ARGUMENT;
    return guarded(sys->ctx, [&]() {
        d2h(fi, sys->fi.get(), sizeof(double) * sys->st.n, sys->ctx->stream);
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_solve(fc_system *sys, int method, int stop_rule, double sor, int nsw,
                                  fc_solve_info *info) {
    if (!sys) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        FC_REQUIRE(nsw > 0 && sor > 0.0, FC_ERR_ARGUMENT, "sor/nsw");
        SolveArgs args;
        args.st = &sys->st;
        args.a = sys->a.as<double>();
        args.su = sys->su.as<double>();
        args.fi = sys->fi.as<double>();
        args.sor = sor;
        args.nsw = nsw;
        args.variant = FC_VARIANT_SRC;
        args.stop_rule = stop_rule;
        // inner tolerance of the FP32 solves: the context's setting, else 1e-2 for flexible CG
        // (restarts help the nonlinear preconditioner) and for the mixed IC(0)-PCG (its cycles
        // keep the search direction: 604 iterations at 192^3 against 909 at 1e-3), and
        // SolveArgs' 1e-3 for the mixed BiCGStab, whose cycles restart
        if (ctx->inner_eta > 0.0 && ctx->inner_eta < 1.0) args.inner_eta = ctx->inner_eta;
        else if (method == FC_METHOD_FCG_LEARNED || method == FC_METHOD_FCG_JACOBI || method == FC_METHOD_ICCG_MIXED)
            args.inner_eta = kFcgInnerEta;
        fc_solve_info local;
        std::memset(&local, 0, sizeof local);
        fc_status s;
        switch (method) {
        case FC_METHOD_ICCG: s = solve_pcg64(ctx, args, true, &local); break;
        case FC_METHOD_DPCG: s = solve_pcg64(ctx, args, false, &local); break;
        case FC_METHOD_BICGSTAB: s = solve_bicgstab64(ctx, args, &local); break;
        case FC_METHOD_ICCG_MIXED: s = solve_mixed(ctx, args, false, &local); break;
        case FC_METHOD_BICGSTAB_MIXED: s = solve_mixed(ctx, args, true, &local); break;
        case FC_METHOD_FCG_LEARNED: {
            FC_REQUIRE(sys->net != nullptr, FC_ERR_ARGUMENT, "no network attached (fc_sys_set_network)");
            PrecondHook P = gnn_hook(sys);
            s = solve_fcg(ctx, args, P, &local);
            break;
        }
        case FC_METHOD_FCG_JACOBI: {
            PrecondHook P = jacobi_hook(sys);
            s = solve_fcg(ctx, args, P, &local);
            break;
        }
        default: throw Error(FC_ERR_ARGUMENT, "unknown method");
        }
        local.rel_l2 = true_rel_l2(ctx, args);
        sys->last_ms[1] = local.solve_ms + local.setup_ms;
        if (info) *info = local;
        return s;
    });
}

extern "C" void *fc_sys_device_ptr(fc_system *sys, int which) {
    if (!sys) return nullptr;
    GeometryDev &g = sys->mesh.g;
    switch (which) {
    case FC_ARR_A: return sys->a.get();
    case FC_ARR_SU: return sys->su.get();
    case FC_ARR_IOFFSET: return sys->st.ioffset.get();
    case FC_ARR_JA: return sys->st.ja.get();
    case FC_ARR_DIAG: return sys->st.diag.get();
    case FC_ARR_IJ_SLOT: return sys->ij.get();
    case FC_ARR_JI_SLOT: return sys->ji.get();
    case FC_ARR_VOL: return g.vol.get();
    case FC_ARR_XC: return g.xc.get();
    case FC_ARR_YC: return g.yc.get();
    case FC_ARR_ZC: return g.zc.get();
    case FC_ARR_FACINT: return g.facint.get();
    case FC_ARR_ARX: return g.arx.get();
    case FC_ARR_ARY: return g.ary.get();
    case FC_ARR_ARZ: return g.arz.get();
    case FC_ARR_XF: return g.xf.get();
    case FC_ARR_YF: return g.yf.get();
    case FC_ARR_ZF: return g.zf.get();
    case FC_ARR_RES: return sys->ctx->ws ? sys->ctx->ws->r.get() : nullptr;
    case FC_ARR_PIVOTS: return sys->ctx->ws ? sys->ctx->ws->d.get() : nullptr;
    default: return nullptr;
    }
}

extern "C" fc_status fc_sys_get(fc_system *sys, int which, void *host_out, size_t bytes) {
    if (!sys || !host_out) return FC_ERR_ARGUMENT;
    fc_context *ctx = sys->ctx;
    return guarded(ctx, [&]() {
        const int n = sys->st.n, nnz = sys->st.nnz, nif = sys->mesh.nif, nf = sys->mesh.nfaces;
        size_t expect = 0;
        bool index = false;
        switch (which) {
        case FC_ARR_A: expect = sizeof(double) * nnz; break;
        case FC_ARR_SU: case FC_ARR_VOL: case FC_ARR_XC: case FC_ARR_YC: case FC_ARR_ZC: case FC_ARR_RES:
        case FC_ARR_PIVOTS: expect = sizeof(double) * n; break;
        case FC_ARR_FACINT: expect = sizeof(double) * nif; break;
        case FC_ARR_ARX: case FC_ARR_ARY: case FC_ARR_ARZ: case FC_ARR_XF: case FC_ARR_YF: case FC_ARR_ZF:
            expect = sizeof(double) * nf; break;
        case FC_ARR_IOFFSET: expect = sizeof(int) * (n + 1); index = true; break;
        case FC_ARR_JA: expect = sizeof(int) * nnz; index = true; break;
        case FC_ARR_DIAG: expect = sizeof(int) * n; index = true; break;
        case FC_ARR_IJ_SLOT: case FC_ARR_JI_SLOT: expect = sizeof(int) * nif; index = true; break;
        default: throw Error(FC_ERR_ARGUMENT, "unknown array");
        }
        FC_REQUIRE(bytes >= expect, FC_ERR_ARGUMENT, "output buffer too small");
        void *src = fc_sys_device_ptr(sys, which);
        FC_REQUIRE(src != nullptr, FC_ERR_ARGUMENT, "array not available");
        if (index) {
            DevBuf tmp(expect);
            FC_HIP_CHECK(hipMemcpyAsync(tmp.get(), src, expect, hipMemcpyDeviceToDevice, ctx->stream));
            to_one_based(ctx, tmp.as<int>(), (long long)(expect / sizeof(int)));
            d2h(host_out, tmp.get(), expect, ctx->stream);
        } else {
            d2h(host_out, src, expect, ctx->stream);
        }
        return FC_OK;
    });
}

extern "C" fc_status fc_sys_sizes(fc_system *sys, int *n, int *nif, int *nnz, int *nlev_fwd, int *nlev_bwd) {
    if (!sys) return FC_ERR_ARGUMENT;
    if (n) *n = sys->st.n;
    if (nif) *nif = sys->mesh.nif;
    if (nnz) *nnz = sys->st.nnz;
    if (nlev_fwd) *nlev_fwd = sys->st.nlev_f;
    if (nlev_bwd) *nlev_bwd = sys->st.nlev_b;
    return FC_OK;
}

extern "C" double fc_sys_last_ms(fc_system *sys, int stage) {
    if (!sys || stage < 0 || stage > 3) return -1.0;
    return sys->last_ms[stage];
}

extern "C" double fc_sys_sweep_bench(fc_system *sys, int dilu, int reps) {
    if (!sys || reps <= 0 || (dilu != 0 && dilu != 1)) return -1.0;
    fc_context *ctx = sys->ctx;
    double out = -1.0;
    fc_status s = guarded(ctx, [&]() {
        Structure &st = sys->st;
        ensure_levels(ctx, st);
        if (dilu) build_tpos(ctx, st);
        const int n = st.n;
        hipStream_t hs = ctx->stream;
        const double *a = sys->a.as<double>();
        const SweepChoice ch = choose_sweeps(ctx, st);
        const bool dep = ch.dep, parts = ch.parts;
        DevBuf d(sizeof(double) * n), z(sizeof(double) * n), err(sizeof(int)), sfs, sft;
        FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), hs));
        SweepCfg cfg;
        if (dep) {
            sfs.alloc(sizeof(SyncFree));
            sft.alloc(sizeof(double) * n);
            syncfree_reset(ctx, hs, sfs.as<SyncFree>());
            cfg.sf = sfs.as<SyncFree>();
            cfg.tmp = sft.get();
            cfg.waves = syncfree_wave_count(ctx, st);   // the pivots' dependency-driven launch
        }
        pivots_launch(hs, st, a, d.as<double>(), dilu != 0, err.as<int>(), dep ? &cfg : nullptr);
        int herr = 0;
        d2h(&herr, err.get(), sizeof(int), hs);
        FC_REQUIRE(!herr, FC_ERR_ZERO_PIVOT, "sweep bench: a pivot is zero or not finite");
        cfg.waves = parts ? part_wave_count(ctx) : syncfree_wave_count(ctx, st);
        cfg.parts = parts;
        DevBuf af, df, ab, db;
        PackedFactor<double> pf;
        if (packed_sweeps(st)) {
            af.alloc(sizeof(double) * (size_t)st.ell_f * n);
            ab.alloc(sizeof(double) * (size_t)st.ell_b * n);
            df.alloc(sizeof(double) * n);
            db.alloc(sizeof(double) * n);
            if (parts)
                pack_factor_parts64(hs, st, a, d.as<double>(), af.as<double>(), df.as<double>(), ab.as<double>(),
                                    db.as<double>());
            else
                pack_factor64(hs, st, a, d.as<double>(), af.as<double>(), df.as<double>(), ab.as<double>(),
                              db.as<double>());
            pf = {af.as<double>(), df.as<double>(), ab.as<double>(), db.as<double>()};
        }
        const SweepCfg *pc = dep && pf.af ? &cfg : nullptr;
        const double small = ctx->variant == FC_VARIANT_TESTS ? kSmallTests : kSmallSrc;
        lu_launch64(hs, st, a, d.as<double>(), small, sys->su.as<double>(), z.as<double>(), nullptr, pc,
                    pf.af ? &pf : nullptr);   // warm-up
        ctx->timer.start(hs);
        for (int r = 0; r < reps; ++r)
            lu_launch64(hs, st, a, d.as<double>(), small, sys->su.as<double>(), z.as<double>(), nullptr, pc,
                        pf.af ? &pf : nullptr);
        FC_HIP_CHECK(hipGetLastError());
        const double ms = ctx->timer.stop_ms(hs);
        if (pc) {
            SyncFree h;
            d2h(&h, sfs.get(), sizeof(SyncFree), hs);
            FC_REQUIRE(!h.abort, FC_ERR_HIP, "dependency-driven sweep: a wait exceeded its time limit");
        }
        out = ms / reps;
        sys->last_ms[3] = out;
        return FC_OK;
    });
    return s == FC_OK ? out : -1.0;
}

extern "C" double fc_sys_spmv_bench(fc_system *sys, int precision, int reps) {
    if (!sys || reps <= 0 || (precision != 32 && precision != 64)) return -1.0;
    fc_context *ctx = sys->ctx;
    double out = -1.0;
    fc_status s = guarded(ctx, [&]() {
        Workspace &ws = workspace(ctx, sys->st.n, sys->st.nnz);
        const Structure &st = sys->st;
        const int n = st.n;
        hipStream_t s = ctx->stream;
        const double *x64 = sys->su.as<double>();
        float *a32 = ws.a32.as<float>(), *x32 = ws.p32.as<float>(), *y32 = ws.q32.as<float>();
        if (precision == 32) {
            cast_d2f(s, st.nnz, sys->a.as<double>(), a32);
            cast_d2f(s, n, x64, x32);
        }
        // warm-up
        if (precision == 64)
            spmv_launch64(s, 0, n, st.ioffset.as<int>(), st.ja.as<int>(), sys->a.as<double>(), x64, nullptr, nullptr,
                          ws.q.as<double>(), nullptr, nullptr, nullptr, nullptr);
        else
            spmv_launch32(s, 0, n, st.ioffset.as<int>(), st.ja.as<int>(), a32, x32, nullptr, nullptr, y32, nullptr,
                          nullptr, nullptr);
        ctx->timer.start(s);
        for (int r = 0; r < reps; ++r) {
            if (precision == 64)
                spmv_launch64(s, 0, n, st.ioffset.as<int>(), st.ja.as<int>(), sys->a.as<double>(), x64, nullptr,
                              nullptr, ws.q.as<double>(), nullptr, nullptr, nullptr, nullptr);
            else
                spmv_launch32(s, 0, n, st.ioffset.as<int>(), st.ja.as<int>(), a32, x32, nullptr, nullptr, y32,
                              nullptr, nullptr, nullptr);
        }
        FC_HIP_CHECK(hipGetLastError());
        double ms = ctx->timer.stop_ms(s);
        out = ms / reps;
        sys->last_ms[2] = out;
        return FC_OK;
    });
    return s == FC_OK ? out : -1.0;
}

// ========================================================= calibration

namespace {
// four independent 16-byte loads in flight per thread
__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long long n16) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        uint4 v0 = src[i], v1 = src[i + stride], v2 = src[i + 2 * stride], v3 = src[i + 3 * stride];
        dst[i] = v0;
        dst[i + stride] = v1;
        dst[i + 2 * stride] = v2;
        dst[i + 3 * stride] = v3;
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" double fc_copy_bench(fc_context *ctx, size_t bytes, int reps) {
    if (!ctx || bytes < 16 || reps <= 0) return -1.0;
    double out = -1.0;
    fc_status s = guarded(ctx, [&]() {
        const long long n16 = (long long)(bytes / 16);
        DevBuf a(n16 * 16), b(n16 * 16);
        FC_HIP_CHECK(hipMemsetAsync(a.get(), 1, n16 * 16, ctx->stream));
        const int grid = std::max(1, std::min<int>((int)((n16 + 255) / 256), ctx->num_cus * 16));
        k_copy16<<<grid, 256, 0, ctx->stream>>>(a.as<uint4>(), b.as<uint4>(), n16);
        ctx->timer.start(ctx->stream);
        for (int r = 0; r < reps; ++r) k_copy16<<<grid, 256, 0, ctx->stream>>>(a.as<uint4>(), b.as<uint4>(), n16);
        FC_HIP_CHECK(hipGetLastError());
        out = ctx->timer.stop_ms(ctx->stream) / reps;
        return FC_OK;
    });
    return s == FC_OK ? out : -1.0;
}

extern "C" fc_status fc_dpcg_noalloc(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                                     const int *diag, const double *a, const double *su, double *fi, double *res,
                                     double sor, int nsw, fc_solve_info *info) {
    return guarded(ctx, [&]() {
        FC_REQUIRE(ioffset && ja && diag && a && su && fi, FC_ERR_ARGUMENT, "arrays");
        const void *dev[7] = {ioffset, ja, diag, a, su, fi, res ? (const void *)res : (const void *)fi};
        for (const void *p : dev)
            FC_REQUIRE(device_resident(p), FC_ERR_ARGUMENT,
                       "fc_dpcg_noalloc takes device pointers on the context's device only");
        FC_REQUIRE(nsw > 0 && sor > 0.0, FC_ERR_ARGUMENT, "sor/nsw");
        return solve_dpcg_noalloc(ctx, n, nnz, ioffset, ja, diag, a, su, fi, res, sor, nsw, info);
    });
}
