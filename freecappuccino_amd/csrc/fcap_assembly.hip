// Face-loop assembly of the CSR matrix on the GPU.
//
// The reference's face loop (src/fvm_laplacian.f90:38-65) writes each face's
// two off-diagonal coefficients into slots no other face touches, and
// subtracts them from the two diagonals in face order.  Here:
//   * a face-parallel kernel computes the coefficient and writes the two
//     off-diagonal slots (one writer per slot: no atomics, no colouring);
//   * a row-parallel kernel rebuilds each diagonal as
//     ((0 - a_f1) - a_f2) - ... over the row's entries taken in face order,
//     then applies O-C cuts and wall faces in their order.  Row i's entry for
//     face f holds exactly the value the reference subtracts from a(diag(i)) for
//     that face, so every diagonal is bit-identical to the sequential loop.
// For meshes in polyMesh (upper-triangular) face order, face order within a
// row equals column order and no permutation is needed.

#include <algorithm>

#include "fcap_internal.h"

namespace fcap {

void scan_dev(fc_context *ctx, const int *in, int *out, int n);
void sort_segments_dev(fc_context *ctx, const int *ptr, int *v, int nseg);

namespace {

constexpr int kBlock = 256;
// slots of one wave's 64 consecutive rows staged in LDS by the row kernels (a
// wave whose rows hold more reads them from memory)
constexpr int kRowStage = 512;

inline int grid_for(long long n, int block = kBlock) {
    long long g = (n + block - 1) / block;
    return (int)std::max<long long>(1, std::min<long long>(g, 1 << 30));
}

// facefluxlaplacian (src/fvm_laplacian.f90:171-221)
__device__ inline double laplacian_coef(int p, int q, double sx, double sy, double sz, double lambda,
                                        const double *__restrict__ mu, const double *__restrict__ xc,
                                        const double *__restrict__ yc, const double *__restrict__ zc) {
    double fxn = lambda;
    double fxp = 1.0 - lambda;
    double xpn = xc[q] - xc[p];
    double ypn = yc[q] - yc[p];
    double zpn = zc[q] - zc[p];
    double num = (sx * sx + sy * sy) + sz * sz;
    double den = (sx * xpn + sy * ypn) + sz * zpn;
    double smdpn = num / den;
    double m = fxp * mu[p] + fxn * mu[q];
    return m * smdpn;
}

__global__ void k_laplacian_faces(int nif, const int *__restrict__ own, const int *__restrict__ nei,
                                  const int *__restrict__ ij, const int *__restrict__ ji,
                                  const double *__restrict__ arx, const double *__restrict__ ary,
                                  const double *__restrict__ arz, const double *__restrict__ facint,
                                  const double *__restrict__ xc, const double *__restrict__ yc,
                                  const double *__restrict__ zc, const double *__restrict__ mu,
                                  double *__restrict__ a) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    double c = laplacian_coef(own[f], nei[f], arx[f], ary[f], arz[f], facint[f], mu, xc, yc, zc);
    a[ij[f]] = c;   // can
    a[ji[f]] = c;   // cap
}

__global__ void k_laplacian_oc(int noc, int oc_face0, const int *__restrict__ ijl, const int *__restrict__ ijr,
                               const double *__restrict__ foc, const double *__restrict__ arx,
                               const double *__restrict__ ary, const double *__restrict__ arz,
                               const double *__restrict__ xc, const double *__restrict__ yc,
                               const double *__restrict__ zc, const double *__restrict__ mu,
                               double *__restrict__ al, double *__restrict__ ar) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= noc) return;
    int f = oc_face0 + i;
    double c = laplacian_coef(ijl[i], ijr[i], arx[f], ary[f], arz[f], foc[i], mu, xc, yc, zc);
    al[i] = c;
    ar[i] = c;
}

// Diagonal in face order, then O-C cuts (:68-86), then wall faces (:140-154).
__global__ void k_laplacian_rows(int n, const int *__restrict__ ptr, const int *__restrict__ diagpos,
                                 const int *__restrict__ kord, const int *__restrict__ co_ptr,
                                 const int *__restrict__ co_idx, const double *__restrict__ al,
                                 const double *__restrict__ ar, const int *__restrict__ cw_ptr,
                                 const int *__restrict__ cw_idx, int wall_face0, int wall_val0,
                                 const double *__restrict__ arx, const double *__restrict__ ary,
                                 const double *__restrict__ arz, const double *__restrict__ xc,
                                 const double *__restrict__ yc, const double *__restrict__ zc,
                                 const double *__restrict__ xf, const double *__restrict__ yf,
                                 const double *__restrict__ zf, const double *__restrict__ mu,
                                 const double *__restrict__ phi, double *__restrict__ a,
                                 double *__restrict__ su) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int dp = diagpos[i];
    double d = 0.0;
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) {
        int k = kord ? kord[p] : p;
        if (k == dp) continue;
        d = d - a[k];
    }
    if (co_ptr) {
        for (int q = co_ptr[i]; q < co_ptr[i + 1]; ++q) {
            int code = co_idx[q];
            int o = code >> 1;
            d = d - ((code & 1) ? al[o] : ar[o]);
        }
    }
    if (cw_ptr && cw_ptr[i] < cw_ptr[i + 1]) {
        double s = su[i];
        double m = mu[i];
        for (int q = cw_ptr[i]; q < cw_ptr[i + 1]; ++q) {
            int w = cw_idx[q];
            int f = wall_face0 + w;
            double fx = arx[f], fy = ary[f], fz = arz[f];
            double are = sqrt((fx * fx + fy * fy) + fz * fz);
            double dx = xc[i] - xf[f], dy = yc[i] - yf[f], dz = zc[i] - zf[f];
            double dpw = sqrt((dx * dx + dy * dy) + dz * dz);
            d = d - (m * are) / dpw;
            s = s + d * phi[wall_val0 + w];
        }
        su[i] = s;
    }
    a[dp] = d;
}

// facefluxuvw coefficient part (src/faceflux_velocity.f90:97-108), upwind
__global__ void k_convdiff_faces(int nif, const int *__restrict__ own, const int *__restrict__ nei,
                                 const int *__restrict__ ij, const int *__restrict__ ji,
                                 const double *__restrict__ arx, const double *__restrict__ ary,
                                 const double *__restrict__ arz, const double *__restrict__ facint,
                                 const double *__restrict__ xc, const double *__restrict__ yc,
                                 const double *__restrict__ zc, const double *__restrict__ vis,
                                 const double *__restrict__ flux, double *__restrict__ a) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int p = own[f], q = nei[f];
    double fxn = facint[f], fxp = 1.0 - facint[f];
    double xpn = xc[q] - xc[p], ypn = yc[q] - yc[p], zpn = zc[q] - zc[p];
    double dpn = sqrt((xpn * xpn + ypn * ypn) + zpn * zpn);
    double sx = arx[f], sy = ary[f], sz = arz[f];
    double are = sqrt((sx * sx + sy * sy) + sz * sz);
    double game = vis[p] * fxp + vis[q] * fxn;
    double de = (game * are) / dpn;
    double fl = flux[f];
    a[ij[f]] = -de + (fl < 0.0 ? fl : 0.0);
    a[ji[f]] = -de - (fl > 0.0 ? fl : 0.0);
}

// calcuvw diagonal (src/calcuvw.f90:417-432) with spu from the wall faces
__global__ void k_convdiff_rows(int n, const int *__restrict__ ptr, const int *__restrict__ diagpos,
                                const int *__restrict__ cw_ptr, const int *__restrict__ cw_idx, int wall_face0,
                                const double *__restrict__ arx, const double *__restrict__ ary,
                                const double *__restrict__ arz, const double *__restrict__ xc,
                                const double *__restrict__ yc, const double *__restrict__ zc,
                                const double *__restrict__ xf, const double *__restrict__ yf,
                                const double *__restrict__ zf, const double *__restrict__ vis, double urfr,
                                double *__restrict__ a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int dp = diagpos[i];
    double spu = 0.0;
    if (cw_ptr) {
        for (int q = cw_ptr[i]; q < cw_ptr[i + 1]; ++q) {
            int f = wall_face0 + cw_idx[q];
            double fx = arx[f], fy = ary[f], fz = arz[f];
            double are = sqrt((fx * fx + fy * fy) + fz * fz);
            double dx = xc[i] - xf[f], dy = yc[i] - yf[f], dz = zc[i] - zf[f];
            double dpw = sqrt((dx * dx + dy * dy) + dz * dz);
            spu = spu + (vis[i] * are) / dpw;
        }
    }
    double s = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) s = s + (k == dp ? 0.0 : a[k]);
    s = s - 0.0;
    double dd = spu - s;
    a[dp] = dd * urfr;
}

__global__ void k_face_flux(int nif, const int *__restrict__ own, const int *__restrict__ nei,
                            const double *__restrict__ facint, const double *__restrict__ arx,
                            const double *__restrict__ ary, const double *__restrict__ arz,
                            const double *__restrict__ u, const double *__restrict__ v,
                            const double *__restrict__ w, double *__restrict__ flux) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int p = own[f], q = nei[f];
    double fxn = facint[f], fxp = 1.0 - facint[f];
    double uf = fxp * u[p] + fxn * u[q];
    double vf = fxp * v[p] + fxn * v[q];
    double wf = fxp * w[p] + fxn * w[q];
    flux[f] = (uf * arx[f] + vf * ary[f]) + wf * arz[f];
}

// su(o) -= F, su(n) += F accumulated per cell in face order
__global__ void k_rhs_divergence(int n, const int *__restrict__ ptr, const int *__restrict__ diagpos,
                                 const int *__restrict__ kord, const int *__restrict__ kface,
                                 const double *__restrict__ flux, double *__restrict__ su) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int dp = diagpos[i];
    double s = su[i];
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) {
        int k = kord ? kord[p] : p;
        if (k == dp) continue;
        const int e = kface[k], f = e >> 1;
        if (!(e & 1)) s = s - flux[f];     // row i is the owner of face f
        else s = s + flux[f];
    }
    su[i] = s;
}

// ----------------------------------------------- pressure-correction assembly

// face_value_central (src/interpolation.f90:145-197); g holds 3 components per cell
__device__ inline double face_value_central(int p, int q, double xf, double yf, double zf,
                                            const double *__restrict__ fi, const double *__restrict__ g,
                                            const RhieChowDev &F) {
    const double xcp = F.xc[p], ycp = F.yc[p], zcp = F.zc[p];
    const double xcn = F.xc[q], ycn = F.yc[q], zcn = F.zc[q];
    const double *gp = g + 3 * (size_t)p, *gn = g + 3 * (size_t)q;
    const double gradfidr = ((((gp[0] * (xf - xcp) + gp[1] * (yf - ycp)) + gp[2] * (zf - zcp)) + gn[0] * (xf - xcn)) +
                             gn[1] * (yf - ycn)) +
                            gn[2] * (zf - zcn);
    return 0.5 * ((fi[p] + fi[q]) + gradfidr);
}

// facefluxmass (src/facefluxmass.f90:34-199): cap (= can) and the Rhie-Chow mass
// flux of the face between cells p and q (0-based), every expression in the
// reference's evaluation order
__device__ inline void flux_mass(int p, int q, double xf, double yf, double zf, double arx, double ary, double arz,
                                 double lambda, const RhieChowDev &F, double &cap, double &fluxmass) {
    const double fxn = lambda;
    const double fxp = 1.0 - lambda;
    const double xcp = F.xc[p], ycp = F.yc[p], zcp = F.zc[p];
    const double xcn = F.xc[q], ycn = F.yc[q], zcn = F.zc[q];
    const double xpn = xcn - xcp, ypn = ycn - ycp, zpn = zcn - zcp;
    const double are = sqrt((arx * arx + ary * ary) + arz * arz);
    const double nxx = arx / are, nyy = ary / are, nzz = arz / are;
    const double vn = F.vol[q], vp = F.vol[p];
    const double Dpu = (fxn * vn) * F.apu[q] + (fxp * vp) * F.apu[p];
    const double Dpv = (fxn * vn) * F.apv[q] + (fxp * vp) * F.apv[p];
    const double Dpw = (fxn * vn) * F.apw[q] + (fxp * vp) * F.apw[p];
    const double dene = F.den[p] * fxp + F.den[q] * fxn;
    const double sfdpnr = 1.0 / (((arx * xpn) * nxx + (ary * ypn) * nyy) + (arz * zpn) * nzz);
    const double smdpn = ((arx * arx + ary * ary) + arz * arz) * sfdpnr;
    cap = -((dene * Dpu) * smdpn);
    const double ui = face_value_central(p, q, xf, yf, zf, F.u, F.dudx, F);
    const double vi = face_value_central(p, q, xf, yf, zf, F.v, F.dvdx, F);
    const double wi = face_value_central(p, q, xf, yf, zf, F.w, F.dwdx, F);
    const double *gpp = F.dpdx + 3 * (size_t)p, *gpn = F.dpdx + 3 * (size_t)q;
    const double dpxi = ((Dpu * (fxn * gpn[0] + fxp * gpp[0])) * xpn) * nxx;
    const double dpyi = ((Dpv * (fxn * gpn[1] + fxp * gpp[1])) * ypn) * nyy;
    const double dpzi = ((Dpw * (fxn * gpn[2] + fxp * gpp[2])) * zpn) * nzz;
    const double xpp = (xf - (xf - xcp) * nxx) - xcp, ypp = (yf - (yf - ycp) * nyy) - ycp,
                 zpp = (zf - (zf - zcp) * nzz) - zcp;
    const double xep = (xf - (xf - xcn) * nxx) - xcn, yep = (yf - (yf - ycn) * nyy) - ycn,
                 zep = (zf - (zf - zcn) * nzz) - zcn;
    // as the reference writes it: only the x term of the owner side is subtracted
    const double dpecorr = ((((gpn[0] * xep + gpn[1] * yep) + gpn[2] * zep) - gpp[0] * xpp) + gpp[1] * ypp) +
                           gpp[2] * zpp;
    const double dpe = (F.p[q] - F.p[p]) + dpecorr;
    const double dpex = ((Dpu * dpe) * sfdpnr) * arx;
    const double dpey = ((Dpv * dpe) * sfdpnr) * ary;
    const double dpez = ((Dpw * dpe) * sfdpnr) * arz;
    const double ue = (ui - dpex) + dpxi;
    const double ve = (vi - dpey) + dpyi;
    const double we = (wi - dpez) + dpzi;
    fluxmass = dene * ((ue * arx + ve * ary) + we * arz);
}

// facefluxmass_piso (src/facefluxmass.f90:294-358): cap (= can) =
// -rho_f (fxp Vol_P apu_P + fxn Vol_N apu_N) |S| / |d_PN| and the mass flux rho_f (u_f . S)
// without the Rhie-Chow pressure terms, in the reference's evaluation order
__device__ inline void flux_mass_piso(int p, int q, double xf, double yf, double zf, double arx, double ary,
                                      double arz, double lambda, const RhieChowDev &F, double &cap,
                                      double &fluxmass) {
    const double fxn = lambda;
    const double fxp = 1.0 - lambda;
    const double xpn = F.xc[q] - F.xc[p], ypn = F.yc[q] - F.yc[p], zpn = F.zc[q] - F.zc[p];
    const double dpn = sqrt((xpn * xpn + ypn * ypn) + zpn * zpn);
    const double are = sqrt((arx * arx + ary * ary) + arz * arz);
    const double dene = F.den[p] * fxp + F.den[q] * fxn;
    cap = -(((dene * ((fxp * F.vol[p]) * F.apu[p] + (fxn * F.vol[q]) * F.apu[q])) * are) / dpn);
    const double ui = face_value_central(p, q, xf, yf, zf, F.u, F.dudx, F);
    const double vi = face_value_central(p, q, xf, yf, zf, F.v, F.dvdx, F);
    const double wi = face_value_central(p, q, xf, yf, zf, F.w, F.dwdx, F);
    fluxmass = dene * ((ui * arx + vi * ary) + wi * arz);
}

template <bool PISO>
__global__ void k_calcp_faces(int nif, const int *__restrict__ own, const int *__restrict__ nei,
                              const int *__restrict__ ij, const int *__restrict__ ji, const double *__restrict__ xf,
                              const double *__restrict__ yf, const double *__restrict__ zf,
                              const double *__restrict__ arx, const double *__restrict__ ary,
                              const double *__restrict__ arz, const double *__restrict__ facint, RhieChowDev F,
                              double *__restrict__ a, double *__restrict__ flmass) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    double cap, fl;
    if (PISO)
        flux_mass_piso(own[f], nei[f], xf[f], yf[f], zf[f], arx[f], ary[f], arz[f], facint[f], F, cap, fl);
    else
        flux_mass(own[f], nei[f], xf[f], yf[f], zf[f], arx[f], ary[f], arz[f], facint[f], F, cap, fl);
    a[ij[f]] = cap;   // can
    a[ji[f]] = cap;   // cap
    flmass[f] = fl;
}

template <bool PISO>
__global__ void k_calcp_oc(int noc, int oc_face0, const int *__restrict__ ijl, const int *__restrict__ ijr,
                           const double *__restrict__ foc, const double *__restrict__ xf,
                           const double *__restrict__ yf, const double *__restrict__ zf,
                           const double *__restrict__ arx, const double *__restrict__ ary,
                           const double *__restrict__ arz, RhieChowDev F, double *__restrict__ al,
                           double *__restrict__ ar, double *__restrict__ fmoc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= noc) return;
    const int f = oc_face0 + i;
    double cap, fl;
    if (PISO)
        flux_mass_piso(ijl[i], ijr[i], xf[f], yf[f], zf[f], arx[f], ary[f], arz[f], foc[i], F, cap, fl);
    else
        flux_mass(ijl[i], ijr[i], xf[f], yf[f], zf[f], arx[f], ary[f], arz[f], foc[i], F, cap, fl);
    al[i] = cap;
    ar[i] = cap;
    fmoc[i] = fl;
}

// Row i's diagonal, 0 - (the row's off-diagonals in face order) - (O-C cuts), and
// its continuity source, 0 -/+ F in face order then the cuts' -/+ fmoc: the
// values the reference's face loops leave in a(diag(i)) and su(i).
__global__ void k_calcp_rows(int n, const int *__restrict__ ptr, const int *__restrict__ diagpos,
                             const int *__restrict__ kord, const int *__restrict__ kface,
                             const double *__restrict__ flmass, const int *__restrict__ co_ptr,
                             const int *__restrict__ co_idx, const double *__restrict__ al,
                             const double *__restrict__ ar, const double *__restrict__ fmoc, double *__restrict__ a,
                             double *__restrict__ su) {
    // The wave's 64 rows are consecutive, so their slots are one contiguous range:
    // with no reordering it is staged in LDS (slot values, face codes and the
    // gathered face fluxes) with coalesced loads, and each lane then walks its row
    // there in the same order.
    __shared__ double sa[kBlock / 64][kRowStage], sv[kBlock / 64][kRowStage];
    __shared__ int se[kBlock / 64][kRowStage];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r0 = (int)(blockIdx.x * blockDim.x) + 64 * w;
    const int i = r0 + lane;
    const int s0 = r0 < n ? ptr[r0] : 0;
    const int cnt = r0 < n ? ptr[min(r0 + 64, n)] - s0 : 0;
    const bool staged = !kord && cnt <= kRowStage;
    if (staged) {
        for (int t = lane; t < cnt; t += 64) {
            const int e = kface[s0 + t];
            sa[w][t] = a[s0 + t];
            se[w][t] = e;
            sv[w][t] = e >= 0 ? flmass[e >> 1] : 0.0;
        }
    }
    __syncthreads();
    if (i >= n) return;
    const int dp = diagpos[i];
    double d = 0.0, s = 0.0;
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
        const int k = kord ? kord[q] : q;
        if (k == dp) continue;
        const int e = staged ? se[w][k - s0] : kface[k];
        const double fv = staged ? sv[w][k - s0] : flmass[e >> 1];
        d = d - (staged ? sa[w][k - s0] : a[k]);
        if (!(e & 1)) s = s - fv;   // row i is the owner of face e >> 1
        else s = s + fv;
    }
    if (co_ptr) {
        for (int q = co_ptr[i]; q < co_ptr[i + 1]; ++q) {
            const int code = co_idx[q];
            const int o = code >> 1;
            if (code & 1) {   // i is the right cell ijr of cut o
                d = d - al[o];
                s = s + fmoc[o];
            } else {
                d = d - ar[o];
                s = s - fmoc[o];
            }
        }
    }
    a[dp] = d;
    su[i] = s;
}

// One velocity component's rows of calcuvw (src/calcuvw.f90:417-437): the
// diagonal slot counts as zero in the row sum (the reference resets it), then
// a(diag) = (sp - (sum(row) - 0)) * urfr, su = src + (urfm a(diag)) phi,
// ap = 1 / (a(diag) + small), in the reference's order.
__global__ void k_momentum_rows(int n, const int *__restrict__ ptr, const int *__restrict__ diagpos,
                                double *__restrict__ a, const double *__restrict__ sp, const double *__restrict__ src,
                                const double *__restrict__ phi, double urfr, double urfm, double small,
                                double *__restrict__ su, double *__restrict__ ap) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int dp = diagpos[i];
    double s = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) s = s + (k == dp ? 0.0 : a[k]);
    const double off = s - 0.0;
    const double d = (sp[i] - off) * urfr;
    a[dp] = d;
    su[i] = src[i] + (urfm * d) * phi[i];
    ap[i] = 1.0 / (d + small);
}

// ------------------------------------------- corrector stage of the pressure correction

// fluxmc (src/facefluxmass.f90:520-607): the non-orthogonal mass flux correction
__device__ inline double fluxmc(int p, int q, double xf, double yf, double zf, double arx, double ary, double arz,
                               double lambda, const double *__restrict__ xc, const double *__restrict__ yc,
                               const double *__restrict__ zc, const double *__restrict__ vol,
                               const double *__restrict__ den, const double *__restrict__ apu,
                               const double *__restrict__ dpdx) {
    const double fxn = lambda;
    const double fxp = 1.0 - lambda;
    const double xcp = xc[p], ycp = yc[p], zcp = zc[p], xcn = xc[q], ycn = yc[q], zcn = zc[q];
    const double xpn = xcn - xcp, ypn = ycn - ycp, zpn = zcn - zcp;
    const double are = sqrt((arx * arx + ary * ary) + arz * arz);
    const double nxx = arx / are, nyy = ary / are, nzz = arz / are;
    const double dppnnr = 1.0 / (((xpn * nxx) + (ypn * nyy)) + (zpn * nzz));
    const double xpp = (xf - (xf - xcp) * nxx) - xcp, ypp = (yf - (yf - ycp) * nyy) - ycp,
                 zpp = (zf - (zf - zcp) * nzz) - zcp;
    const double xep = (xf - (xf - xcn) * nxx) - xcn, yep = (yf - (yf - ycn) * nyy) - ycn,
                 zep = (zf - (zf - zcn) * nzz) - zcn;
    const double rapr = ((apu[p] * den[p]) * vol[p]) * fxp + ((apu[q] * den[q]) * vol[q]) * fxn;
    const double *gp = dpdx + 3 * (size_t)p, *gn = dpdx + 3 * (size_t)q;
    return ((rapr * are) * (((gn[0] * xep - gp[0] * xpp) + (gn[1] * yep - gp[1] * ypp)) + (gn[2] * zep - gp[2] * zpp))) *
           dppnnr;
}

__global__ void k_nonorth_faces(int nif, const int *__restrict__ own, const int *__restrict__ nei,
                                const double *__restrict__ xf, const double *__restrict__ yf,
                                const double *__restrict__ zf, const double *__restrict__ arx,
                                const double *__restrict__ ary, const double *__restrict__ arz,
                                const double *__restrict__ facint, const double *__restrict__ xc,
                                const double *__restrict__ yc, const double *__restrict__ zc,
                                const double *__restrict__ vol, const double *__restrict__ den,
                                const double *__restrict__ apu, const double *__restrict__ dpdx,
                                double *__restrict__ fmcor, double *__restrict__ flmass) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    const double fm = fluxmc(own[f], nei[f], xf[f], yf[f], zf[f], arx[f], ary[f], arz[f], facint[f], xc, yc, zc, vol,
                             den, apu, dpdx);
    fmcor[f] = fm;
    flmass[f] = flmass[f] + fm;
}

__global__ void k_nonorth_oc(int noc, int oc_face0, const int *__restrict__ ijl, const int *__restrict__ ijr,
                             const double *__restrict__ foc, const double *__restrict__ xf,
                             const double *__restrict__ yf, const double *__restrict__ zf,
                             const double *__restrict__ arx, const double *__restrict__ ary,
                             const double *__restrict__ arz, const double *__restrict__ xc,
                             const double *__restrict__ yc, const double *__restrict__ zc,
                             const double *__restrict__ vol, const double *__restrict__ den,
                             const double *__restrict__ apu, const double *__restrict__ dpdx,
                             double *__restrict__ fmcor_oc, double *__restrict__ fmoc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= noc) return;
    const int f = oc_face0 + i;
    const double fm = fluxmc(ijl[i], ijr[i], xf[f], yf[f], zf[f], arx[f], ary[f], arz[f], foc[i], xc, yc, zc, vol,
                             den, apu, dpdx);
    fmcor_oc[i] = fm;
    fmoc[i] = fmoc[i] + fm;
}

// su(i) = 0 -/+ fmcor over the row's faces in face order, then the O-C cuts
__global__ void k_face_source_rows(int n, const int *__restrict__ ptr, const int *__restrict__ diagpos,
                                   const int *__restrict__ kord, const int *__restrict__ kface,
                                   const double *__restrict__ fm,
                                   const int *__restrict__ co_ptr, const int *__restrict__ co_idx,
                                   const double *__restrict__ fm_oc, double *__restrict__ su) {
    // (staging the slots in LDS as k_calcp_rows does made this kernel slower: with
    // one gathered value per slot the strided index loads are not what bounds it)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int dp = diagpos[i];
    double s = 0.0;
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
        const int k = kord ? kord[q] : q;
        if (k == dp) continue;
        const int e = kface[k], f = e >> 1;
        if (!(e & 1)) s = s - fm[f];
        else s = s + fm[f];
    }
    if (co_ptr) {
        for (int q = co_ptr[i]; q < co_ptr[i + 1]; ++q) {
            const int code = co_idx[q];
            if (code & 1) s = s + fm_oc[code >> 1];
            else s = s - fm_oc[code >> 1];
        }
    }
    su[i] = s;
}

// adjustMassFlow (src/adjustMassFlow.f90:29-47): outlet flux den (u.S) from the owner's
// velocity, the value the boundary velocity is extrapolated to
__global__ void k_outlet_flux(int nout, int out_face0, const int *__restrict__ owner, const double *__restrict__ arx,
                              const double *__restrict__ ary, const double *__restrict__ arz,
                              const double *__restrict__ den, const double *__restrict__ u,
                              const double *__restrict__ v, const double *__restrict__ w,
                              double *__restrict__ fmo) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nout) return;
    const int f = out_face0 + i, c = owner[f];
    fmo[i] = den[c] * ((u[c] * arx[f] + v[c] * ary[f]) + w[c] * arz[f]);
}

// flowo = sum of fmo in outlet order, fac = flomas / (flowo + small).  One wave.  The
// values pass through LDS in chunks of kOutChunk: all lanes load the next chunk (coalesced,
// every load unconditional at a clamped index, so the loads stay in flight) while lane 0 adds
// the current chunk's values one by one, in outlet order, from LDS; the adds are the only
// dependent chain.
constexpr int kOutChunk = 2048;
__global__ void __launch_bounds__(64) k_outlet_factor(int nout, const double *__restrict__ fmo, double flomas,
                                                      double *__restrict__ fac) {
    constexpr int PER = kOutChunk / 64;
    __shared__ double buf[2][kOutChunk];
    const int lane = threadIdx.x;
    const int last = max(nout - 1, 0);
    double stage[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) stage[k] = fmo[min(k * 64 + lane, last)];
#pragma unroll
    for (int k = 0; k < PER; ++k) buf[0][k * 64 + lane] = stage[k];
    __syncthreads();
    double flowo = 0.0;
    for (int c0 = 0, cb = 0; c0 < nout; c0 += kOutChunk, cb ^= 1) {
        const int c1 = c0 + kOutChunk;
#pragma unroll
        for (int k = 0; k < PER; ++k) stage[k] = fmo[min(c1 + k * 64 + lane, last)];
        if (lane == 0) {
            const int cnt = min(kOutChunk, nout - c0);
            const double *b = buf[cb];
            int j = 0;
            for (; j + 32 <= cnt; j += 32) {   // 32 reads in flight, then the 32 adds
                double v[32];
#pragma unroll
                for (int t = 0; t < 32; ++t) v[t] = b[j + t];
#pragma unroll
                for (int t = 0; t < 32; ++t) flowo = flowo + v[t];
            }
            for (; j < cnt; ++j) flowo = flowo + b[j];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) buf[cb ^ 1][k * 64 + lane] = stage[k];
        __syncthreads();
    }
    if (lane == 0) fac[0] = flomas / (flowo + kSmallSrc);
}

// :52-69 without the source: fmo *= fac, boundary velocities u(ijb) = u(owner) fac
__global__ void k_outlet_scale(int nout, int out_face0, int out_val0, const int *__restrict__ owner,
                               const double *__restrict__ fac, double *__restrict__ fmo, double *__restrict__ u,
                               double *__restrict__ v, double *__restrict__ w) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nout) return;
    const double k = fac[0];
    const int c = owner[out_face0 + i], b = out_val0 + i;
    fmo[i] = fmo[i] * k;
    u[b] = u[c] * k;
    v[b] = v[c] * k;
    w[b] = w[c] * k;
}

// correctBoundaryConditionsVelocity's symmetry loop
// (src/correctBoundaryConditionsVelocity.f90:63-75): the boundary velocity is the owner's
// minus (u . S) S, with S the face's area vector as the reference writes it
__global__ void k_symmetry_bc(int nsym, int sym_face0, int sym_val0, const int *__restrict__ owner,
                              const double *__restrict__ arx, const double *__restrict__ ary,
                              const double *__restrict__ arz, double *__restrict__ u, double *__restrict__ v,
                              double *__restrict__ w) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsym) return;
    const int f = sym_face0 + i, c = owner[f], b = sym_val0 + i;
    const double un = (u[c] * arx[f] + v[c] * ary[f]) + w[c] * arz[f];
    u[b] = u[c] - un * arx[f];
    v[b] = v[c] - un * ary[f];
    w[b] = w[c] - un * arz[f];
}

// bpres (src/bpres.f90:3-152) on p (cells, then boundary values in group order).  Item t
// of the groups' sequence (inlet, outlet, symmetry, wall, pressure outlet), then the O-C
// cuts.  Stage 1: every boundary value takes its owner cell's value.  Stage >= 2: the wall
// and pressure-outlet values, and the O-C values, are extrapolated from the owner cell
// with dpdx (3, numCells): ((p_P + g_x dx) + g_y dy) + g_z dz, d = x_f - x_P; the O-C
// branch takes the face centre of face number i, the cut's own index (bpres.f90:140-142),
// as the reference does.
__global__ void k_bpres(GradGroups G, int noc, int oc_face0, int oc_val0, const int *__restrict__ owner,
                        const double *__restrict__ xc, const double *__restrict__ yc, const double *__restrict__ zc,
                        const double *__restrict__ xf, const double *__restrict__ yf, const double *__restrict__ zf,
                        const double *__restrict__ dpdx, int stage, double *__restrict__ p) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nq = G.qstart[5];
    if (t >= nq + noc) return;
    int face, b, cf;
    bool extrapolate;
    if (t < nq) {
        int g = 0;
        while (t >= G.qstart[g + 1]) ++g;
        const int i = t - G.qstart[g];
        face = G.fstart[g] + i;
        b = G.vstart[g] + i;
        cf = face;
        extrapolate = g >= 3;        // wall, pressure outlet
    } else {
        const int i = t - nq;
        face = oc_face0 + i;
        b = oc_val0 + i;
        cf = i;
        extrapolate = true;
    }
    const int c = owner[face];
    if (stage <= 1) {
        p[b] = p[c];
        return;
    }
    if (!extrapolate) return;
    const double xpb = xf[cf] - xc[c], ypb = yf[cf] - yc[c], zpb = zf[cf] - zc[c];
    const double *g3 = dpdx + 3 * (size_t)c;
    p[b] = ((p[c] + g3[0] * xpb) + g3[1] * ypb) + g3[2] * zpb;
}

// The reference pressure of PISO (src/PISO_multiple_correction.f90:188-192): row pref of a
// zeroed, its diagonal 1, su(pref) = p(pref)
__global__ void k_pin_reference(const int *__restrict__ ptr, const int *__restrict__ diagpos, int pref,
                                const double *__restrict__ p, double *__restrict__ a, double *__restrict__ su) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int k = ptr[pref]; k < ptr[pref + 1]; ++k) a[k] = 0.0;
    a[diagpos[pref]] = 1.0;
    su[pref] = p[pref];
}

// su(i) -= fmi over the cell's inlet faces in order, then -= fmo over its outlet faces
__global__ void k_adjust_rows(int n, int ninl, const int *__restrict__ cb_ptr, const int *__restrict__ cb_q,
                              const double *__restrict__ fmi, const double *__restrict__ fmo,
                              double *__restrict__ su) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int q0 = cb_ptr[i], q1 = cb_ptr[i + 1];
    if (q0 == q1) return;
    double s = su[i];
    for (int q = q0; q < q1; ++q) {
        const int k = cb_q[q];
        s = s - (k < ninl ? fmi[k] : fmo[k - ninl]);
    }
    su[i] = s;
}

// flmass += a(ij) (pp(N) - pp(P))
__global__ void k_correct_faces(int nif, const int *__restrict__ own, const int *__restrict__ nei,
                                const int *__restrict__ ij, const double *__restrict__ a,
                                const double *__restrict__ pp, double *__restrict__ flmass) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    flmass[f] = flmass[f] + a[ij[f]] * (pp[nei[f]] - pp[own[f]]);
}

__global__ void k_correct_oc(int noc, const int *__restrict__ ijl, const int *__restrict__ ijr,
                             const double *__restrict__ ar, const double *__restrict__ pp, double *__restrict__ fmoc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= noc) return;
    fmoc[i] = fmoc[i] + ar[i] * (pp[ijr[i]] - pp[ijl[i]]);
}

// u -= dPdx Vol apu (v, w alike); p += urf (pp - pp(pref))
__global__ void k_correct_cells(int n, const double *__restrict__ pp, const double *__restrict__ dpdx,
                                const double *__restrict__ vol, const double *__restrict__ apu,
                                const double *__restrict__ apv, const double *__restrict__ apw, double urf, int pref,
                                double *__restrict__ u, double *__restrict__ v, double *__restrict__ w,
                                double *__restrict__ p) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ppref = pp[pref];
    const double *g = dpdx + 3 * (size_t)i;
    u[i] = u[i] - (g[0] * vol[i]) * apu[i];
    v[i] = v[i] - (g[1] * vol[i]) * apv[i];
    w[i] = w[i] - (g[2] * vol[i]) * apw[i];
    p[i] = p[i] + urf * (pp[i] - ppref);
}

// --------------------------------------------------------------- Gauss gradient

// item q of the boundary sequence (groups in the reference's order): its face and value index
__device__ inline void grad_group_item(const GradGroups &G, int q, int &face, int &val) {
    int g = 0;
    while (g < 4 && q >= G.qstart[g + 1]) ++g;
    face = G.fstart[g] + (q - G.qstart[g]);
    val = G.vstart[g] + (q - G.qstart[g]);
}

__global__ void k_bnd_count(int nq, GradGroups G, const int *__restrict__ owner, int n, int *__restrict__ cnt,
                            int *__restrict__ err) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int f, v;
    grad_group_item(G, q, f, v);
    const int c = owner[f];
    if (c < 0 || c >= n) { atomicOr(err, 1); return; }
    atomicAdd(&cnt[c], 1);
}

__global__ void k_bnd_fill(int nq, GradGroups G, const int *__restrict__ owner, int *__restrict__ cursor,
                           int *__restrict__ idx) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int f, v;
    grad_group_item(G, q, f, v);
    idx[atomicAdd(&cursor[owner[f]], 1)] = q;
}

// gradco (src/grad_gauss.f90:128-190): value at the face centre, extrapolated with the
// previous gradient interpolated to the point of the P-N line, times S
__device__ inline void grad_face(int p, int q, double xfc, double yfc, double zfc, double sx, double sy, double sz,
                                 double fif, const double *__restrict__ fi, const double *__restrict__ xc,
                                 const double *__restrict__ yc, const double *__restrict__ zc,
                                 const double *__restrict__ ox, const double *__restrict__ oy,
                                 const double *__restrict__ oz, double *out) {
    const double fxn = fif;
    const double fxp = 1.0 - fxn;
    const double xi = xc[p] * fxp + xc[q] * fxn;
    const double yi = yc[p] * fxp + yc[q] * fxn;
    const double zi = zc[p] * fxp + zc[q] * fxn;
    const double dfxi = ox[p] * fxp + ox[q] * fxn;
    const double dfyi = oy[p] * fxp + oy[q] * fxn;
    const double dfzi = oz[p] * fxp + oz[q] * fxn;
    const double fie = (((fi[p] * fxp + fi[q] * fxn) + dfxi * (xfc - xi)) + dfyi * (yfc - yi)) + dfzi * (zfc - zi);
    out[0] = fie * sx;
    out[1] = fie * sy;
    out[2] = fie * sz;
}

__global__ void k_grad_faces(int nif, const int *__restrict__ own, const int *__restrict__ nei,
                             const double *__restrict__ xf, const double *__restrict__ yf,
                             const double *__restrict__ zf, const double *__restrict__ arx,
                             const double *__restrict__ ary, const double *__restrict__ arz,
                             const double *__restrict__ facint, const double *__restrict__ xc,
                             const double *__restrict__ yc, const double *__restrict__ zc,
                             const double *__restrict__ u, const double *__restrict__ ox,
                             const double *__restrict__ oy, const double *__restrict__ oz, double *__restrict__ gf) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    grad_face(own[f], nei[f], xf[f], yf[f], zf[f], arx[f], ary[f], arz[f], facint[f], u, xc, yc, zc, ox, oy, oz,
              gf + 3 * (size_t)f);
}

__global__ void k_grad_oc(int noc, const int *__restrict__ ocface, const int *__restrict__ ijl,
                          const int *__restrict__ ijr, const double *__restrict__ foc, const double *__restrict__ xf,
                          const double *__restrict__ yf, const double *__restrict__ zf,
                          const double *__restrict__ arx, const double *__restrict__ ary,
                          const double *__restrict__ arz, const double *__restrict__ xc,
                          const double *__restrict__ yc, const double *__restrict__ zc,
                          const double *__restrict__ u, const double *__restrict__ ox, const double *__restrict__ oy,
                          const double *__restrict__ oz, double *__restrict__ gc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= noc) return;
    const int f = ocface[i];
    grad_face(ijl[i], ijr[i], xf[f], yf[f], zf[f], arx[f], ary[f], arz[f], foc[i], u, xc, yc, zc, ox, oy, oz,
              gc + 3 * (size_t)i);
}

// One cell's gradient, summed as grad_gauss sums it: inner faces in face order (+ on the
// owner side, - on the neighbour side), then O-C cuts, then boundary faces in group order,
// then times 1/vol.
__global__ void k_grad_rows(int n, const int *__restrict__ ptr, const int *__restrict__ diagpos,
                            const int *__restrict__ kord, const int *__restrict__ kface,
                            const double *__restrict__ gf, const int *__restrict__ co_ptr,
                            const int *__restrict__ co_idx, const double *__restrict__ gc,
                            const int *__restrict__ cb_ptr, const int *__restrict__ cb_q, GradGroups G,
                            const double *__restrict__ arx, const double *__restrict__ ary,
                            const double *__restrict__ arz, const double *__restrict__ u,
                            const double *__restrict__ vol, double *__restrict__ dudx, double *__restrict__ dudy,
                            double *__restrict__ dudz) {
    // (staging the slots in LDS as k_calcp_rows does made this kernel slower: the
    // 57 KB per block left 8 waves per CU for the scattered face-product gathers)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int dp = diagpos[i];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int p = ptr[i]; p < ptr[i + 1]; ++p) {
        const int k = kord ? kord[p] : p;
        if (k == dp) continue;
        const int e = kface[k];
        const double *g = gf + 3 * (size_t)(e >> 1);
        if (!(e & 1)) {   // row i is the owner of face f
            sx = sx + g[0];
            sy = sy + g[1];
            sz = sz + g[2];
        } else {
            sx = sx - g[0];
            sy = sy - g[1];
            sz = sz - g[2];
        }
    }
    if (co_ptr) {
        for (int q = co_ptr[i]; q < co_ptr[i + 1]; ++q) {
            const int code = co_idx[q];
            const double *g = gc + 3 * (size_t)(code >> 1);
            if (code & 1) {   // right cell of the cut
                sx = sx - g[0];
                sy = sy - g[1];
                sz = sz - g[2];
            } else {
                sx = sx + g[0];
                sy = sy + g[1];
                sz = sz + g[2];
            }
        }
    }
    if (cb_ptr) {
        for (int q = cb_ptr[i]; q < cb_ptr[i + 1]; ++q) {
            int f, v;
            grad_group_item(G, cb_q[q], f, v);
            sx = sx + u[v] * arx[f];
            sy = sy + u[v] * ary[f];
            sz = sz + u[v] * arz[f];
        }
    }
    const double volr = 1.0 / vol[i];
    dudx[i] = sx * volr;
    dudy[i] = sy * volr;
    dudz[i] = sz * volr;
}

// kface[k] = (f << 1) | side: the face of slot k, side 0 in the owner's row, 1 in the
// neighbour's (diagonal slots: -1); the order of the values is the order of the faces
__global__ void k_kface(int nif, const int *__restrict__ ij, const int *__restrict__ ji, int *__restrict__ kface) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    kface[ij[f]] = f << 1;
    kface[ji[f]] = (f << 1) | 1;
}

__global__ void k_face_order_check(int n, const int *__restrict__ ptr, const int *__restrict__ diagpos,
                                   const int *__restrict__ kface, int *__restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int last = -1;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
        if (k == diagpos[i]) continue;
        if (kface[k] < last) { atomicOr(flag, 1); return; }
        last = kface[k];
    }
}

// per row: entry positions ordered by face id (diagonal last)
__global__ void k_kord(int n, const int *__restrict__ ptr, const int *__restrict__ diagpos,
                       const int *__restrict__ kface, int *__restrict__ kord) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = ptr[i], e = ptr[i + 1];
    for (int k = s; k < e; ++k) kord[k] = k;
    auto key = [&](int k) { return k == diagpos[i] ? 0x7fffffff : kface[k]; };
    for (int k = s + 1; k < e; ++k) {
        int v = kord[k], kv = key(v);
        int m = k - 1;
        while (m >= s && key(kord[m]) > kv) { kord[m + 1] = kord[m]; --m; }
        kord[m + 1] = v;
    }
}

// Caller-supplied faces and slot maps are used as indices: check them first.
__global__ void k_check_faces(int nif, int n, int nnz, const int *__restrict__ own, const int *__restrict__ nei,
                              const int *__restrict__ ij, const int *__restrict__ ji, const int *__restrict__ ptr,
                              const int *__restrict__ ja, int *__restrict__ err) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nif) return;
    int o = own[f], q = nei[f], kij = ij[f], kji = ji[f];
    if (o < 0 || o >= n || q < 0 || q >= n) { atomicOr(err, 1); return; }
    if (kij < ptr[o] || kij >= ptr[o + 1] || kji < ptr[q] || kji >= ptr[q + 1] || kij >= nnz || kji >= nnz) {
        atomicOr(err, 2);
        return;
    }
    if (ja[kij] != q || ja[kji] != o) atomicOr(err, 2);
}

__global__ void k_wall_count(int nwal, int wall_face0, int n, const int *__restrict__ own, int *__restrict__ cnt,
                             int *__restrict__ err) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwal) return;
    int c = own[wall_face0 + w];
    if (c < 0 || c >= n) { atomicOr(err, 1); return; }
    atomicAdd(&cnt[c], 1);
}

__global__ void k_wall_fill(int nwal, int wall_face0, const int *__restrict__ own, int *__restrict__ cursor,
                            int *__restrict__ idx) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwal) return;
    idx[atomicAdd(&cursor[own[wall_face0 + w]], 1)] = w;
}

__global__ void k_oc_count(int noc, const int *__restrict__ ijl, const int *__restrict__ ijr, int *__restrict__ cnt) {
    int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= noc) return;
    atomicAdd(&cnt[ijl[o]], 1);
    atomicAdd(&cnt[ijr[o]], 1);
}

__global__ void k_oc_fill(int noc, const int *__restrict__ ijl, const int *__restrict__ ijr, int *__restrict__ cursor,
                          int *__restrict__ idx) {
    int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= noc) return;
    idx[atomicAdd(&cursor[ijl[o]], 1)] = 2 * o;       // left cell: subtract ar
    idx[atomicAdd(&cursor[ijr[o]], 1)] = 2 * o + 1;   // right cell: subtract al
}

__global__ void k_copy(const int *__restrict__ in, int *__restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

}  // namespace

void build_wall_lists(fc_context *ctx, WallDev &w, const int *owner0_d, int n, int nwal, int iwall_faces_start,
                      int iwall_start) {
    w.nwal = nwal;
    w.iwall_faces_start = iwall_faces_start;
    w.iwall_start = iwall_start;
    if (nwal <= 0) { w.cw_ptr.release(); w.cw_idx.release(); return; }
    hipStream_t s = ctx->stream;
    DevBuf cnt(sizeof(int) * (n + 1)), cursor(sizeof(int) * (n + 1)), err(sizeof(int));
    FC_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(int) * (n + 1), s));
    FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), s));
    k_wall_count<<<grid_for(nwal), kBlock, 0, s>>>(nwal, iwall_faces_start, n, owner0_d, cnt.as<int>(),
                                                   err.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    int herr = 0;
    d2h(&herr, err.get(), sizeof(int), s);
    FC_REQUIRE(herr == 0, FC_ERR_STRUCTURE, "wall face owner out of range");
    w.cw_ptr.alloc(sizeof(int) * (n + 1));
    scan_dev(ctx, cnt.as<int>(), w.cw_ptr.as<int>(), n + 1);
    w.cw_idx.alloc(sizeof(int) * nwal);
    k_copy<<<grid_for(n), kBlock, 0, s>>>(w.cw_ptr.as<int>(), cursor.as<int>(), n);
    k_wall_fill<<<grid_for(nwal), kBlock, 0, s>>>(nwal, iwall_faces_start, owner0_d, cursor.as<int>(),
                                                  w.cw_idx.as<int>());
    sort_segments_dev(ctx, w.cw_ptr.as<int>(), w.cw_idx.as<int>(), n);
    FC_HIP_CHECK(hipGetLastError());
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

__global__ void k_check_range(int m, const int *__restrict__ v, int n, int *__restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && (v[i] < 0 || v[i] >= n)) atomicOr(err, 1);
}

void validate_faces(fc_context *ctx, const Structure &st, int nif, const int *own, const int *nei, const int *ij,
                    const int *ji) {
    if (nif <= 0) return;
    hipStream_t s = ctx->stream;
    DevBuf err(sizeof(int));
    FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), s));
    k_check_faces<<<grid_for(nif), kBlock, 0, s>>>(nif, st.n, st.nnz, own, nei, ij, ji, st.ioffset.as<int>(),
                                                   st.ja.as<int>(), err.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    int h = 0;
    d2h(&h, err.get(), sizeof(int), s);
    FC_REQUIRE(!(h & 1), FC_ERR_STRUCTURE, "owner/neighbour out of range");
    FC_REQUIRE(!(h & 2), FC_ERR_STRUCTURE, "face slot maps do not match the CSR structure");
}

void build_oc_lists(fc_context *ctx, OcDev &o, int n) {
    if (o.noc <= 0) { o.co_ptr.release(); o.co_idx.release(); return; }
    hipStream_t s = ctx->stream;
    {
        DevBuf err(sizeof(int));
        FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), s));
        k_check_range<<<grid_for(o.noc), kBlock, 0, s>>>(o.noc, o.ijl.as<int>(), n, err.as<int>());
        k_check_range<<<grid_for(o.noc), kBlock, 0, s>>>(o.noc, o.ijr.as<int>(), n, err.as<int>());
        FC_HIP_CHECK(hipGetLastError());
        int h = 0;
        d2h(&h, err.get(), sizeof(int), s);
        FC_REQUIRE(h == 0, FC_ERR_STRUCTURE, "O-C cell index out of range");
    }
    DevBuf cnt(sizeof(int) * (n + 1)), cursor(sizeof(int) * (n + 1));
    FC_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(int) * (n + 1), s));
    k_oc_count<<<grid_for(o.noc), kBlock, 0, s>>>(o.noc, o.ijl.as<int>(), o.ijr.as<int>(), cnt.as<int>());
    o.co_ptr.alloc(sizeof(int) * (n + 1));
    scan_dev(ctx, cnt.as<int>(), o.co_ptr.as<int>(), n + 1);
    o.co_idx.alloc(sizeof(int) * 2 * o.noc);
    k_copy<<<grid_for(n), kBlock, 0, s>>>(o.co_ptr.as<int>(), cursor.as<int>(), n);
    k_oc_fill<<<grid_for(o.noc), kBlock, 0, s>>>(o.noc, o.ijl.as<int>(), o.ijr.as<int>(), cursor.as<int>(),
                                                 o.co_idx.as<int>());
    sort_segments_dev(ctx, o.co_ptr.as<int>(), o.co_idx.as<int>(), n);
    FC_HIP_CHECK(hipGetLastError());
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

void build_kface(fc_context *ctx, const Structure &st, int nif, const int *ij_slot, const int *ji_slot,
                 DevBuf &kface, DevBuf &kface_order, bool &needs_order) {
    FC_REQUIRE(nif < (1 << 30), FC_ERR_ARGUMENT, "more than 2^30 inner faces");
    hipStream_t s = ctx->stream;
    kface.alloc(sizeof(int) * (size_t)st.nnz);
    FC_HIP_CHECK(hipMemsetAsync(kface.get(), 0xff, sizeof(int) * (size_t)st.nnz, s));
    if (nif) k_kface<<<grid_for(nif), kBlock, 0, s>>>(nif, ij_slot, ji_slot, kface.as<int>());
    DevBuf flag(sizeof(int));
    FC_HIP_CHECK(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
    k_face_order_check<<<grid_for(st.n), kBlock, 0, s>>>(st.n, st.ioffset.as<int>(), st.diag.as<int>(),
                                                         kface.as<int>(), flag.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    int h = 0;
    d2h(&h, flag.get(), sizeof(int), s);
    needs_order = h != 0;
    if (needs_order) {
        kface_order.alloc(sizeof(int) * (size_t)st.nnz);
        k_kord<<<grid_for(st.n), kBlock, 0, s>>>(st.n, st.ioffset.as<int>(), st.diag.as<int>(), kface.as<int>(),
                                                 kface_order.as<int>());
        FC_HIP_CHECK(hipGetLastError());
    } else {
        kface_order.release();
    }
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

void laplacian_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0, const int *neigh0,
                   const int *ij_slot, const int *ji_slot, const double *arx, const double *ary,
                   const double *arz, const double *facint, const double *xc, const double *yc,
                   const double *zc, const double *xf, const double *yf, const double *zf, const double *mu,
                   const double *phi, const WallDev &w, const OcDev &oc, const int *kface_order, double *a,
                   double *su) {
    hipStream_t s = ctx->stream;
    if (nif)
        k_laplacian_faces<<<grid_for(nif), kBlock, 0, s>>>(nif, owner0, neigh0, ij_slot, ji_slot, arx, ary, arz,
                                                           facint, xc, yc, zc, mu, a);
    if (oc.noc)
        k_laplacian_oc<<<grid_for(oc.noc), kBlock, 0, s>>>(oc.noc, oc.ioc_faces_start, oc.ijl.as<int>(),
                                                           oc.ijr.as<int>(), oc.foc.as<double>(), arx, ary, arz,
                                                           xc, yc, zc, mu, oc.al.as<double>(), oc.ar.as<double>());
    k_laplacian_rows<<<grid_for(st.n), kBlock, 0, s>>>(
        st.n, st.ioffset.as<int>(), st.diag.as<int>(), kface_order, oc.noc ? oc.co_ptr.as<int>() : nullptr,
        oc.noc ? oc.co_idx.as<int>() : nullptr, oc.noc ? oc.al.as<double>() : nullptr,
        oc.noc ? oc.ar.as<double>() : nullptr, w.nwal ? w.cw_ptr.as<int>() : nullptr,
        w.nwal ? w.cw_idx.as<int>() : nullptr, w.iwall_faces_start, w.iwall_start, arx, ary, arz, xc, yc, zc, xf,
        yf, zf, mu, phi, a, su);
    FC_HIP_CHECK(hipGetLastError());
}

void convdiff_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0, const int *neigh0,
                  const int *ij_slot, const int *ji_slot, const double *arx, const double *ary, const double *arz,
                  const double *facint, const double *xc, const double *yc, const double *zc, const double *xf,
                  const double *yf, const double *zf, const double *vis, const double *flux, const WallDev &w,
                  double urf, double *a) {
    hipStream_t s = ctx->stream;
    if (nif)
        k_convdiff_faces<<<grid_for(nif), kBlock, 0, s>>>(nif, owner0, neigh0, ij_slot, ji_slot, arx, ary, arz,
                                                          facint, xc, yc, zc, vis, flux, a);
    double urfr = 1.0 / urf;
    k_convdiff_rows<<<grid_for(st.n), kBlock, 0, s>>>(st.n, st.ioffset.as<int>(), st.diag.as<int>(),
                                                      w.nwal ? w.cw_ptr.as<int>() : nullptr,
                                                      w.nwal ? w.cw_idx.as<int>() : nullptr, w.iwall_faces_start,
                                                      arx, ary, arz, xc, yc, zc, xf, yf, zf, vis, urfr, a);
    FC_HIP_CHECK(hipGetLastError());
}

void calcp_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0, const int *neigh0,
               const int *ij_slot, const int *ji_slot, const int *kface, const int *kface_order, const double *xf,
               const double *yf, const double *zf, const double *arx, const double *ary, const double *arz,
               const double *facint, const RhieChowDev &F, const OcDev &oc, double *a, double *su, double *flmass,
               double *fmoc, bool piso) {
    hipStream_t s = ctx->stream;
    if (nif) {
        if (piso)
            k_calcp_faces<true><<<grid_for(nif), kBlock, 0, s>>>(nif, owner0, neigh0, ij_slot, ji_slot, xf, yf, zf,
                                                                 arx, ary, arz, facint, F, a, flmass);
        else
            k_calcp_faces<false><<<grid_for(nif), kBlock, 0, s>>>(nif, owner0, neigh0, ij_slot, ji_slot, xf, yf, zf,
                                                                  arx, ary, arz, facint, F, a, flmass);
    }
    if (oc.noc) {
        if (piso)
            k_calcp_oc<true><<<grid_for(oc.noc), kBlock, 0, s>>>(
                oc.noc, oc.ioc_faces_start, oc.ijl.as<int>(), oc.ijr.as<int>(), oc.foc.as<double>(), xf, yf, zf, arx,
                ary, arz, F, oc.al.as<double>(), oc.ar.as<double>(), fmoc);
        else
            k_calcp_oc<false><<<grid_for(oc.noc), kBlock, 0, s>>>(
                oc.noc, oc.ioc_faces_start, oc.ijl.as<int>(), oc.ijr.as<int>(), oc.foc.as<double>(), xf, yf, zf, arx,
                ary, arz, F, oc.al.as<double>(), oc.ar.as<double>(), fmoc);
    }
    k_calcp_rows<<<grid_for(st.n), kBlock, 0, s>>>(
        st.n, st.ioffset.as<int>(), st.diag.as<int>(), kface_order, kface, flmass,
        oc.noc ? oc.co_ptr.as<int>() : nullptr, oc.noc ? oc.co_idx.as<int>() : nullptr,
        oc.noc ? oc.al.as<double>() : nullptr, oc.noc ? oc.ar.as<double>() : nullptr, oc.noc ? fmoc : nullptr, a, su);
    FC_HIP_CHECK(hipGetLastError());
}

void build_boundary_lists(fc_context *ctx, const GradGroups &G, const int *owner0_all, int n, DevBuf &cb_ptr,
                          DevBuf &cb_q) {
    const int nq = G.qstart[5];
    if (nq <= 0) { cb_ptr.release(); cb_q.release(); return; }
    hipStream_t s = ctx->stream;
    DevBuf cnt(sizeof(int) * (n + 1)), cursor(sizeof(int) * (n + 1)), err(sizeof(int));
    FC_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(int) * (n + 1), s));
    FC_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), s));
    k_bnd_count<<<grid_for(nq), kBlock, 0, s>>>(nq, G, owner0_all, n, cnt.as<int>(), err.as<int>());
    FC_HIP_CHECK(hipGetLastError());
    int herr = 0;
    d2h(&herr, err.get(), sizeof(int), s);
    FC_REQUIRE(herr == 0, FC_ERR_STRUCTURE, "boundary face owner out of range");
    cb_ptr.alloc(sizeof(int) * (n + 1));
    scan_dev(ctx, cnt.as<int>(), cb_ptr.as<int>(), n + 1);
    cb_q.alloc(sizeof(int) * nq);
    k_copy<<<grid_for(n), kBlock, 0, s>>>(cb_ptr.as<int>(), cursor.as<int>(), n);
    k_bnd_fill<<<grid_for(nq), kBlock, 0, s>>>(nq, G, owner0_all, cursor.as<int>(), cb_q.as<int>());
    sort_segments_dev(ctx, cb_ptr.as<int>(), cb_q.as<int>(), n);
    FC_HIP_CHECK(hipGetLastError());
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

// Component arrays (x, y, z; n each) into the reference's (3, numCells) layout, component fastest.
__global__ void k_interleave3(int n, const double *__restrict__ gx, const double *__restrict__ gy,
                              const double *__restrict__ gz, double *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[3 * (size_t)i] = gx[i];
    out[3 * (size_t)i + 1] = gy[i];
    out[3 * (size_t)i + 2] = gz[i];
}

void interleave3_dev(fc_context *ctx, int n, const double *gx, const double *gy, const double *gz, double *out) {
    k_interleave3<<<grid_for(n), kBlock, 0, ctx->stream>>>(n, gx, gy, gz, out);
    FC_HIP_CHECK(hipGetLastError());
}

void grad_gauss_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0, const int *neigh0,
                    const int *kface, const int *kface_order, const double *xf, const double *yf,
                    const double *zf, const double *arx, const double *ary, const double *arz, const double *facint,
                    const double *xc, const double *yc, const double *zc, const double *vol, const double *u,
                    const GradGroups &G, const int *cb_ptr, const int *cb_q, const OcDev &oc, const int *ocface,
                    int nigrad, double *dudx, double *dudy, double *dudz, GradScratch *scratch) {
    hipStream_t s = ctx->stream;
    const int n = st.n;
    GradScratch local;
    GradScratch &S = scratch ? *scratch : local;
    S.old.ensure(sizeof(double) * 3 * (size_t)n);
    S.gf.ensure(sizeof(double) * 3 * (size_t)std::max(nif, 1));
    S.gc.ensure(sizeof(double) * 3 * (size_t)std::max(oc.noc, 1));
    DevBuf &gf = S.gf, &gc = S.gc;
    double *ox = S.old.as<double>(), *oy = ox + n, *oz = oy + n;
    FC_HIP_CHECK(hipMemsetAsync(ox, 0, sizeof(double) * 3 * (size_t)n, s));
    for (int lc = 1; lc <= nigrad; ++lc) {
        if (nif)
            k_grad_faces<<<grid_for(nif), kBlock, 0, s>>>(nif, owner0, neigh0, xf, yf, zf, arx, ary, arz, facint, xc,
                                                          yc, zc, u, ox, oy, oz, gf.as<double>());
        if (oc.noc)
            k_grad_oc<<<grid_for(oc.noc), kBlock, 0, s>>>(oc.noc, ocface, oc.ijl.as<int>(), oc.ijr.as<int>(),
                                                          oc.foc.as<double>(), xf, yf, zf, arx, ary, arz, xc, yc, zc,
                                                          u, ox, oy, oz, gc.as<double>());
        k_grad_rows<<<grid_for(n), kBlock, 0, s>>>(n, st.ioffset.as<int>(), st.diag.as<int>(), kface_order, kface,
                                                   gf.as<double>(), oc.noc ? oc.co_ptr.as<int>() : nullptr,
                                                   oc.noc ? oc.co_idx.as<int>() : nullptr, gc.as<double>(), cb_ptr,
                                                   cb_q, G, arx, ary, arz, u, vol, dudx, dudy, dudz);
        if (lc != nigrad) {
            FC_HIP_CHECK(hipMemcpyAsync(ox, dudx, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            FC_HIP_CHECK(hipMemcpyAsync(oy, dudy, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            FC_HIP_CHECK(hipMemcpyAsync(oz, dudz, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        }
    }
    FC_HIP_CHECK(hipGetLastError());
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

void calcp_nonorth_dev(fc_context *ctx, const Structure &st, int nif, const int *owner0, const int *neigh0,
                       const int *kface, const int *kface_order, const double *xf, const double *yf,
                       const double *zf, const double *arx, const double *ary, const double *arz, const double *facint,
                       const double *xc, const double *yc, const double *zc, const double *vol, const double *den,
                       const double *apu, const double *dpdx, const OcDev &oc, double *su, double *flmass,
                       double *fmoc, GradScratch *scratch) {
    hipStream_t s = ctx->stream;
    GradScratch local;
    GradScratch &S = scratch ? *scratch : local;   // gf holds fmcor per face, gc per cut
    S.gf.ensure(sizeof(double) * (size_t)std::max(nif, 1));
    S.gc.ensure(sizeof(double) * (size_t)std::max(oc.noc, 1));
    DevBuf &fm = S.gf, &fmo = S.gc;
    if (nif)
        k_nonorth_faces<<<grid_for(nif), kBlock, 0, s>>>(nif, owner0, neigh0, xf, yf, zf, arx, ary, arz, facint, xc,
                                                         yc, zc, vol, den, apu, dpdx, fm.as<double>(), flmass);
    if (oc.noc)
        k_nonorth_oc<<<grid_for(oc.noc), kBlock, 0, s>>>(oc.noc, oc.ioc_faces_start, oc.ijl.as<int>(),
                                                         oc.ijr.as<int>(), oc.foc.as<double>(), xf, yf, zf, arx, ary,
                                                         arz, xc, yc, zc, vol, den, apu, dpdx, fmo.as<double>(), fmoc);
    k_face_source_rows<<<grid_for(st.n), kBlock, 0, s>>>(st.n, st.ioffset.as<int>(), st.diag.as<int>(), kface_order,
                                                         kface, fm.as<double>(),
                                                         oc.noc ? oc.co_ptr.as<int>() : nullptr,
                                                         oc.noc ? oc.co_idx.as<int>() : nullptr, fmo.as<double>(), su);
    FC_HIP_CHECK(hipGetLastError());
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

void adjust_mass_flow_dev(fc_context *ctx, int n, const int *owner0_all, const double *arx, const double *ary,
                          const double *arz, int ninl, int inlet_face0, const double *fmi, int nout,
                          int outlet_face0, int outlet_val0, double flomas, const double *den, double *u, double *v,
                          double *w, double *su, double *fmo) {
    hipStream_t s = ctx->stream;
    GradGroups G{};
    G.qstart[0] = 0;
    G.qstart[1] = ninl;
    for (int g = 2; g < 6; ++g) G.qstart[g] = ninl + nout;
    G.fstart[0] = inlet_face0;
    G.fstart[1] = outlet_face0;
    DevBuf cb_ptr, cb_q, fac(sizeof(double));
    build_boundary_lists(ctx, G, owner0_all, n, cb_ptr, cb_q);
    if (nout > 0) {
        k_outlet_flux<<<grid_for(nout), kBlock, 0, s>>>(nout, outlet_face0, owner0_all, arx, ary, arz, den, u, v, w,
                                                        fmo);
        k_outlet_factor<<<1, 64, 0, s>>>(nout, fmo, flomas, fac.as<double>());
        k_outlet_scale<<<grid_for(nout), kBlock, 0, s>>>(nout, outlet_face0, outlet_val0, owner0_all,
                                                         fac.as<double>(), fmo, u, v, w);
    }
    if (ninl + nout > 0)
        k_adjust_rows<<<grid_for(n), kBlock, 0, s>>>(n, ninl, cb_ptr.as<int>(), cb_q.as<int>(), fmi, fmo, su);
    FC_HIP_CHECK(hipGetLastError());
    FC_HIP_CHECK(hipStreamSynchronize(s));
}

void velocity_bc_dev(fc_context *ctx, const int *owner0_all, const double *arx, const double *ary, const double *arz,
                     int nout, int outlet_face0, int outlet_val0, int nsym, int sym_face0, int sym_val0, double flomas,
                     const double *den, double *u, double *v, double *w, double *fmo, double *fac_scratch) {
    hipStream_t s = ctx->stream;
    if (nout > 0) {
        k_outlet_flux<<<grid_for(nout), kBlock, 0, s>>>(nout, outlet_face0, owner0_all, arx, ary, arz, den, u, v, w,
                                                        fmo);
        k_outlet_factor<<<1, 64, 0, s>>>(nout, fmo, flomas, fac_scratch);
        k_outlet_scale<<<grid_for(nout), kBlock, 0, s>>>(nout, outlet_face0, outlet_val0, owner0_all, fac_scratch,
                                                         fmo, u, v, w);
    }
    if (nsym > 0)
        k_symmetry_bc<<<grid_for(nsym), kBlock, 0, s>>>(nsym, sym_face0, sym_val0, owner0_all, arx, ary, arz, u, v,
                                                        w);
    FC_HIP_CHECK(hipGetLastError());
}

void bpres_dev(fc_context *ctx, const GradGroups &G, int noc, int oc_face0, int oc_val0, const int *owner0_all,
               const double *xc, const double *yc, const double *zc, const double *xf, const double *yf,
               const double *zf, const double *dpdx, int stage, double *p) {
    const int total = G.qstart[5] + noc;
    if (total <= 0) return;
    k_bpres<<<grid_for(total), kBlock, 0, ctx->stream>>>(G, noc, oc_face0, oc_val0, owner0_all, xc, yc, zc, xf, yf,
                                                         zf, dpdx, stage, p);
    FC_HIP_CHECK(hipGetLastError());
}

void pin_reference_dev(fc_context *ctx, const Structure &st, int pref0, const double *p, double *a, double *su) {
    k_pin_reference<<<1, 64, 0, ctx->stream>>>(st.ioffset.as<int>(), st.diag.as<int>(), pref0, p, a, su);
    FC_HIP_CHECK(hipGetLastError());
}

void calcp_correct_dev(fc_context *ctx, int n, int nif, const int *owner0, const int *neigh0, const int *ij_slot,
                       const double *a, const double *pp, const double *dpdx, const double *vol, const double *apu,
                       const double *apv, const double *apw, double urf, int pref0, int noc, const int *ijl,
                       const int *ijr, const double *ar, double *fmoc, double *flmass, double *u, double *v,
                       double *w, double *p) {
    hipStream_t s = ctx->stream;
    if (nif) k_correct_faces<<<grid_for(nif), kBlock, 0, s>>>(nif, owner0, neigh0, ij_slot, a, pp, flmass);
    if (noc) k_correct_oc<<<grid_for(noc), kBlock, 0, s>>>(noc, ijl, ijr, ar, pp, fmoc);
    k_correct_cells<<<grid_for(n), kBlock, 0, s>>>(n, pp, dpdx, vol, apu, apv, apw, urf, pref0, u, v, w, p);
    FC_HIP_CHECK(hipGetLastError());
}

void momentum_rows_dev(fc_context *ctx, const Structure &st, double *a, const double *sp, const double *src,
                       const double *phi, double urf, double *su, double *ap) {
    const double urfr = 1.0 / urf, urfm = 1.0 - urf;
    k_momentum_rows<<<grid_for(st.n), kBlock, 0, ctx->stream>>>(st.n, st.ioffset.as<int>(), st.diag.as<int>(), a, sp,
                                                                src, phi, urfr, urfm, kSmallSrc, su, ap);
    FC_HIP_CHECK(hipGetLastError());
}

void face_flux_dev(fc_context *ctx, int nif, const int *owner0, const int *neigh0, const double *facint,
                   const double *arx, const double *ary, const double *arz, const double *u, const double *v,
                   const double *w, double *flux) {
    if (!nif) return;
    k_face_flux<<<grid_for(nif), kBlock, 0, ctx->stream>>>(nif, owner0, neigh0, facint, arx, ary, arz, u, v, w,
                                                           flux);
    FC_HIP_CHECK(hipGetLastError());
}

void rhs_divergence_dev(fc_context *ctx, const Structure &st, const int *kface, const int *kface_order,
                        const double *flux, double *su) {
    k_rhs_divergence<<<grid_for(st.n), kBlock, 0, ctx->stream>>>(st.n, st.ioffset.as<int>(), st.diag.as<int>(),
                                                                 kface_order, kface, flux, su);
    FC_HIP_CHECK(hipGetLastError());
}

}  // namespace fcap
