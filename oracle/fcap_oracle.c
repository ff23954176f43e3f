/*
 * CPU oracle for the pressure-solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates the reference's serial algorithm (cited per function) so that the
 * GPU library can be checked against it.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it; the product never does.
 *
 * Every floating-point expression is evaluated in the order the Fortran source
 * writes it (left to right, products rounded before sums); build with
 * -ffp-contract=off.
 *
 * Pinning (which check fixes each function):
 *   orc_mesh_geometry  -- bitwise vs the reference Fortran (oracle/_ref) on
 *                         uniform, perturbed, graded and tet meshes; closed
 *                         forms on the uniform box.
 *   orc_csr_build      -- bitwise vs the reference; scipy canonical CSR.
 *   orc_laplacian      -- bitwise vs the reference; closed form on the box.
 *   orc_iccg/dpcg/bicgstab -- the 5x5 transcript tests/output.txt of the
 *                         reference; bitwise iteration histories vs the
 *                         reference on assembled mesh systems; dense solves.
 *   orc_face_flux_central / orc_rhs_divergence -- bitwise vs a numpy
 *                         evaluation and the divergence-theorem invariant
 *                         (sum of su over cells equals the net boundary flux,
 *                         zero for inner-face-only fluxes).
 *   orc_calcp_nonorth  -- fmcor bitwise vs the reference's fluxmc (the flux
 *                         harness); zero on an orthogonal box.
 *   orc_calcp_correct  -- closed forms: with pp linear and its exact gradient on
 *                         a uniform box the flux change is a_ij (pp_N - pp_P), the
 *                         velocity change -grad(pp) Vol ap, p shifts by urf (pp - pp_ref).
 *   orc_adjust_mass_flow -- bitwise vs the reference's adjustMassFlow compiled on
 *                         its own (oracle/_ref/ref_adjust_driver) on a hex box and a
 *                         tet box with inlet and outlet patches; global conservation
 *                         (outlet mass flow = flomas) to rounding.
 *   orc_grad_gauss     -- bitwise vs the reference's grad_gauss compiled on its
 *                         own (oracle/_ref/ref_grad_driver) on a box with every
 *                         boundary group and on permuted tets, 1 and 2
 *                         iterations; exact for linear fields on a uniform box.
 *   orc_momentum_rows  -- closed forms: with urf = 1 every row sums to sp and
 *                         su = s; for any urf, urf a_PP = sp - (sum of the row's
 *                         off-diagonals) and su - s = (1 - urf) a_PP phi.
 *   orc_calcp_assemble -- bitwise vs the reference's facefluxmass and calcp
 *                         face loop (oracle/_ref/ref_flux_driver) on perturbed,
 *                         graded and permuted tet meshes; a uniform-flow case
 *                         (constant u, p linear with its exact gradient) whose
 *                         flux is rho u.S in closed form.
 */
#include "fcap_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

double orc_small(void) { return (double)1e-20f; }
double orc_tol(void) { return (double)1e-13f; }
double orc_s0(void) { return (double)1e20f; }
double orc_tiny(void) { return (double)1e-30f; }

/* 1-based accessors */
#define I1(arr, k) ((arr)[(k)-1])

/* ------------------------------------------------------------------------ */
/* Geometry: src/mesh_geometry_and_topology.f90:106-274 (helpers) and
 * :859-1081 (the three face sweeps).                                         */
/* ------------------------------------------------------------------------ */

static double centroid_part(double ax, double bx, double cx, double nx, double vol)
{
    /* 1./(2*vol) * 1./24._dp*nx * ( (ax+bx)**2 + (bx+cx)**2 + (cx+ax)**2 )
     * (src/mesh_geometry_and_topology.f90:181) evaluated left to right. */
    double t = 1.0 / (2.0 * vol);
    t = t * 1.0;
    t = t / 24.0;
    t = t * nx;
    double ab = ax + bx, bc = bx + cx, ca = cx + ax;
    double s = ab * ab + bc * bc;
    s = s + ca * ca;
    return t * s;
}

static void intersection_point(double x1, double y1, double z1, double x2, double y2, double z2,
                               double x3, double y3, double z3, double x4, double y4, double z4,
                               double x5, double y5, double z5, double *xj, double *yj, double *zj)
{
    /* src/mesh_geometry_and_topology.f90:262-272 */
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
    den = den + orc_tiny();
    double t = -(num / den);
    *xj = x4 + dx * t;
    *yj = y4 + dy * t;
    *zj = z4 + dz * t;
}

int orc_mesh_geometry(int nnodes, const double *xyz, int nfaces, int nif,
                      const int *face_ptr, const int *face_nodes1, const int *owner1,
                      const int *neighbour1, int ncells,
                      double *arx, double *ary, double *arz,
                      double *xf, double *yf, double *zf,
                      double *vol, double *xc, double *yc, double *zc, double *facint)
{
    const double half = 0.5, one_third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    (void)nnodes;
    memset(arx, 0, sizeof(double) * nfaces);
    memset(ary, 0, sizeof(double) * nfaces);
    memset(arz, 0, sizeof(double) * nfaces);
    memset(vol, 0, sizeof(double) * ncells);
    memset(xc, 0, sizeof(double) * ncells);
    memset(yc, 0, sizeof(double) * ncells);
    memset(zc, 0, sizeof(double) * ncells);
#define X(nd) xyz[3 * ((nd)-1) + 0]
#define Y(nd) xyz[3 * ((nd)-1) + 1]
#define Z(nd) xyz[3 * ((nd)-1) + 2]
    /* sweep 1: face areas, volumes, face centres (:859-948) */
    for (int f = 0; f < nfaces; ++f) {
        const int *nd = face_nodes1 + face_ptr[f];
        int nn = face_ptr[f + 1] - face_ptr[f];
        int inp = owner1[f];
        for (int i = 0; i < nn - 2; ++i) {
            double px = X(nd[i + 1]) - X(nd[0]), py = Y(nd[i + 1]) - Y(nd[0]), pz = Z(nd[i + 1]) - Z(nd[0]);
            double qx = X(nd[i + 2]) - X(nd[0]), qy = Y(nd[i + 2]) - Y(nd[0]), qz = Z(nd[i + 2]) - Z(nd[0]);
            double nx = py * qz - pz * qy;
            double ny = pz * qx - px * qz;
            double nz = px * qy - py * qx;
            arx[f] = arx[f] + half * nx;
            ary[f] = ary[f] + half * ny;
            arz[f] = arz[f] + half * nz;
            double cx = one_third * ((X(nd[i + 2]) + X(nd[i + 1])) + X(nd[0]));
            double cy = one_third * ((Y(nd[i + 2]) + Y(nd[i + 1])) + Y(nd[0]));
            double cz = one_third * ((Z(nd[i + 2]) + Z(nd[i + 1])) + Z(nd[0]));
            double vp = sixth * ((cx * nx + cy * ny) + cz * nz);
            I1(vol, inp) = I1(vol, inp) + vp;
            if (f < nif) {
                int inn = neighbour1[f];
                double vn = sixth * ((cx * (-nx) + cy * (-ny)) + cz * (-nz));
                I1(vol, inn) = I1(vol, inn) + vn;
            }
        }
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int i = 0; i < nn; ++i) {
            sx = sx + X(nd[i]);
            sy = sy + Y(nd[i]);
            sz = sz + Z(nd[i]);
        }
        xf[f] = sx / (double)nn;
        yf[f] = sy / (double)nn;
        zf[f] = sz / (double)nn;
    }
    /* sweep 2: cell centres (:974-1013) */
    for (int f = 0; f < nfaces; ++f) {
        const int *nd = face_nodes1 + face_ptr[f];
        int nn = face_ptr[f + 1] - face_ptr[f];
        int inp = owner1[f];
        for (int i = 0; i < nn - 2; ++i) {
            double px = X(nd[i + 1]) - X(nd[0]), py = Y(nd[i + 1]) - Y(nd[0]), pz = Z(nd[i + 1]) - Z(nd[0]);
            double qx = X(nd[i + 2]) - X(nd[0]), qy = Y(nd[i + 2]) - Y(nd[0]), qz = Z(nd[i + 2]) - Z(nd[0]);
            double nx = py * qz - pz * qy;
            double ny = pz * qx - px * qz;
            double nz = px * qy - py * qx;
            double vo = I1(vol, inp);
            I1(xc, inp) = I1(xc, inp) + centroid_part(X(nd[0]), X(nd[i + 1]), X(nd[i + 2]), nx, vo);
            I1(yc, inp) = I1(yc, inp) + centroid_part(Y(nd[0]), Y(nd[i + 1]), Y(nd[i + 2]), ny, vo);
            I1(zc, inp) = I1(zc, inp) + centroid_part(Z(nd[0]), Z(nd[i + 1]), Z(nd[i + 2]), nz, vo);
            if (f < nif) {
                int inn = neighbour1[f];
                double vn = I1(vol, inn);
                I1(xc, inn) = I1(xc, inn) + centroid_part(X(nd[0]), X(nd[i + 2]), X(nd[i + 1]), -nx, vn);
                I1(yc, inn) = I1(yc, inn) + centroid_part(Y(nd[0]), Y(nd[i + 2]), Y(nd[i + 1]), -ny, vn);
                I1(zc, inn) = I1(zc, inn) + centroid_part(Z(nd[0]), Z(nd[i + 2]), Z(nd[i + 1]), -nz, vn);
            }
        }
    }
    /* sweep 3: interpolation factor on inner faces (:1040-1081) */
    for (int f = 0; f < nif; ++f) {
        const int *nd = face_nodes1 + face_ptr[f];
        int inp = owner1[f], inn = neighbour1[f];
        double xpn = I1(xc, inn) - I1(xc, inp);
        double ypn = I1(yc, inn) - I1(yc, inp);
        double zpn = I1(zc, inn) - I1(zc, inp);
        double dpn = sqrt((xpn * xpn + ypn * ypn) + zpn * zpn);
        double xj, yj, zj;
        intersection_point(X(nd[0]), Y(nd[0]), Z(nd[0]), X(nd[1]), Y(nd[1]), Z(nd[1]),
                           X(nd[2]), Y(nd[2]), Z(nd[2]), I1(xc, inp), I1(yc, inp), I1(zc, inp),
                           I1(xc, inn), I1(yc, inn), I1(zc, inn), &xj, &yj, &zj);
        xpn = xj - I1(xc, inp);
        ypn = yj - I1(yc, inp);
        zpn = zj - I1(zc, inp);
        double djn = sqrt((xpn * xpn + ypn * ypn) + zpn * zpn);
        facint[f] = djn / dpn;
    }
#undef X
#undef Y
#undef Z
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CSR structure: src/sparse_matrix.f90:42-172, src/utils.f90:76-170.
 * The reference heap-sorts the (row,col) pairs {(i,i)} U {(o,n)} U {(n,o)};
 * the sorted set is unique, so a per-row counting sort gives the same arrays. */
/* ------------------------------------------------------------------------ */

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int orc_csr_build(int n, int nif, const int *owner1, const int *neighbour1,
                  int *ioffset, int *ja, int *diag, int *ij_slot, int *ji_slot)
{
    int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
    if (!cnt) return -1;
    for (int i = 1; i <= n; ++i) cnt[i] = 1;
    for (int f = 0; f < nif; ++f) {
        int o = owner1[f], nb = neighbour1[f];
        if (o < 1 || o > n || nb < 1 || nb > n) { free(cnt); return -1; }
        cnt[o]++;
        cnt[nb]++;
    }
    ioffset[0] = 1;
    for (int i = 1; i <= n; ++i) ioffset[i] = ioffset[i - 1] + cnt[i];
    int *pos = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    for (int i = 1; i <= n; ++i) pos[i] = I1(ioffset, i);
    for (int i = 1; i <= n; ++i) I1(ja, pos[i]++) = i;
    for (int f = 0; f < nif; ++f) {
        int o = owner1[f], nb = neighbour1[f];
        I1(ja, pos[o]++) = nb;
        I1(ja, pos[nb]++) = o;
    }
    for (int i = 1; i <= n; ++i) {
        int s = I1(ioffset, i), e = I1(ioffset, i + 1);
        qsort(&I1(ja, s), (size_t)(e - s), sizeof(int), cmp_int);
        /* find_main_diag_element_positions: first (i,i) in sorted order */
        for (int k = s; k < e; ++k)
            if (I1(ja, k) == i) { I1(diag, i) = k; break; }
    }
    /* csr_to_k: first k in row i with ja(k) == j (src/utils.f90:114-125) */
    for (int f = 0; f < nif; ++f) {
        int o = owner1[f], nb = neighbour1[f], k;
        ij_slot[f] = 0;
        for (k = I1(ioffset, o); k < I1(ioffset, o + 1); ++k)
            if (I1(ja, k) == nb) { ij_slot[f] = k; break; }
        ji_slot[f] = 0;
        for (k = I1(ioffset, nb); k < I1(ioffset, nb + 1); ++k)
            if (I1(ja, k) == o) { ji_slot[f] = k; break; }
    }
    free(pos);
    free(cnt);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Laplacian: src/fvm_laplacian.f90:1-163, facefluxlaplacian :171-221.       */
/* ------------------------------------------------------------------------ */

static double flux_laplacian(int ijp, int ijn, double arx, double ary, double arz, double lambda,
                             const double *mu, const double *xc, const double *yc, const double *zc)
{
    double fxn = lambda;
    double fxp = 1.0 - lambda;
    double xpn = I1(xc, ijn) - I1(xc, ijp);
    double ypn = I1(yc, ijn) - I1(yc, ijp);
    double zpn = I1(zc, ijn) - I1(zc, ijp);
    double num = (arx * arx + ary * ary) + arz * arz;
    double den = (arx * xpn + ary * ypn) + arz * zpn;
    double smdpn = num / den;
    double m = fxp * I1(mu, ijp) + fxn * I1(mu, ijn);
    return m * smdpn;
}

int orc_laplacian(int n, int nif, const int *owner1, const int *neighbour1,
                  const double *arx, const double *ary, const double *arz,
                  const double *facint, const double *xc, const double *yc, const double *zc,
                  const double *xf, const double *yf, const double *zf,
                  const double *mu, const double *phi,
                  int nwal, int iwall_faces_start, int iwall_start,
                  int noc, int ioc_faces_start, const int *ijl1, const int *ijr1, const double *foc,
                  const int *diag, const int *ij_slot, const int *ji_slot, int nnz,
                  double *a, double *su, double *al, double *ar)
{
    (void)n;
    memset(a, 0, sizeof(double) * nnz);
    for (int i = 0; i < nif; ++i) {
        int ijp = owner1[i], ijn = neighbour1[i];
        double cap = flux_laplacian(ijp, ijn, arx[i], ary[i], arz[i], facint[i], mu, xc, yc, zc);
        double can = cap;
        I1(a, ij_slot[i]) = can;
        I1(a, ji_slot[i]) = cap;
        int k = I1(diag, ijp);
        I1(a, k) = I1(a, k) - can;
        k = I1(diag, ijn);
        I1(a, k) = I1(a, k) - cap;
    }
    for (int i = 0; i < noc; ++i) {
        int iface = ioc_faces_start + i; /* 0-based face index of face iOCFacesStart+i+1 */
        int ijp = ijl1[i], ijn = ijr1[i];
        double c = flux_laplacian(ijp, ijn, arx[iface], ary[iface], arz[iface], foc[i], mu, xc, yc, zc);
        al[i] = c;
        ar[i] = c;
        int k = I1(diag, ijp);
        I1(a, k) = I1(a, k) - ar[i];
        k = I1(diag, ijn);
        I1(a, k) = I1(a, k) - al[i];
    }
    for (int i = 1; i <= nwal; ++i) {
        int iface = iwall_faces_start + i; /* 1-based */
        int ijp = I1(owner1, iface);
        int ijb = iwall_start + i;         /* 1-based into phi(numTotal) */
        int k = I1(diag, ijp);
        double fx = I1(arx, iface), fy = I1(ary, iface), fz = I1(arz, iface);
        double are = sqrt((fx * fx + fy * fy) + fz * fz);
        double dx = I1(xc, ijp) - I1(xf, iface), dy = I1(yc, ijp) - I1(yf, iface), dz = I1(zc, ijp) - I1(zf, iface);
        double dpw = sqrt((dx * dx + dy * dy) + dz * dz);
        I1(a, k) = I1(a, k) - (I1(mu, ijp) * are) / dpw;
        I1(su, ijp) = I1(su, ijp) + I1(a, k) * I1(phi, ijb);
    }
    return 0;
}

int orc_convdiff(int n, int nif, const int *owner1, const int *neighbour1,
                 const double *arx, const double *ary, const double *arz,
                 const double *facint, const double *xc, const double *yc, const double *zc,
                 const double *xf, const double *yf, const double *zf,
                 const double *vis, const double *flux, int nwal, int iwall_faces_start,
                 const int *ioffset, const int *diag, const int *ij_slot, const int *ji_slot,
                 int nnz, double urf, double *a)
{
    double *spu = (double *)calloc((size_t)n, sizeof(double));
    if (!spu) return -1;
    memset(a, 0, sizeof(double) * nnz);
    for (int i = 0; i < nif; ++i) {
        int ijp = owner1[i], ijn = neighbour1[i];
        double fxn = facint[i], fxp = 1.0 - facint[i];
        double xpn = I1(xc, ijn) - I1(xc, ijp);
        double ypn = I1(yc, ijn) - I1(yc, ijp);
        double zpn = I1(zc, ijn) - I1(zc, ijp);
        double dpn = sqrt((xpn * xpn + ypn * ypn) + zpn * zpn);
        double are = sqrt((arx[i] * arx[i] + ary[i] * ary[i]) + arz[i] * arz[i]);
        double game = I1(vis, ijp) * fxp + I1(vis, ijn) * fxn;
        double de = (game * are) / dpn;
        double fl = flux[i];
        double can = -de + (fl < 0.0 ? fl : 0.0);
        double cap = -de - (fl > 0.0 ? fl : 0.0);
        I1(a, ij_slot[i]) = can;
        I1(a, ji_slot[i]) = cap;
    }
    for (int i = 1; i <= nwal; ++i) {
        int iface = iwall_faces_start + i;
        int ijp = I1(owner1, iface);
        double fx = I1(arx, iface), fy = I1(ary, iface), fz = I1(arz, iface);
        double are = sqrt((fx * fx + fy * fy) + fz * fz);
        double dx = I1(xc, ijp) - I1(xf, iface), dy = I1(yc, ijp) - I1(yf, iface), dz = I1(zc, ijp) - I1(zf, iface);
        double dpw = sqrt((dx * dx + dy * dy) + dz * dz);
        I1(spu, ijp) = I1(spu, ijp) + (I1(vis, ijp) * are) / dpw;
    }
    double urfr = 1.0 / urf;
    for (int i = 1; i <= n; ++i) {
        double s = 0.0;
        for (int k = I1(ioffset, i); k < I1(ioffset, i + 1); ++k) s = s + I1(a, k);
        s = s - I1(a, I1(diag, i));
        double dd = I1(spu, i) - s;
        I1(a, I1(diag, i)) = dd * urfr;
    }
    free(spu);
    return 0;
}

void orc_face_flux_central(int nif, const int *owner1, const int *neighbour1,
                           const double *facint, const double *arx, const double *ary,
                           const double *arz, const double *u, const double *v,
                           const double *w, double *flux)
{
    for (int f = 0; f < nif; ++f) {
        int p = owner1[f], q = neighbour1[f];
        double fxn = facint[f], fxp = 1.0 - facint[f];
        double uf = fxp * I1(u, p) + fxn * I1(u, q);
        double vf = fxp * I1(v, p) + fxn * I1(v, q);
        double wf = fxp * I1(w, p) + fxn * I1(w, q);
        flux[f] = (uf * arx[f] + vf * ary[f]) + wf * arz[f];
    }
}

void orc_rhs_divergence(int nif, const int *owner1, const int *neighbour1,
                        const double *flux, double *su)
{
    for (int f = 0; f < nif; ++f) {
        I1(su, owner1[f]) = I1(su, owner1[f]) - flux[f];
        I1(su, neighbour1[f]) = I1(su, neighbour1[f]) + flux[f];
    }
}

/* ------------------------------------------------------------------------ */
/* Solver kernels (src/iccg.f90, src/dpcg.f90, src/bicgstab.f90).            */
/* ------------------------------------------------------------------------ */

void orc_spmv(int n, const int *ioffset, const int *ja, const double *a, const double *x, double *y)
{
    /* src/iccg.f90:129-134 */
    for (int i = 1; i <= n; ++i) {
        double s = 0.0;
        for (int k = I1(ioffset, i); k < I1(ioffset, i + 1); ++k) s = s + I1(a, k) * I1(x, I1(ja, k));
        I1(y, i) = s;
    }
}

void orc_residual(int n, const int *ioffset, const int *ja, const double *a, const double *su,
                  const double *fi, int noc, const int *ijl1, const int *ijr1, const double *al,
                  const double *ar, double *res)
{
    /* src/iccg.f90:52-62 */
    for (int i = 1; i <= n; ++i) {
        double r = I1(su, i);
        for (int k = I1(ioffset, i); k < I1(ioffset, i + 1); ++k) r = r - I1(a, k) * I1(fi, I1(ja, k));
        I1(res, i) = r;
    }
    for (int i = 0; i < noc; ++i) {
        I1(res, ijl1[i]) = I1(res, ijl1[i]) - ar[i] * I1(fi, ijr1[i]);
        I1(res, ijr1[i]) = I1(res, ijr1[i]) - al[i] * I1(fi, ijl1[i]);
    }
}

static void oc_add(int noc, const int *ijl1, const int *ijr1, const double *al, const double *ar,
                   const double *p, double *q)
{
    /* src/iccg.f90:137-140 */
    for (int i = 0; i < noc; ++i) {
        I1(q, ijl1[i]) = I1(q, ijl1[i]) + ar[i] * I1(p, ijr1[i]);
        I1(q, ijr1[i]) = I1(q, ijr1[i]) + al[i] * I1(p, ijl1[i]);
    }
}

void orc_dic_pivots(int n, const int *ioffset, const int *ja, const int *diag, const double *a, double *d)
{
    /* src/iccg.f90:77-83 */
    for (int i = 1; i <= n; ++i) {
        double di = I1(a, I1(diag, i));
        for (int k = I1(ioffset, i); k < I1(diag, i); ++k) {
            double ak = I1(a, k);
            di = di - (ak * ak) * I1(d, I1(ja, k));
        }
        I1(d, i) = 1.0 / di;
    }
}

int orc_dilu_pivots(int n, const int *ioffset, const int *ja, const int *diag, const double *a, double *d)
{
    /* src/bicgstab.f90:68-79: a_ji found by a scan of row j from its diagonal */
    int nnz = I1(ioffset, n + 1) - 1;
    for (int i = 1; i <= n; ++i) {
        double di = I1(a, I1(diag, i));
        for (int k = I1(ioffset, i); k < I1(diag, i); ++k) {
            int j = I1(ja, k), l;
            for (l = I1(diag, j); l <= I1(ioffset, j + 1) - 1; ++l)
                if (I1(ja, l) == i) break;
            if (l > nnz) return -1;   /* no transposed entry and nothing after it */
            di = di - (I1(a, k) * I1(d, j)) * I1(a, l);
        }
        I1(d, i) = 1.0 / di;
    }
    return 0;
}

void orc_lu_apply(int n, const int *ioffset, const int *ja, const int *diag, const double *a,
                  const double *d, double small, const double *rhs, double *z)
{
    /* src/iccg.f90:94-111 (identical in bicgstab.f90:117-136) */
    for (int i = 1; i <= n; ++i) {
        double t = I1(rhs, i);
        for (int k = I1(ioffset, i); k < I1(diag, i); ++k) t = t - I1(a, k) * I1(z, I1(ja, k));
        I1(z, i) = t * I1(d, i);
    }
    for (int i = 1; i <= n; ++i) I1(z, i) = I1(z, i) / (I1(d, i) + small);
    for (int i = n; i >= 1; --i) {
        double t = I1(z, i);
        for (int k = I1(diag, i) + 1; k < I1(ioffset, i + 1); ++k) t = t - I1(a, k) * I1(z, I1(ja, k));
        I1(z, i) = t * I1(d, i);
    }
}

static double sum_abs(int n, const double *x)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + fabs(x[i]);
    return s;
}

static double sum_prod(int n, const double *x, const double *y)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + x[i] * y[i];
    return s;
}

typedef struct {
    double small, resmax;
    int ns, early_exit_enabled, use_oc;
} solver_consts;

static solver_consts consts_for(int variant, double sor, int nsw)
{
    solver_consts c;
    if (variant == ORC_VARIANT_TESTS) {
        /* tests/test_sparse_solvers.f90:36-37,99,262-265 */
        c.small = (double)1e-30f;
        c.resmax = (double)1e-13f;
        c.ns = 100;
        c.early_exit_enabled = 0;
        c.use_oc = 0;
    } else {
        c.small = (double)1e-20f;
        c.resmax = sor;
        c.ns = nsw;
        c.early_exit_enabled = 1;
        c.use_oc = 1;
    }
    return c;
}

enum { PREC_DIC = 0, PREC_JACOBI = 1 };

static int pcg_common(int prec, int variant, int n, const int *ioffset, const int *ja,
                      const int *diag, const double *a, const double *su, double *fi, int noc,
                      const int *ijl1, const int *ijr1, const double *al, const double *ar,
                      double sor, int nsw, double *res_out, double *history, orc_stats *st)
{
    solver_consts c = consts_for(variant, sor, nsw);
    if (!c.use_oc) noc = 0;
    double *res = (double *)calloc((size_t)n, sizeof(double));
    double *pk = (double *)calloc((size_t)n, sizeof(double));
    double *zk = (double *)calloc((size_t)n, sizeof(double));
    double *d = (double *)calloc((size_t)n, sizeof(double));
    if (!res || !pk || !zk || !d) { free(res); free(pk); free(zk); free(d); return -1; }
    orc_residual(n, ioffset, ja, a, su, fi, noc, ijl1, ijr1, al, ar, res);
    double res0 = sum_abs(n, res);
    memset(st, 0, sizeof(*st));
    st->res0 = res0;
    st->resl = res0;
    if (c.early_exit_enabled && res0 < orc_tol()) {
        st->early_exit = 1;
        if (res_out) memcpy(res_out, res, sizeof(double) * n);
        free(res); free(pk); free(zk); free(d);
        return 2;
    }
    if (prec == PREC_DIC) orc_dic_pivots(n, ioffset, ja, diag, a, d);
    double s0 = orc_s0(), resl = res0, rsm = 0.0;
    int it = 0;
    for (int l = 1; l <= c.ns; ++l) {
        if (prec == PREC_DIC) {
            orc_lu_apply(n, ioffset, ja, diag, a, d, c.small, res, zk);
        } else {
            for (int i = 1; i <= n; ++i) I1(zk, i) = I1(res, i) / I1(a, I1(diag, i));
        }
        double sk = sum_prod(n, res, zk);
        double bet = sk / s0;
        for (int i = 0; i < n; ++i) pk[i] = zk[i] + bet * pk[i];
        orc_spmv(n, ioffset, ja, a, pk, zk);
        oc_add(noc, ijl1, ijr1, al, ar, pk, zk);
        double pkapk = sum_prod(n, pk, zk);
        double alf = sk / pkapk;
        for (int i = 0; i < n; ++i) fi[i] = fi[i] + alf * pk[i];
        for (int i = 0; i < n; ++i) res[i] = res[i] - alf * zk[i];
        resl = sum_abs(n, res);
        s0 = sk;
        it++;
        if (history) history[l - 1] = resl;
        rsm = resl / (res0 + c.small);
        if (rsm < c.resmax) break;
    }
    st->resl = resl;
    st->rsm = rsm;
    st->iters = it;
    if (res_out) memcpy(res_out, res, sizeof(double) * n);
    free(res); free(pk); free(zk); free(d);
    return (rsm < c.resmax) ? 0 : 1;
}

int orc_iccg(int variant, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
             const double *a, const double *su, double *fi, int noc, const int *ijl1,
             const int *ijr1, const double *al, const double *ar, double sor, int nsw,
             double *res_out, double *history, orc_stats *st)
{
    (void)nnz;
    return pcg_common(PREC_DIC, variant, n, ioffset, ja, diag, a, su, fi, noc, ijl1, ijr1, al, ar,
                      sor, nsw, res_out, history, st);
}

int orc_dpcg(int variant, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
             const double *a, const double *su, double *fi, int noc, const int *ijl1,
             const int *ijr1, const double *al, const double *ar, double sor, int nsw,
             double *res_out, double *history, orc_stats *st)
{
    (void)nnz;
    return pcg_common(PREC_JACOBI, variant, n, ioffset, ja, diag, a, su, fi, noc, ijl1, ijr1, al,
                      ar, sor, nsw, res_out, history, st);
}

int orc_bicgstab(int variant, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
                 const double *a, const double *su, double *fi, int noc, const int *ijl1,
                 const int *ijr1, const double *al, const double *ar, double sor, int nsw,
                 double *res_out, double *history, orc_stats *st)
{
    /* src/bicgstab.f90:1-230 */
    (void)nnz;
    solver_consts c = consts_for(variant, sor, nsw);
    if (!c.use_oc) noc = 0;
    size_t sz = sizeof(double) * (size_t)n;
    double *res = (double *)calloc((size_t)n, sizeof(double));
    double *reso = (double *)calloc((size_t)n, sizeof(double));
    double *pk = (double *)calloc((size_t)n, sizeof(double));
    double *uk = (double *)calloc((size_t)n, sizeof(double));
    double *zk = (double *)calloc((size_t)n, sizeof(double));
    double *vk = (double *)calloc((size_t)n, sizeof(double));
    double *d = (double *)calloc((size_t)n, sizeof(double));
    int rc = 0;
    memset(st, 0, sizeof(*st));
    if (!res || !reso || !pk || !uk || !zk || !vk || !d) { rc = -1; goto out; }
    /* initial residual includes the O-C coupling (:42-52); the SpMVs below do not (:142-147) */
    orc_residual(n, ioffset, ja, a, su, fi, noc, ijl1, ijr1, al, ar, res);
    double res0 = sum_abs(n, res);
    st->res0 = res0;
    st->resl = res0;
    if (c.early_exit_enabled && res0 < orc_tol()) {
        st->early_exit = 1;
        rc = 2;
        if (res_out) memcpy(res_out, res, sz);
        goto out;
    }
    if (orc_dilu_pivots(n, ioffset, ja, diag, a, d) != 0) { rc = -2; goto out; }
    memcpy(reso, res, sz);
    double alf = 1.0, beto = 1.0, gam = 1.0, resl = res0, rsm = 0.0;
    int it = 0;
    for (int l = 1; l <= c.ns; ++l) {
        double bet = sum_prod(n, res, reso);
        double om = (bet * gam) / ((alf * beto) + c.small);
        beto = bet;
        for (int i = 0; i < n; ++i) {
            double t = alf * uk[i];
            t = pk[i] - t;
            t = om * t;
            pk[i] = res[i] + t;
        }
        orc_lu_apply(n, ioffset, ja, diag, a, d, c.small, pk, zk);
        orc_spmv(n, ioffset, ja, a, zk, uk);
        double ukreso = sum_prod(n, uk, reso);
        gam = bet / ukreso;
        for (int i = 0; i < n; ++i) fi[i] = fi[i] + gam * zk[i];
        for (int i = 0; i < n; ++i) res[i] = res[i] - gam * uk[i];
        orc_lu_apply(n, ioffset, ja, diag, a, d, c.small, res, zk);
        orc_spmv(n, ioffset, ja, a, zk, vk);
        double vr = sum_prod(n, vk, res);
        double vv = sum_prod(n, vk, vk);
        alf = vr / (vv + c.small);
        for (int i = 0; i < n; ++i) fi[i] = fi[i] + alf * zk[i];
        for (int i = 0; i < n; ++i) res[i] = res[i] - alf * vk[i];
        resl = sum_abs(n, res);
        it++;
        if (history) history[l - 1] = resl;
        rsm = resl / (res0 + c.small);
        if (rsm < c.resmax) break;
    }
    st->resl = resl;
    st->rsm = rsm;
    st->iters = it;
    rc = (rsm < c.resmax) ? 0 : 1;
    if (res_out) memcpy(res_out, res, sz);
out:
    free(res); free(reso); free(pk); free(uk); free(zk); free(vk); free(d);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Pressure-correction assembly: the face loops of calcp
 * (src/calcp-multiple_correction_SIMPLE.f90:34-107) around facefluxmass
 * (src/facefluxmass.f90:34-199), with face_value_central
 * (src/interpolation.f90:145-197) for the face velocity.                    */
/* ------------------------------------------------------------------------ */

typedef struct {
    const double *xc, *yc, *zc, *vol, *den, *apu, *apv, *apw, *u, *v, *w, *p;
    const double *dudx, *dvdx, *dwdx, *dpdx;   /* 3 per cell, component fastest */
} RcFields;

/* gradient component c (0..2) of cell i (1-based) */
#define G1(g, i, c) ((g)[3 * ((i)-1) + (c)])

/* src/interpolation.f90:145-197 */
static double face_value_central(int p, int q, double xf, double yf, double zf, const double *fi, const double *g,
                                 const RcFields *F)
{
    double phi_p = I1(fi, p), phi_n = I1(fi, q);
    double xcp = I1(F->xc, p), ycp = I1(F->yc, p), zcp = I1(F->zc, p);
    double xcn = I1(F->xc, q), ycn = I1(F->yc, q), zcn = I1(F->zc, q);
    double gradfidr = G1(g, p, 0) * (xf - xcp) + G1(g, p, 1) * (yf - ycp) + G1(g, p, 2) * (zf - zcp)
                    + G1(g, q, 0) * (xf - xcn) + G1(g, q, 1) * (yf - ycn) + G1(g, q, 2) * (zf - zcn);
    return 0.5 * (phi_p + phi_n + gradfidr);
}

/* src/facefluxmass.f90:34-199: cap (= can) and the Rhie-Chow face mass flux */
static void flux_mass(int ijp, int ijn, double xf, double yf, double zf, double arx, double ary, double arz,
                      double lambda, const RcFields *F, double *cap, double *can, double *fluxmass)
{
    double fxn = lambda;
    double fxp = 1.0 - lambda;
    double xpn = I1(F->xc, ijn) - I1(F->xc, ijp);
    double ypn = I1(F->yc, ijn) - I1(F->yc, ijp);
    double zpn = I1(F->zc, ijn) - I1(F->zc, ijp);
    double are = sqrt(arx * arx + ary * ary + arz * arz);
    double nxx = arx / are, nyy = ary / are, nzz = arz / are;
    double Dpu = fxn * I1(F->vol, ijn) * I1(F->apu, ijn) + fxp * I1(F->vol, ijp) * I1(F->apu, ijp);
    double Dpv = fxn * I1(F->vol, ijn) * I1(F->apv, ijn) + fxp * I1(F->vol, ijp) * I1(F->apv, ijp);
    double Dpw = fxn * I1(F->vol, ijn) * I1(F->apw, ijn) + fxp * I1(F->vol, ijp) * I1(F->apw, ijp);
    double dene = I1(F->den, ijp) * fxp + I1(F->den, ijn) * fxn;
    double sfdpnr = 1.0 / (arx * xpn * nxx + ary * ypn * nyy + arz * zpn * nzz);
    double smdpn = (arx * arx + ary * ary + arz * arz) * sfdpnr;
    *cap = -(dene * Dpu * smdpn);
    *can = *cap;
    double ui = face_value_central(ijp, ijn, xf, yf, zf, F->u, F->dudx, F);
    double vi = face_value_central(ijp, ijn, xf, yf, zf, F->v, F->dvdx, F);
    double wi = face_value_central(ijp, ijn, xf, yf, zf, F->w, F->dwdx, F);
    double dpxi = Dpu * (fxn * G1(F->dpdx, ijn, 0) + fxp * G1(F->dpdx, ijp, 0)) * xpn * nxx;
    double dpyi = Dpv * (fxn * G1(F->dpdx, ijn, 1) + fxp * G1(F->dpdx, ijp, 1)) * ypn * nyy;
    double dpzi = Dpw * (fxn * G1(F->dpdx, ijn, 2) + fxp * G1(F->dpdx, ijp, 2)) * zpn * nzz;
    double xcp = I1(F->xc, ijp), ycp = I1(F->yc, ijp), zcp = I1(F->zc, ijp);
    double xcn = I1(F->xc, ijn), ycn = I1(F->yc, ijn), zcn = I1(F->zc, ijn);
    double xpp = xf - (xf - xcp) * nxx, ypp = yf - (yf - ycp) * nyy, zpp = zf - (zf - zcp) * nzz;
    double xep = xf - (xf - xcn) * nxx, yep = yf - (yf - ycn) * nyy, zep = zf - (zf - zcn) * nzz;
    xpp = xpp - xcp;
    ypp = ypp - ycp;
    zpp = zpp - zcp;
    xep = xep - xcn;
    yep = yep - ycn;
    zep = zep - zcn;
    double dpe = I1(F->p, ijn) - I1(F->p, ijp);
    /* as written in the reference: only the x term of the owner side is subtracted */
    double dpecorr = G1(F->dpdx, ijn, 0) * xep + G1(F->dpdx, ijn, 1) * yep + G1(F->dpdx, ijn, 2) * zep
                   - G1(F->dpdx, ijp, 0) * xpp + G1(F->dpdx, ijp, 1) * ypp + G1(F->dpdx, ijp, 2) * zpp;
    dpe = dpe + dpecorr;
    double dpex = Dpu * dpe * sfdpnr * arx;
    double dpey = Dpv * dpe * sfdpnr * ary;
    double dpez = Dpw * dpe * sfdpnr * arz;
    double ue = ui - dpex + dpxi;
    double ve = vi - dpey + dpyi;
    double we = wi - dpez + dpzi;
    *fluxmass = dene * (ue * arx + ve * ary + we * arz);
}

int orc_calcp_assemble(int n, int nif, const int *owner1, const int *neighbour1,
                       const double *xc, const double *yc, const double *zc, const double *vol,
                       const double *xf, const double *yf, const double *zf,
                       const double *arx, const double *ary, const double *arz, const double *facint,
                       const double *den, const double *apu, const double *apv, const double *apw,
                       const double *u, const double *v, const double *w, const double *p,
                       const double *dudx, const double *dvdx, const double *dwdx, const double *dpdx,
                       int noc, int ioc_faces_start, const int *ijl1, const int *ijr1, const double *foc,
                       const int *diag, const int *ij_slot, const int *ji_slot, int nnz,
                       double *a, double *su, double *flmass, double *cap_out, double *al, double *ar,
                       double *fmoc)
{
    RcFields F = {xc, yc, zc, vol, den, apu, apv, apw, u, v, w, p, dudx, dvdx, dwdx, dpdx};
    memset(a, 0, sizeof(double) * nnz);
    memset(su, 0, sizeof(double) * n);
    for (int i = 0; i < nif; ++i) {
        int ijp = owner1[i], ijn = neighbour1[i];
        double cap, can;
        flux_mass(ijp, ijn, xf[i], yf[i], zf[i], arx[i], ary[i], arz[i], facint[i], &F, &cap, &can, &flmass[i]);
        if (cap_out) cap_out[i] = cap;
        I1(a, ij_slot[i]) = can;
        I1(a, ji_slot[i]) = cap;
        int k = I1(diag, ijp);
        I1(a, k) = I1(a, k) - can;
        k = I1(diag, ijn);
        I1(a, k) = I1(a, k) - cap;
        I1(su, ijp) = I1(su, ijp) - flmass[i];
        I1(su, ijn) = I1(su, ijn) + flmass[i];
    }
    for (int i = 0; i < noc; ++i) {
        int iface = ioc_faces_start + i; /* 0-based face index of face iOCFacesStart+i+1 */
        int ijp = ijl1[i], ijn = ijr1[i];
        flux_mass(ijp, ijn, xf[iface], yf[iface], zf[iface], arx[iface], ary[iface], arz[iface], foc[i], &F,
                  &al[i], &ar[i], &fmoc[i]);
        int k = I1(diag, ijp);
        I1(a, k) = I1(a, k) - ar[i];
        k = I1(diag, ijn);
        I1(a, k) = I1(a, k) - al[i];
        I1(su, ijp) = I1(su, ijp) - fmoc[i];
        I1(su, ijn) = I1(su, ijn) + fmoc[i];
    }
    return 0;
}

/* facefluxmass_piso (src/facefluxmass.f90:294-358) */
static void flux_mass_piso(int ijp, int ijn, double xf, double yf, double zf, double arx, double ary, double arz,
                           double lambda, const RcFields *F, double *cap, double *can, double *fluxmass)
{
    double fxn = lambda;
    double fxp = 1.0 - lambda;
    double xpn = I1(F->xc, ijn) - I1(F->xc, ijp);
    double ypn = I1(F->yc, ijn) - I1(F->yc, ijp);
    double zpn = I1(F->zc, ijn) - I1(F->zc, ijp);
    double dpn = sqrt(xpn * xpn + ypn * ypn + zpn * zpn);
    double are = sqrt(arx * arx + ary * ary + arz * arz);
    double dene = I1(F->den, ijp) * fxp + I1(F->den, ijn) * fxn;
    *cap = -dene * (fxp * I1(F->vol, ijp) * I1(F->apu, ijp) + fxn * I1(F->vol, ijn) * I1(F->apu, ijn)) * are / dpn;
    *can = *cap;
    double ui = face_value_central(ijp, ijn, xf, yf, zf, F->u, F->dudx, F);
    double vi = face_value_central(ijp, ijn, xf, yf, zf, F->v, F->dvdx, F);
    double wi = face_value_central(ijp, ijn, xf, yf, zf, F->w, F->dwdx, F);
    *fluxmass = dene * (ui * arx + vi * ary + wi * arz);
}

/* PISO's pressure assembly (src/PISO_multiple_correction.f90:107-171): the face loop of
 * orc_calcp_assemble with facefluxmass_piso, then (pref_cell > 0) the reference pressure
 * (:188-192): row pref_cell of a zeroed, a(diag) = 1, su(pref_cell) = p(pref_cell). */
int orc_piso_assemble(int n, int nif, const int *owner1, const int *neighbour1,
                      const double *xc, const double *yc, const double *zc, const double *vol,
                      const double *xf, const double *yf, const double *zf,
                      const double *arx, const double *ary, const double *arz, const double *facint,
                      const double *den, const double *apu, const double *u, const double *v, const double *w,
                      const double *dudx, const double *dvdx, const double *dwdx,
                      int noc, int ioc_faces_start, const int *ijl1, const int *ijr1, const double *foc,
                      const int *ioffset, const int *diag, const int *ij_slot, const int *ji_slot, int nnz,
                      int pref_cell, const double *p,
                      double *a, double *su, double *flmass, double *cap_out, double *al, double *ar,
                      double *fmoc)
{
    RcFields F = {xc, yc, zc, vol, den, apu, NULL, NULL, u, v, w, NULL, dudx, dvdx, dwdx, NULL};
    memset(a, 0, sizeof(double) * nnz);
    memset(su, 0, sizeof(double) * n);
    for (int i = 0; i < nif; ++i) {
        int ijp = owner1[i], ijn = neighbour1[i];
        double cap, can;
        flux_mass_piso(ijp, ijn, xf[i], yf[i], zf[i], arx[i], ary[i], arz[i], facint[i], &F, &cap, &can, &flmass[i]);
        if (cap_out) cap_out[i] = cap;
        I1(a, ij_slot[i]) = can;
        I1(a, ji_slot[i]) = cap;
        int k = I1(diag, ijp);
        I1(a, k) = I1(a, k) - can;
        k = I1(diag, ijn);
        I1(a, k) = I1(a, k) - cap;
        I1(su, ijp) = I1(su, ijp) - flmass[i];
        I1(su, ijn) = I1(su, ijn) + flmass[i];
    }
    for (int i = 0; i < noc; ++i) {
        int iface = ioc_faces_start + i;
        int ijp = ijl1[i], ijn = ijr1[i];
        flux_mass_piso(ijp, ijn, xf[iface], yf[iface], zf[iface], arx[iface], ary[iface], arz[iface], foc[i], &F,
                       &al[i], &ar[i], &fmoc[i]);
        int k = I1(diag, ijp);
        I1(a, k) = I1(a, k) - ar[i];
        k = I1(diag, ijn);
        I1(a, k) = I1(a, k) - al[i];
        I1(su, ijp) = I1(su, ijp) - fmoc[i];
        I1(su, ijn) = I1(su, ijn) + fmoc[i];
    }
    if (pref_cell > 0) {
        for (int k = I1(ioffset, pref_cell); k <= I1(ioffset, pref_cell + 1) - 1; ++k) I1(a, k) = 0.0;
        I1(a, I1(diag, pref_cell)) = 1.0;
        I1(su, pref_cell) = I1(p, pref_cell);
    }
    return 0;
}

/* bpres (src/bpres.f90:3-152).  Boundary values follow the cells in the groups' order (inlet,
 * outlet, symmetry, wall, pressure outlet; src/mesh_geometry_and_topology.f90:593-605), then
 * the O-C values.  owner1 covers every face; face arrays are 0-based per face. */
void orc_bpres(int n, const int *owner1, const double *xc, const double *yc, const double *zc,
               const double *xf, const double *yf, const double *zf, const int *bnd, int noc,
               int ioc_faces_start, const double *dpdx, int istage, double *p)
{
    int fstart[5], vstart[6];
    int v = n;
    for (int g = 0; g < 5; ++g) {
        fstart[g] = bnd[2 * g + 1];
        vstart[g] = v;
        v += bnd[2 * g];
    }
    vstart[5] = v; /* iOCStart */
    if (istage == 1) {
        for (int g = 0; g < 5; ++g)
            for (int i = 1; i <= bnd[2 * g]; ++i) {
                int ijp = I1(owner1, fstart[g] + i);
                I1(p, vstart[g] + i) = I1(p, ijp);
            }
        for (int i = 1; i <= noc; ++i) {
            int ijp = I1(owner1, ioc_faces_start + i);
            I1(p, vstart[5] + i) = I1(p, ijp);
        }
        return;
    }
    for (int g = 3; g < 5; ++g)   /* wall, then pressure outlet */
        for (int i = 1; i <= bnd[2 * g]; ++i) {
            int iface = fstart[g] + i;
            int ijp = I1(owner1, iface);
            double xpb = I1(xf, iface) - I1(xc, ijp);
            double ypb = I1(yf, iface) - I1(yc, ijp);
            double zpb = I1(zf, iface) - I1(zc, ijp);
            I1(p, vstart[g] + i) = I1(p, ijp) + dpdx[3 * (ijp - 1)] * xpb + dpdx[3 * (ijp - 1) + 1] * ypb +
                                   dpdx[3 * (ijp - 1) + 2] * zpb;
        }
    for (int i = 1; i <= noc; ++i) {
        int iface = ioc_faces_start + i;
        int ijp = I1(owner1, iface);
        /* the face centre of face number i, as src/bpres.f90:140-142 index it */
        double xpb = I1(xf, i) - I1(xc, ijp);
        double ypb = I1(yf, i) - I1(yc, ijp);
        double zpb = I1(zf, i) - I1(zc, ijp);
        I1(p, vstart[5] + i) = I1(p, ijp) + dpdx[3 * (ijp - 1)] * xpb + dpdx[3 * (ijp - 1) + 1] * ypb +
                               dpdx[3 * (ijp - 1) + 2] * zpb;
    }
}

/* correctBoundaryConditionsVelocity (src/correctBoundaryConditionsVelocity.f90:1-79) */
void orc_correct_velocity_bc(int n, const int *owner1, const double *arx, const double *ary, const double *arz,
                             const int *bnd, double flomas, const double *den, double *u, double *v, double *w,
                             double *fmo)
{
    int ninl = bnd[0], nout = bnd[2], nsym = bnd[4];
    int out_faces = bnd[3], sym_faces = bnd[5];
    int out_start = n + ninl, sym_start = n + ninl + nout;
    double flowo = 0.0;
    for (int i = 1; i <= nout; ++i) {
        int iface = out_faces + i;
        int ijp = I1(owner1, iface);
        int ijb = out_start + i;
        I1(u, ijb) = I1(u, ijp);
        I1(v, ijb) = I1(v, ijp);
        I1(w, ijb) = I1(w, ijp);
        I1(fmo, i) = I1(den, ijp) * (I1(u, ijb) * I1(arx, iface) + I1(v, ijb) * I1(ary, iface) + I1(w, ijb) * I1(arz, iface));
        flowo = flowo + I1(fmo, i);
    }
    double fac = flomas / (flowo + orc_small());
    for (int i = 1; i <= nout; ++i) {
        int ijb = out_start + i;
        I1(fmo, i) = I1(fmo, i) * fac;
        I1(u, ijb) = I1(u, ijb) * fac;
        I1(v, ijb) = I1(v, ijb) * fac;
        I1(w, ijb) = I1(w, ijb) * fac;
    }
    for (int i = 1; i <= nsym; ++i) {
        int iface = sym_faces + i;
        int ijp = I1(owner1, iface);
        int ijb = sym_start + i;
        double unmag = I1(u, ijp) * I1(arx, iface) + I1(v, ijp) * I1(ary, iface) + I1(w, ijp) * I1(arz, iface);
        I1(u, ijb) = I1(u, ijp) - unmag * I1(arx, iface);
        I1(v, ijb) = I1(v, ijp) - unmag * I1(ary, iface);
        I1(w, ijb) = I1(w, ijp) - unmag * I1(arz, iface);
    }
}

/* ------------------------------------------------------------------------ */
/* Momentum component rows of calcuvw (src/calcuvw.f90:417-437, and the v/w
 * twins :468-489, :522-543): the diagonal slot is reset to zero, then
 *   a(diag) = (sp - (sum(a(row)) - a(diag))) * urfr
 *   su      = s + urfm * a(diag) * phi
 *   ap      = 1 / (a(diag) + small)
 * with urfr = 1/urf and urfm = 1 - urf (src/init.f90:80-81).               */
/* ------------------------------------------------------------------------ */

void orc_momentum_rows(int n, const int *ioffset, const int *diag, double *a, const double *sp, const double *src,
                       const double *phi, double urf, double *su, double *ap)
{
    double urfr = 1.0 / urf;
    double urfm = 1.0 - urf;
    double small = orc_small();
    for (int i = 1; i <= n; ++i) {
        int dp = I1(diag, i);
        I1(a, dp) = 0.0;
        double s = 0.0;
        for (int k = I1(ioffset, i); k < I1(ioffset, i + 1); ++k) s = s + I1(a, k);
        double off = s - I1(a, dp);
        I1(a, dp) = I1(sp, i) - off;
        I1(a, dp) = I1(a, dp) * urfr;
        I1(su, i) = I1(src, i) + urfm * I1(a, dp) * I1(phi, i);
        I1(ap, i) = 1.0 / (I1(a, dp) + small);
    }
}

/* ------------------------------------------------------------------------ */
/* Gauss gradient: src/grad_gauss.f90:1-124, gradco :128-190, gradbc :194-211. */
/* ------------------------------------------------------------------------ */

static void gradco(int ijp, int ijn, double xfc, double yfc, double zfc, double sx, double sy, double sz,
                   double fif, const double *fi, const double *xc, const double *yc, const double *zc,
                   const double *dfxo, const double *dfyo, const double *dfzo, double *dfx, double *dfy, double *dfz)
{
    double fxn = fif;
    double fxp = 1.0 - fxn;
    double xi = I1(xc, ijp) * fxp + I1(xc, ijn) * fxn;
    double yi = I1(yc, ijp) * fxp + I1(yc, ijn) * fxn;
    double zi = I1(zc, ijp) * fxp + I1(zc, ijn) * fxn;
    double dfxi = I1(dfxo, ijp) * fxp + I1(dfxo, ijn) * fxn;
    double dfyi = I1(dfyo, ijp) * fxp + I1(dfyo, ijn) * fxn;
    double dfzi = I1(dfzo, ijp) * fxp + I1(dfzo, ijn) * fxn;
    double fie = I1(fi, ijp) * fxp + I1(fi, ijn) * fxn + dfxi * (xfc - xi) + dfyi * (yfc - yi) + dfzi * (zfc - zi);
    double dfxe = fie * sx, dfye = fie * sy, dfze = fie * sz;
    I1(dfx, ijp) = I1(dfx, ijp) + dfxe;
    I1(dfy, ijp) = I1(dfy, ijp) + dfye;
    I1(dfz, ijp) = I1(dfz, ijp) + dfze;
    I1(dfx, ijn) = I1(dfx, ijn) - dfxe;
    I1(dfy, ijn) = I1(dfy, ijn) - dfye;
    I1(dfz, ijn) = I1(dfz, ijn) - dfze;
}

int orc_grad_gauss(int n, int nif, const int *owner1, const int *neighbour1,
                   const double *arx, const double *ary, const double *arz, const double *facint,
                   const double *xc, const double *yc, const double *zc,
                   const double *xf, const double *yf, const double *zf, const double *vol,
                   const double *u, const int *bnd, int noc, const int *ijlface1, const int *ijl1,
                   const int *ijr1, const double *foc, int nigrad, double *dudx, double *dudy, double *dudz)
{
    double *o = (double *)calloc(3 * (size_t)n, sizeof(double));
    if (!o) return -1;
    double *dfxo = o, *dfyo = o + n, *dfzo = o + 2 * n;
    for (int lc = 1; lc <= nigrad; ++lc) {
        memset(dudx, 0, sizeof(double) * n);
        memset(dudy, 0, sizeof(double) * n);
        memset(dudz, 0, sizeof(double) * n);
        for (int i = 0; i < nif; ++i)
            gradco(owner1[i], neighbour1[i], xf[i], yf[i], zf[i], arx[i], ary[i], arz[i], facint[i], u, xc, yc, zc,
                   dfxo, dfyo, dfzo, dudx, dudy, dudz);
        for (int i = 0; i < noc; ++i) {
            int iface = ijlface1[i];
            gradco(ijl1[i], ijr1[i], I1(xf, iface), I1(yf, iface), I1(zf, iface), I1(arx, iface), I1(ary, iface),
                   I1(arz, iface), foc[i], u, xc, yc, zc, dfxo, dfyo, dfzo, dudx, dudy, dudz);
        }
        /* boundary groups in the reference's order: inlet, outlet, symmetry, wall, pressure
         * outlet; values of group g follow those of the groups before it, after the cells */
        int vstart = n;
        for (int g = 0; g < 5; ++g) {
            int cnt = bnd[2 * g], fstart = bnd[2 * g + 1];
            for (int i = 1; i <= cnt; ++i) {
                int iface = fstart + i;
                int ijp = I1(owner1, iface);
                int ijb = vstart + i;
                I1(dudx, ijp) = I1(dudx, ijp) + I1(u, ijb) * I1(arx, iface);
                I1(dudy, ijp) = I1(dudy, ijp) + I1(u, ijb) * I1(ary, iface);
                I1(dudz, ijp) = I1(dudz, ijp) + I1(u, ijb) * I1(arz, iface);
            }
            vstart += cnt;
        }
        for (int i = 0; i < n; ++i) {
            double volr = 1.0 / vol[i];
            dudx[i] = dudx[i] * volr;
            dudy[i] = dudy[i] * volr;
            dudz[i] = dudz[i] * volr;
        }
        if (lc != nigrad) {
            memcpy(dfxo, dudx, sizeof(double) * n);
            memcpy(dfyo, dudy, sizeof(double) * n);
            memcpy(dfzo, dudz, sizeof(double) * n);
        }
    }
    free(o);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Corrector stage of calcp (src/calcp-multiple_correction_SIMPLE.f90:145-231):
 * fluxmc (src/facefluxmass.f90:520-607) and the flux, velocity and pressure
 * corrections.                                                               */
/* ------------------------------------------------------------------------ */

static double fluxmc(int ijp, int ijn, double xf, double yf, double zf, double arx, double ary, double arz,
                     double lambda, const double *xc, const double *yc, const double *zc, const double *vol,
                     const double *den, const double *apu, const double *dpdx)
{
    double fxn = lambda;
    double fxp = 1.0 - lambda;
    double xpn = I1(xc, ijn) - I1(xc, ijp);
    double ypn = I1(yc, ijn) - I1(yc, ijp);
    double zpn = I1(zc, ijn) - I1(zc, ijp);
    double are = sqrt(arx * arx + ary * ary + arz * arz);
    double nxx = arx / are, nyy = ary / are, nzz = arz / are;
    double dppnnr = 1.0 / ((xpn * nxx) + (ypn * nyy) + (zpn * nzz));
    double xpp = xf - (xf - I1(xc, ijp)) * nxx, ypp = yf - (yf - I1(yc, ijp)) * nyy, zpp = zf - (zf - I1(zc, ijp)) * nzz;
    double xep = xf - (xf - I1(xc, ijn)) * nxx, yep = yf - (yf - I1(yc, ijn)) * nyy, zep = zf - (zf - I1(zc, ijn)) * nzz;
    xpp = xpp - I1(xc, ijp);
    ypp = ypp - I1(yc, ijp);
    zpp = zpp - I1(zc, ijp);
    xep = xep - I1(xc, ijn);
    yep = yep - I1(yc, ijn);
    zep = zep - I1(zc, ijn);
    double rapr = I1(apu, ijp) * I1(den, ijp) * I1(vol, ijp) * fxp + I1(apu, ijn) * I1(den, ijn) * I1(vol, ijn) * fxn;
    return rapr * are * ((G1(dpdx, ijn, 0) * xep - G1(dpdx, ijp, 0) * xpp) + (G1(dpdx, ijn, 1) * yep - G1(dpdx, ijp, 1) * ypp) +
                         (G1(dpdx, ijn, 2) * zep - G1(dpdx, ijp, 2) * zpp)) * dppnnr;
}

/* the non-orthogonal corrector's source (:190-229): su = 0, then per inner face and
 * per O-C cut flmass/fmoc += fmcor, su(P) -= fmcor, su(N) += fmcor */
void orc_calcp_nonorth(int n, int nif, const int *owner1, const int *neighbour1,
                       const double *xc, const double *yc, const double *zc, const double *vol,
                       const double *xf, const double *yf, const double *zf,
                       const double *arx, const double *ary, const double *arz, const double *facint,
                       const double *den, const double *apu, const double *dpdx,
                       int noc, int ioc_faces_start, const int *ijl1, const int *ijr1, const double *foc,
                       double *su, double *flmass, double *fmoc, double *fmcor_out)
{
    memset(su, 0, sizeof(double) * n);
    for (int i = 0; i < nif; ++i) {
        int ijp = owner1[i], ijn = neighbour1[i];
        double fm = fluxmc(ijp, ijn, xf[i], yf[i], zf[i], arx[i], ary[i], arz[i], facint[i], xc, yc, zc, vol, den,
                           apu, dpdx);
        if (fmcor_out) fmcor_out[i] = fm;
        flmass[i] = flmass[i] + fm;
        I1(su, ijp) = I1(su, ijp) - fm;
        I1(su, ijn) = I1(su, ijn) + fm;
    }
    for (int i = 0; i < noc; ++i) {
        int iface = ioc_faces_start + i;
        int ijp = ijl1[i], ijn = ijr1[i];
        double fm = fluxmc(ijp, ijn, xf[iface], yf[iface], zf[iface], arx[iface], ary[iface], arz[iface], foc[i],
                           xc, yc, zc, vol, den, apu, dpdx);
        fmoc[i] = fmoc[i] + fm;
        I1(su, ijp) = I1(su, ijp) - fm;
        I1(su, ijn) = I1(su, ijn) + fm;
    }
}

/* adjustMassFlow (src/adjustMassFlow.f90:1-71): su(owner) -= fmi over the inlet faces; outlet
 * boundary velocities extrapolated from the owner cells, fmo = den (u.S) summed in outlet order
 * into flowo, scaled by fac = flomas / (flowo + small) together with the outlet velocities, and
 * su(owner) -= fmo.  Faces and boundary values are 1-based positions after the given starts. */
void orc_adjust_mass_flow(const int *owner1, const double *arx, const double *ary, const double *arz,
                          int ninl, int inlet_faces_start, const double *fmi, int nout,
                          int outlet_faces_start, int outlet_start, double flomas, const double *den,
                          double *u, double *v, double *w, double *su, double *fmo)
{
    for (int i = 1; i <= ninl; ++i) {
        int ijp = owner1[inlet_faces_start + i - 1];
        I1(su, ijp) = I1(su, ijp) - I1(fmi, i);
    }
    double flowo = 0.0;
    for (int i = 1; i <= nout; ++i) {
        int iface = outlet_faces_start + i;
        int ijp = I1(owner1, iface);
        int ijb = outlet_start + i;
        I1(u, ijb) = I1(u, ijp);
        I1(v, ijb) = I1(v, ijp);
        I1(w, ijb) = I1(w, ijp);
        I1(fmo, i) = I1(den, ijp) * (I1(u, ijb) * I1(arx, iface) + I1(v, ijb) * I1(ary, iface) + I1(w, ijb) * I1(arz, iface));
        flowo = flowo + I1(fmo, i);
    }
    double fac = flomas / (flowo + orc_small());
    for (int i = 1; i <= nout; ++i) {
        int iface = outlet_faces_start + i;
        int ijp = I1(owner1, iface);
        int ijb = outlet_start + i;
        I1(fmo, i) = I1(fmo, i) * fac;
        I1(u, ijb) = I1(u, ijb) * fac;
        I1(v, ijb) = I1(v, ijb) * fac;
        I1(w, ijb) = I1(w, ijb) * fac;
        I1(su, ijp) = I1(su, ijp) - I1(fmo, i);
    }
}

/* the corrections (:145-183): ppref = pp(pRefCell); flmass += a(ij) (pp(N) - pp(P));
 * fmoc += ar (pp(ijr) - pp(ijl)); u -= dPdx Vol apu (v, w alike); p += urf (pp - ppref) */
void orc_calcp_correct(int n, int nif, const int *owner1, const int *neighbour1, const int *ij_slot,
                       const double *a, const double *pp, const double *dpdx, const double *vol,
                       const double *apu, const double *apv, const double *apw, double urf, int pref_cell,
                       int noc, const int *ijl1, const int *ijr1, const double *ar, double *fmoc,
                       double *flmass, double *u, double *v, double *w, double *p)
{
    double ppref = I1(pp, pref_cell);
    for (int i = 0; i < nif; ++i) {
        int ijp = owner1[i], ijn = neighbour1[i];
        flmass[i] = flmass[i] + I1(a, ij_slot[i]) * (I1(pp, ijn) - I1(pp, ijp));
    }
    for (int i = 0; i < noc; ++i) fmoc[i] = fmoc[i] + ar[i] * (I1(pp, ijr1[i]) - I1(pp, ijl1[i]));
    for (int i = 1; i <= n; ++i) {
        I1(u, i) = I1(u, i) - G1(dpdx, i, 0) * I1(vol, i) * I1(apu, i);
        I1(v, i) = I1(v, i) - G1(dpdx, i, 1) * I1(vol, i) * I1(apv, i);
        I1(w, i) = I1(w, i) - G1(dpdx, i, 2) * I1(vol, i) * I1(apw, i);
        I1(p, i) = I1(p, i) + urf * (I1(pp, i) - ppref);
    }
}
