/*
 * CPU oracle for the pressure-solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's serial Fortran algorithm, used by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the
 * checker.  The product (freecappuccino_amd/libfcap.so) never links, loads or
 * calls anything here.
 *
 * Index conventions follow the reference: every index array (owner, neighbour,
 * ioffset, ja, diag, slot maps, ijl, ijr) is 1-based.  face_ptr/face_nodes are
 * the CSR list of face vertices with 0-based pointers and 1-based node ids.
 *
 * Compile with -O2 -ffp-contract=off: the reference's evaluation order is
 * reproduced operation by operation and must not be contracted into FMAs.
 */
#ifndef FCAP_ORACLE_H
#define FCAP_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Constants written as default-REAL literals in the reference and widened to
 * real(dp) (src/modules_allocatable.f90:27, src/iccg.f90:37,85). */
double orc_small(void);      /* (double)1e-20f */
double orc_tol(void);        /* (double)1e-13f */
double orc_s0(void);         /* (double)1e20f  */
double orc_tiny(void);       /* (double)1e-30f (src/mesh_geometry_and_topology.f90:58) */

/* src/mesh_geometry_and_topology.f90:859-1081 (O-C / cyclic faces not handled). */
int orc_mesh_geometry(int nnodes, const double *xyz /* nnodes*3 */, int nfaces, int nif,
                      const int *face_ptr, const int *face_nodes1, const int *owner1,
                      const int *neighbour1, int ncells,
                      double *arx, double *ary, double *arz,
                      double *xf, double *yf, double *zf,
                      double *vol, double *xc, double *yc, double *zc, double *facint);

/* src/sparse_matrix.f90:42-172 + src/utils.f90:76-170. Returns 0 or -1 on bad input. */
int orc_csr_build(int n, int nif, const int *owner1, const int *neighbour1,
                  int *ioffset, int *ja, int *diag, int *ij_slot, int *ji_slot);

/* src/fvm_laplacian.f90:1-163 (+ facefluxlaplacian :171-221).
 * phi has numTotal = n + nboundary entries; wall values at iwall_start+i (1-based).
 * a is zeroed here, su is accumulated (never zeroed), as in the reference. */
int orc_laplacian(int n, int nif, const int *owner1, const int *neighbour1,
                  const double *arx, const double *ary, const double *arz,
                  const double *facint, const double *xc, const double *yc, const double *zc,
                  const double *xf, const double *yf, const double *zf,
                  const double *mu, const double *phi,
                  int nwal, int iwall_faces_start, int iwall_start,
                  int noc, int ioc_faces_start, const int *ijl1, const int *ijr1, const double *foc,
                  const int *diag, const int *ij_slot, const int *ji_slot, int nnz,
                  double *a, double *su, double *al, double *ar);

/* Momentum-like convection-diffusion operator (the config-3 system):
 * coefficients of facefluxuvw (src/faceflux_velocity.f90:97-108) with upwind
 * convection, diagonal = (spu - sum of the row's off-diagonals) / urf
 * (src/calcuvw.f90:417-432); spu holds the Dirichlet-wall diffusion
 * vis*|S|/|x_P - x_f| of each wall face, accumulated in wall-face order. */
int orc_convdiff(int n, int nif, const int *owner1, const int *neighbour1,
                 const double *arx, const double *ary, const double *arz,
                 const double *facint, const double *xc, const double *yc, const double *zc,
                 const double *xf, const double *yf, const double *zf,
                 const double *vis, const double *flux, int nwal, int iwall_faces_start,
                 const int *ioffset, const int *diag, const int *ij_slot, const int *ji_slot,
                 int nnz, double urf, double *a);

/* Central face flux F_f = u_f . S_f with u_f = (1-facint) u_P + facint u_N
 * (the interpolation of src/facefluxmass.f90 without Rhie-Chow), inner faces. */
void orc_face_flux_central(int nif, const int *owner1, const int *neighbour1,
                           const double *facint, const double *arx, const double *ary,
                           const double *arz, const double *u, const double *v,
                           const double *w, double *flux);

/* Pressure-correction assembly of calcp (src/calcp-multiple_correction_SIMPLE.f90:34-107):
 * a = 0, su = 0, then per inner face facefluxmass (src/facefluxmass.f90:34-199) --
 * cap = can and the Rhie-Chow mass flux flmass -- scattered as a(ij) = can,
 * a(ji) = cap, a(diag(P)) -= can, a(diag(N)) -= cap, su(P) -= F, su(N) += F; then
 * the O-C cuts (al = cap, ar = can, fmoc).  Gradients are 3 x n, component
 * fastest (dUdxi(3, numCells)); cap_out (nif, may be NULL) receives cap. */
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
                       double *fmoc);

/* PISO's pressure assembly (src/PISO_multiple_correction.f90:107-171, facefluxmass_piso
 * src/facefluxmass.f90:294-358) and, for pref_cell > 0, its reference pressure (:188-192). */
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
                      double *fmoc);

/* bpres(p, istage) (src/bpres.f90:3-152) on p (numTotal); bnd as for orc_grad_gauss. */
void orc_bpres(int n, const int *owner1, const double *xc, const double *yc, const double *zc,
               const double *xf, const double *yf, const double *zf, const int *bnd, int noc,
               int ioc_faces_start, const double *dpdx, int istage, double *p);

/* correctBoundaryConditionsVelocity (src/correctBoundaryConditionsVelocity.f90:1-79). */
void orc_correct_velocity_bc(int n, const int *owner1, const double *arx, const double *ary, const double *arz,
                             const int *bnd, double flomas, const double *den, double *u, double *v, double *w,
                             double *fmo);

/* Gauss gradient of u (numTotal values: cells, then the boundary groups' values) as
 * src/grad_gauss.f90 computes it, nigrad iterations.  owner1 covers every face; the
 * face arrays are 0-based per face.  bnd: (count, face start) of the inlet, outlet,
 * symmetry, wall and pressure-outlet groups (face start = faces before the group);
 * O-C cuts through 1-based face numbers ijlface1. */
int orc_grad_gauss(int n, int nif, const int *owner1, const int *neighbour1,
                   const double *arx, const double *ary, const double *arz, const double *facint,
                   const double *xc, const double *yc, const double *zc,
                   const double *xf, const double *yf, const double *zf, const double *vol,
                   const double *u, const int *bnd, int noc, const int *ijlface1, const int *ijl1,
                   const int *ijr1, const double *foc, int nigrad, double *dudx, double *dudy, double *dudz);

/* Non-orthogonal corrector source of calcp (src/calcp-multiple_correction_SIMPLE.f90:190-229)
 * with fluxmc (src/facefluxmass.f90:520-607): su = 0; per inner face and O-C cut
 * flmass/fmoc += fmcor, su(P) -= fmcor, su(N) += fmcor.  fmcor_out (nif) may be NULL. */
void orc_calcp_nonorth(int n, int nif, const int *owner1, const int *neighbour1,
                       const double *xc, const double *yc, const double *zc, const double *vol,
                       const double *xf, const double *yf, const double *zf,
                       const double *arx, const double *ary, const double *arz, const double *facint,
                       const double *den, const double *apu, const double *dpdx,
                       int noc, int ioc_faces_start, const int *ijl1, const int *ijr1, const double *foc,
                       double *su, double *flmass, double *fmoc, double *fmcor_out);

/* adjustMassFlow (src/adjustMassFlow.f90:1-71); owner1 holds every face's owner (1-based). */
void orc_adjust_mass_flow(const int *owner1, const double *arx, const double *ary, const double *arz,
                          int ninl, int inlet_faces_start, const double *fmi, int nout,
                          int outlet_faces_start, int outlet_start, double flomas, const double *den,
                          double *u, double *v, double *w, double *su, double *fmo);
/* Corrections of calcp (src/calcp-multiple_correction_SIMPLE.f90:145-183). */
void orc_calcp_correct(int n, int nif, const int *owner1, const int *neighbour1, const int *ij_slot,
                       const double *a, const double *pp, const double *dpdx, const double *vol,
                       const double *apu, const double *apv, const double *apw, double urf, int pref_cell,
                       int noc, const int *ijl1, const int *ijr1, const double *ar, double *fmoc,
                       double *flmass, double *u, double *v, double *w, double *p);

/* Momentum component rows of calcuvw (src/calcuvw.f90:417-437): diagonal reset to zero, then
 * a(diag) = (sp - sum of the row's off-diagonals) / urf, su = src + (1 - urf) a(diag) phi,
 * ap = 1/(a(diag) + small).  a is updated in place (its diagonal slots). */
void orc_momentum_rows(int n, const int *ioffset, const int *diag, double *a, const double *sp, const double *src,
                       const double *phi, double urf, double *su, double *ap);

/* su(o) -= F, su(n) += F in face order (src/calcp-multiple_correction_SIMPLE.f90:74-75). */
void orc_rhs_divergence(int nif, const int *owner1, const int *neighbour1,
                        const double *flux, double *su);

/* Kernels of the solvers, each with the reference's loop order. */
void orc_spmv(int n, const int *ioffset, const int *ja, const double *a, const double *x, double *y);
void orc_residual(int n, const int *ioffset, const int *ja, const double *a, const double *su,
                  const double *fi, int noc, const int *ijl1, const int *ijr1, const double *al,
                  const double *ar, double *res);
void orc_dic_pivots(int n, const int *ioffset, const int *ja, const int *diag, const double *a, double *d);
int  orc_dilu_pivots(int n, const int *ioffset, const int *ja, const int *diag, const double *a, double *d);
void orc_lu_apply(int n, const int *ioffset, const int *ja, const int *diag, const double *a,
                  const double *d, double small, const double *rhs, double *z);

/* Solver variants: the module solvers of src/ and the explicit-argument
 * copies of tests/test_sparse_solvers.f90 differ in constants. */
enum { ORC_VARIANT_SRC = 0, ORC_VARIANT_TESTS = 1 };

typedef struct {
    double res0, resl, rsm;
    int iters;
    int early_exit;      /* res0 < tol */
} orc_stats;

/* History (optional, may be NULL): resl after each iteration, capacity nsw. */
int orc_iccg(int variant, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
             const double *a, const double *su, double *fi, int noc, const int *ijl1,
             const int *ijr1, const double *al, const double *ar, double sor, int nsw,
             double *res_out, double *history, orc_stats *st);
int orc_dpcg(int variant, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
             const double *a, const double *su, double *fi, int noc, const int *ijl1,
             const int *ijr1, const double *al, const double *ar, double sor, int nsw,
             double *res_out, double *history, orc_stats *st);
int orc_bicgstab(int variant, int n, int nnz, const int *ioffset, const int *ja, const int *diag,
                 const double *a, const double *su, double *fi, int noc, const int *ijl1,
                 const int *ijr1, const double *al, const double *ar, double sor, int nsw,
                 double *res_out, double *history, orc_stats *st);

#ifdef __cplusplus
}
#endif
#endif
