/*
 * fcap.h -- C ABI of the MI355X pressure-solve library (libfcap.so, gfx950).
 *
 * Drop-in replacements for the reference's hot path.  The reference passes
 * everything through Fortran module globals; these entry points take the same
 * arrays as explicit arguments, in the reference's layout:
 *
 *   - every index array is 1-based (owner, neighbour, ioffset, ja, diag, the
 *     face slot maps, ijl, ijr, face node ids);
 *   - CSR as in src/sparse_matrix.f90:8-22: ioffset(n+1), ja(nnz), diag(n),
 *     a(nnz), nnz = 2*numInnerFaces + numCells;
 *   - real(dp) = double, integer = int (32-bit).
 *
 * Host-pointer entry points ("fc_<name>") copy their inputs to the GPU,
 * compute there and copy results back; the caller owns every array and nothing
 * is allocated across the boundary.  A context caches the device copy of the
 * CSR structure (keyed by n, nnz and a checksum of ioffset/ja) together with
 * its level schedules, so repeated solves on one mesh upload only values.
 *
 * Device-resident entry points ("fc_sys_*") keep a mesh, its CSR structure,
 * the assembled system and the solver state in HBM for callers that loop
 * assembly + solve without host round trips.
 *
 * Every function returns an fc_status; none aborts.  One host thread per
 * context.
 */
#ifndef FCAP_H
#define FCAP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FC_OK = 0,                /* converged (rsm < sor)                                  */
    FC_NOT_CONVERGED = 1,     /* nsw iterations used without reaching sor (the reference
                                 returns silently; src/iccg.f90:89-171)                 */
    FC_EARLY_EXIT = 2,        /* initial L1 residual below 1e-13: no iteration
                                 (src/iccg.f90:67-71)                                   */
    FC_FALLBACK_USED = 3,     /* flexible CG stagnated; IC(0)-PCG finished the solve    */
    FC_ERR_ARGUMENT = -1,
    FC_ERR_HIP = -2,
    FC_ERR_NONFINITE = -3,    /* NaN/Inf in the matrix, RHS or an iterate               */
    FC_ERR_ZERO_PIVOT = -4,   /* DIC/DILU pivot is zero or not finite                   */
    FC_ERR_STRUCTURE = -5,    /* malformed owner/neighbour or CSR (duplicates, range)   */
    FC_ERR_ALLOC = -6,
    FC_ERR_NO_DEVICE = -7
} fc_status;

typedef struct fc_context fc_context;
typedef struct fc_system fc_system;

/* Per-solve report.  res0/resl/rsm are the reference's L1 quantities
 * (src/iccg.f90:65,154,164); rel_l2 = ||su - A fi||_2 / ||su||_2 recomputed
 * in FP64 from scratch after the solve. */
typedef struct {
    double res0, resl, rsm;
    double rel_l2;
    int iters;                /* inner Krylov iterations in total                        */
    int outer;                /* refinement passes (mixed-precision methods), else 0     */
    int status;               /* fc_status of the solve                                 */
    int fallback_iters;       /* IC(0)-PCG iterations after a stagnation fallback        */
    double setup_ms;          /* pivots / schedules / network set-up, device time        */
    double solve_ms;          /* iteration loop, device time                             */
} fc_solve_info;

/* ---------------------------------------------------------------- context */
fc_status fc_init(fc_context **ctx, int device);
fc_status fc_finalize(fc_context *ctx);
/* Run all work of this context on an existing hipStream_t (NULL: the context's own). */
fc_status fc_set_stream(fc_context *ctx, void *hip_stream);
const char *fc_last_error(const fc_context *ctx);
/* Print the reference's one-line solver report after each host-pointer solve
 * (src/iccg.f90:174-175): 0 off (default), 1 on. */
fc_status fc_set_report(fc_context *ctx, int on, const char *variable_name);

/* -------------------------------------------------- mesh geometry and CSR */

/* Face areas, face centres, cell volumes, cell centres and the interpolation
 * factor, as src/mesh_geometry_and_topology.f90:859-1081 computes them (fan
 * triangulation from node 1; accumulation in face order, bit-identical).
 * face_ptr is 0-based (numFaces+1), face_nodes 1-based.  O-C and cyclic
 * faces are not handled. */
fc_status fc_mesh_geometry(fc_context *ctx, int numNodes, const double *xyz /* numNodes*3 */,
                           int numFaces, int numInnerFaces, const int *face_ptr,
                           const int *face_nodes, const int *owner, const int *neighbour,
                           int numCells,
                           double *arx, double *ary, double *arz,
                           double *xf, double *yf, double *zf,
                           double *vol, double *xc, double *yc, double *zc, double *facint);

/* create_CSR_matrix_from_mesh_data (src/sparse_matrix.f90:42-172): ioffset,
 * ja, diag and the face slot maps icell_jcell_csr_value_index (ij_slot) and
 * jcell_icell_csr_value_index (ji_slot).  Bit-identical to the reference. */
fc_status fc_csr_build(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                       const int *neighbour, int *ioffset, int *ja, int *diag, int *ij_slot,
                       int *ji_slot);

/* ------------------------------------------------------------- assembly */

/* laplacian(mu, phi) (src/fvm_laplacian.f90:1-163): a is overwritten, su is
 * accumulated (the reference never zeroes it).  Wall faces are
 * iWallFacesStart+1 .. iWallFacesStart+nwal (1-based face numbers), their
 * boundary values phi(iWallStart+i) (phi has numTotal entries).  O-C faces
 * (noc > 0): ijl/ijr/foc of the cuts, faces iOCFacesStart+1..; al/ar out. */
fc_status fc_laplacian(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                       const int *neighbour, const double *arx, const double *ary,
                       const double *arz, const double *facint, const double *xc,
                       const double *yc, const double *zc, const double *xf, const double *yf,
                       const double *zf, const double *mu, const double *phi, int numTotal,
                       int nwal, int iWallFacesStart, int iWallStart, int noc, int iOCFacesStart,
                       const int *ijl, const int *ijr, const double *foc, const int *ioffset,
                       const int *ja, const int *diag, const int *ij_slot, const int *ji_slot,
                       int nnz, double *a, double *su, double *al, double *ar);

/* Momentum-like convection-diffusion matrix: facefluxuvw coefficients with
 * upwind convection (src/faceflux_velocity.f90:97-108) and the calcuvw
 * diagonal (src/calcuvw.f90:417-432): a(diag) = (spu - sum of off-diagonals)
 * / urf, spu = vis|S|/|x_P - x_f| summed over the cell's wall faces. */
fc_status fc_convdiff(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                      const int *neighbour, const double *arx, const double *ary,
                      const double *arz, const double *facint, const double *xc,
                      const double *yc, const double *zc, const double *xf, const double *yf,
                      const double *zf, const double *vis, const double *flux, int nwal,
                      int iWallFacesStart, const int *ioffset, const int *ja, const int *diag,
                      const int *ij_slot, const int *ji_slot, int nnz, double urf, double *a);

/* Pressure-correction assembly: the face loops of calcp
 * (src/calcp-multiple_correction_SIMPLE.f90:34-107) around facefluxmass
 * (src/facefluxmass.f90:34-199).  a and su are overwritten (the reference
 * zeroes both); per inner face the coefficient cap = can =
 * -rho_f (Vol/a_P)_f |S|^2 / (S . d_PN n) goes to a(ij), a(ji) and off the two
 * diagonals, and the Rhie-Chow mass flux flmass to su(P) -= F, su(N) += F,
 * all in face order; then the O-C cuts (iOCFacesStart+1.., al = cap, ar = can,
 * fmoc).  Cell fields (den, apu/apv/apw = 1/a_P, u, v, w, p) have at least
 * numCells entries; dUdxi .. dPdxi are the reference's gradient arrays
 * (3, numCells).  Face arrays cover every face the loops touch. */
fc_status fc_calcp_assemble(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                            const int *neighbour, const double *xc, const double *yc, const double *zc,
                            const double *vol, const double *xf, const double *yf, const double *zf,
                            const double *arx, const double *ary, const double *arz, const double *facint,
                            const double *den, const double *apu, const double *apv, const double *apw,
                            const double *u, const double *v, const double *w, const double *p,
                            const double *dUdxi, const double *dVdxi, const double *dWdxi,
                            const double *dPdxi, int noc, int iOCFacesStart, const int *ijl,
                            const int *ijr, const double *foc, const int *ioffset, const int *ja,
                            const int *diag, const int *ij_slot, const int *ji_slot, int nnz, double *a,
                            double *su, double *flmass, double *al, double *ar, double *fmoc);

/* Non-orthogonal corrector source of calcp
 * (src/calcp-multiple_correction_SIMPLE.f90:185-229) with fluxmc
 * (src/facefluxmass.f90:520-607): per inner face, then per O-C cut,
 * fmcor = rho_f (Vol/a_P)_f |S| ((grad p_N . r_N') - (grad p_P . r_P')) / (d_PN . n),
 * r' the offset of the cell centre's projection onto the face normal line;
 * flmass (fmoc) += fmcor, and su, overwritten with zero first, gets
 * su(P) -= fmcor, su(N) += fmcor in face order.  dPdxi is the gradient of the
 * pressure correction (3, numCells). */
fc_status fc_calcp_nonorth(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                           const int *neighbour, const double *xc, const double *yc, const double *zc,
                           const double *vol, const double *xf, const double *yf, const double *zf,
                           const double *arx, const double *ary, const double *arz, const double *facint,
                           const double *den, const double *apu, const double *dPdxi, int noc,
                           int iOCFacesStart, const int *ijl, const int *ijr, const double *foc,
                           const int *ioffset, const int *ja, const int *diag, const int *ij_slot,
                           const int *ji_slot, int nnz, double *su, double *flmass, double *fmoc);

/* PISO's pressure assembly: the face loop of src/PISO_multiple_correction.f90:
 * 107-148 around facefluxmass_piso (src/facefluxmass.f90:294-358), with the
 * same scatter and arguments as fc_calcp_assemble, but the coefficient cap =
 * can = -rho_f (fxp Vol_P apu_P + fxn Vol_N apu_N) |S| / |d_PN| and the mass
 * flux rho_f (u_f . S) without Rhie-Chow terms (so no apv/apw, p or dPdxi).
 * The reference-pressure pin (PISO:188-192) is fc_sys_pin_reference, or the
 * caller's own loop on the host arrays. */
fc_status fc_piso_assemble(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                           const int *neighbour, const double *xc, const double *yc, const double *zc,
                           const double *vol, const double *xf, const double *yf, const double *zf,
                           const double *arx, const double *ary, const double *arz, const double *facint,
                           const double *den, const double *apu, const double *u, const double *v,
                           const double *w, const double *dUdxi, const double *dVdxi,
                           const double *dWdxi, int noc, int iOCFacesStart, const int *ijl,
                           const int *ijr, const double *foc, const int *ioffset, const int *ja,
                           const int *diag, const int *ij_slot, const int *ji_slot, int nnz, double *a,
                           double *su, double *flmass, double *al, double *ar, double *fmoc);

/* bpres(p, istage) (src/bpres.f90:3-152): boundary values of p (numTotal:
 * cells, then the boundary values in the groups' order of bnd, then the O-C
 * values).  istage 1: every boundary value takes its owner cell's value;
 * istage >= 2: the wall and pressure-outlet values, and the O-C values, are
 * extrapolated linearly from the owner with dPdxi (3, numCells), the other
 * values are left as they are.  The O-C branch reads the face centre of face
 * number i for cut i, as the reference does (bpres.f90:140-142).  bnd as for
 * fc_grad_gauss; owner and the face centres cover numFaces. */
fc_status fc_bpres(fc_context *ctx, int numCells, int numTotal, int numFaces, int numInnerFaces,
                   const int *owner, const double *xc, const double *yc, const double *zc,
                   const double *xf, const double *yf, const double *zf, const int *bnd, int noc,
                   int iOCFacesStart, const double *dPdxi, int istage, double *p);

/* correctBoundaryConditionsVelocity (src/correctBoundaryConditionsVelocity.f90:
 * 1-79), called by calcp after the corrections (line 183): over the outlet
 * faces the boundary velocity takes the owner's, fmo = den (u . S), flowo =
 * their sum in outlet order, fac = flomas / (flowo + small), then fmo and the
 * boundary velocities are scaled by fac; over the symmetry faces the boundary
 * velocity is the owner's minus (u . S) S (S the area vector, as written in
 * the reference).  u, v, w have numTotal values; den numCells; fmo (nout) out. */
fc_status fc_correct_velocity_bc(fc_context *ctx, int numCells, int numTotal, int numFaces,
                                 int numInnerFaces, const int *owner, const double *arx,
                                 const double *ary, const double *arz, const int *bnd, double flomas,
                                 const double *den, double *u, double *v, double *w, double *fmo);

/* adjustMassFlow (src/adjustMassFlow.f90:1-71), called by calcp between the
 * face loops and the solve (src/calcp-multiple_correction_SIMPLE.f90:107):
 * su(owner) -= fmi over the inlet faces iInletFacesStart+1..+ninl; over the
 * outlet faces iOutletFacesStart+1..+nout, fmo = den (u.S) of the owner cell,
 * flowo = their sum in outlet order, fac = flomas / (flowo + small); then
 * fmo *= fac, the outlet boundary values u(iOutletStart+i) = u(owner) fac
 * (v and w alike) and su(owner) -= fmo.  owner and the face areas cover
 * numFaces; u, v and w hold numTotal values. */
fc_status fc_adjust_mass_flow(fc_context *ctx, int numCells, int numTotal, int numFaces, const int *owner,
                              const double *arx, const double *ary, const double *arz, int ninl,
                              int iInletFacesStart, const double *fmi, int nout, int iOutletFacesStart,
                              int iOutletStart, double flomas, const double *den, double *u, double *v,
                              double *w, double *su, double *fmo);

/* Corrections after a pressure-correction solve
 * (src/calcp-multiple_correction_SIMPLE.f90:145-183): ppref = pp(pRefCell);
 * flmass += a(ij) (pp(N) - pp(P)) per inner face; fmoc += ar (pp(ijr) - pp(ijl))
 * per O-C cut; u -= dPdxi(1) Vol apu, v and w alike; p += urf (pp - ppref).
 * a is the assembled pressure-correction matrix, dPdxi the gradient of pp. */
fc_status fc_calcp_correct(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                           const int *neighbour, const int *ioffset, const int *ja, const int *diag,
                           const int *ij_slot, const int *ji_slot, int nnz, const double *a,
                           const double *pp, const double *dPdxi, const double *vol, const double *apu,
                           const double *apv, const double *apw, double urf, int pRefCell, int noc,
                           const int *ijl, const int *ijr, const double *ar, double *fmoc, double *flmass,
                           double *u, double *v, double *w, double *p);

/* Gauss gradient (src/grad_gauss.f90:1-124 with gradco :128-190 and gradbc
 * :194-211): nigrad iterations, each summing per cell the inner faces' value
 * times S (the value at the face centre extrapolated with the previous
 * iteration's gradient), then the O-C cuts (faces ijlFace, 1-based), then the
 * boundary faces with their boundary values, then dividing by the volume.
 * u has numTotal entries: cells, then the boundary values in the reference's
 * group order.  bnd (10 ints): (count, faces before the group) for the inlet,
 * outlet, symmetry, wall and pressure-outlet groups (ninl, iInletFacesStart,
 * nout, iOutletFacesStart, nsym, iSymmetryFacesStart, nwal, iWallFacesStart,
 * npru, iPressOutletFacesStart).  owner and the face arrays cover numFaces. */
fc_status fc_grad_gauss(fc_context *ctx, int numCells, int numFaces, int numInnerFaces, const int *owner,
                        const int *neighbour, const double *arx, const double *ary, const double *arz,
                        const double *facint, const double *xc, const double *yc, const double *zc,
                        const double *xf, const double *yf, const double *zf, const double *vol,
                        const double *u, int numTotal, const int *bnd, int noc, const int *ijlFace,
                        const int *ijl, const int *ijr, const double *foc, int nigrad, double *dudx,
                        double *dudy, double *dudz);

/* Face mass flux F_f = u_f . S_f, u_f = (1-facint) u_P + facint u_N, and the
 * continuity source su(o) -= F, su(n) += F of
 * src/calcp-multiple_correction_SIMPLE.f90:74-75 (accumulated in face order). */
fc_status fc_face_flux_central(fc_context *ctx, int numInnerFaces, const int *owner,
                               const int *neighbour, const double *facint, const double *arx,
                               const double *ary, const double *arz, const double *u,
                               const double *v, const double *w, int numCells, double *flux);
fc_status fc_rhs_divergence(fc_context *ctx, int numCells, int numInnerFaces, const int *owner,
                            const int *neighbour, const double *flux, double *su);

/* ------------------------------------------------------------- kernels */

/* y = A x in the reference's row order (src/iccg.f90:129-134), FP64. */
fc_status fc_spmv(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                  const double *a, const double *x, double *y);

/* DIC (src/iccg.f90:77-83) / DILU (src/bicgstab.f90:68-79) reciprocal pivots d. */
fc_status fc_dic_pivots(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                        const int *diag, const double *a, double *d);
fc_status fc_dilu_pivots(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                         const int *diag, const double *a, double *d);
/* z = M^-1 r, M = (D+L) D^-1 (D+U): forward sweep, z/(d+small), backward
 * sweep (src/iccg.f90:94-111), level-scheduled. */
fc_status fc_lu_apply(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                      const int *diag, const double *a, const double *d, const double *r,
                      double *z);

/* ------------------------------------------------------------- solvers */

/* Explicit-argument forms of iccg(fi,ifi), dpcg(fi,ifi), bicgstab(fi,ifi)
 * (src/iccg.f90, src/dpcg.f90, src/bicgstab.f90).  fi holds the initial
 * guess on entry (numCells entries are read and written).  res (may be NULL)
 * receives the final residual vector, like the module array res.
 * sor = relative L1 tolerance sor(ifi), nsw = iteration cap nsw(ifi).
 * noc/ijl/ijr/al/ar: O-C couplings (noc = 0 when none).  info may be NULL. */
fc_status fc_iccg(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                  const int *diag, const double *a, const double *su, double *fi, double *res,
                  int noc, const int *ijl, const int *ijr, const double *al, const double *ar,
                  double sor, int nsw, fc_solve_info *info);
fc_status fc_dpcg(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                  const int *diag, const double *a, const double *su, double *fi, double *res,
                  int noc, const int *ijl, const int *ijr, const double *al, const double *ar,
                  double sor, int nsw, fc_solve_info *info);
fc_status fc_bicgstab(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                      const int *diag, const double *a, const double *su, double *fi, double *res,
                      int noc, const int *ijl, const int *ijr, const double *al, const double *ar,
                      double sor, int nsw, fc_solve_info *info);

/* dpcg with no device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  As fc_dpcg, but every
 * array is a device pointer on the context's device (res may be NULL), and
 * there are no O-C couplings.  numCells is at most 1,048,576 (larger:
 * FC_ERR_ARGUMENT); nnz is not limited, the CSR arrays are read in place.
 * Every other buffer lives in one static device array of about 32 MiB that
 * each GPU the library runs on holds until the process ends; calls that
 * overlap on one GPU share it and race.  Its kernels and sums are those of
 * fc_dpcg, so the results are bitwise equal.  info->solve_ms is host wall
 * time, setup_ms 0. */
fc_status fc_dpcg_noalloc(fc_context *ctx, int n, int nnz, const int *ioffset, const int *ja,
                          const int *diag, const double *a, const double *su, double *fi, double *res,
                          double sor, int nsw, fc_solve_info *info);

/* Solver variants.  FC_VARIANT_SRC: the module solvers of src/ (small = 1e-20,
 * early exit below 1e-13, O-C terms).  FC_VARIANT_TESTS: the explicit-argument
 * copies of tests/test_sparse_solvers.f90 (small = 1e-30, sor = 1e-13,
 * nsw = 100, no early exit).  Default SRC. */
enum { FC_VARIANT_SRC = 0, FC_VARIANT_TESTS = 1 };
fc_status fc_set_variant(fc_context *ctx, int variant);

/* Order of the sums of the FP64 solvers (iccg, dpcg, bicgstab: fc_iccg,
 * fc_dpcg, fc_bicgstab, fc_sys_solve with FC_METHOD_ICCG / _DPCG / _BICGSTAB,
 * fc_sys_solve_momentum with FC_METHOD_BICGSTAB).
 *   FC_REDUCE_TREE        per-block partials added in a fixed tree (default;
 *                         deterministic, full speed).
 *   FC_REDUCE_SEQUENTIAL  every dot product and norm formed in index order, as
 *                         the reference's sum(...) forms it: the iterates,
 *                         residuals, L1 history and iteration count are then
 *                         bit-identical to the reference's serial solvers.  One
 *                         wave adds the terms one by one (about 10 ns per entry
 *                         and sum): a verification mode, not a fast path.
 * No counterpart in the reference. */
enum { FC_REDUCE_TREE = 0, FC_REDUCE_SEQUENTIAL = 1 };
fc_status fc_set_reduction_order(fc_context *ctx, int order);

/* Tuning of the FP32-in-FP64 solvers; no counterpart in the reference.
 * fc_set_inner_tolerance: relative residual reduction of each inner FP32 solve
 * (0 = the method's default: 1e-2 for FC_METHOD_FCG_LEARNED and FC_METHOD_ICCG_MIXED,
 * 1e-3 for FC_METHOD_BICGSTAB_MIXED).  FC_METHOD_ICCG_MIXED keeps its search direction
 * from one refinement cycle to the next: each cycle replaces the FP32 residual by the
 * refined FP64 one and continues the same conjugate-gradient sequence.
 * fc_set_fcg_window: directions kept by flexible CG, 1..16, rounded up to a
 * multiple of 4; 0 (default): 4 with the coarse correction (fc_set_fcg_coarse),
 * 12 without it. */
fc_status fc_set_inner_tolerance(fc_context *ctx, double eta);
fc_status fc_set_fcg_window(fc_context *ctx, int m);
/* Preconditioner of FC_METHOD_FCG_LEARNED: the network, then `steps` damped-Jacobi
 * steps z <- z + omega (r - A z) / a_ii (0..8 steps, 0 < omega < 2; default 3
 * steps, omega 0.8).  No counterpart in the reference. */
fc_status fc_set_fcg_smoothing(fc_context *ctx, int steps, double omega);
/* Coarse correction of FC_METHOD_FCG_LEARNED's preconditioner, applied after the
 * Jacobi steps: z <- z + P A_c^{-1} P^T (r - A z), P the indicator of aggregates
 * of `cells_per_aggregate` consecutive cells of the network's Morton order
 * (compact bricks of the mesh; cell order where the system has no cluster
 * layout), A_c = P^T A P factored by dense Cholesky once per assembled matrix.
 * A mesh of more than 4,096 x cells_per_aggregate cells gets larger aggregates
 * (at most 4,096 of them).  0 switches it off; otherwise a multiple of 256
 * (default 4,096).  The correction is skipped when A_c is not positive
 * definite.  No counterpart in the reference. */
fc_status fc_set_fcg_coarse(fc_context *ctx, int cells_per_aggregate);
/* How often FC_METHOD_FCG_LEARNED's preconditioner runs the network: on the
 * first of every `period` iterations (1..4); the others start from the Jacobi
 * guess r / a_ii instead, with the same Jacobi steps and coarse correction
 * (flexible CG allows a preconditioner that changes between iterations).
 * 0 (default): 2 with the coarse correction, 1 without it.  No counterpart in
 * the reference. */
fc_status fc_set_fcg_network_period(fc_context *ctx, int period);
/* Layout of the network's inference layers (FC_METHOD_FCG_LEARNED, fc_gnn_apply).
 *   FC_NET_CLUSTERS   (default) the cells renumbered into compact clusters of
 *                     256 (Morton order of their per-axis centroid ranks); one
 *                     workgroup per cluster stages the cluster's rows and its
 *                     halo (at most 384 outside neighbours) in LDS and reads
 *                     every message from there.  Used for ELL widths 4, 6 and 8
 *                     when every halo fits; else cell order.
 *   FC_NET_CELL_ORDER rows in cell order, each neighbour row gathered from
 *                     global memory by four lanes.
 * Both give bitwise the same z.  No counterpart in the reference. */
enum { FC_NET_CLUSTERS = 0, FC_NET_CELL_ORDER = 1 };
fc_status fc_set_network_layout(fc_context *ctx, int layout);
/* Stagnation test of FC_METHOD_FCG_LEARNED's inner solves: the first `window`
 * iterations only set the reference; after every later window the window's
 * smallest ||r||_2^2 must be below factor x the smallest before it, else plain
 * FP64 IC(0)-PCG finishes the solve from the last refined iterate
 * (FC_FALLBACK_USED) with what is left of nsw and the caller's stopping rule.
 * window 0 switches the test off; 0 <= factor <= 1 (factor 0 forces the
 * fallback after the second window).  Default 64 iterations, 0.81.  No
 * counterpart in the reference. */
fc_status fc_set_fcg_stagnation(fc_context *ctx, int window, double factor);

/* Schedule of the DIC/DILU triangular sweeps (src/iccg.f90:94-111,
 * src/bicgstab.f90:107-124); every schedule computes each row with the same
 * operations in the same order, so results are bit-identical.
 *   FC_SWEEP_LEVELS    one launch per level;
 *   FC_SWEEP_DEPENDENT one launch per sweep, each row waiting only for the rows
 *                      it reads (packed structures of width 3, 4 or 8; other
 *                      structures keep the level launches);
 *   FC_SWEEP_PARTS     (default) one launch per sweep over parts of 4,096
 *                      consecutive rows, one persistent wave per part at a time,
 *                      64 rows per window in a schedule fixed at analysis:
 *                      inputs inside the part are read from the wave's LDS, only
 *                      inputs from other parts wait on global memory.  Used when
 *                      at least half of each triangle's entries lie inside their
 *                      row's part (meshes numbered with locality, as the
 *                      reference's are); else FC_SWEEP_DEPENDENT where levels
 *                      are narrow (fewer than 65,536 rows on average) and
 *                      FC_SWEEP_LEVELS where they are wide.
 * A wait that runs past its time limit fails the call with FC_ERR_HIP.
 * No counterpart in the reference. */
enum { FC_SWEEP_LEVELS = 0, FC_SWEEP_DEPENDENT = 2, FC_SWEEP_PARTS = 3 };
fc_status fc_set_sweep_mode(fc_context *ctx, int mode);
/* Persistent waves of the dependency-driven sweeps and pivots (0: automatic --
 * two per compute unit for the part-wise sweeps, for the row-wise ones one
 * 64-row block per wave for an average level's rows, within 1 to 16 per compute
 * unit; at most 65,536).  No counterpart in the reference. */
fc_status fc_set_sweep_waves(fc_context *ctx, int waves);
/* Wall-clock limit of one wait of the dependency-driven sweeps, in
 * microseconds (0: the built-in 2 s).  A wait that runs past it ends the
 * sweep and the call returns FC_ERR_HIP.  No counterpart in the reference. */
fc_status fc_set_sweep_wait_limit(fc_context *ctx, int microseconds);
/* Replay each solver iteration as a captured HIP graph (1, default) or launch
 * its kernels one by one (0; e.g. under a profiler).  Same results either way.
 * No counterpart in the reference. */
fc_status fc_set_graphs(fc_context *ctx, int on);

/* ------------------------------------------------ device-resident system */

enum {
    FC_METHOD_ICCG = 1,            /* DIC-PCG, FP64 (reference semantics)                 */
    FC_METHOD_DPCG = 2,            /* Jacobi-PCG, FP64                                    */
    FC_METHOD_BICGSTAB = 3,        /* DILU-BiCGStab, FP64                                 */
    FC_METHOD_ICCG_MIXED = 11,     /* FP32 DIC-PCG inside FP64 iterative refinement       */
    FC_METHOD_BICGSTAB_MIXED = 13, /* FP32 DILU-BiCGStab inside FP64 iterative refinement */
    FC_METHOD_FCG_LEARNED = 21,    /* flexible CG, BF16 message-passing preconditioner,
                                      FP32 inner / FP64 refinement, IC(0)-PCG fallback    */
    FC_METHOD_FCG_JACOBI = 22      /* FC_METHOD_FCG_LEARNED without the network: the Jacobi
                                      guess, the same damped-Jacobi steps and coarse
                                      correction (no network needed); no counterpart in
                                      the reference                                      */
};

/* Stopping rule of fc_sys_solve: FC_STOP_L1 is the reference's
 * ||r||_1/||r0||_1 < sor; FC_STOP_L2 is ||su - A fi||_2/||su||_2 < sor (FP64). */
enum { FC_STOP_L1 = 0, FC_STOP_L2 = 1 };

/* Create a device-resident system from a mesh (arrays as for fc_mesh_geometry;
 * patches: wall faces iWallFacesStart+1..+nwal, wall values at iWallStart+i).
 * Geometry, CSR structure, level schedules and slot maps are built on the GPU. */
fc_status fc_sys_create(fc_context *ctx, fc_system **sys, int numNodes, const double *xyz,
                        int numFaces, int numInnerFaces, const int *face_ptr,
                        const int *face_nodes, const int *owner, const int *neighbour,
                        int numCells, int nwal, int iWallFacesStart, int iWallStart);
fc_status fc_sys_destroy(fc_system *sys);

/* Assemble the Laplacian (mu: numCells, phi: numTotal, su0: numCells or NULL
 * for zero) / the convection-diffusion matrix (vis, flux: host arrays; su0 as
 * for the Laplacian) into the system, on the GPU. */
fc_status fc_sys_assemble_laplacian(fc_system *sys, const double *mu, const double *phi,
                                    const double *su0);
fc_status fc_sys_assemble_convdiff(fc_system *sys, const double *vis, const double *flux,
                                   double urf, const double *su0);
/* Pressure-correction assembly of calcp (src/calcp-multiple_correction_SIMPLE.f90:34-107,
 * src/facefluxmass.f90:34-199) on the system's own mesh: as fc_calcp_assemble, with
 * the geometry, CSR structure and face-to-slot maps the system already holds, and
 * a and su written into the system, so that fc_sys_solve then solves for the
 * pressure correction (calcp's iccg call, line 119).  Cell fields have numCells
 * entries, gradients (3, numCells); flmass (numInnerFaces) is written.  Each of
 * these pointers may be host or device memory: device pointers are read and
 * written in place, host arrays pass through buffers the system keeps.  The
 * system's solution vector (the pressure correction pp) is set to zero, as
 * calcp sets pp = 0 before its solve (line 112). */
fc_status fc_sys_calcp_assemble(fc_system *sys, const double *den, const double *apu,
                                const double *apv, const double *apw, const double *u,
                                const double *v, const double *w, const double *p,
                                const double *dUdxi, const double *dVdxi, const double *dWdxi,
                                const double *dPdxi, double *flmass);
/* Gauss gradient on the system's own mesh (src/grad_gauss.f90:1-124, the
 * calcp calls grad(U,dUdxi) .. grad(pp,dPdxi) at
 * src/calcp-multiple_correction_SIMPLE.f90:38-41,138): as fc_grad_gauss with the
 * system's geometry and face order, written in the reference's (3, numCells)
 * layout that fc_sys_calcp_assemble reads.  u has the system's numTotal
 * entries (cells, then the boundary values in group order); bnd as for
 * fc_grad_gauss (the per-cell boundary lists are kept until bnd changes).  u
 * and dudxi may be host or device memory, as for fc_sys_calcp_assemble. */
fc_status fc_sys_grad_gauss(fc_system *sys, const int *bnd, const double *u, int nigrad, double *dudxi);
/* Corrections of calcp (src/calcp-multiple_correction_SIMPLE.f90:145-183) on
 * the resident system: as fc_calcp_correct with a the system's matrix (the
 * last fc_sys_calcp_assemble) and the system's geometry and slot maps.  pp
 * (numCells) may be NULL for the system's solution (the last fc_sys_solve);
 * dPdxi is (3, numCells), e.g. from fc_sys_grad_gauss.  flmass
 * (numInnerFaces), u, v, w, p (numCells) are updated in place.  Host or
 * device memory, as for fc_sys_calcp_assemble. */
fc_status fc_sys_calcp_correct(fc_system *sys, const double *pp, const double *dPdxi,
                               const double *apu, const double *apv, const double *apw,
                               double urf, int pRefCell, double *flmass, double *u, double *v,
                               double *w, double *p);
/* Source of calcp's non-orthogonal corrector
 * (src/calcp-multiple_correction_SIMPLE.f90:185-229, fluxmc
 * src/facefluxmass.f90:520-607) on the resident system: as fc_calcp_nonorth,
 * with su written into the system for the corrector's next fc_sys_solve and
 * flmass (numInnerFaces) updated in place.  Host or device memory, as for
 * fc_sys_calcp_assemble.  The system's solution vector is then set to zero
 * for the next corrector pass (calcp:112). */
fc_status fc_sys_calcp_nonorth(fc_system *sys, const double *den, const double *apu,
                               const double *dPdxi, double *flmass);
/* PISO's pressure assembly (fc_piso_assemble) on the resident system; a and su
 * are written into the system and its solution vector is left as it is (PISO
 * starts each solve from the last p, src/PISO_multiple_correction.f90:198-199).
 * Host or device memory, as for fc_sys_calcp_assemble. */
fc_status fc_sys_piso_assemble(fc_system *sys, const double *den, const double *apu, const double *u,
                               const double *v, const double *w, const double *dUdxi, const double *dVdxi,
                               const double *dWdxi, double *flmass);
/* PISO's reference pressure (src/PISO_multiple_correction.f90:188-192) in the
 * system's matrix: row pRefCell (1-based) zeroed, its diagonal 1, su(pRefCell)
 * = p(pRefCell).  p (numCells) host or device memory. */
fc_status fc_sys_pin_reference(fc_system *sys, int pRefCell, const double *p);
/* bpres on the resident system (fc_bpres without O-C cuts): p has the system's
 * numTotal values and is updated in place; dPdxi (3, numCells) is read from
 * istage 2 on.  Host or device memory. */
fc_status fc_sys_bpres(fc_system *sys, const int *bnd, int istage, const double *dPdxi, double *p);
/* The corrector's pressure-correction gradient (src/calcp-multiple_correction_
 * SIMPLE.f90:131-136) on the system's solution pp: for istage = 1..nipgrad,
 * bpres(pp, istage) then the Gauss gradient with nigrad iterations into dPdxi
 * (3, numCells, host or device), each stage extrapolating with the previous
 * stage's gradient.  pp's boundary values are kept in the system. */
fc_status fc_sys_grad_pp(fc_system *sys, const int *bnd, int nipgrad, int nigrad, double *dPdxi);
/* correctBoundaryConditionsVelocity (fc_correct_velocity_bc) on the resident
 * system's mesh: u, v, w (numTotal) updated in place, fmo (nout) written; host
 * or device memory. */
fc_status fc_sys_correct_velocity_bc(fc_system *sys, const int *bnd, double flomas, const double *den,
                                     double *u, double *v, double *w, double *fmo);
/* Upload a RHS (numCells) replacing su. */
fc_status fc_sys_set_rhs(fc_system *sys, const double *su);
/* Momentum solves of calcuvw (src/calcuvw.f90:414-556): the off-diagonal
 * coefficients of the last fc_sys_assemble_convdiff are shared by the three
 * components; for c = u, v, w in turn the row kernel rebuilds a(diag) =
 * (sp[c] - sum of the row's off-diagonals) / urf[c], su = src[c] + (1 -
 * urf[c]) a(diag) phi[c] and ap[c] = 1/(a(diag) + small) (apu/apv/apw of the
 * reference), then phi[c] (numCells, in: initial guess, out: solution) is
 * solved with BiCGStab-DILU (method FC_METHOD_BICGSTAB, the reference's L1
 * rule with sor/nsw) or its FP32-in-FP64 form (FC_METHOD_BICGSTAB_MIXED).
 * ap may be NULL; info (3 entries) may be NULL.  All three components are
 * solved whatever each solve's outcome, as calcuvw does; the call returns
 * FC_NOT_CONVERGED if any component ran out of nsw, and stops only on an
 * error (status < 0). */
fc_status fc_sys_solve_momentum(fc_system *sys, const double *const *sp, const double *const *src, const double *urf,
                                double *const *phi, double *const *ap, int method, double sor, int nsw,
                                fc_solve_info *info);
/* Upload / download the solution vector (numCells). */
fc_status fc_sys_set_solution(fc_system *sys, const double *fi);
fc_status fc_sys_get_solution(fc_system *sys, double *fi);

/* Device-side assembly and solve on the resident inputs (no host copies).
 * fc_sys_assemble_laplacian_dev uses mu/phi/su0 already on the device
 * (uploaded by a previous fc_sys_assemble_laplacian). */
fc_status fc_sys_assemble_laplacian_dev(fc_system *sys);
fc_status fc_sys_solve(fc_system *sys, int method, int stop_rule, double sor, int nsw,
                       fc_solve_info *info);

/* Read back arrays of the resident system.  which: */
enum {
    FC_ARR_A = 1, FC_ARR_SU = 2, FC_ARR_IOFFSET = 3, FC_ARR_JA = 4, FC_ARR_DIAG = 5,
    FC_ARR_IJ_SLOT = 6, FC_ARR_JI_SLOT = 7, FC_ARR_VOL = 8, FC_ARR_XC = 9, FC_ARR_YC = 10,
    FC_ARR_ZC = 11, FC_ARR_FACINT = 12, FC_ARR_ARX = 13, FC_ARR_ARY = 14, FC_ARR_ARZ = 15,
    FC_ARR_XF = 16, FC_ARR_YF = 17, FC_ARR_ZF = 18, FC_ARR_RES = 19, FC_ARR_PIVOTS = 20
};
fc_status fc_sys_get(fc_system *sys, int which, void *host_out, size_t bytes);
/* Device pointer of an array of the resident system (for callers that share
 * the device memory, e.g. a training loop); NULL if unknown. */
void *fc_sys_device_ptr(fc_system *sys, int which);
/* Sizes: numCells, numInnerFaces, nnz, forward and backward level counts. */
fc_status fc_sys_sizes(fc_system *sys, int *n, int *nif, int *nnz, int *nlev_fwd, int *nlev_bwd);

/* Device time of the last call of each stage, milliseconds (HIP events on the
 * context's stream): 0 assembly, 1 solve, 2 SpMV (last fc_sys_spmv_bench),
 * 3 sweep pair (last fc_sys_sweep_bench). */
double fc_sys_last_ms(fc_system *sys, int stage);
/* Time `reps` FP64 (precision 64) or FP32 (32) SpMVs y = A x on the resident
 * matrix; returns the average device time per SpMV in milliseconds, or < 0. */
double fc_sys_spmv_bench(fc_system *sys, int precision, int reps);
/* Time `reps` applications of the FP64 IC(0) (dilu 0, src/iccg.f90:94-111) or
 * ILU(0)/DILU (dilu 1, src/bicgstab.f90:107-124) preconditioner of the resident
 * matrix -- a forward and a backward sweep in the context's sweep schedule, each
 * after the fill of its output with the "not ready" mark -- after one untimed
 * application; returns the average device time per pair in milliseconds, or < 0.
 * No counterpart in the reference. */
double fc_sys_sweep_bench(fc_system *sys, int dilu, int reps);

/* --------------------------------------------- learned preconditioner */

/* A message-passing network over the cell graph (hidden width 32, 4 layers,
 * BF16 weights and activations, MFMA node GEMMs with FP32 accumulation) used
 * as the preconditioner of flexible CG (FC_METHOD_FCG_LEARNED).  No
 * counterpart in the reference.  Parameters are a flat FP32 vector of
 * fc_gnn_num_params() entries (layout in oracle/gnn_oracle.py and
 * csrc/fcap_gnn.h). */
typedef struct fc_gnn fc_gnn;
int fc_gnn_num_params(void);
fc_status fc_gnn_create(fc_context *ctx, fc_gnn **net, const float *params /* NULL: zeros */);
fc_status fc_gnn_destroy(fc_gnn *net);
fc_status fc_gnn_set_params(fc_gnn *net, const float *params);
fc_status fc_gnn_get_params(fc_gnn *net, float *params);
/* z = P(r) on a system's graph (host arrays of numCells floats). */
fc_status fc_gnn_apply(fc_gnn *net, fc_system *sys, const float *r, float *z);
/* Average milliseconds of one network application on the system (device-resident
 * input), `reps` applications timed with events on the library's stream; -1 on error. */
double fc_gnn_bench(fc_gnn *net, fc_system *sys, int reps);
/* Average milliseconds of one launch of a hidden message-passing layer (layer 2 of 4,
 * rows in HBM) on the system, `reps` launches timed with events on the library's
 * stream; -1 on error.  No counterpart in the reference (performance measurement). */
double fc_gnn_bench_layer(fc_gnn *net, fc_system *sys, int reps);
/* Edge features of the system's CSR entries (nnz each; 0 on the diagonal) and its FP32 diagonal. */
fc_status fc_gnn_edge_features(fc_gnn *net, fc_system *sys, float *e1, float *e2, float *diag);
/* The cluster layout of the system's network graph (fc_set_network_layout),
 * built on first use: the number of clusters (0: the graph has none -- ELL width
 * other than 4, 6 or 8, or a halo above 384), the largest halo found and the
 * halo rows of all clusters together. */
fc_status fc_sys_network_layout(fc_system *sys, int *clusters, int *halo_max, long long *halo_total);
/* The coarse space FC_METHOD_FCG_LEARNED uses on the system (fc_set_fcg_coarse),
 * built for its current matrix on first use: the number of aggregates, cells per
 * aggregate and whether A_c was factored (all 0 when the correction is off);
 * optionally each cell's aggregate (numCells ints, 0-based) and A_c = P^T A P
 * (aggregates^2 doubles, column-major).  No counterpart in the reference. */
fc_status fc_sys_coarse_space(fc_system *sys, int *aggregates, int *cells_per_aggregate, int *factored,
                              int *cell_aggregate, double *matrix);
/* One application of FC_METHOD_FCG_LEARNED's preconditioner to r (numCells
 * floats): the attached network, the Jacobi steps and the coarse correction,
 * as the solve applies them.  No counterpart in the reference. */
fc_status fc_sys_precondition(fc_system *sys, const float *r, float *z);
/* Loss (z-e)^T A (z-e) / (e^T A e) of one pair and its gradient (fc_gnn_num_params floats). */
fc_status fc_gnn_loss_grad(fc_gnn *net, fc_system *sys, const float *r, const float *e, double *loss, float *grad);
/* Add (residual, error) pairs of IC(0)-PCG iterates on the homogeneous system
 * A x = 0 of sys's matrix, after iters[0..nk-1] iterations from seeded
 * standard-normal starts (e = -x, r = -A x).  The system must outlive the
 * samples (or be passed to fc_gnn_drop_system first). */
fc_status fc_gnn_add_samples(fc_gnn *net, fc_system *sys, int nk, const int *iters, unsigned seed);
int fc_gnn_num_samples(fc_gnn *net);
/* Copy sample `index` (residual and error, n floats each; n = its system's numCells). */
fc_status fc_gnn_get_sample(fc_gnn *net, int index, float *r, float *e, int n);
fc_status fc_gnn_clear_samples(fc_gnn *net);
/* Drop the samples that point at `sys` (call before fc_sys_destroy of a sample source). */
fc_status fc_gnn_drop_system(fc_gnn *net, fc_system *sys);
/* Adam over `steps` mini-batches of `batch` samples; loss_out (steps floats, may be NULL). */
fc_status fc_gnn_train(fc_gnn *net, int steps, int batch, float lr, unsigned seed, float *loss_out);
/* Training with the preconditioner's damped-Jacobi steps in the loss (fc_gnn_train,
 * fc_gnn_loss_grad): the loss is taken on z after `steps` steps
 * z <- z + omega (r - A z) / a_ii, and the gradient passes back through them
 * (0..8 steps, 0 < omega < 2; default 0: the network's own output).  Set it to
 * the solver's fc_set_fcg_smoothing to train what the solver applies. */
fc_status fc_gnn_set_training_smoothing(fc_gnn *net, int steps, double omega);
/* Training with the coarse correction after the Jacobi steps in the loss (on = 1;
 * fc_gnn_train, fc_gnn_loss_grad): z <- z + P A_c^-1 P^T (r - A z) with each sample
 * system's coarse space (fc_set_fcg_coarse of its context), the gradient through
 * its adjoint dz - A P A_c^-1 P^T dz.  Default 0. */
fc_status fc_gnn_set_training_coarse(fc_gnn *net, int on);
/* Attach a network as the preconditioner of FC_METHOD_FCG_LEARNED solves of sys. */
fc_status fc_sys_set_network(fc_system *sys, fc_gnn *net);

/* ---------------------------------------------------------- polyMesh reader */

/* Host-side reader of the polyMesh files the reference reads in
 * src/mesh_geometry_and_topology.f90:409-605 (points, faces, owner,
 * neighbour, boundary), OpenFOAM ASCII (FoamFile headers, 0-based labels,
 * nCells from the owner note as at :516-527) or the reference's native
 * layout (:473-480), detected from the owner file.  Lists are parsed by
 * `nthreads` threads (0: all cores, at most 16); faces may have any vertex
 * count.  The boundary file is the reference's simplified one: '#' comment,
 * then `bctype nFaces startFace` lines (bctypes of :417-468).  On error the
 * message goes to err (errlen bytes). */
typedef struct fc_polymesh fc_polymesh;
fc_status fc_polymesh_read(const char *dir, int nthreads, fc_polymesh **mesh, char *err, int errlen);
fc_status fc_polymesh_sizes(const fc_polymesh *mesh, int *numNodes, int *numFaces, int *numInnerFaces,
                            int *numCells, int *numFaceNodes, int *numPatches);
/* Caller-owned outputs: xyz (3*numNodes, x y z per node), face_ptr
 * (numFaces+1, 0-based offsets), face_nodes (numFaceNodes), owner (numFaces),
 * neighbour (numInnerFaces); labels 1-based, as fc_sys_create takes them. */
fc_status fc_polymesh_arrays(const fc_polymesh *mesh, double *xyz, int *face_ptr, int *face_nodes, int *owner,
                             int *neighbour);
fc_status fc_polymesh_patch(const fc_polymesh *mesh, int index, char *bctype, int len, int *nFaces, int *startFace);
void fc_polymesh_free(fc_polymesh *mesh);

/* ------------------------------------------------------------- restart */

/* Restart records of src/write_restart_files.f90:22-47 (read back by
 * src/readfiles.f90:19-43): one unformatted sequential record per Fortran
 * write, each framed by its 4-byte byte count.  Records: (itime, time),
 * gradpcmf when has_gradpcmf (constant mass flow), then nrec float64 arrays of
 * len[k] values -- the reference's order is flmass (numInnerFaces); u, v, w, p,
 * te, ed, t, vis (numTotal); uu, vv, ww, uv, uw, vw (numCells); uo, vo, wo,
 * teo, edo (numTotal).  Host only; errors go to err (errlen bytes). */
fc_status fc_restart_write(const char *path, int itime, double time, int has_gradpcmf, double gradpcmf, int nrec,
                           const long long *len, const double *const *rec, char *err, int errlen);
/* itime, time, gradpcmf and the number and lengths of the field records
 * (len has room for max_rec; *nrec is the count in the file). */
fc_status fc_restart_probe(const char *path, int has_gradpcmf, int *itime, double *time, double *gradpcmf,
                           int max_rec, int *nrec, long long *len, char *err, int errlen);
/* The first nrec field records into rec[k] (len[k] values each; a null rec[k]
 * skips that record); a length that differs from the file is an error. */
fc_status fc_restart_read(const char *path, int has_gradpcmf, int nrec, const long long *len, double *const *rec,
                          char *err, int errlen);

/* No counterpart in the reference: a plain device-to-device copy kernel of
 * `bytes` bytes, `reps` times (calibration of the HBM traffic counters and the
 * achievable-bandwidth figure); returns the average milliseconds per copy. */
double fc_copy_bench(fc_context *ctx, size_t bytes, int reps);

#ifdef __cplusplus
}
#endif
#endif /* FCAP_H */
