! Test harness around the reference's facefluxmass (Rhie-Chow face mass flux
! and pressure-correction coefficient) -- TEST INFRASTRUCTURE ONLY.  It is
! linked against the reference sources compiled in place from
! /root/reference/src by oracle/Makefile and writes golden vectors for
! tests/golden/.  Nothing here is shipped or called by the product.
!
! Reads, from the current directory:
!   polyMesh/{points,faces,owner,neighbour,boundary}   (mesh_geometry)
!   flux_in.bin (stream, float64): den, apu, apv, apw, u, v, w, p (numCells
!               each), then dUdxi, dVdxi, dWdxi, dPdxi (3 x numCells each,
!               component fastest)
! Writes flux_out.bin (stream):
!   int32   numCells, numInnerFaces, nnz
!   float64 cap(nif), can(nif), flmass(nif), a(nnz), su(numCells), fmcor(nif)
! where a and su come from the face loop of the pressure-correction assembly
! (a = 0 and su = 0, then per inner face in order: a(ij) = can, a(ji) = cap,
! a(diag(P)) -= can, a(diag(N)) -= cap, su(P) -= F, su(N) += F), and fmcor of
! the reference's fluxmc (non-orthogonal mass flux correction) for every face.
! Writes piso_out.bin (stream): float64 cap(nif), flmass(nif), a(nnz), su(numCells)
! of the same face loop around facefluxmass_piso (PISO's assembly,
! src/PISO_multiple_correction.f90:107-148) followed by its reference pressure
! (:188-192) at pRefCell = numCells/2 + 1.
program ref_flux_driver
  use types
  use parameters
  use geometry
  use sparse_matrix
  use variables
  use faceflux_mass
  implicit none
  integer :: un, i, k, ijp, ijn, pref
  real(dp) :: capf, canf
  real(dp), allocatable :: capv(:), canv(:), fmc(:)

  call mesh_geometry
  call create_CSR_matrix_from_mesh_data

  allocate(den(numTotal), u(numTotal), v(numTotal), w(numTotal), p(numTotal))
  allocate(dUdxi(3,numCells), dVdxi(3,numCells), dWdxi(3,numCells), dPdxi(3,numCells))
  allocate(flmass(numInnerFaces), capv(numInnerFaces), canv(numInnerFaces), fmc(numInnerFaces))
  den = 0.0_dp
  u = 0.0_dp
  v = 0.0_dp
  w = 0.0_dp
  p = 0.0_dp

  open(newunit=un, file='flux_in.bin', access='stream', form='unformatted', status='old')
  read(un) den(1:numCells)
  read(un) apu
  read(un) apv
  read(un) apw
  read(un) u(1:numCells)
  read(un) v(1:numCells)
  read(un) w(1:numCells)
  read(un) p(1:numCells)
  read(un) dUdxi
  read(un) dVdxi
  read(un) dWdxi
  read(un) dPdxi
  close(un)

  a = 0.0_dp
  su = 0.0_dp
  do i = 1, numInnerFaces
    ijp = owner(i)
    ijn = neighbour(i)
    call facefluxmass(ijp, ijn, xf(i), yf(i), zf(i), arx(i), ary(i), arz(i), facint(i), capf, canf, flmass(i))
    capv(i) = capf
    canv(i) = canf
    k = icell_jcell_csr_value_index(i)
    a(k) = canf
    k = jcell_icell_csr_value_index(i)
    a(k) = capf
    k = diag(ijp)
    a(k) = a(k) - canf
    k = diag(ijn)
    a(k) = a(k) - capf
    su(ijp) = su(ijp) - flmass(i)
    su(ijn) = su(ijn) + flmass(i)
  end do

  do i = 1, numInnerFaces
    call fluxmc(owner(i), neighbour(i), xf(i), yf(i), zf(i), arx(i), ary(i), arz(i), facint(i), fmc(i))
  end do

  open(newunit=un, file='flux_out.bin', access='stream', form='unformatted', status='replace')
  write(un) numCells, numInnerFaces, nnz
  write(un) capv, canv, flmass, a, su, fmc
  close(un)

  a = 0.0_dp
  su = 0.0_dp
  do i = 1, numInnerFaces
    ijp = owner(i)
    ijn = neighbour(i)
    call facefluxmass_piso(ijp, ijn, xf(i), yf(i), zf(i), arx(i), ary(i), arz(i), facint(i), capf, canf, flmass(i))
    capv(i) = capf
    k = icell_jcell_csr_value_index(i)
    a(k) = canf
    k = jcell_icell_csr_value_index(i)
    a(k) = capf
    k = diag(ijp)
    a(k) = a(k) - canf
    k = diag(ijn)
    a(k) = a(k) - capf
    su(ijp) = su(ijp) - flmass(i)
    su(ijn) = su(ijn) + flmass(i)
  end do
  pref = numCells / 2 + 1
  a(ioffset(pref):ioffset(pref+1)-1) = 0.0_dp
  a(diag(pref)) = 1.0_dp
  su(pref) = p(pref)
  open(newunit=un, file='piso_out.bin', access='stream', form='unformatted', status='replace')
  write(un) capv, flmass, a, su
  close(un)
end program ref_flux_driver
