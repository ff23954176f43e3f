! Test harness around the reference's own hot-path subroutines -- TEST
! INFRASTRUCTURE ONLY.  It is linked against the reference sources compiled in
! place from /root/reference/src by oracle/Makefile and writes golden vectors
! for tests/golden/.  Nothing here is shipped or called by the product.
!
! Reads, from the current directory:
!   polyMesh/{points,faces,owner,neighbour,boundary}   (mesh_geometry)
!   drv.txt   : solver (0 none, 1 iccg, 2 dpcg, 3 bicgstab), sor, nsw, amode
!   drv_in.bin: mu(numCells), phi(numTotal), su0(numCells), fi0(numCells)  [stream, float64]
!               and, when amode = 1, a(nnz) and su(numCells): the system to
!               solve instead of the assembled Laplacian
! Writes drv_out.bin (stream):
!   int32  numCells, numFaces, numInnerFaces, nnz
!   float64 arx,ary,arz,xf,yf,zf (numFaces), vol,xc,yc,zc (numCells), facint (numInnerFaces)
!   int32  ioffset(numCells+1), ja(nnz), diag(numCells), ijslot(nif), jislot(nif)
!   float64 a(nnz), su(numCells)            (after laplacian)
!   float64 fi(numCells), res(numCells), resor(ip)   (after the solve)
! The iteration count is taken from the solver's report line on stdout, and the
! solver call's wall time from the 'solve_seconds' line.
program ref_driver
  use types
  use parameters
  use geometry
  use sparse_matrix
  implicit none
  integer :: solver, nswv, u, uinp, i, amode
  integer(8) :: c0, c1, crate
  real(dp) :: sorv
  real(dp), allocatable :: mu(:), phi(:), fi(:), su0(:)

  open(newunit=u, file='drv.txt', status='old')
  read(u,*) solver, sorv, nswv, amode
  close(u)

  call mesh_geometry
  call create_CSR_matrix_from_mesh_data

  allocate(mu(numCells), phi(numTotal), fi(numTotal), su0(numCells))
  open(newunit=uinp, file='drv_in.bin', access='stream', form='unformatted', status='old')
  read(uinp) mu
  read(uinp) phi
  read(uinp) su0
  read(uinp) fi(1:numCells)
  do i = numCells+1, numTotal
    fi(i) = phi(i)
  end do

  su = su0
  call laplacian(mu, phi)

  open(newunit=u, file='drv_out.bin', access='stream', form='unformatted', status='replace')
  write(u) numCells, numFaces, numInnerFaces, nnz
  write(u) arx, ary, arz, xf, yf, zf
  write(u) vol, xc, yc, zc, facint
  write(u) ioffset, ja, diag, icell_jcell_csr_value_index, jcell_icell_csr_value_index
  write(u) a, su

  if (amode == 1) then
    read(uinp) a
    read(uinp) su
  end if
  close(uinp)

  ltest = .true.
  sor(ip) = sorv
  nsw(ip) = nswv
  resor(ip) = 0.0_dp
  call system_clock(c0, crate)
  select case (solver)
  case (1)
    call iccg(fi, ip)
  case (2)
    call dpcg(fi, ip)
  case (3)
    call bicgstab(fi, ip)
  end select
  call system_clock(c1)
  ! wall time of the solver call alone (factorisation included), for the CPU calibration
  write(*,'(a,es16.8)') ' solve_seconds ', real(c1 - c0, dp) / real(crate, dp)
  write(u) fi(1:numCells), res, resor(ip)
  close(u)
end program ref_driver
