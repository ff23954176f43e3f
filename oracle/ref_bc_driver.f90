! Test harness around the reference's bpres (src/bpres.f90) and
! correctBoundaryConditionsVelocity (src/correctBoundaryConditionsVelocity.f90),
! both compiled on their own -- TEST INFRASTRUCTURE ONLY.  It is linked against
! the reference sources compiled in place from /root/reference/src by
! oracle/Makefile and writes golden vectors for tests/golden/.  Nothing here is
! shipped or called by the product.
!
! Reads, from the current directory:
!   polyMesh/{points,faces,owner,neighbour,boundary}   (mesh_geometry)
!   bc_in.bin (stream, float64): flomas, p(numTotal), dPdxi (3 x numCells,
!             component fastest), den(numTotal), u, v, w (numTotal each)
! Writes bc_out.bin (stream):
!   int32   numCells, numTotal, nout
!   float64 p1(numTotal)  -- p after bpres(p, 1)
!           p2(numTotal)  -- p1 after bpres(p1, 2) with dPdxi
!           u, v, w (numTotal each), fmo(nout) after correctBoundaryConditionsVelocity
program ref_bc_driver
  use types
  use parameters
  use geometry
  use variables
  implicit none
  integer :: un
  real(dp), allocatable :: pbc(:)
  external :: bpres, correctBoundaryConditionsVelocity

  call mesh_geometry
  allocate(pbc(numTotal), dPdxi(3, numCells))
  allocate(den(numTotal), u(numTotal), v(numTotal), w(numTotal))
  allocate(fmo(max(nout, 1)))
  open(newunit=un, file='bc_in.bin', access='stream', form='unformatted', status='old')
  read(un) flomas
  read(un) pbc
  read(un) dPdxi
  read(un) den
  read(un) u
  read(un) v
  read(un) w
  close(un)
  open(newunit=un, file='bc_out.bin', access='stream', form='unformatted', status='replace')
  write(un) numCells, numTotal, nout
  call bpres(pbc, 1)
  write(un) pbc
  call bpres(pbc, 2)
  write(un) pbc
  fmo = 0.0_dp
  call correctBoundaryConditionsVelocity
  write(un) u, v, w, fmo(1:nout)
  close(un)
end program ref_bc_driver
