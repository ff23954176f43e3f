! Test harness around the reference's adjustMassFlow (src/adjustMassFlow.f90,
! compiled on its own) -- TEST INFRASTRUCTURE ONLY.  It is linked against the
! reference sources compiled in place from /root/reference/src by
! oracle/Makefile and writes golden vectors for tests/golden/.  Nothing here is
! shipped or called by the product.
!
! Reads, from the current directory:
!   polyMesh/{points,faces,owner,neighbour,boundary}   (mesh_geometry)
!   adjust_in.bin (stream, float64): flomas, den(numTotal), u, v, w (numTotal
!                 each), su(numCells), fmi(ninl)
! Writes adjust_out.bin (stream):
!   int32   numCells, numTotal, ninl, nout
!   float64 su(numCells), u, v, w (numTotal each), fmo(nout)
program ref_adjust_driver
  use types
  use parameters
  use geometry
  use sparse_matrix
  use variables
  implicit none
  integer :: un
  external :: adjustMassFlow

  call mesh_geometry
  allocate(den(numTotal), u(numTotal), v(numTotal), w(numTotal), su(numCells))
  allocate(fmi(max(ninl, 1)), fmo(max(nout, 1)))
  open(newunit=un, file='adjust_in.bin', access='stream', form='unformatted', status='old')
  read(un) flomas
  read(un) den
  read(un) u
  read(un) v
  read(un) w
  read(un) su
  if (ninl > 0) read(un) fmi(1:ninl)
  close(un)
  fmo = 0.0_dp
  call adjustMassFlow
  open(newunit=un, file='adjust_out.bin', access='stream', form='unformatted', status='replace')
  write(un) numCells, numTotal, ninl, nout
  write(un) su, u, v, w, fmo(1:nout)
  close(un)
end program ref_adjust_driver
