! Test harness around the reference's Gauss gradient (src/grad_gauss.f90,
! compiled on its own) -- TEST INFRASTRUCTURE ONLY.  It is linked against the
! reference sources compiled in place from /root/reference/src by
! oracle/Makefile and writes golden vectors for tests/golden/.  Nothing here is
! shipped or called by the product.
!
! Reads, from the current directory:
!   polyMesh/{points,faces,owner,neighbour,boundary}   (mesh_geometry)
!   grad.txt   : nigrad
!   grad_in.bin: u(numTotal) [stream, float64]
! Writes grad_out.bin (stream): int32 numCells, numTotal; float64 dudx, dudy, dudz (numCells each)
program ref_grad_driver
  use types
  use parameters
  use geometry
  implicit none
  integer :: un
  real(dp), allocatable :: u(:), dudx(:), dudy(:), dudz(:)
  external :: grad_gauss

  open(newunit=un, file='grad.txt', status='old')
  read(un,*) nigrad
  close(un)
  call mesh_geometry
  allocate(u(numTotal), dudx(numCells), dudy(numCells), dudz(numCells))
  open(newunit=un, file='grad_in.bin', access='stream', form='unformatted', status='old')
  read(un) u
  close(un)
  dudx = 0.0_dp
  dudy = 0.0_dp
  dudz = 0.0_dp
  call grad_gauss(u, dudx, dudy, dudz)
  open(newunit=un, file='grad_out.bin', access='stream', form='unformatted', status='replace')
  write(un) numCells, numTotal
  write(un) dudx, dudy, dudz
  close(un)
end program ref_grad_driver
