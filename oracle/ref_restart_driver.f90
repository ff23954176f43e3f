! Test harness around the reference's write_restart_files -- TEST
! INFRASTRUCTURE ONLY.  It is linked against the reference sources compiled in
! place from /root/reference/src by oracle/Makefile and writes golden restart
! files for tests/golden/.  Nothing here is shipped or called by the product.
!
! Reads, from the current directory:
!   restart_in.txt : numCells numInnerFaces numTotal const_mflux(0/1) itime
!   restart_in.bin (stream, float64): time, gradpcmf, then flmass (numInnerFaces);
!                  u v w p te ed t vis (numTotal); uu vv ww uv uw vw (numCells);
!                  uo vo wo teo edo (numTotal)
! and calls write_restart_files, which writes the file 'restart_ref'.
program ref_restart_driver
  use types
  use parameters
  use geometry
  use variables
  use title_mod
  use statistics
  implicit none
  integer :: un, nc, nif, ntot, cm

  open(newunit=un, file='restart_in.txt', status='old')
  read(un,*) nc, nif, ntot, cm, itime
  close(un)
  const_mflux = (cm == 1)
  ltransient = .false.
  restart_file = 'restart_ref'
  allocate(flmass(nif), u(ntot), v(ntot), w(ntot), p(ntot), te(ntot), ed(ntot), t(ntot), vis(ntot))
  allocate(uu(nc), vv(nc), ww(nc), uv(nc), uw(nc), vw(nc))
  allocate(uo(ntot), vo(ntot), wo(ntot), teo(ntot), edo(ntot))
  open(newunit=un, file='restart_in.bin', access='stream', form='unformatted', status='old')
  read(un) time, gradpcmf
  read(un) flmass, u, v, w, p, te, ed, t, vis, uu, vv, ww, uv, uw, vw, uo, vo, wo, teo, edo
  close(un)
  call write_restart_files
end program ref_restart_driver
