! A Fortran caller of DSTEDC_DC_BR written the way the reference's callers
! call it (scripts/cpu/smoke_br.f90:24-27: LP64, every argument by
! reference): linked against libbrdc.so with
! -Wl,--defsym,dstedc_dc_br_=brdc_dstebr_, so no source line names this
! library.  It runs the workspace queries and the argument errors, which need
! no GPU, and prints one result per line.
program fortran_caller
  implicit none
  integer :: n, info, lw, liw
  double precision :: d(100), e(100), wq(1), work(2000)
  integer :: iwq(1), iwork(1000)
  external :: dstedc_dc_br

  d = 1.0d0
  e = 0.5d0
  n = 4096
  call dstedc_dc_br(n, d, e, wq, -1, iwq, -1, info)
  print '(A,I0,1X,I0,1X,I0)', 'query4096 ', info, int(wq(1)), iwq(1)
  n = 25
  call dstedc_dc_br(n, d, e, wq, -1, iwq, 7, info)
  print '(A,I0,1X,I0,1X,I0)', 'query25 ', info, int(wq(1)), iwq(1)
  n = -1
  call dstedc_dc_br(n, d, e, work, 2000, iwork, 1000, info)
  print '(A,I0)', 'badn ', info
  n = 100
  lw = 10
  liw = 1000
  call dstedc_dc_br(n, d, e, work, lw, iwork, liw, info)
  print '(A,I0)', 'smalllwork ', info
  lw = 2000
  liw = 10
  call dstedc_dc_br(n, d, e, work, lw, iwork, liw, info)
  print '(A,I0)', 'smallliwork ', info
  n = 1
  d(1) = 3.5d0
  call dstedc_dc_br(n, d, e, work, 1, iwork, 1, info)
  print '(A,I0,1X,F6.3)', 'n1 ', info, d(1)
end program fortran_caller
