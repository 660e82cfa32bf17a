! stafft.f90 -- drop-in for lib/stafft/stafft.f90's 1D real FFT
! (initfft/forfft/revfft, stafft.f90:63-406) on the GPU through
! libhydra_amd.  Same names, arguments and layouts: x(0:m*n-1) holds m
! vectors of length n, the vector index fastest; transforms are in place.
! Lengths are powers of two (the radix-4/2 part of stafft); other lengths
! stop the program, as the reference stops on a failed factorisation.
module stafft

use iso_c_binding
use hydra_amd_c
use hydra_amd_ctx

implicit none

contains

subroutine initfft(n, factors, trig)
  integer, intent(in) :: n
  integer, intent(out) :: factors(5)
  double precision, intent(out) :: trig(2*n)
  integer(c_int32_t) :: f(5)
  call hyd_check(hyd_initfft(int(n, c_int32_t), f, trig), 'initfft')
  factors = f
end subroutine

subroutine forfft(m, n, x, trig, factors)
  integer, intent(in) :: m, n, factors(5)
  double precision, intent(inout) :: x(0:m*n-1)
  double precision, intent(in) :: trig(2*n)
  integer(c_int32_t) :: f(5)
  call hyd_open
  f = factors
  call hyd_check(hyd_forfft(hctx, int(m, c_int32_t), int(n, c_int32_t), x, trig, f), 'forfft')
end subroutine

subroutine revfft(m, n, x, trig, factors)
  integer, intent(in) :: m, n, factors(5)
  double precision, intent(inout) :: x(0:m*n-1)
  double precision, intent(in) :: trig(2*n)
  integer(c_int32_t) :: f(5)
  call hyd_open
  f = factors
  call hyd_check(hyd_revfft(hctx, int(m, c_int32_t), int(n, c_int32_t), x, trig, f), 'revfft')
end subroutine

end module stafft
