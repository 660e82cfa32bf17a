! contours.f90 -- drop-in for ca/plane/qg/caps/contours.f90 (module
! contours) on the GPU through libhydra_amd: the same public arrays,
! parameters and routines (contours.f90:12-42, init_contours 48-132, velint
! 136-169, con2grid 173-346, surgery 350-683, renode 687-933).  The
! contour store stays in these module arrays (host memory), as the
! reference's callers (congen.f90, common.f90, evolution.f90) expect; each
! routine copies what it needs to the device and back.
module contours

use iso_c_binding
use hydra_amd_c
use hydra_amd_ctx
use constants
use variables
use generic

implicit none

 !Contour store (contours.f90:12-15):
double precision :: xq(npm), yq(npm), qjump
integer :: indq(nm), npq(nm), i1q(nm), i2q(nm)
integer :: nextq(npm)
integer :: nptq, nq

 !Fine grid of con2grid, ultra-fine grid of congen (contours.f90:17-25):
double precision :: xgf(nxf), ygf(nyf)
double precision, parameter :: glxf = ellx/dble(nxf), glxfi = dble(nxf)/ellx
double precision, parameter :: glyf = elly/dble(nyf), glyfi = dble(nyf)/elly
double precision :: xgu(nxu), ygu(nyu)
double precision, parameter :: glxu = ellx/dble(nxu), glxui = dble(nxu)/ellx
double precision, parameter :: glyu = elly/dble(nyu), glyui = dble(nyu)/elly

 !Interpolation neighbours, congen weights and indices (contours.f90:27-35):
integer :: ixp(nx), iyp(ny)
double precision :: w00(mgu,mgu), w10(mgu,mgu), w01(mgu,mgu), w11(mgu,mgu)
integer :: ixfw(nxu), ix0w(nxu), ix1w(nxu)
integer :: iyfw(nyu), iy0w(nyu), iy1w(nyu)
integer :: ibx(nxu+1), iby(nyu+1)

 !Surgery and node-redistribution scales (contours.f90:37-41):
double precision, parameter :: amu = 0.2d0, ell = 6.25d0*glx
double precision, parameter :: dm = amu**2*ell/four, dm2 = dm**2, d4small = small*glx**4
double precision, parameter :: dmsq = four*dm2, dmi = two/dm
double precision, parameter :: elf = one/ell**2, densf = one/(amu*sqrt(ell))

contains

subroutine init_contours
  double precision :: w(4*mgu*mgu), t(3*max(nxu, nyu)+1)
  call hyd_open
  call hyd_check(hyd_get_table(hctx, 18_c_int32_t, xgf), 'xgf')
  call hyd_check(hyd_get_table(hctx, 19_c_int32_t, ygf), 'ygf')
  call hyd_check(hyd_get_table(hctx, 20_c_int32_t, xgu), 'xgu')
  call hyd_check(hyd_get_table(hctx, 21_c_int32_t, ygu), 'ygu')
  call hyd_check(hyd_get_table(hctx, 22_c_int32_t, w), 'weights')
  w00 = reshape(w(1:mgu*mgu), [mgu, mgu])
  w10 = reshape(w(mgu*mgu+1:2*mgu*mgu), [mgu, mgu])
  w01 = reshape(w(2*mgu*mgu+1:3*mgu*mgu), [mgu, mgu])
  w11 = reshape(w(3*mgu*mgu+1:4*mgu*mgu), [mgu, mgu])
  call hyd_check(hyd_get_table(hctx, 23_c_int32_t, t), 'ixp')
  ixp = nint(t(1:nx))
  call hyd_check(hyd_get_table(hctx, 24_c_int32_t, t), 'iyp')
  iyp = nint(t(1:ny))
  call hyd_check(hyd_get_table(hctx, 25_c_int32_t, t), 'ixw')
  ixfw = nint(t(1:nxu))
  ix0w = nint(t(nxu+1:2*nxu))
  ix1w = nint(t(2*nxu+1:3*nxu))
  call hyd_check(hyd_get_table(hctx, 26_c_int32_t, t), 'iyw')
  iyfw = nint(t(1:nyu))
  iy0w = nint(t(nyu+1:2*nyu))
  iy1w = nint(t(2*nyu+1:3*nyu))
  call hyd_check(hyd_get_table(hctx, 27_c_int32_t, t), 'ibx')
  ibx = nint(t(1:nxu+1))
  call hyd_check(hyd_get_table(hctx, 28_c_int32_t, t), 'iby')
  iby = nint(t(1:nyu+1))
end subroutine

subroutine velint(uu, vv, uq, vq)
  double precision, intent(in) :: uu(ny,nx), vv(ny,nx)
  double precision, intent(out) :: uq(nptq), vq(nptq)
  call hyd_check(hyd_velint(hctx, uu, vv, xq, yq, int(nptq, c_int32_t), uq, vq), 'velint')
end subroutine

subroutine con2grid(qc)
  double precision, intent(out) :: qc(ny,nx)
  call hyd_check(hyd_con2grid(hctx, xq, yq, nextq, int(nptq, c_int32_t), qjump, qc, &
                              c_null_ptr), 'con2grid')
end subroutine

subroutine surgery
  integer(c_int32_t) :: n, m
  call hyd_check(hyd_set_contours(hctx, int(nptq, c_int32_t), int(nq, c_int32_t), xq, yq, &
                                  nextq, i1q, i2q, npq, indq, qjump), 'surgery (store)')
  call hyd_check(hyd_surgery(hctx), 'surgery')
  call hyd_check(hyd_get_counts(hctx, n, m), 'surgery (counts)')
  if (n > npm .or. m > nm) then
    write(*, '(a)') ' hydra_amd: surgery: the new contours exceed npm/nm'
    stop 1
  endif
  nptq = n
  nq = m
  call hyd_check(hyd_get_contours(hctx, xq, yq, nextq, i1q, i2q, npq, indq), 'surgery (load)')
end subroutine

subroutine renode(xd, yd, npd, xr, yr, npr)
  integer, intent(in) :: npd
  double precision, intent(in) :: xd(npd), yd(npd)
  double precision, intent(out) :: xr(nprm), yr(nprm)
  integer, intent(out) :: npr
  integer(c_int32_t) :: np
  call hyd_check(hyd_renode(hctx, xd, yd, int(npd, c_int32_t), xr, yr, int(nprm, c_int32_t), &
                            np), 'renode')
  npr = np
end subroutine

end module contours
