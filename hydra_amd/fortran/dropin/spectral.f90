! spectral.f90 -- drop-in for ca/plane/qg/caps/spectral.f90 (module
! spectral) on the GPU through libhydra_amd: the same public arrays and
! routines (spectral.f90:6-32, init_spectral 42-158, main_invert 304-342,
! gradient 345-366, spec1d 370-424).  init_spectral fills the arrays from
! the library's tables, which are bit-identical to the reference's; the
! inversion, gradient and spectrum run on the device.  Forcing (init_forcing,
! nb_forcing, myrand) is outside this path and stops the run if called.
module spectral

use iso_c_binding
use hydra_amd_c
use hydra_amd_ctx
use constants
use sta2dfft

implicit none

integer, parameter :: nwx = nx/2, nwy = ny/2
integer, parameter :: nx2 = nx+2, ny2 = ny+2
integer, parameter :: nkfmax = nwx*nwy

double precision :: rksq(nx,ny), qdiss(nx,ny)
double precision :: green(nx,ny), flo(nx,ny), fhi(nx,ny)
double precision :: rk(nx,ny)
double precision :: rkx(nx), hrkx(nx)
double precision :: rky(ny), hrky(ny)
double precision :: bety(ny)

double precision :: xtrig(2*nx), ytrig(2*ny)
integer :: xfactors(5), yfactors(5)

double precision :: dqdt(nx,ny)
double precision :: kxf(nkfmax), kyf(nkfmax), forfac
double precision :: kxfxo(nwx), kyfyo(nwy)
integer :: nkf, nkfxo, nkfyo

double precision :: spmf(0:max(nx,ny)), alk(max(nx,ny))
integer :: kmag(nx,ny), kmax, kmaxred

contains

subroutine init_spectral
  double precision :: wk(nx*ny), sh(3), sp(0:max(nx,ny))
  call hyd_open
  call initfft(nx, xfactors, xtrig)
  call initfft(ny, yfactors, ytrig)
  call table(1, green)
  call table(2, flo)
  call table(3, fhi)
  call table(4, qdiss)
  call table(5, rksq)
  call table(16, rk)
  call hyd_check(hyd_get_table(hctx, 6_c_int32_t, hrkx), 'hrkx')
  call hyd_check(hyd_get_table(hctx, 7_c_int32_t, rkx), 'rkx')
  call hyd_check(hyd_get_table(hctx, 8_c_int32_t, hrky), 'hrky')
  call hyd_check(hyd_get_table(hctx, 9_c_int32_t, rky), 'rky')
  call hyd_check(hyd_get_table(hctx, 15_c_int32_t, bety), 'bety')
  call hyd_check(hyd_get_table(hctx, 14_c_int32_t, sh), 'shells')
  kmax = nint(sh(1))
  kmaxred = nint(sh(2))
  sp = 0.d0
  call hyd_check(hyd_get_table(hctx, 12_c_int32_t, sp), 'spmf')
  spmf = sp
  sp = 0.d0
  call hyd_check(hyd_get_table(hctx, 13_c_int32_t, sp), 'alk')
  alk = sp(1:max(nx,ny))
  call hyd_check(hyd_get_table(hctx, 17_c_int32_t, wk), 'kmag')
  kmag = reshape(nint(wk), [nx, ny])
  dqdt = 0.d0
end subroutine

subroutine table(id, a)
  integer, intent(in) :: id
  double precision, intent(out) :: a(nx,ny)
  call hyd_check(hyd_get_table(hctx, int(id, c_int32_t), a), 'table')
end subroutine

subroutine init_forcing
  write(*, '(a)') ' hydra_amd: spectral forcing is not part of this build'
  stop 1
end subroutine

subroutine nb_forcing(dt)
  double precision :: dt
  write(*, '(a)') ' hydra_amd: spectral forcing is not part of this build'
  stop 1
end subroutine

subroutine myrand(uni)
  double precision :: uni
  write(*, '(a)') ' hydra_amd: spectral forcing is not part of this build'
  stop 1
end subroutine

subroutine main_invert(qq, uu, vv, pp)
  double precision, intent(in) :: qq(nx,ny)
  double precision, intent(out) :: uu(ny,nx), vv(ny,nx), pp(nx,ny)
  call hyd_check(hyd_main_invert(hctx, qq, uu, vv, pp), 'main_invert')
end subroutine

subroutine gradient(ff, ffx, ffy)
  double precision, intent(in) :: ff(nx,ny)
  double precision, intent(out) :: ffx(ny,nx), ffy(ny,nx)
  call hyd_check(hyd_gradient(hctx, ff, ffx, ffy), 'gradient')
end subroutine

subroutine spec1d(ss, spec, iopt)
  double precision, intent(inout) :: ss(nx,ny)
  double precision, intent(out) :: spec(0:max(nx,ny))
  integer, intent(in) :: iopt
  call hyd_check(hyd_spec1d(hctx, ss, spec, int(iopt, c_int32_t)), 'spec1d')
end subroutine

end module spectral
