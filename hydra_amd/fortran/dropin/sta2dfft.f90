! sta2dfft.f90 -- drop-in for the doubly periodic part of
! lib/stafft/sta2dfft.f90 (init2dfft, ptospc, spctop, xderiv, yderiv,
! sta2dfft.f90:40-245) on the GPU through libhydra_amd.  Same names,
! arguments and layouts: physical (ny,nx), spectral (nx,ny) in stafft's
! half-complex layout.  The grid is the run's (constants.f90); ptospc and
! spctop leave their input unchanged (the reference overwrites it).
! xderiv/yderiv take the module's filtered wavenumbers (hrkx, hrky of
! spectral.f90), the only ones the path passes; other arrays stop the run.
module sta2dfft

use iso_c_binding
use hydra_amd_c
use hydra_amd_ctx
use stafft

implicit none

double precision, save, private :: hkx(nx), hky(ny)
logical, save, private :: have_hk = .false.

contains

subroutine check_grid(mx, my, what)
  integer, intent(in) :: mx, my
  character(len=*), intent(in) :: what
  if (mx /= nx .or. my /= ny) then
    write(*, '(a,a,a,2i6)') ' hydra_amd: ', what, ': grid differs from the run''s ', nx, ny
    stop 1
  endif
end subroutine

subroutine init2dfft(mx, my, lx, ly, xfactors, yfactors, xtrig, ytrig, kx, ky)
  integer, intent(in) :: mx, my
  double precision, intent(in) :: lx, ly
  integer, intent(out) :: xfactors(5), yfactors(5)
  double precision, intent(out) :: xtrig(2*mx), ytrig(2*my), kx(mx), ky(my)
  integer :: k
  call check_grid(mx, my, 'init2dfft')
  if (lx /= ellx .or. ly /= elly) then
    write(*, '(a)') ' hydra_amd: init2dfft: domain differs from the run''s'
    stop 1
  endif
  call initfft(mx, xfactors, xtrig)
  call initfft(my, yfactors, ytrig)
  ! the wavenumbers pi/l * k, k = 1..n (sta2dfft.f90:64-88)
  do k = 1, mx
    kx(k) = (pi / lx) * dble(k)
  enddo
  do k = 1, my
    ky(k) = (pi / ly) * dble(k)
  enddo
end subroutine

subroutine ptospc(mx, my, rvar, svar, xfactors, yfactors, xtrig, ytrig)
  integer, intent(in) :: mx, my, xfactors(5), yfactors(5)
  double precision, intent(inout) :: rvar(my, mx)
  double precision, intent(out) :: svar(mx, my)
  double precision, intent(in) :: xtrig(2*mx), ytrig(2*my)
  call check_grid(mx, my, 'ptospc')
  call hyd_open
  call hyd_check(hyd_ptospc(hctx, rvar, svar), 'ptospc')
end subroutine

subroutine spctop(mx, my, svar, rvar, xfactors, yfactors, xtrig, ytrig)
  integer, intent(in) :: mx, my, xfactors(5), yfactors(5)
  double precision, intent(inout) :: svar(mx, my)
  double precision, intent(out) :: rvar(my, mx)
  double precision, intent(in) :: xtrig(2*mx), ytrig(2*my)
  call check_grid(mx, my, 'spctop')
  call hyd_open
  call hyd_check(hyd_spctop(hctx, svar, rvar), 'spctop')
end subroutine

subroutine load_hk
  if (have_hk) return
  call hyd_open
  call hyd_check(hyd_get_table(hctx, 6_c_int32_t, hkx), 'hrkx')
  call hyd_check(hyd_get_table(hctx, 8_c_int32_t, hky), 'hrky')
  have_hk = .true.
end subroutine

subroutine xderiv(mx, my, rkx, var, der)
  integer, intent(in) :: mx, my
  double precision, intent(in) :: rkx(mx), var(mx, my)
  double precision, intent(out) :: der(mx, my)
  call check_grid(mx, my, 'xderiv')
  call load_hk
  if (any(rkx /= hkx)) then
    write(*, '(a)') ' hydra_amd: xderiv: wavenumbers are not the module''s hrkx'
    stop 1
  endif
  call hyd_check(hyd_xderiv(hctx, var, der), 'xderiv')
end subroutine

subroutine yderiv(mx, my, rky, var, der)
  integer, intent(in) :: mx, my
  double precision, intent(in) :: rky(my), var(mx, my)
  double precision, intent(out) :: der(mx, my)
  call check_grid(mx, my, 'yderiv')
  call load_hk
  if (any(rky /= hky)) then
    write(*, '(a)') ' hydra_amd: yderiv: wavenumbers are not the module''s hrky'
    stop 1
  endif
  call hyd_check(hyd_yderiv(hctx, var, der), 'yderiv')
end subroutine

end module sta2dfft
