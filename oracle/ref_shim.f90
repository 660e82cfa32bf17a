! ref_shim.f90 -- bind(c) accessors around the reference's own Fortran
! modules (contours, spectral, sta2dfft, stafft, congen), compiled together
! with them into oracle/_ref/libhydra_ref_ngNN.so by oracle/Makefile.
!
! TEST INFRASTRUCTURE ONLY: used by tests/ to pin the C oracle bit-for-bit
! against the reference itself.  Nothing here is part of the product.
!
! Every entry point passes plain arrays; contour index arrays are 1-based as
! in the reference.  Grid sizes are the compile-time nx=ny of the build.

module ref_shim

use iso_c_binding
use constants
use contours
use spectral
use congen
use generic
use common, only: readcont, qs_common => qs, qr_common => qr
use variables, only: t_var => t

implicit none

contains

subroutine ref_init() bind(c, name='ref_init')
  call init_spectral
  call init_contours
end subroutine

subroutine ref_sizes(ngo, npmo, nmo) bind(c, name='ref_sizes')
  integer(c_int):: ngo, npmo, nmo
  ngo = nx
  npmo = npm
  nmo = nm
end subroutine

subroutine ref_set_contours(np, nc, x, y, nxt, i1, i2, npc, ind, qj) &
    bind(c, name='ref_set_contours')
  integer(c_int), value:: np, nc
  real(c_double):: x(np), y(np), qj
  integer(c_int):: nxt(np), i1(nc), i2(nc), npc(nc), ind(nc)
  nptq = np
  nq = nc
  xq(1:np) = x
  yq(1:np) = y
  nextq(1:np) = nxt
  i1q(1:nc) = i1
  i2q(1:nc) = i2
  npq(1:nc) = npc
  indq(1:nc) = ind
  qjump = qj
end subroutine

subroutine ref_counts(np, nc) bind(c, name='ref_counts')
  integer(c_int):: np, nc
  np = nptq
  nc = nq
end subroutine

subroutine ref_get_contours(x, y, nxt, i1, i2, npc, ind) bind(c, name='ref_get_contours')
  real(c_double):: x(nptq), y(nptq)
  integer(c_int):: nxt(nptq), i1(nq), i2(nq), npc(nq), ind(nq)
  x = xq(1:nptq)
  y = yq(1:nptq)
  nxt = nextq(1:nptq)
  i1 = i1q(1:nq)
  i2 = i2q(1:nq)
  npc = npq(1:nq)
  ind = indq(1:nq)
end subroutine

subroutine ref_velint(uu, vv, uq, vq) bind(c, name='ref_velint')
  real(c_double):: uu(ny,nx), vv(ny,nx), uq(nptq), vq(nptq)
  call velint(uu, vv, uq, vq)
end subroutine

subroutine ref_con2grid(qc) bind(c, name='ref_con2grid')
  real(c_double):: qc(ny,nx)
  call con2grid(qc)
end subroutine

subroutine ref_surgery() bind(c, name='ref_surgery')
  call surgery
end subroutine

subroutine ref_renode(xdi, ydi, npd, xro, yro, npr) bind(c, name='ref_renode')
  integer(c_int), value:: npd
  integer(c_int):: npr
  real(c_double):: xdi(npd), ydi(npd), xro(*), yro(*)
  double precision:: xd(nprm), yd(nprm), xr(nprm), yr(nprm)
  xd(1:npd) = xdi
  yd(1:npd) = ydi
  call renode(xd, yd, npd, xr, yr, npr)
  if (npr .gt. 0) then
    xro(1:npr) = xr(1:npr)
    yro(1:npr) = yr(1:npr)
  endif
end subroutine

subroutine ref_ptospc(rvar, svar) bind(c, name='ref_ptospc')
  real(c_double):: rvar(ny,nx), svar(nx,ny)
  call ptospc(nx, ny, rvar, svar, xfactors, yfactors, xtrig, ytrig)
end subroutine

subroutine ref_spctop(svar, rvar) bind(c, name='ref_spctop')
  real(c_double):: rvar(ny,nx), svar(nx,ny)
  call spctop(nx, ny, svar, rvar, xfactors, yfactors, xtrig, ytrig)
end subroutine

subroutine ref_forfft(m, x) bind(c, name='ref_forfft')
  integer(c_int), value:: m
  real(c_double):: x(m*nx)
  call forfft(m, nx, x, xtrig, xfactors)
end subroutine

subroutine ref_revfft(m, x) bind(c, name='ref_revfft')
  integer(c_int), value:: m
  real(c_double):: x(m*nx)
  call revfft(m, nx, x, xtrig, xfactors)
end subroutine

subroutine ref_xderiv(var, der) bind(c, name='ref_xderiv')
  real(c_double):: var(nx,ny), der(nx,ny)
  call xderiv(nx, ny, hrkx, var, der)
end subroutine

subroutine ref_yderiv(var, der) bind(c, name='ref_yderiv')
  real(c_double):: var(nx,ny), der(nx,ny)
  call yderiv(nx, ny, hrky, var, der)
end subroutine

subroutine ref_main_invert(qqi, uu, vv, pp) bind(c, name='ref_main_invert')
  real(c_double):: qqi(nx,ny), uu(ny,nx), vv(ny,nx), pp(nx,ny)
  call main_invert(qqi, uu, vv, pp)
end subroutine

subroutine ref_gradient(ff, ffx, ffy) bind(c, name='ref_gradient')
  real(c_double):: ff(nx,ny), ffx(ny,nx), ffy(ny,nx)
  call gradient(ff, ffx, ffy)
end subroutine

! id: 1 green, 2 flo, 3 fhi, 4 qdiss, 5 rksq, 6 rk  (nx*ny)
!     7 hrkx, 8 rkx (nx); 9 hrky, 10 rky (ny); 11 xtrig (2nx); 12 ytrig (2ny)
subroutine ref_table(id, out) bind(c, name='ref_table')
  integer(c_int), value:: id
  real(c_double):: out(*)
  select case (id)
  case (1); out(1:nx*ny) = reshape(green, [nx*ny])
  case (2); out(1:nx*ny) = reshape(flo, [nx*ny])
  case (3); out(1:nx*ny) = reshape(fhi, [nx*ny])
  case (4); out(1:nx*ny) = reshape(qdiss, [nx*ny])
  case (5); out(1:nx*ny) = reshape(rksq, [nx*ny])
  case (6); out(1:nx*ny) = reshape(rk, [nx*ny])
  case (7); out(1:nx) = hrkx
  case (8); out(1:nx) = rkx
  case (9); out(1:ny) = hrky
  case (10); out(1:ny) = rky
  case (11); out(1:2*nx) = xtrig
  case (12); out(1:2*ny) = ytrig
  end select
end subroutine

subroutine ref_consts(out) bind(c, name='ref_consts')
  real(c_double):: out(12)
  out(1) = dm
  out(2) = dm2
  out(3) = d4small
  out(4) = dmsq
  out(5) = dmi
  out(6) = elf
  out(7) = densf
  out(8) = ell
  out(9) = glx
  out(10) = oms
  out(11) = small3
  out(12) = kdsq
end subroutine

! spectral.f90:370-424 (ss is modified when iopt = 1)
subroutine ref_spec1d(ss, spec, iopt) bind(c, name='ref_spec1d')
  real(c_double):: ss(nx,ny), spec(0:max(nx,ny))
  integer(c_int), value:: iopt
  call spec1d(ss, spec, iopt)
end subroutine

! generic.f90:30-64
subroutine ref_norms(qq, ql1, ql2) bind(c, name='ref_norms')
  real(c_double):: qq(ny,nx), ql1, ql2
  call l1norm(qq, ql1)
  call l2norm(qq, ql2)
end subroutine

! generic.f90:66-79
subroutine ref_contint(qq, n, dq) bind(c, name='ref_contint')
  real(c_double):: qq(ny,nx), dq
  integer(c_int), value:: n
  call contint(qq, n, dq)
end subroutine

subroutine ref_kmax(k) bind(c, name='ref_kmax')
  integer(c_int):: k
  k = kmax
end subroutine

! common.f90:30-100, readcont(iind) in the current working directory
! (cont/qqsynopsis.asc, cont/qqindexNNN, cont/qqnodesNNN, cont/qqresi.r4)
subroutine ref_readcont(iind, qs_out, qr_out, tout) bind(c, name='ref_readcont')
  integer(c_int), value:: iind
  real(c_double):: qs_out(nx,ny), qr_out(ny,nx), tout
  call readcont(iind)
  qs_out = qs_common
  qr_out = qr_common
  tout = t_var
end subroutine

! The open and read statements of caps.f90:150-153 (qq_init.r8 in the
! current working directory).
subroutine ref_read_qq_init(tout, qq) bind(c, name='ref_read_qq_init')
  real(c_double):: tout, qq(ny,nx)
  double precision:: tt, qa(ny,nx)
  open(11,file='qq_init.r8',form='unformatted', &
      & access='direct',status='old',recl=2*nbytes)
  read(11,rec=1) tt,qa
  close(11)
  tout = tt
  qq = qa
end subroutine

! The formats of evolution.f90:676-728 and 750, written to line (len 128):
! kind 1 ene.asc, 2 norms.asc, 3 spectra.asc header, 4 spectra.asc line,
! 5 qqsynopsis.asc, 6 complexity.asc (evolution.f90:72, 111), 7 monitor.asc
! (evolution.f90:583-584).  v holds the reals, iv the integers in statement
! order.
subroutine ref_format(kind, v, iv, line) bind(c, name='ref_format')
  integer(c_int), value:: kind
  real(c_double):: v(8)
  integer(c_int):: iv(2)
  character(kind=c_char):: line(128)
  character(len=128):: buf
  integer:: i
  buf = ' '
  select case (kind)
  case (1); write(buf,'(f9.2,3(1x,f17.11))') v(1),v(2),v(3),v(2)+v(3)
  case (2); write(buf,'(f9.2,1x,f16.9)') v(1),v(2)
  case (3); write(buf,'(f9.2,2(1x,f16.9),1x,i5)') v(1),v(2),v(3),iv(1)
  case (4); write(buf,'(2(1x,f12.8))') v(1),v(2)
  case (5); write(buf,'(i9,1x,i10,1x,f12.5,1x,f16.12)') iv(1),iv(2),v(1),v(2)
  case (6); write(buf,'(1x,f12.5,1x,i9,1x,i10)') v(1),iv(1),iv(2)
  case (7); write(buf,'(1x,f12.5,1x,f6.4,4(1x,f12.6),1x,f5.3)') v(1),v(2),v(3),v(4),v(5),v(6),v(7)
  end select
  do i = 1, 128
    line(i) = buf(i:i)
  enddo
end subroutine

! congen.f90 recontour: qqr is the residual (or the full field when there are
! no contours); the new contours replace the module's store.
subroutine ref_recontour(qqr) bind(c, name='ref_recontour')
  real(c_double):: qqr(ny,nx)
  double precision:: qw(ny,nx)
  qw = qqr
  call recontour(qw)
end subroutine

end module
