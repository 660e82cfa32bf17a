! hydra_amd_c.f90 -- iso_c_binding interface to libhydra_amd.so.
!
! A Fortran host (the CAPS driver of ca/plane/qg/caps) uses this module to
! call the MI355X step through the C ABI of include/hydra_amd.h.  Arrays are
! passed by reference exactly as the reference keeps them: physical fields
! (ny,nx), spectral fields (nx,ny), contour arrays 1-based.  Every function
! returns 0 on success; hyd_check stops the program on an error, which is how
! the reference reacts to the same conditions (stafft.f90:85-93).
module hydra_amd_c
use, intrinsic :: iso_c_binding
implicit none

integer(c_int), parameter :: HYD_OK = 0

type, bind(c) :: hyd_params
  integer(c_int32_t) :: ng, ncontq
  real(c_double) :: ellx, elly, kd, beta, rtherm, rekman, cdamp, nnu
  real(c_double) :: tgsave, tcsave, dt_fixed
  integer(c_int64_t) :: node_capacity
  integer(c_int32_t) :: precision, device
  real(c_double) :: ell_cells
end type hyd_params

interface
  function hyd_last_error() bind(c, name="hyd_last_error")
    import :: c_ptr
    type(c_ptr) :: hyd_last_error
  end function

  integer(c_int) function hyd_init(p, ctx) bind(c, name="hyd_init")
    import :: c_int, c_ptr, hyd_params
    type(hyd_params), intent(in) :: p
    type(c_ptr), intent(out) :: ctx
  end function

  integer(c_int) function hyd_destroy(ctx) bind(c, name="hyd_destroy")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function

  ! stafft.f90:63-125 (no context)
  integer(c_int) function hyd_initfft(n, factors, trig) bind(c, name="hyd_initfft")
    import :: c_int, c_int32_t, c_double
    integer(c_int32_t), value :: n
    integer(c_int32_t), intent(out) :: factors(5)
    real(c_double), intent(out) :: trig(*)
  end function

  ! stafft.f90:187-406
  integer(c_int) function hyd_forfft(ctx, m, n, x, trig, factors) bind(c, name="hyd_forfft")
    import :: c_int, c_int32_t, c_ptr, c_double
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: m, n
    real(c_double), intent(inout) :: x(*)
    real(c_double), intent(in) :: trig(*)
    integer(c_int32_t), intent(in) :: factors(5)
  end function

  integer(c_int) function hyd_revfft(ctx, m, n, x, trig, factors) bind(c, name="hyd_revfft")
    import :: c_int, c_int32_t, c_ptr, c_double
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: m, n
    real(c_double), intent(inout) :: x(*)
    real(c_double), intent(in) :: trig(*)
    integer(c_int32_t), intent(in) :: factors(5)
  end function

  ! init_spectral / init_contours tables (spectral.f90:42-158, contours.f90:48-132)
  integer(c_int) function hyd_get_table(ctx, id, out) bind(c, name="hyd_get_table")
    import :: c_int, c_int32_t, c_ptr, c_double
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: id
    real(c_double), intent(out) :: out(*)
  end function

  ! sta2dfft.f90:103-165
  integer(c_int) function hyd_ptospc(ctx, rvar, svar) bind(c, name="hyd_ptospc")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: rvar(*)
    real(c_double), intent(out) :: svar(*)
  end function

  integer(c_int) function hyd_spctop(ctx, svar, rvar) bind(c, name="hyd_spctop")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: svar(*)
    real(c_double), intent(out) :: rvar(*)
  end function

  ! sta2dfft.f90:169-245
  integer(c_int) function hyd_xderiv(ctx, var, der) bind(c, name="hyd_xderiv")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: var(*)
    real(c_double), intent(out) :: der(*)
  end function

  integer(c_int) function hyd_yderiv(ctx, var, der) bind(c, name="hyd_yderiv")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: var(*)
    real(c_double), intent(out) :: der(*)
  end function

  ! spectral.f90:304-366
  integer(c_int) function hyd_main_invert(ctx, qq, uu, vv, pp) bind(c, name="hyd_main_invert")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: qq(*)
    real(c_double), intent(out) :: uu(*), vv(*), pp(*)
  end function

  integer(c_int) function hyd_gradient(ctx, ff, ffx, ffy) bind(c, name="hyd_gradient")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: ff(*)
    real(c_double), intent(out) :: ffx(*), ffy(*)
  end function

  ! spectral.f90:370-424 (ss modified when iopt = 1)
  integer(c_int) function hyd_spec1d(ctx, ss, spec, iopt) bind(c, name="hyd_spec1d")
    import :: c_int, c_int32_t, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(inout) :: ss(*)
    real(c_double), intent(inout) :: spec(*)
    integer(c_int32_t), value :: iopt
  end function

  ! generic.f90:30-64
  integer(c_int) function hyd_norms(ctx, qq, ql1, ql2) bind(c, name="hyd_norms")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: qq(*)
    real(c_double), intent(out) :: ql1, ql2
  end function

  ! evolution.f90:612-644
  integer(c_int) function hyd_energy(ctx, ekin, epot) bind(c, name="hyd_energy")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(out) :: ekin, epot
  end function

  ! contours.f90:136-169
  integer(c_int) function hyd_velint(ctx, uu, vv, xq, yq, nptq, uq, vq) &
      bind(c, name="hyd_velint")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: uu(*), vv(*), xq(*), yq(*)
    integer(c_int32_t), value :: nptq
    real(c_double), intent(out) :: uq(*), vq(*)
  end function

  ! contours.f90:173-346
  integer(c_int) function hyd_con2grid(ctx, xq, yq, nextq, nptq, qjump, qc, counts) &
      bind(c, name="hyd_con2grid")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: xq(*), yq(*)
    integer(c_int32_t), intent(in) :: nextq(*)
    integer(c_int32_t), value :: nptq
    real(c_double), value :: qjump
    real(c_double), intent(out) :: qc(*)
    type(c_ptr), value :: counts
  end function

  ! contours.f90:687-933
  integer(c_int) function hyd_renode(ctx, xd, yd, npd, xr, yr, cap_out, npr) &
      bind(c, name="hyd_renode")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: xd(*), yd(*)
    integer(c_int32_t), value :: npd, cap_out
    real(c_double), intent(out) :: xr(*), yr(*)
    integer(c_int32_t), intent(out) :: npr
  end function

  ! device-resident state: fields, contours, scalars
  integer(c_int) function hyd_set_field(ctx, qq, t0) bind(c, name="hyd_set_field")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(in) :: qq(*)
    real(c_double), value :: t0
  end function

  integer(c_int) function hyd_set_contours(ctx, nptq, nq, xq, yq, nextq, i1q, i2q, npq, &
      indq, qjump) bind(c, name="hyd_set_contours")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: nptq, nq
    real(c_double), intent(in) :: xq(*), yq(*)
    integer(c_int32_t), intent(in) :: nextq(*), i1q(*), i2q(*), npq(*), indq(*)
    real(c_double), value :: qjump
  end function

  integer(c_int) function hyd_get_counts(ctx, nptq, nq) bind(c, name="hyd_get_counts")
    import :: c_int, c_ptr, c_int32_t
    type(c_ptr), value :: ctx
    integer(c_int32_t), intent(out) :: nptq, nq
  end function

  integer(c_int) function hyd_get_contours(ctx, xq, yq, nextq, i1q, i2q, npq, indq) &
      bind(c, name="hyd_get_contours")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value :: ctx
    real(c_double), intent(out) :: xq(*), yq(*)
    integer(c_int32_t), intent(out) :: nextq(*), i1q(*), i2q(*), npq(*), indq(*)
  end function

  integer(c_int) function hyd_get_field(ctx, id, out) bind(c, name="hyd_get_field")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: id
    real(c_double), intent(out) :: out(*)
  end function

  integer(c_int) function hyd_get_scalar(ctx, id, v) bind(c, name="hyd_get_scalar")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: id
    real(c_double), intent(out) :: v
  end function

  ! contours.f90:350-683, congen.f90:24-128, evolution.f90:28-326
  integer(c_int) function hyd_surgery(ctx) bind(c, name="hyd_surgery")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function

  integer(c_int) function hyd_recontour(ctx) bind(c, name="hyd_recontour")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function

  integer(c_int) function hyd_evolution_init(ctx) bind(c, name="hyd_evolution_init")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function

  integer(c_int) function hyd_prepare(ctx) bind(c, name="hyd_prepare")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function

  integer(c_int) function hyd_inversion(ctx) bind(c, name="hyd_inversion")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function

  integer(c_int) function hyd_advance(ctx) bind(c, name="hyd_advance")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function

  integer(c_int) function hyd_advect_step(ctx) bind(c, name="hyd_advect_step")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function

  integer(c_int) function hyd_run(ctx, nsteps) bind(c, name="hyd_run")
    import :: c_int, c_ptr, c_int32_t
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: nsteps
  end function

  ! advances taken and their node-updates (throughput accounting)
  integer(c_int) function hyd_advance_counts(ctx, advances, updates, reset) &
      bind(c, name="hyd_advance_counts")
    import :: c_int, c_ptr, c_int32_t, c_int64_t
    type(c_ptr), value :: ctx
    integer(c_int64_t), intent(out) :: advances, updates
    integer(c_int32_t), value :: reset
  end function

  ! ensemble of independent flows (array of contexts)
  integer(c_int) function hyd_run_batch(ctxs, n, nsteps) bind(c, name="hyd_run_batch")
    import :: c_int, c_ptr, c_int32_t
    type(c_ptr), intent(in) :: ctxs(*)
    integer(c_int32_t), value :: n, nsteps
  end function

  ! the save schedule decided inside adapt (evolution.f90:582-599)
  integer(c_int) function hyd_set_saves(ctx, on, tinit, igrec, icrec) bind(c, name="hyd_set_saves")
    import :: c_int, c_ptr, c_int32_t, c_double
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: on, igrec, icrec
    real(c_double), value :: tinit
  end function

  integer(c_int) function hyd_pending_saves(ctx, ngrid, ncont) bind(c, name="hyd_pending_saves")
    import :: c_int, c_ptr, c_int32_t
    type(c_ptr), value :: ctx
    integer(c_int32_t), intent(out) :: ngrid, ncont
  end function

  ! savegrid's record (evolution.f90:646-676): info = (tgrid, ekin, epot,
  ! igrec), qqs(ny,nx) interpolated to the save time
  integer(c_int) function hyd_take_grid_save(ctx, info, qqs) bind(c, name="hyd_take_grid_save")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double), intent(out) :: info(4), qqs(*)
  end function

  integer(c_int) function hyd_contour_save_size(ctx, nq, nptq) bind(c, name="hyd_contour_save_size")
    import :: c_int, c_ptr, c_int32_t
    type(c_ptr), value :: ctx
    integer(c_int32_t), intent(out) :: nq, nptq
  end function

  ! savecont's record (evolution.f90:680-778): info = (t, qjump, icrec,
  ! contex), the residual and the contours
  integer(c_int) function hyd_take_contour_save(ctx, info, residual, xq, yq, npq, i1q, indq) &
       bind(c, name="hyd_take_contour_save")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value :: ctx
    real(c_double), intent(out) :: info(4), residual(*), xq(*), yq(*)
    integer(c_int32_t), intent(out) :: npq(*), i1q(*), indq(*)
  end function

  ! monitor.asc (kind 0, 7 values a row) and complexity.asc (kind 1, 3 a row)
  integer(c_int) function hyd_take_records(ctx, kind, max_rows, rows) bind(c, name="hyd_take_records")
    import :: c_int, c_ptr, c_int32_t, c_double
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: kind, max_rows
    real(c_double), intent(out) :: rows(*)
  end function

  ! the learned velocity correction (no reference counterpart)
  integer(c_int) function hyd_set_correction(ctx, net, su, sq, sd) bind(c, name="hyd_set_correction")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx, net
    real(c_double), value :: su, sq, sd
  end function

  integer(c_int) function hyd_contour_velocity(ctx, nptq, xq, yq, nextq, qjump, m, xs, ys, us, vs, &
       fields) bind(c, name="hyd_contour_velocity")
    import :: c_int, c_ptr, c_int32_t, c_double
    type(c_ptr), value :: ctx, fields
    integer(c_int32_t), value :: nptq, m
    real(c_double), value :: qjump
    real(c_double), intent(in) :: xq(*), yq(*), xs(*), ys(*)
    integer(c_int32_t), intent(in) :: nextq(*)
    real(c_double), intent(out) :: us(*), vs(*)
  end function
end interface

contains

! Text of the last error as a Fortran string.
function hyd_error_text() result(msg)
  character(len=:), allocatable :: msg
  character(kind=c_char), pointer :: s(:)
  type(c_ptr) :: p
  integer :: n
  p = hyd_last_error()
  call c_f_pointer(p, s, [4096])
  n = 0
  do while (n < 4096)
    if (s(n + 1) == c_null_char) exit
    n = n + 1
  enddo
  allocate(character(len=n) :: msg)
  msg = transfer(s(1:n), msg)
end function

! Stop on a non-zero status, as the reference stops on its errors.
subroutine hyd_check(rc, what)
  integer(c_int), intent(in) :: rc
  character(len=*), intent(in) :: what
  if (rc /= HYD_OK) then
    write(*, '(a,a,a,i0,a,a)') ' hydra_amd: ', what, ' failed (', rc, '): ', hyd_error_text()
    stop 1
  endif
end subroutine

end module hydra_amd_c
