! Calls the library through hydra_amd_c: with a device it advances a small
! flow from a zero-mean field and prints the counts; without one it prints
! the error that hyd_init returns.
program check_binding
use hydra_amd_c
implicit none
integer, parameter :: ng = 32
type(hyd_params) :: p
type(c_ptr) :: ctx
integer(c_int) :: rc
integer(c_int32_t) :: nptq, nq
double precision :: qq(ng, ng), t, pi
integer :: ix, iy
pi = 4.d0 * atan(1.d0)
p%ng = ng; p%ncontq = 20; p%ellx = 2.d0 * pi; p%elly = p%ellx
p%kd = 0.d0; p%beta = 0.d0; p%rtherm = 0.d0; p%rekman = 0.d0
p%cdamp = 2.d0; p%nnu = 3.d0; p%tgsave = 1000.d0; p%tcsave = 1000.d0
p%dt_fixed = 0.d0; p%node_capacity = 0; p%precision = 0; p%device = 0; p%ell_cells = 0.d0
rc = hyd_init(p, ctx)
if (rc /= HYD_OK) then
  write(*, '(a,i0,a,a)') 'init rc=', rc, ' msg=', hyd_error_text()
  stop
endif
do ix = 1, ng
  do iy = 1, ng
    qq(iy, ix) = 4.d0 * pi * cos(2.d0 * pi * (ix - 1) / ng) * cos(2.d0 * pi * (iy - 1) / ng)
  enddo
enddo
call hyd_check(hyd_set_field(ctx, qq, 0.d0), 'set_field')
call hyd_check(hyd_run(ctx, 10), 'run')
call hyd_check(hyd_get_counts(ctx, nptq, nq), 'get_counts')
call hyd_check(hyd_get_scalar(ctx, 0, t), 'get_scalar')
write(*, '(a,i0,a,i0,a,es24.16)') 'nptq=', nptq, ' nq=', nq, ' t=', t
rc = hyd_destroy(ctx)
end program
