! hydra_amd_ctx.f90 -- the library context that the drop-in modules
! (stafft, sta2dfft, spectral, contours) share: one per program, opened on
! first use with the run's compile-time parameters (parameters.f90,
! constants.f90 of ca/plane/qg/caps), as the reference's modules share
! their module state.
module hydra_amd_ctx

use iso_c_binding
use hydra_amd_c
use constants

implicit none

type(c_ptr), save :: hctx = c_null_ptr

contains

subroutine hyd_open
  type(hyd_params) :: p
  if (c_associated(hctx)) return
  p%ng = nx
  p%ncontq = ncontq
  p%ellx = ellx
  p%elly = elly
  p%kd = kd
  p%beta = beta
  p%rtherm = rtherm
  p%rekman = rekman
  p%cdamp = cdamp
  p%nnu = nnu
  p%tgsave = tgsave
  p%tcsave = tcsave
  p%dt_fixed = 0.d0
  p%node_capacity = int(npm, c_int64_t)
  p%precision = 0
  p%device = 0
  p%ell_cells = 0.d0
  call hyd_check(hyd_init(p, hctx), 'hyd_init')
end subroutine

end module hydra_amd_ctx
