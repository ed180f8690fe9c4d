!------------------------------------------------------------------------------
! ISO_C_BINDING interfaces of the C-ABI in include/ufx.h.
!
! Every routine returns an integer status (UFX_OK = 0); ufx_message() gives
! the text of the last failure.  Index arrays are 1-based, as the reference
! holds them.
!------------------------------------------------------------------------------
module ufx_api
use, intrinsic :: iso_c_binding
implicit none

integer(c_int), parameter :: UFX_OK = 0, UFX_ERR_ARG = 1, UFX_ERR_MESH = 2, &
                             UFX_ERR_HIP = 3, UFX_ERR_STATE = 4, UFX_ERR_UNSUPPORTED = 5
integer(c_int), parameter :: UFX_FLUX_ROE = 1, UFX_FLUX_AUSM = 2, UFX_FLUX_VANLEER = 3, &
                             UFX_FLUX_RUSANOV = 4
integer(c_int), parameter :: UFX_GRAD_LSQR = 1, UFX_GRAD_GG = 2, UFX_GRAD_GGFC = 3
integer(c_int), parameter :: UFX_FLOW_INVISCID = 1, UFX_FLOW_LAMINAR = 2
integer(c_int), parameter :: UFX_TIME_RK3 = 1, UFX_TIME_LUSGS = 2
integer(c_int), parameter :: UFX_PREC_FP64 = 0, UFX_PREC_MIXED = 1

integer(c_int), parameter :: UFX_NC_LIM_VKN = 0, UFX_NC_LIM_BJ = 1
integer(c_int), parameter :: UFX_LUSGS_REFERENCE = 0, UFX_LUSGS_MULTICOLOUR = 1

! ufx_nc_params
type, bind(C) :: ufx_nc_params
   real(c_double) :: mach, aoa_deg, cfl_max
   integer(c_int) :: grad, timemode, precision, flux, limiter, lusgs_order
end type ufx_nc_params

! ufx_cc_params
type, bind(C) :: ufx_cc_params
   real(c_double) :: mach, aoa_deg, reynolds, t_wall
   integer(c_int) :: flow_type, iwall, flux, grad, ilimit, irs, timemode, precision
   real(c_double) :: cfl_max, xref, yref
   integer(c_int) :: lusgs_order
end type ufx_cc_params

interface
   function ufx_last_error() bind(C, name="ufx_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
   end function

   integer(c_int) function ufx_cc_geometric(nop, noc, nfi, nbf, px, py, ccx, ccy, cv, &
         fi_p1, fi_p2, fi_in, fi_out, fi_bc, fb_p1, fb_p2, fb_in, fb_nx, fb_ny, fb_bc, mesh) &
         bind(C, name="ufx_cc_geometric")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nop, noc, nfi, nbf
      real(c_double), intent(in) :: px(*), py(*), ccx(*), ccy(*), cv(*)
      integer(c_int), intent(in) :: fi_p1(*), fi_p2(*), fi_in(*), fi_out(*)
      type(c_ptr), value :: fi_bc
      integer(c_int), intent(in) :: fb_p1(*), fb_p2(*), fb_in(*)
      real(c_double), intent(in) :: fb_nx(*), fb_ny(*)
      integer(c_int), intent(in) :: fb_bc(*)
      type(c_ptr) :: mesh
   end function

   subroutine ufx_mesh_destroy(mesh) bind(C, name="ufx_mesh_destroy")
      import :: c_ptr
      type(c_ptr), value :: mesh
   end subroutine

   integer(c_int) function ufx_cc_create(mesh, prm, ctx) bind(C, name="ufx_cc_create")
      import :: c_int, c_ptr, ufx_cc_params
      type(c_ptr), value :: mesh
      type(ufx_cc_params), intent(in) :: prm
      type(c_ptr) :: ctx
   end function

   integer(c_int) function ufx_cc_destroy(ctx) bind(C, name="ufx_cc_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function

   integer(c_int) function ufx_cc_set(ctx, name, src) bind(C, name="ufx_cc_set")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(in) :: src(*)
   end function

   integer(c_int) function ufx_cc_get(ctx, name, dst) bind(C, name="ufx_cc_get")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: dst(*)
   end function

   integer(c_int) function ufx_cc_prim(ctx) bind(C, name="ufx_cc_prim")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function
   integer(c_int) function ufx_cc_avg_c2v(ctx) bind(C, name="ufx_cc_avg_c2v")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function
   integer(c_int) function ufx_cc_time_step(ctx, cfl) bind(C, name="ufx_cc_time_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfl
   end function
   integer(c_int) function ufx_cc_save_old(ctx) bind(C, name="ufx_cc_save_old")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function
   integer(c_int) function ufx_cc_gradient(ctx) bind(C, name="ufx_cc_gradient")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function
   integer(c_int) function ufx_cc_limit(ctx) bind(C, name="ufx_cc_limit")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function
   integer(c_int) function ufx_cc_fvresidual(ctx) bind(C, name="ufx_cc_fvresidual")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function
   integer(c_int) function ufx_cc_rk_update(ctx, stage) bind(C, name="ufx_cc_rk_update")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: stage
   end function
   integer(c_int) function ufx_cc_rk_stage(ctx, stage) bind(C, name="ufx_cc_rk_stage")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: stage
   end function
   integer(c_int) function ufx_cc_lusgs(ctx) bind(C, name="ufx_cc_lusgs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function
   integer(c_int) function ufx_cc_residue(ctx, iter, fres, fres_raw) bind(C, name="ufx_cc_residue")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: iter
      real(c_double), intent(out) :: fres, fres_raw
   end function
   integer(c_int) function ufx_cc_check_positivity(ctx, bad) bind(C, name="ufx_cc_check_positivity")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: bad
   end function
   integer(c_int) function ufx_cc_forces(ctx, out, cp, cf) bind(C, name="ufx_cc_forces")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(8)
      type(c_ptr), value :: cp, cf
   end function
   integer(c_int) function ufx_cc_iterate(ctx, iter0, niter, cfl, hist) bind(C, name="ufx_cc_iterate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: iter0, niter
      real(c_double), intent(in) :: cfl(*)
      real(c_double), intent(out) :: hist(*)
   end function
   ! many cases per launch (no reference counterpart: the batched-cases scheduler)
   integer(c_int) function ufx_cc_batch_create(cases, n, batch) bind(C, name="ufx_cc_batch_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: cases(*)
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: batch
   end function
   integer(c_int) function ufx_cc_batch_set_active(batch, active) bind(C, name="ufx_cc_batch_set_active")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), intent(in) :: active(*)
   end function
   integer(c_int) function ufx_cc_batch_iterate(batch, niter, cfl) bind(C, name="ufx_cc_batch_iterate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), value :: niter
      real(c_double), value :: cfl
   end function
   integer(c_int) function ufx_cc_batch_residual_norms(batch, var, out) &
         bind(C, name="ufx_cc_batch_residual_norms")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), value :: var
      real(c_double), intent(out) :: out(*)
   end function
   integer(c_int) function ufx_cc_batch_destroy(batch) bind(C, name="ufx_cc_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
   end function
   integer(c_int) function ufx_cc_sync(ctx) bind(C, name="ufx_cc_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function
   integer(c_int) function ufx_cc_history(ctx, n, hist) bind(C, name="ufx_cc_history")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      real(c_double), intent(out) :: hist(*)
   end function
   integer(c_int) function ufx_cc_error(ctx, kind, face, cell) bind(C, name="ufx_cc_error")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: kind, face, cell
   end function

   ! ---------------------------------------------- vertex-centred (NCFVM)
   integer(c_int) function ufx_nc_geometric(nop, noc, nof, px, py, pbc, ccx, ccy, cv, &
         e_p1, e_p2, e_in, e_out, e_bc, mesh) bind(C, name="ufx_nc_geometric")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nop, noc, nof
      real(c_double), intent(in) :: px(*), py(*)
      integer(c_int), intent(in) :: pbc(*)
      real(c_double), intent(in) :: ccx(*), ccy(*), cv(*)
      integer(c_int), intent(in) :: e_p1(*), e_p2(*), e_in(*), e_out(*), e_bc(*)
      type(c_ptr) :: mesh
   end function

   integer(c_int) function ufx_nc_create(mesh, prm, ctx) bind(C, name="ufx_nc_create")
      import :: c_int, c_ptr, ufx_nc_params
      type(c_ptr), value :: mesh
      type(ufx_nc_params), intent(in) :: prm
      type(c_ptr) :: ctx
   end function

   integer(c_int) function ufx_nc_destroy(ctx) bind(C, name="ufx_nc_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function

   integer(c_int) function ufx_nc_save_old(ctx) bind(C, name="ufx_nc_save_old")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function

   integer(c_int) function ufx_nc_fvresidual(ctx, cfl) bind(C, name="ufx_nc_fvresidual")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfl
   end function

   integer(c_int) function ufx_nc_rk_stage(ctx, stage, cfl) bind(C, name="ufx_nc_rk_stage")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: stage
      real(c_double), value :: cfl
   end function

   integer(c_int) function ufx_nc_lusgs(ctx, cfl) bind(C, name="ufx_nc_lusgs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfl
   end function

   integer(c_int) function ufx_nc_residue(ctx, iter, out) bind(C, name="ufx_nc_residue")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: iter
      real(c_double), intent(out) :: out(5)
   end function

   integer(c_int) function ufx_nc_iterate(ctx, iter0, niter, hist) bind(C, name="ufx_nc_iterate")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: iter0, niter
      type(c_ptr), value :: hist
   end function

   integer(c_int) function ufx_nc_set(ctx, name, src) bind(C, name="ufx_nc_set")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(in) :: src(*)
   end function

   integer(c_int) function ufx_nc_get(ctx, name, dst) bind(C, name="ufx_nc_get")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: dst(*)
   end function

   integer(c_int) function ufx_nc_forces(ctx, out, cp) bind(C, name="ufx_nc_forces")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(4)
      type(c_ptr), value :: cp
   end function

   integer(c_int) function ufx_nc_sync(ctx) bind(C, name="ufx_nc_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
   end function
end interface

contains

! Text of the last failing call.
function ufx_message() result(msg)
   character(len=:), allocatable :: msg
   character(kind=c_char), pointer :: s(:)
   type(c_ptr) :: p
   integer :: n
   p = ufx_last_error()
   call c_f_pointer(p, s, [4096])
   n = 0
   do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
   enddo
   allocate(character(len=n) :: msg)
   msg = transfer(s(1:n), msg)
end function ufx_message

! C string of a Fortran string.
function cstr(s) result(c)
   character(len=*), intent(in) :: s
   character(kind=c_char) :: c(len_trim(s) + 1)
   integer :: i
   do i = 1, len_trim(s)
      c(i) = s(i:i)
   enddo
   c(len_trim(s) + 1) = c_null_char
end function cstr

end module ufx_api
