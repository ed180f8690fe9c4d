!------------------------------------------------------------------------------
! Fortran host of the vertex-centred hot path on the GPU.
!
! Module procedures named after the reference's NCFVM routines, each
! performing that routine on the device through the C-ABI (module ufx_api);
! the state stays resident on the GPU between calls.  read_input and
! geometric read the reference's own input files (flo.inp, geometry.inp).
!------------------------------------------------------------------------------
module ufx_nc
use ufx_api
implicit none

type(c_ptr) :: ctx = c_null_ptr, mesh = c_null_ptr
type(ufx_nc_params) :: prm

! deck, NCFVM/read_input.f90:23-40
character(len=24) :: flow_type, timemode, flux_type
character(len=64) :: gridfile
integer :: istart, iterlast, maxiter, saveinterval, scrinterval, niso, ilimit, vortex
integer :: gmaxiter, prectype
real(8) :: m_inf, aoa_deg, rey, cfl_max, gerrtol, minres, xref, yref

! iteration state (NCFVM/All_modules.f90, module param)
integer :: iter = 0
real(8) :: cfl = 0.0d0, fres = 1.0d0, fresi = 0.0d0
integer :: iresi = 0
integer :: nop, noc, nof

contains

subroutine ufx_check(rc, where)
   integer(c_int), intent(in) :: rc
   character(len=*), intent(in) :: where
   if (rc /= UFX_OK) then
      print *, where, ': ', ufx_message()
      stop 1
   endif
end subroutine ufx_check

! NCFVM/read_input.f90:4-86 (same labels, same validation).  grad and
! limiter select the two sides of the merge conflict at
! NCFVM/fvresidual.f90:11-17 and the reference's uncalled limiter_min_max.
subroutine read_input(deck, grad, limiter)
   character(len=*), intent(in) :: deck
   integer(c_int), intent(in) :: grad, limiter
   character(len=32) :: sdummy
   integer :: inp
   logical :: ok
   open(newunit=inp, file=deck, status='old')
   read(inp,*) sdummy, istart
   read(inp,*) sdummy, flow_type
   read(inp,*) sdummy, m_inf
   read(inp,*) sdummy, aoa_deg
   read(inp,*) sdummy, rey
   read(inp,*) sdummy, cfl_max
   read(inp,*) sdummy, timemode
   read(inp,*) sdummy, gmaxiter, prectype, gerrtol
   read(inp,*) sdummy, iterlast
   read(inp,*) sdummy, maxiter
   read(inp,*) sdummy, minres
   read(inp,*) sdummy, saveinterval
   read(inp,*) sdummy, scrinterval
   read(inp,*) sdummy, niso
   read(inp,*) sdummy, flux_type
   read(inp,*) sdummy, ilimit
   read(inp,*) sdummy, vortex, xref, yref
   read(inp,*) sdummy, gridfile
   close(inp)
   ok = .true.
   if (istart /= 1 .and. istart /= 2) then
      print *, 'Unknown start option', istart; ok = .false.
   endif
   if (flow_type /= 'inviscid' .and. flow_type /= 'laminar' .and. flow_type /= 'rans') then
      print *, 'Unknown flow type', flow_type; ok = .false.
   endif
   if (timemode /= 'rk3' .and. timemode /= 'lusgs') then
      print *, 'Unknown time integration scheme ', timemode; ok = .false.
   endif
   if (ilimit /= 0 .and. ilimit /= 1) then
      print *, 'Unknown limiter option', ilimit; ok = .false.
   endif
   if (vortex /= 0 .and. vortex /= 1) then
      print *, 'Unknown vortex option', vortex; ok = .false.
   endif
   if (.not. ok) stop 1
   ! NCFVM/main.f90:21-23
   if (timemode == 'rk3' .and. cfl_max > 1.0d0) stop 'CFL should be less than 1 for rk3 !!!'
   prm%mach = m_inf
   prm%aoa_deg = aoa_deg
   prm%cfl_max = cfl_max
   prm%grad = grad
   prm%timemode = merge(UFX_TIME_RK3, UFX_TIME_LUSGS, timemode == 'rk3')
   prm%precision = UFX_PREC_FP64
   ! the reference hard-wires van Leer; "ausm" selects the build-defined option
   prm%flux = merge(UFX_FLUX_AUSM, UFX_FLUX_VANLEER, flux_type == 'ausm')
   prm%limiter = limiter
   prm%lusgs_order = UFX_LUSGS_REFERENCE
end subroutine read_input

! read_grid (NCFVM/read_input.f90:89-126), then geometric
! (NCFVM/geometric.f90:1-46) and initialize (NCFVM/common.f90:21-58) on the
! device.
subroutine geometric(path)
   character(len=*), intent(in) :: path
   real(8), allocatable :: px(:), py(:), ccx(:), ccy(:), cv(:)
   real(8) :: nx, ny
   integer(c_int), allocatable :: pbc(:), e1(:), e2(:), ei(:), eo(:), ebc(:)
   integer :: u, i, j
   open(newunit=u, file=path, status='old')
   read(u,*) nop, noc, nof
   allocate(px(nop), py(nop), pbc(nop), ccx(noc), ccy(noc), cv(noc))
   allocate(e1(nof), e2(nof), ei(nof), eo(nof), ebc(nof))
   do i = 1, nop
      read(u,*) j, px(i), py(i), pbc(i)
   enddo
   do i = 1, noc
      read(u,*) j, ccx(i), ccy(i), cv(i)
   enddo
   do i = 1, nof
      read(u,*) j, e1(i), e2(i), ei(i), eo(i), nx, ny, ebc(i)
   enddo
   close(u)
   call ufx_check(ufx_nc_geometric(nop, noc, nof, px, py, pbc, ccx, ccy, cv, e1, e2, ei, eo, &
                  ebc, mesh), 'geometric')
   call ufx_check(ufx_nc_create(mesh, prm, ctx), 'ufx_nc_create')
end subroutine geometric

! NCFVM/common.f90:88-104
subroutine save_old()
   call ufx_check(ufx_nc_save_old(ctx), 'save_old')
end subroutine
! NCFVM/fvresidual.f90:4-85 followed by the stage update of
! NCFVM/main.f90:59-70 (airk, birk of :35-43)
subroutine rk_stage(irk)
   integer, intent(in) :: irk
   call ufx_check(ufx_nc_rk_stage(ctx, irk, cfl), 'fvresidual')
end subroutine
! NCFVM/lusgs.f90:4-131
subroutine lusgs()
   call ufx_check(ufx_nc_lusgs(ctx, cfl), 'lusgs')
end subroutine
! NCFVM/common.f90:108-154
subroutine residue()
   real(c_double) :: out(5)
   call ufx_check(ufx_nc_residue(ctx, iter, out), 'residue')
   fres = out(1)
   fresi = out(3)
   iresi = nint(out(4))
end subroutine
! Wall force (build-defined: NCFVM has no force routine): cl, cd, fx, fy
subroutine find_cl_cd(out)
   real(8), intent(out) :: out(4)
   call ufx_check(ufx_nc_forces(ctx, out, c_null_ptr), 'find_cl_cd')
end subroutine
! NCFVM/common.f90:58-70
subroutine save_flow()
   real(c_double), allocatable :: q(:)
   integer :: u, i, j
   allocate(q(4 * nop))
   call ufx_check(ufx_nc_get(ctx, cstr('qc'), q), 'save_flow')
   open(newunit=u, file='FLO.DAT')
   do i = 1, nop
      write(u,'(4e20.10)') (q(4 * (i - 1) + j), j = 1, 4)
   enddo
   close(u)
end subroutine
! NCFVM/common.f90:73-85
subroutine read_flow()
   real(c_double), allocatable :: q(:)
   integer :: u, i, j
   allocate(q(4 * nop))
   open(newunit=u, file='FLO.DAT', status='old')
   do i = 1, nop
      read(u,*) (q(4 * (i - 1) + j), j = 1, 4)
   enddo
   close(u)
   call ufx_check(ufx_nc_set(ctx, cstr('qc'), q), 'read_flow')
end subroutine

subroutine finish()
   integer(c_int) :: rc
   if (c_associated(ctx)) rc = ufx_nc_destroy(ctx)
   if (c_associated(mesh)) call ufx_mesh_destroy(mesh)
   ctx = c_null_ptr
   mesh = c_null_ptr
end subroutine finish

end module ufx_nc
