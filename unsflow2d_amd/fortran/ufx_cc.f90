!------------------------------------------------------------------------------
! Fortran host of the cell-centred hot path on the GPU.
!
! Module procedures named after the reference's routines, each performing
! that routine on the device through the C-ABI (module ufx_api); the state
! stays resident on the GPU between calls.  read_input and geometric read the
! reference's own input files (flo.inp deck, geometry.inp mesh).
!------------------------------------------------------------------------------
module ufx_cc
use ufx_api
implicit none

type(c_ptr) :: ctx = c_null_ptr, mesh = c_null_ptr
type(ufx_cc_params) :: prm

! deck, CCFVM/read_input.F90:31-50
character(len=24) :: flow_type, timemode, grad_type, flux_type, cfl_type
character(len=64) :: gridfile
integer :: istart, iwall, iterlast, maxiter, saveinterval, scrinterval, niso
integer :: ilimit, irs, vortex, cfl_ramp_steps
real(8) :: t_wall, m_inf, aoa_deg, rey, cfl_min, cfl_max, minres, xref, yref
integer :: precision_mode = UFX_PREC_FP64

! iteration state (CCFVM/types/param.F90)
integer :: iter = 0
real(8) :: cfl = 0.0d0, fres = 1.0d0, fres_raw = 0.0d0
integer :: nop, noc, nfi, nbf

contains

subroutine ufx_check(rc, where)
   integer(c_int), intent(in) :: rc
   character(len=*), intent(in) :: where
   if (rc /= UFX_OK) then
      print *, where, ': ', ufx_message()
      stop 1
   endif
end subroutine ufx_check

! CCFVM/read_input.F90:4-121 (same labels, same validation)
subroutine read_input(deck)
   character(len=*), intent(in) :: deck
   character(len=32) :: sdummy
   integer :: inp
   logical :: ok
   open(newunit=inp, file=deck, status='old')
   read(inp,*) sdummy, istart
   read(inp,*) sdummy, flow_type, iwall, t_wall
   read(inp,*) sdummy, m_inf
   read(inp,*) sdummy, aoa_deg
   read(inp,*) sdummy, rey
   read(inp,*) sdummy, cfl_type, cfl_min, cfl_max, cfl_ramp_steps
   read(inp,*) sdummy, timemode
   read(inp,*) sdummy, grad_type
   read(inp,*) sdummy, iterlast
   read(inp,*) sdummy, maxiter
   read(inp,*) sdummy, minres
   read(inp,*) sdummy, saveinterval
   read(inp,*) sdummy, scrinterval
   read(inp,*) sdummy, niso
   read(inp,*) sdummy, flux_type
   read(inp,*) sdummy, ilimit
   read(inp,*) sdummy, irs
   read(inp,*) sdummy, vortex, xref, yref
   read(inp,*) sdummy, gridfile
   close(inp)
   ok = .true.
   if (istart /= 1 .and. istart /= 2) then
      print *, 'Unknown start option', istart; ok = .false.
   endif
   if (flow_type /= 'inviscid' .and. flow_type /= 'laminar') then
      print *, 'Unknown flow type', flow_type; ok = .false.
   endif
   if (iwall /= 0 .and. iwall /= 1) then
      print *, 'Unknown wall conditions :', iwall; ok = .false.
   endif
   if (timemode /= 'rk3' .and. timemode /= 'lusgs') then
      print *, 'Unknown time integration scheme ', timemode; ok = .false.
   endif
   if (cfl_type /= 'tanh' .and. cfl_type /= 'ramp') then
      print *, 'Unknown flow type', cfl_type; ok = .false.
   endif
   if (grad_type /= 'gg' .and. grad_type /= 'lsqr' .and. grad_type /= 'ggfc') then
      print *, 'Unknown  gradient type :', grad_type; ok = .false.
   endif
   if (flux_type /= 'roe' .and. flux_type /= 'ausm' .and. flux_type /= 'vanleer' .and. &
       flux_type /= 'rusanov') then
      print *, 'Unknown flux ', flux_type; ok = .false.
   endif
   if (irs /= 0 .and. irs /= 1) then
      print *, 'Unknown Implicit Residual Smoothening option', irs; ok = .false.
   endif
   if (ilimit /= 0 .and. ilimit /= 1) then
      print *, 'Unknown limiter option', ilimit; ok = .false.
   endif
   if (.not. ok) stop 1
   if (timemode == 'rk3' .and. cfl_max >= 1.0d0) stop 'CFL should be less than 1 for rk3 !!!'
   prm%mach = m_inf
   prm%aoa_deg = aoa_deg
   prm%reynolds = rey
   prm%t_wall = t_wall
   prm%flow_type = merge(UFX_FLOW_INVISCID, UFX_FLOW_LAMINAR, flow_type == 'inviscid')
   prm%iwall = iwall
   select case (trim(flux_type))
   case ('roe');     prm%flux = UFX_FLUX_ROE
   case ('ausm');    prm%flux = UFX_FLUX_AUSM
   case ('vanleer'); prm%flux = UFX_FLUX_VANLEER
   case default;     prm%flux = UFX_FLUX_RUSANOV
   end select
   select case (trim(grad_type))
   case ('lsqr'); prm%grad = UFX_GRAD_LSQR
   case ('gg');   prm%grad = UFX_GRAD_GG
   case default;  prm%grad = UFX_GRAD_GGFC
   end select
   prm%ilimit = ilimit
   prm%irs = irs
   prm%timemode = merge(UFX_TIME_RK3, UFX_TIME_LUSGS, timemode == 'rk3')
   prm%precision = precision_mode
   prm%cfl_max = cfl_max
   prm%xref = xref
   prm%yref = yref
   prm%lusgs_order = UFX_LUSGS_REFERENCE
end subroutine read_input

! Read geometry.inp (CCFVM/read_input.F90:138-173), preprocess it on the host
! as the reference's geometric does, and create the device context (which
! holds the state `initialize` sets, CCFVM/solution.F90:78-81).
subroutine geometric(path)
   character(len=*), intent(in) :: path
   real(8), allocatable :: px(:), py(:), ccx(:), ccy(:), cv(:), bnx(:), bny(:)
   integer(c_int), allocatable :: pbc(:), a1(:), a2(:), ai(:), ao(:), abc(:), b1(:), b2(:), bi(:), bbc(:)
   integer :: u, i, j
   open(newunit=u, file=path, status='old')
   read(u,*) nop, noc, nfi, nbf
   allocate(px(nop), py(nop), pbc(nop), ccx(noc), ccy(noc), cv(noc))
   allocate(a1(nfi), a2(nfi), ai(nfi), ao(nfi), abc(nfi))
   allocate(b1(nbf), b2(nbf), bi(nbf), bnx(nbf), bny(nbf), bbc(nbf))
   do i = 1, nop
      read(u,*) j, px(i), py(i), pbc(i)
   enddo
   do i = 1, noc
      read(u,*) j, ccx(i), ccy(i), cv(i)
   enddo
   do i = 1, nfi
      read(u,*) j, a1(i), a2(i), ai(i), ao(i), bnx(1), bny(1), abc(i)
   enddo
   do i = 1, nbf
      read(u,*) j, b1(i), b2(i), bi(i), bnx(i), bny(i), bbc(i)
   enddo
   close(u)
   call ufx_check(ufx_cc_geometric(nop, noc, nfi, nbf, px, py, ccx, ccy, cv, a1, a2, ai, ao, &
                  c_null_ptr, b1, b2, bi, bnx, bny, bbc, mesh), 'geometric')
   call ufx_check(ufx_cc_create(mesh, prm, ctx), 'ufx_cc_create')
end subroutine geometric

! CCFVM/main.F90:148-168
subroutine cfl_number(cfl_no)
   real(8), intent(inout) :: cfl_no
   real(8) :: s
   if (cfl_type == 'tanh') then
      s = real(iter, 8) / real(cfl_ramp_steps - 1, 8)
      cfl_no = 0.5d0 * cfl_max * (1.0d0 + tanh(s * 7.0d0 - 3.5d0))
   elseif (cfl_type == 'ramp' .and. iter <= cfl_ramp_steps) then
      s = real(iter, 8) / real(cfl_ramp_steps, 8)
      cfl_no = cfl_min + (cfl_max - cfl_min) * (1.0d0 - exp(-s * 0.5d0)) / (1.0d0 - exp(-0.5d0))
   endif
end subroutine cfl_number

! CCFVM/main.F90:62-67
subroutine prim_refresh()
   call ufx_check(ufx_cc_prim(ctx), 'prim')
end subroutine
! CCFVM/average.F90:1-39
subroutine avg_c2v()
   call ufx_check(ufx_cc_avg_c2v(ctx), 'avg_c2v')
end subroutine
! CCFVM/timestep.F90:98-121
subroutine time_step02()
   call ufx_check(ufx_cc_time_step(ctx, cfl), 'time_step02')
end subroutine
! CCFVM/solution.F90:118-132
subroutine save_old()
   call ufx_check(ufx_cc_save_old(ctx), 'save_old')
end subroutine
! CCFVM/gradient.F90 (the deck's method; the limiter follows when ILIMIT=1)
subroutine gradient()
   call ufx_check(ufx_cc_gradient(ctx), 'gradient')
end subroutine
! CCFVM/limiter.F90:3-17
subroutine limit()
   call ufx_check(ufx_cc_limit(ctx), 'limit')
end subroutine
! CCFVM/fvresidual.F90:4-187
subroutine fvresidual()
   call ufx_check(ufx_cc_fvresidual(ctx), 'fvresidual')
end subroutine
! CCFVM/main.F90:92-101
subroutine rk_update(irk)
   integer, intent(in) :: irk
   call ufx_check(ufx_cc_rk_update(ctx, irk), 'rk_update')
end subroutine
! CCFVM/lusgs.F90:5-10
subroutine lusgs()
   call ufx_check(ufx_cc_lusgs(ctx), 'lusgs')
end subroutine
! CCFVM/solution.F90:136-180
subroutine residue()
   call ufx_check(ufx_cc_residue(ctx, iter, fres, fres_raw), 'residue')
end subroutine
! CCFVM/main.F90:358-414
subroutine check_positivity()
   integer(c_int) :: bad, rc
   rc = ufx_cc_check_positivity(ctx, bad)
   if (bad /= 0) then
      print *
      print *, 'Density/pressure is negative at point ', bad
      stop
   endif
   call ufx_check(rc, 'check_positivity')
end subroutine
! CCFVM/main.F90:263-355: cl, cd (pressure integral), cl1, cd1 (printed), cm
subroutine find_cl_cd(out)
   real(8), intent(out) :: out(8)
   call ufx_check(ufx_cc_forces(ctx, out, c_null_ptr, c_null_ptr), 'find_cl_cd')
end subroutine

subroutine finish()
   integer(c_int) :: rc
   if (c_associated(ctx)) rc = ufx_cc_destroy(ctx)
   if (c_associated(mesh)) call ufx_mesh_destroy(mesh)
   ctx = c_null_ptr
   mesh = c_null_ptr
end subroutine finish

end module ufx_cc
