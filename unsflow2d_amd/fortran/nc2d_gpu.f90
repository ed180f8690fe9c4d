!------------------------------------------------------------------------------
! Vertex-centred flow solver on the GPU: the reference's main program
! (NCFVM/main.f90:1-110) with every hot-path routine executed on the device.
!
! Usage: nc2d_gpu DECK [MESH] [lsqr] [bj]
!   DECK  the reference's flo.inp deck;  MESH  geometry.inp (default)
!   lsqr  Gradient_LSQR instead of Gradient_GG (the two sides of the merge
!         conflict at NCFVM/fvresidual.f90:11-17)
!   bj    limiter_min_max instead of limiter_VKN
! Writes FLO.RES as the reference does ('(i6,3e16.6)': iteration, log10 L2,
! log10 Linf, CFL) and FLO.DAT at the end; start option 2 begins from FLO.DAT.
!------------------------------------------------------------------------------
program nc2d_gpu
use ufx_nc
implicit none
character(len=256) :: deck, meshfile, arg
integer :: irk, u, k
integer(c_int) :: grad, limiter
real(8) :: forces(4)

if (command_argument_count() < 1) then
   print *, 'You must specify an input file.'
   stop 1
endif
call get_command_argument(1, deck)
meshfile = 'geometry.inp'
grad = UFX_GRAD_GG
limiter = UFX_NC_LIM_VKN
do k = 2, command_argument_count()
   call get_command_argument(k, arg)
   if (trim(arg) == 'lsqr') then
      grad = UFX_GRAD_LSQR
   elseif (trim(arg) == 'bj') then
      limiter = UFX_NC_LIM_BJ
   else
      meshfile = arg
   endif
enddo

call read_input(trim(deck), grad, limiter)
call geometric(trim(meshfile))
if (istart == 2) call read_flow()

iter = 0
fres = 1.0d0
open(newunit=u, file='FLO.RES')
do while (iter < maxiter .and. fres > minres)
   iter = iter + 1
   cfl = 0.5d0 * cfl_max * (1 + dtanh(iter * 7.d0 / 500.0 - 3.5))
   call save_old()
   if (timemode == 'rk3') then
      do irk = 1, 3
         call rk_stage(irk)
      enddo
   else
      call lusgs()
   endif
   call residue()
   write(u,'(i6,3e16.6)') iter, dlog10(fres), dlog10(fresi), cfl
enddo
close(u)
call save_flow()
call find_cl_cd(forces)
print '(a,es25.16)', ' Final Residue :', fres
print '(a,2es25.16)', ' Cl, Cd :', forces(1), forces(2)
call finish()
end program nc2d_gpu
