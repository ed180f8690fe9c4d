!------------------------------------------------------------------------------
! Cell-centred flow solver on the GPU: the reference's main program
! (CCFVM/main.F90:4-145) with every hot-path routine executed on the device.
!
! Usage: flo2d_gpu DECK [MESH] [mixed]
!   DECK  the reference's flo.inp deck;  MESH  geometry.inp (default)
!   mixed run the bulk flux/gradient/limiter arithmetic in FP32
! Writes FLO.RES with iteration, log10 of the normalised residual and CFL
! (full precision), and prints the final residue and force coefficients.
!------------------------------------------------------------------------------
program flo2d_gpu
use ufx_cc
implicit none
character(len=256) :: deck, meshfile, arg
integer :: irk, u
real(8) :: forces(8)

if (command_argument_count() < 1) then
   print *, 'You must specify an input file.'
   stop 1
endif
call get_command_argument(1, deck)
meshfile = 'geometry.inp'
if (command_argument_count() >= 2) call get_command_argument(2, meshfile)
if (command_argument_count() >= 3) then
   call get_command_argument(3, arg)
   if (trim(arg) == 'mixed') precision_mode = UFX_PREC_MIXED
endif

call read_input(trim(deck))
call geometric(trim(meshfile))

iter = 0
fres = 1.0d0
open(newunit=u, file='FLO.RES')
do while (iter < maxiter .and. fres > minres)
   iter = iter + 1
   call cfl_number(cfl)
   call prim_refresh()
   call avg_c2v()
   call time_step02()
   call save_old()
   call gradient()
   if (timemode == 'rk3') then
      do irk = 1, 3
         call fvresidual()
         call rk_update(irk)
      enddo
   else
      call lusgs()
   endif
   call check_positivity()
   call residue()
   write(u, '(i8,2es25.16)') iter, log10(fres), cfl
enddo
close(u)
call find_cl_cd(forces)
print '(a,es25.16)', ' Final Residue :', fres
print '(a,3es25.16)', ' Cl1, Cd1, Cm :', forces(3), forces(4), forces(5)
call finish()
end program flo2d_gpu
