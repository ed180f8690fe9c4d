!------------------------------------------------------------------------------
! Oracle harness for the cell-centred reference solver (test infrastructure).
!
! Links every object of the reference CCFVM solver except its main program
! (CCFVM/main.F90, whose loop and cfl_number are internal) and restates the
! iteration loop of CCFVM/main.F90:57-128 with cfl_number from :148-168, so
! that every hot-path routine of the reference (avg_c2v, time_step02,
! save_old, Gradient_*, limit, fvresidual, lusgs, residue) runs unmodified and
! its state can be dumped in full double precision between calls.
!
! Usage (run in the directory holding geometry.inp):
!   ref_cc_driver DECK OUTDIR NITER DUMPITERS PHI0 [TIMEWARM]
!     DECK       input deck (read by the reference's read_input as argument 1)
!     OUTDIR     directory for the raw little-endian dumps
!     NITER      iterations to run (the deck's MaxIter is ignored)
!     DUMPITERS  per-routine state is dumped for iterations 1..DUMPITERS
!     PHI0       if >= 0, every cell's limiter value starts at PHI0 instead of
!                1 (PHI0=0 gives the first-order scheme: the limiter only
!                ratchets downward, CCFVM/limiter.F90:153)
!     TIMEWARM   if present, no dumps: iterations 1..TIMEWARM are untimed and
!                the rest are timed with system_clock; prints ms/iteration
!------------------------------------------------------------------------------
program ref_cc_driver
use param
use pri
use grid
use output
use inf
use visc
implicit none
integer(kind=i4) :: i, j, irk, niter, ndump, narg, twarm, ierr
real(kind=dp)    :: c1(nvar), c2(nvar), c3(nvar), dtbyarea, phi0
character(len=512) :: outdir, arg
integer(kind=8)  :: t0, t1, rate
logical          :: timing
real(kind=dp), allocatable :: hist(:,:)

narg = command_argument_count()
if (narg < 5) then
   print*, 'usage: ref_cc_driver DECK OUTDIR NITER DUMPITERS PHI0 [TIMEWARM]'
   stop 2
endif
call get_command_argument(2, outdir)
call get_command_argument(3, arg); read(arg,*) niter
call get_command_argument(4, arg); read(arg,*) ndump
call get_command_argument(5, arg); read(arg,*) phi0
timing = .false.
twarm = 0
if (narg >= 6) then
   call get_command_argument(6, arg); read(arg,*) twarm
   timing = .true.
   ndump = 0
endif

call math
call read_input
call geometric
call initialize
if (phi0 >= 0.0_dp) then
   do j = 1, size(cell)
      cell(j)%phi(:) = phi0
   enddo
endif

! Runge-Kutta coefficients, CCFVM/main.F90:32-45
if (timemode == "rk3") then
   if (cfl_max >= 1.0d0) stop 'rk3 needs cfl_max < 1 (CCFVM/main.F90:32-45)'
   NIRK = 3
   allocate(airk(nirk), birk(nirk))
   airk(1) = 0.0d0
   airk(2) = 3.0d0/4.0d0
   airk(3) = 1.0d0/3.0d0
   birk(1) = 1.0d0
   birk(2) = 1.0d0/4.0d0
   birk(3) = 2.0d0/3.0d0
endif

if (.not. timing) call dump_mesh()

allocate(hist(8, niter))
hist = 0.0_dp
iter = 0
fres = 1.0d0
call system_clock(t0, rate)
do while (iter < niter)
   iter = iter + 1
   if (timing .and. iter == twarm + 1) call system_clock(t0, rate)
   call cfl_number(CFL)

   do i = 1, noc
      call con2prim(cell(i)%qc(1:nvar))
      do j = 1, npvar
         cell(i)%qp(j) = prim(j)
      enddo
   enddo
   if (iter <= ndump) call dump_cell5('qp', 0)

   call avg_c2v
   if (iter <= ndump) call dump_avg()
   call time_step02
   if (iter <= ndump) call dump_dt()
   call save_old

   if (trim(grad_type) == 'gg') then
      call Gradient_GG
   elseif (trim(grad_type) == 'ggfc') then
      call Gradient_GG_FC
   elseif (trim(grad_type) == 'lsqr') then
      call Gradient_LSQR
   else
      stop 'unknown gradient'
   endif
   if (iter <= ndump) call dump_grad(0)

   if (timemode == 'rk3') then
      do irk = 1, nirk
         call fvresidual
         if (iter <= ndump) call dump_stage(irk)
         do i = 1, noc
            c1(:) = cell(i)%qold(:)
            c2(:) = cell(i)%qc(:)
            dtbyarea = cell(i)%dt/cell(i)%cv
            do j = 1, nvar
               c3(j) = airk(irk)*c1(j) + birk(irk)*(c2(j) - dtbyarea*cell(i)%res(j))
            enddo
            cell(i)%qc(:) = c3(:)
         enddo
         if (iter <= ndump) call dump_qc(irk)
      enddo
   elseif (timemode == 'lusgs') then
      call lusgs
      if (iter <= ndump) then
         call dump_stage(1)
         call dump_qc(1)
      endif
   endif

   call residue
   hist(1, iter) = real(iter, dp)
   hist(2, iter) = fres
   hist(3, iter) = fres1
   hist(4, iter) = cfl
   hist(5, iter) = Fx1
   hist(6, iter) = Fy1
   hist(7, iter) = dtglobal
   hist(8, iter) = 0.0_dp
   if (fres /= fres) exit
enddo
call system_clock(t1, rate)

if (timing) then
   write(*,'(a,i8,a,es16.8)') 'timed_iterations ', niter - twarm, ' ms_per_iteration ', &
        1000.0_dp*real(t1 - t0, dp)/real(rate, dp)/real(max(1, niter - twarm), dp)
else
   call dumpd('hist', reshape(hist, (/ 8*niter /)))
   call dump_final()
endif

contains

subroutine cfl_number(cfl_no)
! restated from CCFVM/main.F90:148-168
real(kind=dp) :: cfl1, cfl2, exp_factor, s, cfl_no
if (cfl_type == 'tanh') then
   s = real(iter,dp)/real(CFL_ramp_steps-1,dp)
   cfl_no = 0.5_dp*cfl_max*(1.0_dp + dtanh(s*7.0_dp - 3.5_dp))
elseif (cfl_type == 'ramp' .and. iter <= CFL_ramp_steps) then
   CFL1 = CFL_min
   CFL2 = CFL_max
   exp_factor = 0.5_dp
   s = real(iter,dp)/real(CFL_ramp_steps,dp)
   CFL_no = CFL1 + (CFL2-CFL1)*(1.0_dp - dexp(-s*exp_factor))/(1.0_dp - dexp(-exp_factor))
endif
end subroutine cfl_number

subroutine dumpd(name, a)
character(len=*), intent(in) :: name
real(kind=dp), intent(in) :: a(:)
integer :: u
open(newunit=u, file=trim(outdir)//'/'//name//'.f64', access='stream', &
     form='unformatted', status='replace')
write(u) a
close(u)
end subroutine dumpd

subroutine dumpi(name, a)
character(len=*), intent(in) :: name
integer(kind=i4), intent(in) :: a(:)
integer :: u
open(newunit=u, file=trim(outdir)//'/'//name//'.i32', access='stream', &
     form='unformatted', status='replace')
write(u) a
close(u)
end subroutine dumpi

function tag(prefix, k) result(s)
character(len=*), intent(in) :: prefix
integer, intent(in) :: k
character(len=64) :: s
write(s, '(a,"_it",i4.4,"_",i1)') prefix, iter, k
end function tag

subroutine dump_mesh()
integer(kind=i4) :: k, m, tot
integer(kind=i4), allocatable :: ib(:)
real(kind=dp), allocatable :: db(:)
allocate(ib(nof), db(nof))
ib = fc(:)%pt(1); call dumpi('fc_p1', ib)
ib = fc(:)%pt(2); call dumpi('fc_p2', ib)
ib = fc(:)%in;    call dumpi('fc_in', ib)
ib = fc(:)%out;   call dumpi('fc_out', ib)
ib = fc(:)%bc;    call dumpi('fc_bc', ib)
db = fc(:)%nx;    call dumpd('fc_nx', db)
db = fc(:)%ny;    call dumpd('fc_ny', db)
db = fc(:)%area;  call dumpd('fc_area', db)
db = fc(:)%cen(1); call dumpd('fc_cenx', db)
db = fc(:)%cen(2); call dumpd('fc_ceny', db)
db = fc(:)%cov;   call dumpd('fc_cov', db)
deallocate(ib, db)
allocate(db(size(cell)))
db = cell(:)%cen(1); call dumpd('cell_cenx', db)
db = cell(:)%cen(2); call dumpd('cell_ceny', db)
db = cell(:)%cov;    call dumpd('cell_cov', db)
deallocate(db)
allocate(db(noc))
db = cell(1:noc)%cv;  call dumpd('cell_cv', db)
db = cell(1:noc)%r11; call dumpd('cell_r11', db)
db = cell(1:noc)%r12; call dumpd('cell_r12', db)
db = cell(1:noc)%r22; call dumpd('cell_r22', db)
deallocate(db)
allocate(ib(noc))
ib = cell(1:noc)%nc2c; call dumpi('cell_nc2c', ib)
ib = cell(1:noc)%nc2f; call dumpi('cell_nc2f', ib)
tot = sum(cell(1:noc)%nc2f)
deallocate(ib)
allocate(ib(tot))
m = 0
do k = 1, noc
   ib(m+1:m+cell(k)%nc2f) = cell(k)%c2f(1:cell(k)%nc2f)
   m = m + cell(k)%nc2f
enddo
call dumpi('cell_c2f', ib)
deallocate(ib)
allocate(db(nop))
db = pt(:)%x; call dumpd('pt_x', db)
db = pt(:)%y; call dumpd('pt_y', db)
deallocate(db)
allocate(ib(nop))
ib = pt(:)%nv2c; call dumpi('pt_nv2c', ib)
tot = sum(pt(:)%nv2c)
deallocate(ib)
allocate(ib(tot), db(tot))
m = 0
do k = 1, nop
   ib(m+1:m+pt(k)%nv2c) = pt(k)%v2c(1:pt(k)%nv2c)
   db(m+1:m+pt(k)%nv2c) = pt(k)%wt(1:pt(k)%nv2c)
   m = m + pt(k)%nv2c
enddo
call dumpi('pt_v2c', ib)
call dumpd('pt_wt', db)
deallocate(ib, db)
allocate(ib(8))
ib = (/ nop, noc, nof, nbf, startFC, endFC, startBC, endBC /)
call dumpi('counts', ib)
deallocate(ib)
end subroutine dump_mesh

subroutine dump_cell5(name, k)
character(len=*), intent(in) :: name
integer, intent(in) :: k
real(kind=dp), allocatable :: a(:,:)
integer(kind=i4) :: c
allocate(a(npvar, size(cell)))
do c = 1, size(cell)
   a(:, c) = cell(c)%qp(:)
enddo
call dumpd(trim(tag(name, k)), reshape(a, (/ npvar*size(cell) /)))
end subroutine dump_cell5

subroutine dump_avg()
real(kind=dp), allocatable :: a(:,:)
integer(kind=i4) :: c
allocate(a(npvar, nop))
do c = 1, nop
   a(:, c) = pt(c)%prim(:)
enddo
call dumpd(trim(tag('ptprim', 0)), reshape(a, (/ npvar*nop /)))
deallocate(a)
allocate(a(npvar, nof))
do c = 1, nof
   a(:, c) = fc(c)%qp(:)
enddo
call dumpd(trim(tag('fcqp', 0)), reshape(a, (/ npvar*nof /)))
end subroutine dump_avg

subroutine dump_dt()
real(kind=dp), allocatable :: a(:)
allocate(a(noc))
a = cell(1:noc)%la; call dumpd(trim(tag('la', 0)), a)
a = cell(1:noc)%dt; call dumpd(trim(tag('dt', 0)), a)
deallocate(a)
allocate(a(nof))
a = fc(:)%la; call dumpd(trim(tag('fcla', 0)), a)
end subroutine dump_dt

subroutine dump_grad(k)
integer, intent(in) :: k
real(kind=dp), allocatable :: a(:,:), b(:,:)
integer(kind=i4) :: c
allocate(a(2*ngrad, noc), b(nvar, noc))
do c = 1, noc
   a(:, c) = reshape(cell(c)%grad, (/ 2*ngrad /))
   b(:, c) = cell(c)%phi(:)
enddo
call dumpd(trim(tag('grad', k)), reshape(a, (/ 2*ngrad*noc /)))
call dumpd(trim(tag('phi', k)), reshape(b, (/ nvar*noc /)))
end subroutine dump_grad

subroutine dump_stage(k)
integer, intent(in) :: k
real(kind=dp), allocatable :: a(:,:), g(:,:), gp(:,:)
integer(kind=i4) :: c
allocate(a(nvar, noc), g(nvar, nbf), gp(npvar, nbf))
do c = 1, noc
   a(:, c) = cell(c)%res(:)
enddo
do c = 1, nbf
   g(:, c) = cell(noc + c)%qc(:)
   gp(:, c) = cell(noc + c)%qp(:)
enddo
call dumpd(trim(tag('res', k)), reshape(a, (/ nvar*noc /)))
call dumpd(trim(tag('gqc', k)), reshape(g, (/ nvar*nbf /)))
call dumpd(trim(tag('gqp', k)), reshape(gp, (/ npvar*nbf /)))
call dumpd(trim(tag('force', k)), (/ Fx1, Fy1 /))
if (flow_type /= 'inviscid') then
   deallocate(a)
   allocate(a(2*ngrad, noc))
   do c = 1, noc
      a(:, c) = reshape(cell(c)%grad, (/ 2*ngrad /))
   enddo
   call dumpd(trim(tag('sgrad', k)), reshape(a, (/ 2*ngrad*noc /)))
   deallocate(a)
   allocate(a(nvar, noc))
   do c = 1, noc
      a(:, c) = cell(c)%phi(:)
   enddo
   call dumpd(trim(tag('sphi', k)), reshape(a, (/ nvar*noc /)))
endif
end subroutine dump_stage

subroutine dump_qc(k)
integer, intent(in) :: k
real(kind=dp), allocatable :: a(:,:)
integer(kind=i4) :: c
allocate(a(nvar, noc))
do c = 1, noc
   a(:, c) = cell(c)%qc(:)
enddo
call dumpd(trim(tag('qc', k)), reshape(a, (/ nvar*noc /)))
end subroutine dump_qc

subroutine dump_final()
real(kind=dp), allocatable :: a(:,:)
integer(kind=i4) :: c
allocate(a(nvar, size(cell)))
do c = 1, size(cell)
   a(:, c) = cell(c)%qc(:)
enddo
call dumpd('final_qc', reshape(a, (/ nvar*size(cell) /)))
deallocate(a)
allocate(a(npvar, size(cell)))
do c = 1, size(cell)
   a(:, c) = cell(c)%qp(:)
enddo
call dumpd('final_qp', reshape(a, (/ npvar*size(cell) /)))
deallocate(a)
allocate(a(nvar, noc))
do c = 1, noc
   a(:, c) = cell(c)%phi(:)
enddo
call dumpd('final_phi', reshape(a, (/ nvar*noc /)))
do c = 1, noc
   a(:, c) = cell(c)%res(:)
enddo
call dumpd('final_res', reshape(a, (/ nvar*noc /)))
end subroutine dump_final

end program ref_cc_driver
