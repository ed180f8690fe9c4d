!------------------------------------------------------------------------------
! Oracle harness for the vertex-centred (median-dual) reference solver (test
! infrastructure).
!
! Links every object of the reference NCFVM solver except its main program
! and fvresidual (NCFVM/main.f90, NCFVM/fvresidual.f90: both hold unresolved
! merge markers and do not compile as shipped).  This file restates
!   * the iteration loop of NCFVM/main.f90:51-93 (tanh CFL ramp, save_old,
!     RK3 stages with the inline update, or lusgs; residue), and
!   * fvresidual, NCFVM/fvresidual.f90:4-85, with the gradient of either side
!     of the merge conflict at :11-17 (GRAD=gg: Gradient_GG, the side the
!     reference's history ends on; GRAD=lsqr: Gradient_LSQR),
! so that every other routine (geometric, save_old, Gradient_GG/LSQR,
! limiter_VKN, vanleer_flux, solid_flux, farfield_flux1, time_step2, lusgs,
! residue) runs unmodified, and dumps the state in full precision.
!
! NCFVM/geometric.f90:101-104 reads newly allocated arrays before writing
! them, and alloc_int_ptr (:906-951) tests the association of pointer
! components nothing nullified; run the harness with MALLOC_MMAP_THRESHOLD_=0
! so every allocation is fresh zeroed memory, which gives both the intended
! meaning (zero and not associated).
!
! Usage (in the directory holding geometry.inp):
!   ref_nc_driver DECK OUTDIR NITER DUMPITERS GRAD[:bj] [TIMEWARM]
! (":bj" calls the reference's limiter_min_max, NCFVM/limiter.f90:192-260,
! which its own fvresidual never calls, in place of LinearReconLimiter)
!------------------------------------------------------------------------------
module nc_harness
use data_type
use grid
use param, only: iter
implicit none
integer :: grad_kind = 1          ! 1 = Gradient_GG, 2 = Gradient_LSQR
integer :: lim_kind = 0           ! 0 = LinearReconLimiter (limiter_VKN), 1 = limiter_min_max
integer :: ndump = 0, stage = 0
character(len=512) :: outdir = '.'
contains
subroutine dumpd(name, a)
implicit none
character(len=*), intent(in) :: name
real(kind=dp), intent(in) :: a(:)
integer :: u
open(newunit=u, file=trim(outdir)//'/'//name//'.f64', access='stream', &
     form='unformatted', status='replace')
write(u) a
close(u)
end subroutine dumpd

subroutine dumpi(name, a)
implicit none
character(len=*), intent(in) :: name
integer(kind=i4), intent(in) :: a(:)
integer :: u
open(newunit=u, file=trim(outdir)//'/'//name//'.i32', access='stream', &
     form='unformatted', status='replace')
write(u) a
close(u)
end subroutine dumpi

function tag(prefix, k) result(s)
implicit none
character(len=*), intent(in) :: prefix
integer, intent(in) :: k
character(len=64) :: s
write(s, '(a,"_it",i4.4,"_",i1)') prefix, iter, k
end function tag

subroutine dump_grad(name)
use grid
implicit none
character(len=*), intent(in) :: name
real(kind=dp), allocatable :: a(:,:)
integer(kind=i4) :: c
allocate(a(2*nvar, nop))
do c = 1, nop
   a(:, c) = reshape(pt(c)%grad, (/ 2*nvar /))
enddo
call dumpd(trim(tag(name, stage)), reshape(a, (/ 2*nvar*nop /)))
end subroutine dump_grad

subroutine dump_phi()
use grid
implicit none
real(kind=dp), allocatable :: a(:,:)
integer(kind=i4) :: c
allocate(a(3*nvar, nop))
do c = 1, nop
   a(1:nvar, c) = pt(c)%phi(:)
   a(nvar+1:2*nvar, c) = pt(c)%dumin(:)
   a(2*nvar+1:3*nvar, c) = pt(c)%dumax(:)
enddo
call dumpd(trim(tag('phi', stage)), reshape(a, (/ 3*nvar*nop /)))
end subroutine dump_phi

end module nc_harness

! NCFVM/fvresidual.f90:4-85 (Recon = 1 branch; the merge conflict at :11-17
! resolved by grad_kind)
subroutine fvresidual
use grid
use param, only: iter
use nc_harness
implicit none
integer(kind=i4) :: i, j, ie, p1, p2
real(kind=dp) :: left(nvar), right(nvar), dist(ndim), grad(nvar)

if (grad_kind == 2) then
   call Gradient_LSQR
else
   call Gradient_GG
endif
if (iter <= ndump) call dump_grad('grad')

do i = 1, nop
   do j = 1, nvar
      pt(i)%res(j) = 0.0d0
   enddo
enddo
pt(:)%la = 0.d0
fc(:)%la = 0.d0

if (lim_kind == 1) then
   call limiter_min_max
else
   call LinearReconLimiter
endif
if (iter <= ndump) call dump_phi()

do ie = 1, nof
   p1 = fc(ie)%pt(1)
   p2 = fc(ie)%pt(2)
   dist(1) = pt(p2)%x - pt(p1)%x
   dist(2) = pt(p2)%y - pt(p1)%y

   grad(:) = 0.d0
   do i = 1, ndim
      grad(:) = grad(:) + pt(p1)%grad(i,:)*dist(i)
   enddo
   left(:) = pt(p1)%qp(:) + 0.5d0*grad(:)*pt(p1)%phi(:)

   grad(:) = 0.d0
   do i = 1, ndim
      grad(:) = grad(:) + pt(p2)%grad(i,:)*dist(i)
   enddo
   right(:) = pt(p2)%qp(:) - 0.5d0*grad(:)*pt(p2)%phi(:)

   call vanleer_flux(ie, p1, p2, left, right)
   if (fc(ie)%bc == 1001) call solid_flux(ie, p1, p2)
   if (fc(ie)%bc == 2001) call farfield_flux1(ie, p1, p2)
enddo

call time_step2
end subroutine fvresidual

program ref_nc_driver
use param
use pri
use grid
use inf
use nc_harness
implicit none
integer(kind=i4) :: i, j, irk, niter, narg, twarm
real(kind=dp)    :: c1(nvar), c2(nvar), c3(nvar), dtbyarea
character(len=512) :: arg
integer(kind=8)  :: t0, t1, rate
logical          :: timing
real(kind=dp), allocatable :: hist(:,:)

narg = command_argument_count()
if (narg < 5) then
   print*, 'usage: ref_nc_driver DECK OUTDIR NITER DUMPITERS GRAD [TIMEWARM]'
   stop 2
endif
call get_command_argument(2, outdir)
call get_command_argument(3, arg); read(arg,*) niter
call get_command_argument(4, arg); read(arg,*) ndump
call get_command_argument(5, arg)
if (index(arg, ':bj') > 0) then
   lim_kind = 1
   arg = arg(1:index(arg, ':bj') - 1)
endif
if (trim(arg) == 'lsqr') then
   grad_kind = 2
elseif (trim(arg) == 'gg') then
   grad_kind = 1
else
   stop 'GRAD must be gg or lsqr, optionally followed by :bj'
endif
timing = .false.
twarm = 0
if (narg >= 6) then
   call get_command_argument(6, arg); read(arg,*) twarm
   timing = .true.
   ndump = 0
endif

call math
call read_input
call read_grid
call geometric
call initialize

! NCFVM/main.f90:31-44
if (timemode == "rk3" .and. cfl_max > 1.0d0) then
   stop 'CFL should be less than 1 for rk3'
elseif (timemode == "rk3" .and. cfl_max < 1.0d0) then
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
   ! NCFVM/main.f90:54
   cfl = 0.5d0*cfl_max*(1+dtanh(iter*7.d0/500.0 - 3.5))
   call save_old
   if (iter <= ndump) call dump_pt4('qp', 0, 1)

   if (timemode == 'rk3') then
      do irk = 1, nirk
         stage = irk
         call fvresidual
         if (iter <= ndump) call dump_stage(irk)
         do i = 1, nop
            c1(:) = pt(i)%qold(:)
            c2(:) = pt(i)%qc(:)
            dtbyarea = pt(i)%dt/pt(i)%cv
            do j = 1, nvar
               c3(j) = airk(irk)*c1(j) + birk(irk)*(c2(j) - dtbyarea*pt(i)%res(j))
            enddo
            pt(i)%qc(:) = c3(:)
         enddo
         if (iter <= ndump) call dump_pt4('qc', irk, 2)
      enddo
   elseif (timemode == 'lusgs') then
      stage = 1
      call lusgs
      if (iter <= ndump) then
         call dump_stage(1)
         call dump_pt4('qc', 1, 2)
      endif
   endif

   call residue
   hist(1, iter) = real(iter, dp)
   hist(2, iter) = fres
   hist(3, iter) = fres1
   hist(4, iter) = cfl
   hist(5, iter) = fresi
   hist(6, iter) = real(iresi, dp)
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
   call dump_pt4('final_qc', 0, 2)
endif

contains

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
deallocate(ib, db)
allocate(db(nop), ib(nop))
db = pt(:)%x;  call dumpd('pt_x', db)
db = pt(:)%y;  call dumpd('pt_y', db)
db = pt(:)%cv; call dumpd('pt_cv', db)
ib = pt(:)%nv2v; call dumpi('pt_nv2v', ib)
ib = pt(:)%nv2f; call dumpi('pt_nv2f', ib)
deallocate(db, ib)
tot = sum(pt(:)%nv2v)
allocate(ib(tot), db(2*tot))
m = 0
do k = 1, nop
   ib(m+1:m+pt(k)%nv2v) = pt(k)%v2v(1:pt(k)%nv2v)
   db(m+1:m+pt(k)%nv2v) = pt(k)%wx(1:pt(k)%nv2v)
   db(tot+m+1:tot+m+pt(k)%nv2v) = pt(k)%wy(1:pt(k)%nv2v)
   m = m + pt(k)%nv2v
enddo
call dumpi('pt_v2v', ib)
call dumpd('pt_wxy', db)
deallocate(ib, db)
tot = sum(pt(:)%nv2f)
allocate(ib(tot))
m = 0
do k = 1, nop
   ib(m+1:m+pt(k)%nv2f) = pt(k)%v2f(1:pt(k)%nv2f)
   m = m + pt(k)%nv2f
enddo
call dumpi('pt_v2f', ib)
deallocate(ib)
allocate(ib(3))
ib = (/ nop, noc, nof /)
call dumpi('counts', ib)
end subroutine dump_mesh

subroutine dump_pt4(name, k, what)
character(len=*), intent(in) :: name
integer, intent(in) :: k, what
real(kind=dp), allocatable :: a(:,:)
integer(kind=i4) :: c
allocate(a(nvar, nop))
do c = 1, nop
   if (what == 1) then
      a(:, c) = pt(c)%qp(:)
   else
      a(:, c) = pt(c)%qc(:)
   endif
enddo
if (name == 'final_qc') then
   call dumpd('final_qc', reshape(a, (/ nvar*nop /)))
else
   call dumpd(trim(tag(name, k)), reshape(a, (/ nvar*nop /)))
endif
end subroutine dump_pt4

subroutine dump_stage(k)
integer, intent(in) :: k
real(kind=dp), allocatable :: a(:,:), b(:)
integer(kind=i4) :: c
allocate(a(nvar, nop), b(nop))
do c = 1, nop
   a(:, c) = pt(c)%res(:)
enddo
call dumpd(trim(tag('res', k)), reshape(a, (/ nvar*nop /)))
b = pt(:)%la; call dumpd(trim(tag('la', k)), b)
b = pt(:)%dt; call dumpd(trim(tag('dt', k)), b)
deallocate(b)
allocate(b(nof))
b = fc(:)%la; call dumpd(trim(tag('fcla', k)), b)
end subroutine dump_stage

end program ref_nc_driver
