! Drop-in harness of the device hot path (test infrastructure, not product).
!
! Project-written replacement for the reference's "time_step_loop"
! (src/Main_algorithm/time_step_loop.f90:25-510), linked with the reference
! objects compiled by oracle/Makefile.ref and with the device library.  The
! reference's own Gest_Input/Gest_Trans set the case up from a deck; this unit
! then hands the state to the device through the Fortran bridge
! (sphera_amd/fortran/sphera_hip_bridge.f90), runs the pre-loop and the time
! steps there, and writes the state with the same dump layout as
! time_step_loop_dump.f90, so the reference program itself drives the device
! and its output can be compared with the reference's.
!
! Environment:
!   SPH_HARNESS_OUT    output file (required)
!   SPH_HARNESS_STEPS  number of time steps (default input_any_t%itmax)
!   SPH_HARNESS_EVERY  dump every K steps (default 1)
!   SPH_HARNESS_IC     optional int32 n + n*3 float64 positions (pg order)
!   SPH_HARNESS_ICV    optional int32 n + n*3 float64 velocities (pg order)
!   SPH_HIP_PRECISION  0 FP64 (default), 1 FP32-mixed
!   SPH_HIP_PHASES     1: call the per-phase entry points in the reference's
!                      order instead of sph_step
!   SPH_HARNESS_PELO   1: at every dumped step also run the free-surface
!                      monitor on the device (sph_hip_calc_pelo), which
!                      appends to the case's .plb and .zlft files as the
!                      reference's calc_pelo does
subroutine time_step_loop
use iso_c_binding
use I_O_file_module
use Static_allocation_module
use Hybrid_allocation_module
use Dynamic_allocation_module
use Memory_I_O_interface_module
use sphera_hip_c
use sphera_hip_bridge
use harness_dump_io
implicit none
integer(4) :: it,npi,nsteps,every,u,ios,n_ic,k,phases,pelo
integer(c_int32_t) :: prec
double precision,allocatable :: xyz(:)
character(len=512) :: fout,aux
type(c_ptr) :: ctx

on_going_time_step = -999
it_eff = it_start
if (it_start/=0) then
   write(*,*) "hip harness: restarts are not handled"
   stop 2
endif
call get_environment_variable("SPH_HARNESS_OUT",fout)
if (len_trim(fout)==0) then
   write(*,*) "hip harness: SPH_HARNESS_OUT not set"
   stop 2
endif
nsteps = input_any_t%itmax
call get_environment_variable("SPH_HARNESS_STEPS",aux)
if (len_trim(aux)>0) read(aux,*) nsteps
every = 1
call get_environment_variable("SPH_HARNESS_EVERY",aux)
if (len_trim(aux)>0) read(aux,*) every
prec = SPH_FP64
call get_environment_variable("SPH_HIP_PRECISION",aux)
if (len_trim(aux)>0) read(aux,*) prec
phases = 0
call get_environment_variable("SPH_HIP_PHASES",aux)
if (len_trim(aux)>0) read(aux,*) phases
pelo = 0
call get_environment_variable("SPH_HARNESS_PELO",aux)
if (len_trim(aux)>0) read(aux,*) pelo
call get_environment_variable("SPH_HARNESS_ICV",aux)
if (len_trim(aux)>0) then
   open(newunit=u,file=trim(aux),access="stream",form="unformatted",          &
      status="old",iostat=ios)
   if (ios/=0) stop 3
   read(u) n_ic
   if (n_ic/=nag) stop 3
   allocate(xyz(3*nag))
   read(u) xyz
   close(u)
   do npi=1,nag
      pg(npi)%vel(1:3) = xyz(3*npi-2:3*npi)
      pg(npi)%var(1:3) = pg(npi)%vel(1:3)
   enddo
   deallocate(xyz)
endif
call get_environment_variable("SPH_HARNESS_IC",aux)
if (len_trim(aux)>0) then
   open(newunit=u,file=trim(aux),access="stream",form="unformatted",          &
      status="old",iostat=ios)
   if (ios/=0) stop 3
   read(u) n_ic
   if (n_ic/=nag) stop 3
   allocate(xyz(3*nag))
   read(u) xyz
   close(u)
   do npi=1,nag
      pg(npi)%coord(1:3) = xyz(3*npi-2:3*npi)
      pg(npi)%CoordOld(1:3) = pg(npi)%coord(1:3)
   enddo
   deallocate(xyz)
   call OrdGrid1
endif
open(newunit=u,file=trim(fout),access="stream",form="unformatted",            &
   status="replace")
call write_header(u)
! Pre-loop (time_step_loop.f90:82-173): lists, boundary table, first dt
call sph_hip_open(prec,ctx)
call sph_hip_check(sph_preloop(ctx),"sph_preloop")
call sph_hip_pull(ctx)
if (pelo==1) call sph_hip_monitor_lines(ctx)
it = 0
call write_step(u,it,0)
if (pelo==1) call sph_hip_calc_pelo(ctx)
do while (it<nsteps)
   k = min(every,nsteps-it)
   if (phases==1) then
      do npi=1,k
         call one_step_by_phases
      enddo
      else
         call sph_hip_check(sph_step(ctx,int(k,c_int32_t)),"sph_step")
   endif
   it = it + k
   call sph_hip_pull(ctx)
   call write_step(u,it,0)
   if (pelo==1) call sph_hip_calc_pelo(ctx)
enddo
close(u)
call sph_hip_close(ctx)
return

contains

! One time step through the per-phase entry points, in the order of
! time_step_loop.f90:175-489 (semi, leapfrog).
subroutine one_step_by_phases
call sph_hip_check(sph_dt(ctx),"sph_dt")
call sph_hip_check(sph_momentum(ctx),"sph_momentum")
call sph_hip_check(sph_leapfrog_mom(ctx),"sph_leapfrog_mom")
call sph_hip_check(sph_velocity_smoothing(ctx),"sph_velocity_smoothing")
call sph_hip_check(sph_leapfrog_traj(ctx),"sph_leapfrog_traj")
call sph_hip_check(sph_sort(ctx),"sph_sort")
call sph_hip_check(sph_neighbours(ctx),"sph_neighbours")
call sph_hip_check(sph_boundary_tab(ctx),"sph_boundary_tab")
call sph_hip_check(sph_continuity(ctx),"sph_continuity")
call sph_hip_check(sph_leapfrog_cont(ctx),"sph_leapfrog_cont")
call sph_hip_check(sph_eos(ctx),"sph_eos")
call sph_hip_check(sph_pressure_smoothing(ctx),"sph_pressure_smoothing")
end subroutine one_step_by_phases

end subroutine time_step_loop
