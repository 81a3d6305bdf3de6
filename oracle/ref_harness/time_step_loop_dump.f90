! Parity harness for the SPHERA v10 reference (test infrastructure, not product).
!
! Project-written replacement for the reference's "time_step_loop"
! (src/Main_algorithm/time_step_loop.f90:25-510).  It is linked, in place of
! that unit, with the reference objects compiled by oracle/Makefile.ref, so
! Gest_Input/Gest_Trans (src/Main_algorithm/Gest_Trans.f90:769) set everything
! up from a deck exactly as the reference does, and this unit then runs the
! same per-step call sequence for the branch the dam-break decks use
! (Domain%tipo=="semi", time_split=1 leapfrog, no bodies, no KTGF; 3D, or 2D
! when compiled with -DSPACE_2D) and writes the FP64 state the hot path
! produces to a binary stream file.  With time_split=0 and RKscheme=1 it runs
! the explicit RK1 branch of time_step_loop.f90:175-489 instead, where the
! reference's time_integration.f90:54 calls the 3D Euler(BC_zmax_flag) with no
! argument (its BC_zmax_flag then reads an undefined address and the 3D
! program stops with a segmentation fault when TetaV > 0); this harness calls
! Euler with the loop's BC_zmax_flag (.false.), then EoS_fluid_loop, which is
! the rest of time_integration.  With Domain%tipo=="bsph" (DB-SPH walls, 3D,
! RK1) it runs the DB-SPH pre-loop of time_step_loop.f90:82-135 (CalcVarLength
! with on_going_time_step = -2, removing_DBSPH_fictitious_reservoirs,
! OrdGrid1, CalcVarLength with -1, Gradients_to_MUSCL and BC_wall_elements)
! and the DB-SPH stage (no stage-start CalcVarLength; momentum, continuity,
! Euler with its second OrdGrid1 and CalcVarLength, EoS_fluid_loop, the
! pressure smoothing, then Gradients_to_MUSCL and BC_wall_elements, :470-473).
!
! Environment:
!   SPH_HARNESS_OUT    output file (required)
!   SPH_HARNESS_STEPS  number of time steps to run (default input_any_t%itmax)
!   SPH_HARNESS_EVERY  dump every K steps (default 1); step 0 is the state
!                      after the pre-loop (first lists and boundary table)
!   SPH_HARNESS_IC     optional file: int32 n, then n*3 float64 coordinates
!                      (x,y,z per particle, pg order).  They replace the
!                      generated positions before the pre-loop (jittered ICs).
!   SPH_HARNESS_ICV    optional file: int32 n, then n*3 float64 velocities
!                      (pg order); they replace the initial vel and var
!                      (near-wall cases that approach a frontier at once).
!   SPH_HARNESS_LISTS  0 to omit neighbour lists / boundary tables from dumps
!   SPH_HARNESS_PELO   1: at every dumped step also run the reference's
!                      free-surface monitor (CalcVarp, then calc_pelo, as
!                      time_step_post_processing.f90:73,111 call them), which
!                      appends to the case's .plb and .zlft files
!
! File layout: harness_dump_io.f90 (write_header/write_step).
subroutine time_step_loop
use I_O_file_module
use Static_allocation_module
use Hybrid_allocation_module
use Dynamic_allocation_module
use Memory_I_O_interface_module
use harness_dump_io
implicit none
logical :: BC_zmax_flag
integer(4) :: it,npi,nsteps,every,u,ios,n_ic,i_zone,Ncbf_Max,lists,pelo
double precision :: dt_previous_step,dtvel
double precision,allocatable :: xyz(:)
character(len=512) :: fout,aux

dtvel = 0.d0
on_going_time_step = -999
EpOrdGrid = 0
#ifdef SPACE_2D
SourceSide = 0
#else
SourceFace = 0
#endif
it_eff = it_start
BC_zmax_flag = .false.
if (((Domain%tipo/="semi").and.(Domain%tipo/="bsph")).or.                   &
   ((Domain%time_split/=1).and.                                               &
   ((Domain%time_split/=0).or.(Domain%RKscheme/=1))).or.                      &
   ((Domain%tipo=="bsph").and.(Domain%time_split/=0))) then
   write(*,*) "harness: only tipo=semi with leapfrog (time_split=1) or RK1 ",   &
      "(time_split=0, RKscheme=1), or tipo=bsph with RK1, is handled"
   stop 2
endif
if (Domain%time_split==0) Domain%time_stage = 1
call get_environment_variable("SPH_HARNESS_OUT",fout)
if (len_trim(fout)==0) then
   write(*,*) "harness: SPH_HARNESS_OUT not set"
   stop 2
endif
nsteps = input_any_t%itmax
call get_environment_variable("SPH_HARNESS_STEPS",aux)
if (len_trim(aux)>0) read(aux,*) nsteps
every = 1
call get_environment_variable("SPH_HARNESS_EVERY",aux)
if (len_trim(aux)>0) read(aux,*) every
lists = 1
call get_environment_variable("SPH_HARNESS_LISTS",aux)
if (len_trim(aux)>0) read(aux,*) lists
pelo = 0
call get_environment_variable("SPH_HARNESS_PELO",aux)
if (len_trim(aux)>0) read(aux,*) pelo
! Optional replacement of the particle velocities (pg order)
call get_environment_variable("SPH_HARNESS_ICV",aux)
if (len_trim(aux)>0) then
   open(newunit=u,file=trim(aux),access="stream",form="unformatted",          &
      status="old",iostat=ios)
   if (ios/=0) stop 3
   read(u) n_ic
   if (n_ic/=nag) then
      write(*,*) "harness: ICV file has",n_ic," particles, deck has",nag
      stop 3
   endif
   allocate(xyz(3*nag))
   read(u) xyz
   close(u)
   do npi=1,nag
      pg(npi)%vel(1:3) = xyz(3*npi-2:3*npi)
      pg(npi)%var(1:3) = pg(npi)%vel(1:3)
   enddo
   deallocate(xyz)
endif
! Optional replacement of the particle positions (same pg order)
call get_environment_variable("SPH_HARNESS_IC",aux)
if (len_trim(aux)>0) then
   open(newunit=u,file=trim(aux),access="stream",form="unformatted",          &
      status="old",iostat=ios)
   if (ios/=0) stop 3
   read(u) n_ic
   if (n_ic/=nag) then
      write(*,*) "harness: IC file has",n_ic," particles, deck has",nag
      stop 3
   endif
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
#ifdef SPACE_3D
if (Domain%tipo=="bsph") then
! DB-SPH pre-loop (time_step_loop.f90:82-135)
   call liquid_particle_ID_array
   call inlet_sections
   on_going_time_step = -2
   call CalcVarLength
   on_going_time_step = it_start
   call removing_DBSPH_fictitious_reservoirs
   call OrdGrid1
   call liquid_particle_ID_array
   on_going_time_step = -1
   call CalcVarLength
   on_going_time_step = it_start
   if ((nag>0).and.(DBSPH%n_w>0)) then
      call Gradients_to_MUSCL
      call BC_wall_elements
   endif
   if (it_start==0) call time_step_duration
   it = it_start
   call write_step(u,it,lists)
   do while (it<nsteps)
      it = it + 1
      on_going_time_step = it
      it_eff = it
      dt_previous_step = dt
      if (nag>0) call time_step_duration
      simulation_time = simulation_time + dt
      call liquid_particle_ID_array
      call RHS_momentum_equation
!$omp parallel do default(none) shared(nag,pg) private(npi)
      do npi=1,nag
         pg(npi)%dden = 0.d0
         pg(npi)%dden_ALE12 = 0.d0
         pg(npi)%dden_ALE12_frozen = 0.d0
         pg(npi)%koddens = 0
         call Continuity_Equation(npi)
      enddo
!$omp end parallel do
      call Euler(BC_zmax_flag)
      call EoS_fluid_loop
      if (input_any_t%TetaP>zero) call PressureSmoothing_3D
      if ((nag>0).and.(DBSPH%n_w>0)) then
         call Gradients_to_MUSCL
         call BC_wall_elements
      endif
      dtvel = dt
      if (mod(it,every)==0) call write_step(u,it,lists)
      if (simulation_time>=input_any_t%tmax) exit
   enddo
   close(u)
   return
endif
#endif
! Pre-loop (time_step_loop.f90:82-173 for the semi branch)
call liquid_particle_ID_array
call inlet_sections
#ifdef SPACE_3D
call BC_zmax_t0
#endif
call CalcVarLength
#ifdef SPACE_2D
call ComputeBoundaryDataTab_2D
#else
call ComputeBoundaryDataTab_3D
#endif
if (Domain%NormFix) call NormFix
if (it_start==0) call time_step_duration
it = it_start
call write_step(u,it,lists)
if ((pelo==1).and.(nlines>0)) then
   call CalcVarp
   call calc_pelo
endif
do while (it<nsteps)
   it = it + 1
   on_going_time_step = it
   it_eff = it
   dt_previous_step = dt
   if (nag>0) call time_step_duration
   simulation_time = simulation_time + dt
   if (Domain%time_split==0) then
! Explicit RK1 (time_step_loop.f90:190-489 with time_stage = RKscheme = 1)
      call CalcVarLength
#ifdef SPACE_2D
      call ComputeBoundaryDataTab_2D
#else
      call ComputeBoundaryDataTab_3D
#endif
      call fluid_particle_imposed_kinematics
      call liquid_particle_ID_array
      call RHS_momentum_equation
#ifdef SPACE_3D
      Ncbf_Max = 0
!$omp parallel do default(none) shared(nag,pg,Domain,Ncbf_Max) private(npi)
#else
!$omp parallel do default(none) shared(nag,pg,Domain) private(npi)
#endif
      do npi=1,nag
         pg(npi)%dden = 0.d0
         pg(npi)%dden_ALE12 = 0.d0
         pg(npi)%dden_ALE12_frozen = 0.d0
         pg(npi)%koddens = 0
#ifdef SPACE_3D
         if (pg(npi)%state=="flu") call SASPH_continuity(npi,Ncbf_Max)
#else
         if (pg(npi)%state=="flu") call SASPH_continuity(npi)
#endif
         call Continuity_Equation(npi)
      enddo
!$omp end parallel do
#ifdef SPACE_3D
      if (Ncbf_Max>input_any_t%MAXCLOSEBOUNDFACES) then
         write(*,*) "harness: MAXCLOSEBOUNDFACES exceeded",Ncbf_Max
         stop 4
      endif
      call Euler(BC_zmax_flag)
#else
      call Euler
#endif
      call EoS_fluid_loop
#ifdef SPACE_2D
      if (input_any_t%TetaP>zero) call PressureSmoothing_2D
      if (NumOpenSides>0) call CancelOutgoneParticles_2D
#else
      if (input_any_t%TetaP>zero) call PressureSmoothing_3D
      if (NumOpenFaces>0) call CancelOutgoneParticles_3D
      call BC_zmax_anyt
#endif
      call OrdGrid1
      call liquid_particle_ID_array
      dtvel = dt
      if (mod(it,every)==0) call write_step(u,it,lists)
      if (simulation_time>=input_any_t%tmax) exit
      cycle
   endif
   call fluid_particle_imposed_kinematics
   call RHS_momentum_equation
   call Leapfrog_momentum(dt_previous_step,dtvel)
#ifdef SPACE_2D
   if (input_any_t%TetaV>1.d-9) call velocity_smoothing
   call velocity_smoothing_2
   call Leapfrog_trajectories
   if (NumOpenSides>0) call CancelOutgoneParticles_2D
   if (SourceSide/=0) then
      call inlet_sections
      call GenerateSourceParticles
   endif
#else
   if (input_any_t%TetaV>1.d-9) call velocity_smoothing(BC_zmax_flag)
   call velocity_smoothing_2(BC_zmax_flag)
   call Leapfrog_trajectories
   if (NumOpenFaces>0) call CancelOutgoneParticles_3D
   if (SourceFace/=0) then
      call inlet_sections
      call GenerateSourceParticles
   endif
   call BC_zmax_anyt
#endif
   call OrdGrid1
   if (Domain%NormFix) call NormFix
   call CalcVarLength
#ifdef SPACE_2D
   call ComputeBoundaryDataTab_2D
   call liquid_particle_ID_array
!$omp parallel do default(none) shared(nag,pg,Domain) private(npi)
   do npi=1,nag
      pg(npi)%dden = 0.d0
      pg(npi)%dden_ALE12 = 0.d0
      pg(npi)%dden_ALE12_frozen = 0.d0
      pg(npi)%koddens = 0
      if (pg(npi)%state=="flu") call SASPH_continuity(npi)
      call Continuity_Equation(npi)
   enddo
!$omp end parallel do
#else
   do i_zone=1,NPartZone
      if (Partz(i_zone)%tipo=="zmax") then
         BC_zmax_flag = .true.
         exit
      endif
   enddo
   if (BC_zmax_flag) then
      call liquid_particle_ID_array
      call velocity_smoothing(BC_zmax_flag)
      call velocity_smoothing_2(BC_zmax_flag)
   endif
   BC_zmax_flag = .false.
   call ComputeBoundaryDataTab_3D
   call liquid_particle_ID_array
   Ncbf_Max = 0
!$omp parallel do default(none) shared(nag,pg,Domain,Ncbf_Max) private(npi)
   do npi=1,nag
      pg(npi)%dden = 0.d0
      pg(npi)%dden_ALE12 = 0.d0
      pg(npi)%dden_ALE12_frozen = 0.d0
      pg(npi)%koddens = 0
      if (pg(npi)%state=="flu") call SASPH_continuity(npi,Ncbf_Max)
      call Continuity_Equation(npi)
   enddo
!$omp end parallel do
   if (Ncbf_Max>input_any_t%MAXCLOSEBOUNDFACES) then
      write(*,*) "harness: MAXCLOSEBOUNDFACES exceeded",Ncbf_Max
      stop 4
   endif
#endif
   call Leapfrog_continuity
   call EoS_fluid_loop
#ifdef SPACE_2D
   if (input_any_t%TetaP>zero) call PressureSmoothing_2D
#else
   if (input_any_t%TetaP>zero) call PressureSmoothing_3D
#endif
   if (mod(it,every)==0) then
      call write_step(u,it,lists)
      if ((pelo==1).and.(nlines>0)) then
         call CalcVarp
         call calc_pelo
      endif
   endif
   if (simulation_time>=input_any_t%tmax) exit
enddo
close(u)
return

end subroutine time_step_loop
