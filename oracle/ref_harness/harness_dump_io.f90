! Dump writer of the parity harnesses (test infrastructure, not product).
!
! write_header / write_step write the FP64 state of the reference's module
! globals to a binary stream file (little endian, no record markers); the
! layout is read by oracle/decks.py:read_dump.  Used by the harness units
! time_step_loop_dump.f90 (reference hot path) and time_step_loop_hip_dump.f90
! (device hot path through the Fortran bridge).
module harness_dump_io
use Static_allocation_module
use Hybrid_allocation_module
use Dynamic_allocation_module
implicit none

contains

subroutine write_header(u)
integer(4),intent(in) :: u
integer(4) :: kf,nf,st,n,ic
write(u) 1213417555                                  ! 'SPHH'
#ifdef SPACE_2D
write(u) 2, nag, NMAXPARTJ, Grid%nmax, Grid%ncd(1:3)
#else
if (Domain%tipo=="bsph") then
   write(u) 3, nag, NMAXPARTJ, Grid%nmax, Grid%ncd(1:3)
   else
      write(u) 1, nag, NMAXPARTJ, Grid%nmax, Grid%ncd(1:3)
endif
#endif
write(u) Grid%extr(1:3,1), Grid%extr(1:3,2), Grid%dcd(1:3)
write(u) Domain%h, Domain%dx, Domain%coefke, Domain%coefkacl, Domain%prif,  &
   Domain%grav(1:3), eta, eta2, Domain%PVolume
write(u) Med(1)%den0, Med(1)%eps, Med(1)%celerita, Med(1)%alfaMon,          &
   Med(1)%kin_visc, Med(1)%mumx
write(u) input_any_t%CFL, input_any_t%vsc_coeff, pesodt, input_any_t%TetaP, &
   input_any_t%TetaV
ic = 0
if (dt_alfa_Mon) ic = 1
#ifdef SPACE_2D
! 2D: the boundary sides (DefineBoundarySideGeometry2D.f90) with their
! Tratto data and end-vertex coordinates
write(u) ic, MAXCLOSEBOUNDSIDES, NumBSides, 0
do kf=1,NumBSides
   st = BoundarySide(kf)%stretch
   write(u) st, BoundarySide(kf)%previous_side, BoundarySide(kf)%vertex(1:2), &
      Tratto(st)%tipo
   write(u) BoundarySide(kf)%length, BoundarySide(kf)%angle,                 &
      BoundarySide(kf)%T(1:3,1:3), BoundarySide(kf)%R(1:3,1:3),               &
      BoundarySide(kf)%RN(1:3,1:3), BoundarySide(kf)%velocity(1:3),           &
      Tratto(st)%velocity(1:3), Vertice(1:3,BoundarySide(kf)%vertex(1)),      &
      Vertice(1:3,BoundarySide(kf)%vertex(2))
   ic = 0
   if (Tratto(st)%laminar_no_slip_check) ic = 1
   write(u) Partz(Tratto(st)%zone)%slip_coefficient_mode,                   &
      Partz(Tratto(st)%zone)%BC_shear_stress_input, ic
enddo
#else
if (Domain%tipo=="bsph") then
! DB-SPH: the wall elements (DBSPH_IC_surface_elements.f90,
! semi_particle_volumes.f90) and the DBSPH section's parameters
   write(u) ic, NMAXPARTJ, DBSPH%n_w, 0
   write(u) DBSPH%dx_dxw, DBSPH%k_w
   write(u) DBSPH%slip_ID, merge(1,0,DBSPH%MUSCL_boundary_flag),              &
      merge(1,0,DBSPH%Gamma_limiter_flag),                                    &
      merge(1,0,DBSPH%negative_wall_p_allowed), merge(1,0,DBSPH%FS_allowed),  &
      DBSPH%n_inlet, DBSPH%n_outlet
   do kf=1,DBSPH%n_w
      write(u) pg_w(kf)%coord(1:3), pg_w(kf)%normal(1:3), pg_w(kf)%vel(1:3),  &
         pg_w(kf)%weight, pg_w(kf)%volume, pg_w(kf)%mass, pg_w(kf)%k_d
      write(u) pg_w(kf)%adjacent_faces(1:3), pg_w(kf)%cella
   enddo
   return
endif
write(u) ic, input_any_t%MAXCLOSEBOUNDFACES, NumFacce, NumBEdges
do kf=1,NumFacce
   nf = BFaceList(kf)
   write(u) nf
   if (nf==0) cycle
   st = BoundaryFace(nf)%stretch
   write(u) BoundaryFace(nf)%nodes, st, Tratto(st)%tipo
   write(u) BoundaryFace(nf)%area, BoundaryFace(nf)%T(1:3,1:3),              &
      BoundaryFace(nf)%velocity(1:3), Tratto(st)%velocity(1:3)
   do n=1,MAXFACENODES
      write(u) BoundaryFace(nf)%Node(n)%name, BoundaryFace(nf)%Node(n)%GX(1:3),&
         BoundaryFace(nf)%Node(n)%LX(1:3)
   enddo
   ic = 0
   if (Tratto(st)%laminar_no_slip_check) ic = 1
   write(u) Partz(Tratto(st)%zone)%slip_coefficient_mode,                   &
      Partz(Tratto(st)%zone)%BC_shear_stress_input, ic
enddo
write(u) GCBFPointers(1:Grid%nmax,1), GCBFPointers(1:Grid%nmax,2)
write(u) size(GCBFVector)
write(u) GCBFVector(:)
write(u) kerneltab(0:ktrows,0:ktcols), ktdelta
write(u) BoundIntegralTab(1:BITrows,1:BITcols)
#endif
end subroutine write_header

subroutine write_step(u,it,lists)
integer(4),intent(in) :: u,it,lists
integer(4) :: i,k,nc,ib
write(u) 1347704147, it, nag                         ! 'STEP'
write(u) simulation_time, dt, dt_average
write(u) (pg(i)%coord(1:3),i=1,nag)
write(u) (pg(i)%vel(1:3),i=1,nag)
write(u) (pg(i)%var(1:3),i=1,nag)
write(u) (pg(i)%acc(1:3),i=1,nag)
write(u) (pg(i)%dens,i=1,nag), (pg(i)%pres,i=1,nag), (pg(i)%dden,i=1,nag),  &
   (pg(i)%vpres,i=1,nag), (pg(i)%mass,i=1,nag)
write(u) (pg(i)%cella,i=1,nag), (pg(i)%kodvel,i=1,nag),                     &
   (pg(i)%koddens,i=1,nag), (pg(i)%laminar_flag,i=1,nag)
#ifdef SPACE_3D
if (Domain%tipo=="bsph") then
! DB-SPH per-particle and wall-element state
   write(u) (pg(i)%Gamma,i=1,nag), (pg(i)%sigma,i=1,nag),                     &
      (pg(i)%rhoSPH_new,i=1,nag), (pg(i)%dShep,i=1,nag)
   write(u) (pg(i)%drho(1:3),i=1,nag)
   write(u) (pg(i)%dvel(1:3,1:3),i=1,nag)
   write(u) (pg(i)%FS,i=1,nag)
   write(u) (pg_w(i)%dens,i=1,DBSPH%n_w), (pg_w(i)%pres,i=1,DBSPH%n_w)
   write(u) (pg_w(i)%wet,i=1,DBSPH%n_w)
endif
#endif
write(u) lists
if (lists==0) return
write(u) nPartIntorno(1:nag)
do i=1,nag
   write(u) PartIntorno((i-1)*NMAXPARTJ+1:(i-1)*NMAXPARTJ+nPartIntorno(i))
enddo
#ifdef SPACE_3D
if (Domain%tipo=="bsph") then
! the fluid-wall lists (CalcVarLength.f90:462-540)
   write(u) nPartIntorno_fw(1:nag)
   do i=1,nag
      write(u) PartIntorno_fw((i-1)*NMAXPARTJ+1:(i-1)*NMAXPARTJ+              &
         nPartIntorno_fw(i))
   enddo
   return
endif
#endif
#ifdef SPACE_2D
! 2D: Ncbs and IntNcbs per particle, then the IntNcbs records
! (ComputeBoundaryDataTab_2D.f90 fills IntNcbs records)
write(u) BoundaryDataPointer(1,1:nag), BoundaryDataPointer(2,1:nag)
do i=1,nag
   nc = BoundaryDataPointer(2,i)
#else
write(u) BoundaryDataPointer(1,1:nag)
do i=1,nag
   nc = BoundaryDataPointer(1,i)
#endif
   do k=1,nc
      ib = BoundaryDataPointer(3,i) + k - 1
      write(u) BoundaryDataTab(ib)%CloBoNum, BoundaryDataTab(ib)%LocXYZ(1:3),  &
         BoundaryDataTab(ib)%BoundaryIntegral(1:8),                          &
         BoundaryDataTab(ib)%IntGiWrRdV(1:3,1:3)
   enddo
enddo
end subroutine write_step

end module harness_dump_io
