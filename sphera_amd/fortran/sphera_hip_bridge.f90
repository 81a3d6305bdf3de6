! Bridge between SPHERA v10's module globals and the C ABI of the MI355X hot
! path (include/sphera_hip.h, interfaces in sphera_hip_c.f90).
!
! It is compiled against the reference's own modules and replaces, inside
! time_step_loop (src/Main_algorithm/time_step_loop.f90:175-489), the calls of
! the hot-path subroutines by device calls:
!
!   sph_hip_open(prec,ctx)  packs Domain, Grid, Med(1), input_any_t and the
!                           SA-SPH frontier tables (BoundaryFace, BFaceList,
!                           Tratto, Partz, GCBFPointers, GCBFVector, kerneltab,
!                           BoundIntegralTab; in the -DSPACE_2D build the
!                           boundary sides, BoundarySide, and
!                           BoundIntegralTab2D) and uploads pg(1:nag)
!   sph_hip_pull(ctx)       device state -> pg(1:nag), nag, dt, dt_average,
!                           simulation_time (at output steps)
!   sph_hip_check(rc,where) stops like diagnostic() on a non-zero status
!   sph_hip_monitor_lines(ctx)  control_lines / control_points -> device
!   sph_hip_calc_pelo(ctx)  calc_pelo on the device: writes the same .plb and
!                           .zlft records (calc_pelo.f90:132-133)
!
! Scope the device implements (checked by sph_hip_open): Domain%tipo="semi",
! leapfrog (time_split=1) or explicit RK1, one medium, no NormFix, no source
! faces or sides, open faces (3D) and open sides (2D) of type "open" only, no
! dt_alfa_Mon, fluid particles only; or Domain%tipo="bsph" (3D DB-SPH walls,
! pg_w passed to sph_dbsph_walls) with explicit RK1, FP64, slip_ID = 0, no
! DB-SPH inlets, outlets or fictitious reservoirs; kinematics records for one
! mesh.
module sphera_hip_bridge
use iso_c_binding
use sphera_hip_c
use Static_allocation_module
use Hybrid_allocation_module
use Dynamic_allocation_module
#ifdef SPACE_2D
use SA_SPH_module, only: BoundIntegralTab2D
#endif
implicit none
private
public :: sph_hip_open, sph_hip_pull, sph_hip_check, sph_hip_close
public :: sph_hip_monitor_lines, sph_hip_calc_pelo

! Host copies of the frontier tables in the C layout; they only need to live
! for the sph_ctx_create call, but are kept so a caller may inspect them.
real(c_double),allocatable,target,save :: fT(:), fLX(:), fGX(:), farea(:), fvel(:)
real(c_double),allocatable,target,save :: ftvel(:), fslip(:), fkt(:), fbit(:)
integer(c_int32_t),allocatable,target,save :: fnodes(:), fbtype(:), fmode(:), flam(:)
integer(c_int32_t),allocatable,target,save :: gcount(:), gstart(:), gvec(:)
! convex frontier edges (BoundaryConvexEdge, in the reference's order)
real(c_double),allocatable,target,save :: egx(:), ecomp(:), elen(:)
integer(c_int32_t),allocatable,target,save :: eface(:)
! 2D: boundary sides
real(c_double),allocatable,target,save :: sRN(:), sv1(:), sv2(:), slen(:), sang(:)
! DB-SPH: wall elements (pg_w)
real(c_double),allocatable,target,save :: wx(:), wn(:), wv(:), ww(:), wvol(:), wm(:)
integer(c_int32_t),allocatable,target,save :: sprev(:)

contains

subroutine sph_hip_check(rc,where)
integer(c_int),intent(in) :: rc
character(len=*),intent(in) :: where
if (rc==0) return
write(*,'(a,a,a,i5,a,a)') "sphera_hip: ",where," failed, status ",rc,": ",         &
   sph_error_message()
stop 1
end subroutine sph_hip_check

#ifdef SPACE_3D
! True when the DB-SPH case uses what the device does not build: the VSL
! wall model (slip_ID > 0), inlet or outlet sections, kinematics records
! with more than one mesh, or fictitious reservoirs / reservoirs
! from topography (removed in the pre-loop by
! removing_DBSPH_fictitious_reservoirs.f90)
logical function dbsph_outside_scope()
integer(4) :: k
dbsph_outside_scope = (DBSPH%slip_ID/=0).or.                                  &
   (DBSPH%n_inlet/=0).or.(DBSPH%n_outlet/=0).or.(Domain%time_split/=0).or.      &
   (Domain%RKscheme/=1)
if (allocated(DBSPH%n_kinematics_records)) then
   if ((DBSPH%surface_mesh_files>1).and.any(DBSPH%n_kinematics_records/=0))  &
      dbsph_outside_scope = .true.
endif
do k=1,NPartZone
   if (Partz(k)%DBSPH_fictitious_reservoir_flag.or.(Partz(k)%IC_source_type==2))&
      dbsph_outside_scope = .true.
enddo
end function dbsph_outside_scope

! True when a face of OpenFace (Gest_Input.f90:617-629) is of type "leve",
! "crit", "velo" or "flow": only "open" ones are on the device.
logical function other_open_faces()
integer(4) :: k
other_open_faces = .false.
do k=1,NumOpenFaces
   if (Tratto(BoundaryFace(OpenFace(k))%stretch)%tipo/="open") other_open_faces = .true.
enddo
end function other_open_faces
#else
! True when a side of OpenSide (Gest_Input.f90:600-613) is of type "leve",
! "crit", "velo" or "flow": only "open" ones are on the device.
logical function other_open_sides()
integer(4) :: k
other_open_sides = .false.
do k=1,NumOpenSides
   if (BoundarySide(OpenSide(k))%tipo/="open") other_open_sides = .true.
enddo
end function other_open_sides
#endif

integer(c_int32_t) function btype_code(tipo)
character(len=4),intent(in) :: tipo
select case (tipo)
   case ("fixe")
      btype_code = 0
   case ("tapi")
      btype_code = 1
   case ("open")
      btype_code = 2
   case ("sour")
      btype_code = 3
   case ("velo")
      btype_code = 4
   case ("flow")
      btype_code = 5
   case default
      btype_code = 6
end select
end function btype_code

subroutine pack_params(prec,p)
integer(c_int32_t),intent(in) :: prec
type(sph_params_t),intent(out) :: p
p%precision = prec
p%capacity = max(nag,1)
p%nmaxpartj = NMAXPARTJ
#ifdef SPACE_2D
p%maxclose = MAXCLOSEBOUNDSIDES
#else
p%maxclose = input_any_t%MAXCLOSEBOUNDFACES
#endif
p%ncd(1:3) = Grid%ncd(1:3)
p%extr_lo(1:3) = Grid%extr(1:3,1)
p%dcd(1:3) = Grid%dcd(1:3)
p%dx = Domain%dx
p%h = Domain%h
p%coefke = Domain%coefke
p%coefkacl = Domain%coefkacl
p%eta = eta
p%eta2 = eta2
p%grav(1:3) = Domain%grav(1:3)
p%prif = Domain%prif
p%den0 = Med(1)%den0
p%eps = Med(1)%eps
p%celerita = Med(1)%celerita
p%alfamon = Med(1)%alfaMon
p%kin_visc = Med(1)%kin_visc
p%cfl = input_any_t%CFL
p%vsc_coeff = input_any_t%vsc_coeff
p%pesodt = pesodt
p%tetap = input_any_t%TetaP
p%tetav = input_any_t%TetaV
p%semi = 0
if (Domain%tipo=="semi") p%semi = 1
p%wg_size = 0
p%chunk = 0
p%chunk_v = 0
p%chunk_c = 0
p%chunk_p = 0
p%nb_shape = 0
! Deck options the hot path branches on (SPH_OPT_* of sphera_hip.h): the
! density clamp, RK1 and convex edges are built; the device refuses ALE3,
! C1, KTGF and RK2 with status 95
p%options = 0
if (input_any_t%density_thresholds==1) p%options = ior(p%options,1)
if (input_any_t%ALE3) p%options = ior(p%options,2)
if (input_any_t%C1_BE.or.input_any_t%C1_divu) p%options = ior(p%options,4)
if (Granular_flows_options%KTGF_config>0) p%options = ior(p%options,8)
#ifdef SPACE_3D
if (NumBEdges>0) p%options = ior(p%options,16)
#endif
! explicit RK (time_split=0): RK1 is built, RK2 (Heun) is refused by the device
if ((Domain%time_split==0).and.(Domain%RKscheme==1)) p%options = ior(p%options,32)
if ((Domain%time_split==0).and.(Domain%RKscheme/=1)) p%options = ior(p%options,64)
end subroutine pack_params

#ifdef SPACE_2D
! Boundary sides in BoundarySide order (DefineBoundarySideGeometry2D.f90),
! with their Tratto data and BoundIntegralTab2D.
subroutine pack_sides(sd)
type(sph_sides2d_t),intent(out) :: sd
integer(4) :: k,ns,st,r,c
ns = NumBSides
allocate(fT(9*max(ns,1)),sRN(9*max(ns,1)),sv1(3*max(ns,1)),sv2(3*max(ns,1)),slen(max(ns,1)))
allocate(sang(max(ns,1)))
allocate(fvel(3*max(ns,1)),ftvel(3*max(ns,1)),fslip(max(ns,1)),fbtype(max(ns,1)))
allocate(sprev(max(ns,1)),fmode(max(ns,1)),flam(max(ns,1)),fbit(22*6))
do k=0,ns-1
   st = BoundarySide(k+1)%stretch
   do c=1,3
      do r=1,3
         fT(9*k+3*(c-1)+r) = BoundarySide(k+1)%T(r,c)
         sRN(9*k+3*(c-1)+r) = BoundarySide(k+1)%RN(r,c)
      enddo
   enddo
   sv1(3*k+1:3*k+3) = Vertice(1:3,BoundarySide(k+1)%vertex(1))
   sv2(3*k+1:3*k+3) = Vertice(1:3,BoundarySide(k+1)%vertex(2))
   slen(k+1) = BoundarySide(k+1)%length
   sang(k+1) = BoundarySide(k+1)%angle
   sprev(k+1) = BoundarySide(k+1)%previous_side
   fvel(3*k+1:3*k+3) = BoundarySide(k+1)%velocity(1:3)
   ftvel(3*k+1:3*k+3) = Tratto(st)%velocity(1:3)
   fbtype(k+1) = btype_code(Tratto(st)%tipo)
   fmode(k+1) = Partz(Tratto(st)%zone)%slip_coefficient_mode
   fslip(k+1) = Partz(Tratto(st)%zone)%BC_shear_stress_input
   flam(k+1) = 0
   if (Tratto(st)%laminar_no_slip_check) flam(k+1) = 1
enddo
do r=1,22
   do c=0,5
      fbit(6*(r-1)+c+1) = BoundIntegralTab2D(r,c)
   enddo
enddo
sd%ns = ns
sd%T = c_loc(fT)
sd%RN = c_loc(sRN)
sd%v1 = c_loc(sv1)
sd%length = c_loc(slen)
sd%angle = c_loc(sang)
sd%velocity = c_loc(fvel)
sd%tvel = c_loc(ftvel)
sd%slip_coeff = c_loc(fslip)
sd%btype = c_loc(fbtype)
sd%previous = c_loc(sprev)
sd%slip_mode = c_loc(fmode)
sd%lam_check = c_loc(flam)
sd%bit2d = c_loc(fbit)
sd%v2 = c_loc(sv2)
end subroutine pack_sides
#else
! Frontier faces in BFaceList order; GCBFVector entries (BoundaryFace
! indices) become 0-based positions in that order.
subroutine pack_faces(f)
type(sph_faces_t),intent(out) :: f
integer(4) :: kf,nf,st,k,n,c,r,nv
integer(4),allocatable :: pos(:)
allocate(pos(size(BoundaryFace)))
pos(:) = -1
k = 0
do kf=1,NumFacce
   nf = BFaceList(kf)
   if (nf==0) cycle
   pos(nf) = k
   k = k + 1
enddo
n = k
if (allocated(fT)) call free_tables
allocate(fT(9*max(n,1)),fLX(8*max(n,1)),fGX(12*max(n,1)),farea(max(n,1)))
allocate(fvel(3*max(n,1)),ftvel(3*max(n,1)),fslip(max(n,1)))
allocate(fnodes(max(n,1)),fbtype(max(n,1)),fmode(max(n,1)),flam(max(n,1)))
do kf=1,NumFacce
   nf = BFaceList(kf)
   if (nf==0) cycle
   k = pos(nf)
   st = BoundaryFace(nf)%stretch
   do c=1,3
      do r=1,3
         fT(9*k+3*(c-1)+r) = BoundaryFace(nf)%T(r,c)
      enddo
   enddo
   do c=1,4
      fLX(8*k+2*(c-1)+1:8*k+2*(c-1)+2) = BoundaryFace(nf)%Node(c)%LX(1:2)
      fGX(12*k+3*(c-1)+1:12*k+3*(c-1)+3) = BoundaryFace(nf)%Node(c)%GX(1:3)
   enddo
   farea(k+1) = BoundaryFace(nf)%area
   fvel(3*k+1:3*k+3) = BoundaryFace(nf)%velocity(1:3)
   ftvel(3*k+1:3*k+3) = Tratto(st)%velocity(1:3)
   fnodes(k+1) = BoundaryFace(nf)%nodes
   fbtype(k+1) = btype_code(Tratto(st)%tipo)
   fmode(k+1) = Partz(Tratto(st)%zone)%slip_coefficient_mode
   fslip(k+1) = Partz(Tratto(st)%zone)%BC_shear_stress_input
   flam(k+1) = 0
   if (Tratto(st)%laminar_no_slip_check) flam(k+1) = 1
enddo
allocate(gcount(Grid%nmax),gstart(Grid%nmax))
gcount(:) = GCBFPointers(1:Grid%nmax,1)
gstart(:) = GCBFPointers(1:Grid%nmax,2)
nv = size(GCBFVector)
allocate(gvec(max(nv,1)))
gvec(:) = 0
do k=1,nv
   if (GCBFVector(k)>0) gvec(k) = pos(GCBFVector(k))
enddo
allocate(fkt(41*5),fbit(256*7))
do r=0,40
   do c=0,4
      fkt(5*r+c+1) = kerneltab(r,c)
   enddo
enddo
do r=1,256
   do c=1,7
      fbit(7*(r-1)+c) = BoundIntegralTab(r,c)
   enddo
enddo
f%nf = n
f%T = c_loc(fT)
f%LX = c_loc(fLX)
f%GX = c_loc(fGX)
f%area = c_loc(farea)
f%velocity = c_loc(fvel)
f%tvel = c_loc(ftvel)
f%slip_coeff = c_loc(fslip)
f%nodes = c_loc(fnodes)
f%btype = c_loc(fbtype)
f%slip_mode = c_loc(fmode)
f%lam_check = c_loc(flam)
f%gcbf_count = c_loc(gcount)
f%gcbf_start = c_loc(gstart)
f%gcbf_vec = c_loc(gvec)
f%kerneltab = c_loc(fkt)
f%ktdelta = ktdelta
f%bitab = c_loc(fbit)
! FindBoundaryConvexEdges3D's edges, faces as 0-based positions
allocate(eface(2*max(NumBEdges,1)),egx(3*max(NumBEdges,1)),ecomp(3*max(NumBEdges,1)))
allocate(elen(max(NumBEdges,1)))
eface(:) = 0
egx(:) = 0.d0
ecomp(:) = 0.d0
elen(:) = 0.d0
do k=1,NumBEdges
   eface(2*k-1) = pos(BoundaryConvexEdge(k)%face(1))
   eface(2*k) = pos(BoundaryConvexEdge(k)%face(2))
   egx(3*k-2:3*k) = BoundaryConvexEdge(k)%node(1)%GX(1:3)
   ecomp(3*k-2:3*k) = BoundaryConvexEdge(k)%component(1:3)
   elen(k) = BoundaryConvexEdge(k)%length
enddo
f%nedges = NumBEdges
f%edge_face = c_loc(eface)
f%edge_gx = c_loc(egx)
f%edge_comp = c_loc(ecomp)
f%edge_len = c_loc(elen)
deallocate(pos)
end subroutine pack_faces
#endif

#ifdef SPACE_3D
! No frontier faces (a DB-SPH context: the walls come from sph_dbsph_walls)
subroutine empty_faces(f)
type(sph_faces_t),intent(out) :: f
f%nf = 0
f%T = c_null_ptr
f%LX = c_null_ptr
f%GX = c_null_ptr
f%area = c_null_ptr
f%velocity = c_null_ptr
f%tvel = c_null_ptr
f%slip_coeff = c_null_ptr
f%nodes = c_null_ptr
f%btype = c_null_ptr
f%slip_mode = c_null_ptr
f%lam_check = c_null_ptr
f%gcbf_count = c_null_ptr
f%gcbf_start = c_null_ptr
f%gcbf_vec = c_null_ptr
f%kerneltab = c_null_ptr
f%ktdelta = 0.d0
f%bitab = c_null_ptr
f%nedges = 0
f%edge_face = c_null_ptr
f%edge_gx = c_null_ptr
f%edge_comp = c_null_ptr
f%edge_len = c_null_ptr
end subroutine empty_faces

! pg_w(1:n_w) (DBSPH_IC_surface_elements.f90, semi_particle_volumes.f90) and
! the DBSPH section's flags in the C layout
subroutine pack_walls(w)
type(sph_dbsph_walls_t),intent(out) :: w
integer(4) :: k,nw
nw = DBSPH%n_w
if (allocated(wx)) deallocate(wx,wn,wv,ww,wvol,wm)
allocate(wx(3*max(nw,1)),wn(3*max(nw,1)),wv(3*max(nw,1)),ww(max(nw,1)),wvol(max(nw,1)))
allocate(wm(max(nw,1)))
do k=1,nw
   wx(3*k-2:3*k) = pg_w(k)%coord(1:3)
   wn(3*k-2:3*k) = pg_w(k)%normal(1:3)
   wv(3*k-2:3*k) = pg_w(k)%vel(1:3)
   ww(k) = pg_w(k)%weight
   wvol(k) = pg_w(k)%volume
   wm(k) = pg_w(k)%mass
enddo
w%nw = nw
w%coord = c_loc(wx)
w%normal = c_loc(wn)
w%vel = c_loc(wv)
w%weight = c_loc(ww)
w%volume = c_loc(wvol)
w%mass = c_loc(wm)
w%gamma_limiter = merge(1,0,DBSPH%Gamma_limiter_flag)
w%fs_allowed = merge(1,0,DBSPH%FS_allowed)
w%negative_wall_p_allowed = merge(1,0,DBSPH%negative_wall_p_allowed)
w%muscl_boundary = merge(1,0,DBSPH%MUSCL_boundary_flag)
end subroutine pack_walls

! The mesh's kinematics records (DBSPH%kinematics(1,:,1:7)) and rotation
! centre to sph_dbsph_kinematics (DBSPH_kinematics.f90 on the device)
subroutine pass_kinematics(ctx)
type(c_ptr),intent(in) :: ctx
integer(4) :: j,nr
double precision,allocatable :: kr(:)
double precision :: cr(3)
nr = 0
if (allocated(DBSPH%n_kinematics_records)) nr = DBSPH%n_kinematics_records(1)
if (nr<=0) return
allocate(kr(7*nr))
do j=1,nr
   kr(7*(j-1)+1:7*j) = DBSPH%kinematics(1,j,1:7)
enddo
cr(:) = DBSPH%rotation_centre(1,1:3)
call sph_hip_check(sph_dbsph_kinematics(ctx,int(nr,c_int32_t),kr,cr),"sph_dbsph_kinematics")
deallocate(kr)
end subroutine pass_kinematics
#endif

subroutine free_tables
if (allocated(fLX)) deallocate(fLX,fGX,farea,fnodes)
if (allocated(sRN)) deallocate(sRN,sv1,sv2,slen,sang,sprev)
deallocate(fT,fvel,ftvel,fslip,fbtype,fmode,flam)
if (allocated(gcount)) deallocate(gcount,gstart,gvec)
if (allocated(fkt)) deallocate(fkt)
if (allocated(fbit)) deallocate(fbit)
if (allocated(eface)) deallocate(eface,egx,ecomp,elen)
end subroutine free_tables

! Creates the device context for the current case and uploads pg(1:nag).
subroutine sph_hip_open(prec,ctx)
integer(c_int32_t),intent(in) :: prec
type(c_ptr),intent(out) :: ctx
type(sph_params_t) :: p
#ifdef SPACE_2D
type(sph_sides2d_t) :: sd
#else
type(sph_faces_t) :: f
type(sph_dbsph_walls_t) :: w
#endif
real(c_double),allocatable :: xyz(:),uvw(:),rho(:),pp(:),m(:)
integer(4) :: i
#ifdef SPACE_2D
if ((Domain%tipo/="semi").or.(Domain%time_split<0).or.(Domain%time_split>1).or.    &
   (NMedium/=1).or.Domain%NormFix.or.(SourceSide/=0).or.dt_alfa_Mon.or.             &
   other_open_sides()) then
#else
if (((Domain%tipo/="semi").and.(Domain%tipo/="bsph")).or.(Domain%time_split<0).or. &
   (Domain%time_split>1).or.(NMedium/=1).or.Domain%NormFix.or.(SourceFace/=0).or.  &
   dt_alfa_Mon.or.other_open_faces()) then
#endif
   write(*,*) "sphera_hip: case outside the device scope (semi, leapfrog or ",     &
      "explicit RK, one medium, fixed/tapis/open frontiers)"
   stop 1
endif
call pack_params(prec,p)
#ifdef SPACE_2D
call pack_sides(sd)
call sph_hip_check(sph_ctx_create2d(p,sd,ctx),"sph_ctx_create2d")
#else
if (Domain%tipo=="bsph") then
   if (dbsph_outside_scope()) then
      write(*,*) "sphera_hip: DB-SPH case outside the device scope (RK1, ",         &
         "slip_ID 0, static walls, no inlets/outlets)"
      stop 1
   endif
   call empty_faces(f)
   call sph_hip_check(sph_ctx_create(p,f,ctx),"sph_ctx_create")
   call pack_walls(w)
   call sph_hip_check(sph_dbsph_walls(ctx,w),"sph_dbsph_walls")
   call pass_kinematics(ctx)
   else
      call pack_faces(f)
      call sph_hip_check(sph_ctx_create(p,f,ctx),"sph_ctx_create")
endif
#endif
allocate(xyz(3*max(nag,1)),uvw(3*max(nag,1)),rho(max(nag,1)),pp(max(nag,1)),m(max(nag,1)))
do i=1,nag
   xyz(3*i-2:3*i) = pg(i)%coord(1:3)
   uvw(3*i-2:3*i) = pg(i)%vel(1:3)
   rho(i) = pg(i)%dens
   pp(i) = pg(i)%pres
   m(i) = pg(i)%mass
enddo
call sph_hip_check(sph_state_upload(ctx,int(nag,c_int32_t),xyz,uvw,rho,pp,m),          &
   "sph_state_upload")
deallocate(xyz,uvw,rho,pp,m)
end subroutine sph_hip_open

! Device state -> pg(1:nag).  Particles that left the grid are removed on the
! device; the remaining rows come back in pg order and fill pg(1:n).
subroutine sph_hip_pull(ctx)
type(c_ptr),intent(in) :: ctx
real(c_double),allocatable :: xyz(:),uvw(:),a(:),rho(:),pp(:),dd(:),vp(:)
integer(c_int32_t),allocatable :: ids(:)
integer(c_int32_t) :: n
real(c_double) :: sc(5)
integer(4) :: i,k
call sph_hip_check(sph_particle_count(ctx,n),"sph_particle_count")
allocate(xyz(3*max(n,1)),uvw(3*max(n,1)),a(3*max(n,1)),rho(max(n,1)),pp(max(n,1)))
allocate(dd(max(n,1)),vp(max(n,1)),ids(max(n,1)))
call sph_hip_check(sph_state_download(ctx,xyz,uvw,a,rho,pp,dd,vp,ids),                &
   "sph_state_download")
do k=1,n
   i = ids(k) + 1
   if (i/=k) pg(k) = pg(i)
   pg(k)%coord(1:3) = xyz(3*k-2:3*k)
   pg(k)%vel(1:3) = uvw(3*k-2:3*k)
   pg(k)%var(1:3) = uvw(3*k-2:3*k)
   pg(k)%acc(1:3) = a(3*k-2:3*k)
   pg(k)%dens = rho(k)
   pg(k)%pres = pp(k)
   pg(k)%dden = dd(k)
   pg(k)%vpres = vp(k)
enddo
do k=n+1,nag
   pg(k) = PgZero
enddo
nag = n
call sph_hip_check(sph_get_scalars(ctx,sc),"sph_get_scalars")
dt = sc(1)
dt_average = sc(2)
simulation_time = sc(3)
deallocate(xyz,uvw,a,rho,pp,dd,vp,ids)
end subroutine sph_hip_pull

! Monitoring lines of the deck (ReadInputControlLines) -> device.
subroutine sph_hip_monitor_lines(ctx)
type(c_ptr),intent(in) :: ctx
integer(c_int32_t),allocatable :: ic(:)
real(c_double),allocatable :: xyz(:)
integer(4) :: i
allocate(ic(2*max(nlines,1)),xyz(3*max(npointst,1)))
do i=1,nlines
   ic(2*i-1:2*i) = control_lines(i)%icont(1:2)
enddo
do i=1,npointst
   xyz(3*i-2:3*i) = control_points(i)%coord(1:3)
enddo
call sph_hip_check(sph_monitor_lines(ctx,int(nlines,c_int32_t),ic,                    &
   int(npointst,c_int32_t),xyz),"sph_monitor_lines")
deallocate(ic,xyz)
end subroutine sph_hip_monitor_lines

! calc_pelo (calc_pelo.f90:82-133) with the lines on the device: the time
! series records go to the same units in the same format.  Call after
! sph_hip_pull (simulation_time).
subroutine sph_hip_calc_pelo(ctx)
use I_O_file_module
type(c_ptr),intent(in) :: ctx
real(c_double) :: pelolib(3,max(nlines,1)),z_lower_fluid_top(3,max(nlines,1))
if (nlines<1) return
call sph_hip_check(sph_free_surface(ctx,pelolib,z_lower_fluid_top),"sph_free_surface")
write(nplb,'(30g14.7)') simulation_time,pelolib(:,1:nlines)
write(uzlft,'(30g14.7)') simulation_time,z_lower_fluid_top(:,1:nlines)
end subroutine sph_hip_calc_pelo

subroutine sph_hip_close(ctx)
type(c_ptr),intent(inout) :: ctx
call sph_hip_check(sph_ctx_destroy(ctx),"sph_ctx_destroy")
ctx = c_null_ptr
if (allocated(fT)) call free_tables
end subroutine sph_hip_close

end module sphera_hip_bridge
