! ISO_C_BINDING interfaces of the C ABI in include/sphera_hip.h.
!
! The bind(C) types follow the C structs member for member, so the companion C
! processor gives them the same layout (padding included).  Every function
! returns the status code of the C ABI (0 = success).
module sphera_hip_c
use iso_c_binding
implicit none

integer(c_int32_t),parameter :: SPH_FP64 = 0, SPH_FP32_MIXED = 1

type,bind(C) :: sph_params_t
   integer(c_int32_t) :: precision, capacity, nmaxpartj, maxclose
   integer(c_int32_t) :: ncd(3)
   real(c_double) :: extr_lo(3), dcd(3)
   real(c_double) :: dx, h
   real(c_double) :: coefke, coefkacl, eta, eta2
   real(c_double) :: grav(3), prif
   real(c_double) :: den0, eps, celerita, alfamon, kin_visc
   real(c_double) :: cfl, vsc_coeff, pesodt, tetap, tetav
   integer(c_int32_t) :: semi
   integer(c_int32_t) :: wg_size
   integer(c_int32_t) :: chunk
   integer(c_int32_t) :: chunk_v, chunk_c, chunk_p
   integer(c_int32_t) :: nb_shape
   integer(c_int32_t) :: options
end type sph_params_t

type,bind(C) :: sph_faces_t
   integer(c_int32_t) :: nf
   type(c_ptr) :: T, LX, GX, area, velocity, tvel, slip_coeff
   type(c_ptr) :: nodes, btype, slip_mode, lam_check
   type(c_ptr) :: gcbf_count, gcbf_start, gcbf_vec
   type(c_ptr) :: kerneltab
   real(c_double) :: ktdelta
   type(c_ptr) :: bitab
   integer(c_int32_t) :: nedges
   type(c_ptr) :: edge_face, edge_gx, edge_comp, edge_len
end type sph_faces_t

! DB-SPH wall elements (sph_dbsph_walls_t of sphera_hip.h)
type,bind(C) :: sph_dbsph_walls_t
   integer(c_int32_t) :: nw
   type(c_ptr) :: coord, normal, vel, weight, volume, mass
   integer(c_int32_t) :: gamma_limiter, fs_allowed, negative_wall_p_allowed
   integer(c_int32_t) :: muscl_boundary
end type sph_dbsph_walls_t

! 2D boundary sides (sph_sides2d of sphera_hip.h)
type,bind(C) :: sph_sides2d_t
   integer(c_int32_t) :: ns
   type(c_ptr) :: T, RN, v1, length, angle, velocity, tvel, slip_coeff
   type(c_ptr) :: btype, previous, slip_mode, lam_check
   type(c_ptr) :: bit2d
   type(c_ptr) :: v2
end type sph_sides2d_t

interface
   integer(c_int) function sph_ctx_create(p,f,ctx) bind(C,name="sph_ctx_create")
      import :: c_int, c_ptr, sph_params_t, sph_faces_t
      type(sph_params_t),intent(in) :: p
      type(sph_faces_t),intent(in) :: f
      type(c_ptr),intent(out) :: ctx
   end function
   integer(c_int) function sph_ctx_create2d(p,sd,ctx) bind(C,name="sph_ctx_create2d")
      import :: c_int, c_ptr, sph_params_t, sph_sides2d_t
      type(sph_params_t),intent(in) :: p
      type(sph_sides2d_t),intent(in) :: sd
      type(c_ptr),intent(out) :: ctx
   end function
   integer(c_int) function sph_ctx_destroy(ctx) bind(C,name="sph_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_state_upload(ctx,n,coord,vel,dens,pres,mass)           &
      bind(C,name="sph_state_upload")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr),value :: ctx
      integer(c_int32_t),value :: n
      real(c_double),intent(in) :: coord(*), vel(*), dens(*), pres(*), mass(*)
   end function
   integer(c_int) function sph_particle_count(ctx,n) bind(C,name="sph_particle_count")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr),value :: ctx
      integer(c_int32_t),intent(out) :: n
   end function
   integer(c_int) function sph_state_download(ctx,coord,vel,acc,dens,pres,dden,vpres,ids) &
      bind(C,name="sph_state_download")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr),value :: ctx
      real(c_double),intent(out) :: coord(*), vel(*), acc(*), dens(*), pres(*)
      real(c_double),intent(out) :: dden(*), vpres(*)
      integer(c_int32_t),intent(out) :: ids(*)
   end function
   integer(c_int) function sph_get_laminar(ctx,lam) bind(C,name="sph_get_laminar")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr),value :: ctx
      integer(c_int32_t),intent(out) :: lam(*)
   end function
   integer(c_int) function sph_get_scalars(ctx,out) bind(C,name="sph_get_scalars")
      import :: c_int, c_ptr, c_double
      type(c_ptr),value :: ctx
      real(c_double),intent(out) :: out(5)
   end function
   integer(c_int) function sph_pair_counter(ctx,total) bind(C,name="sph_pair_counter")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr),value :: ctx
      integer(c_long_long),intent(out) :: total
   end function
   integer(c_int) function sph_get_neighbour_lists(ctx,cnt,ids,stride)                 &
      bind(C,name="sph_get_neighbour_lists")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr),value :: ctx
      integer(c_int32_t),intent(out) :: cnt(*), ids(*)
      integer(c_int32_t),value :: stride
   end function
   ! phases, in time_step_loop order
   integer(c_int) function sph_dt(ctx) bind(C,name="sph_dt")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_momentum(ctx) bind(C,name="sph_momentum")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_leapfrog_mom(ctx) bind(C,name="sph_leapfrog_mom")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_velocity_smoothing(ctx) bind(C,name="sph_velocity_smoothing")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_leapfrog_traj(ctx) bind(C,name="sph_leapfrog_traj")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_sort(ctx) bind(C,name="sph_sort")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_neighbours(ctx) bind(C,name="sph_neighbours")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_boundary_tab(ctx) bind(C,name="sph_boundary_tab")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_continuity(ctx) bind(C,name="sph_continuity")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_leapfrog_cont(ctx) bind(C,name="sph_leapfrog_cont")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_eos(ctx) bind(C,name="sph_eos")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_pressure_smoothing(ctx) bind(C,name="sph_pressure_smoothing")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_dbsph_walls(ctx,w) bind(C,name="sph_dbsph_walls")
      import :: c_int, c_ptr, sph_dbsph_walls_t
      type(c_ptr),value :: ctx
      type(sph_dbsph_walls_t),intent(in) :: w
   end function
   integer(c_int) function sph_dbsph_kinematics(ctx,nrec,records,centre)                 &
      bind(C,name="sph_dbsph_kinematics")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr),value :: ctx
      integer(c_int32_t),value :: nrec
      real(c_double),intent(in) :: records(*), centre(3)
   end function
   integer(c_int) function sph_preloop(ctx) bind(C,name="sph_preloop")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   integer(c_int) function sph_step(ctx,nsteps) bind(C,name="sph_step")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr),value :: ctx
      integer(c_int32_t),value :: nsteps
   end function
   integer(c_int) function sph_step_async(ctx,nsteps) bind(C,name="sph_step_async")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr),value :: ctx
      integer(c_int32_t),value :: nsteps
   end function
   integer(c_int) function sph_step_graph(ctx,nsteps) bind(C,name="sph_step_graph")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr),value :: ctx
      integer(c_int32_t),value :: nsteps
   end function
   integer(c_int) function sph_global_quantities(ctx,out)                              &
      bind(C,name="sph_global_quantities")
      import :: c_int, c_ptr, c_double
      type(c_ptr),value :: ctx
      real(c_double),intent(out) :: out(8)
   end function
   ! the same sums with no device allocation (a fixed library array per GPU)
   integer(c_int) function sph_global_quantities_noalloc(ctx,out)                      &
      bind(C,name="sph_global_quantities_noalloc")
      import :: c_int, c_ptr, c_double
      type(c_ptr),value :: ctx
      real(c_double),intent(out) :: out(8)
   end function
   integer(c_int) function sph_monitor_lines(ctx,nlines,icont,npoints,points)          &
      bind(C,name="sph_monitor_lines")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr),value :: ctx
      integer(c_int32_t),value :: nlines, npoints
      integer(c_int32_t),intent(in) :: icont(*)
      real(c_double),intent(in) :: points(*)
   end function
   integer(c_int) function sph_free_surface(ctx,pelolib,zlft)                          &
      bind(C,name="sph_free_surface")
      import :: c_int, c_ptr, c_double
      type(c_ptr),value :: ctx
      real(c_double),intent(out) :: pelolib(*), zlft(*)
   end function
   integer(c_int) function sph_query(ctx,busy) bind(C,name="sph_query")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr),value :: ctx
      integer(c_int32_t),intent(out) :: busy
   end function
   integer(c_int) function sph_device_memory(free_bytes,total_bytes)                  &
      bind(C,name="sph_device_memory")
      import :: c_int, c_long_long
      integer(c_long_long),intent(out) :: free_bytes, total_bytes
   end function
   integer(c_int) function sph_ctx_bytes(p,bytes) bind(C,name="sph_ctx_bytes")
      import :: c_int, c_long_long, sph_params_t
      type(sph_params_t),intent(in) :: p
      integer(c_long_long),intent(out) :: bytes
   end function
   integer(c_int) function sph_get_launch(ctx,out) bind(C,name="sph_get_launch")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr),value :: ctx
      integer(c_int32_t),intent(out) :: out(3)
   end function
   integer(c_int) function sph_set_timing(ctx,on) bind(C,name="sph_set_timing")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr),value :: ctx
      integer(c_int32_t),value :: on
   end function
   integer(c_int) function sph_sync(ctx) bind(C,name="sph_sync")
      import :: c_int, c_ptr
      type(c_ptr),value :: ctx
   end function
   type(c_ptr) function sph_last_error() bind(C,name="sph_last_error")
      import :: c_ptr
   end function
   integer(c_size_t) function c_strlen(s) bind(C,name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr),value :: s
   end function
end interface

contains

! Message of the last failed call, as a Fortran string.
function sph_error_message() result(msg)
character(len=:),allocatable :: msg
type(c_ptr) :: p
character(kind=c_char),pointer :: s(:)
integer :: n,k
p = sph_last_error()
if (.not.c_associated(p)) then
   msg = ""
   return
endif
n = int(c_strlen(p))
call c_f_pointer(p,s,[n])
allocate(character(len=n) :: msg)
do k=1,n
   msg(k:k) = s(k)
enddo
end function sph_error_message

end module sphera_hip_c
