/* sphera_hip.h — C ABI of the MI355X (gfx950) particle-interaction hot path of
 * SPHERA v10.  Plain pointers and sizes only; no C++ or torch types.
 *
 * The reference has no FFI: its hot path is a set of argument-less Fortran
 * subroutines acting on module globals, called in a fixed order by
 * time_step_loop (src/Main_algorithm/time_step_loop.f90:175-489).  Each
 * per-phase entry point below replaces one of those subroutines (cited), with
 * the same meaning, so that a Fortran ISO_C_BINDING bridge can call them in
 * the reference's own order; sph_step runs that whole order on the device.
 *
 * Conventions
 *  - Return value: 0 on success, otherwise a status code; codes 10..99 map to
 *    the reference's diagnostic(arg1,arg2) pairs as 10*arg1+arg2 where one
 *    exists (e.g. 86 "too many close boundary faces"), codes >= 100 are
 *    device/runtime errors.  sph_last_error() gives the message.  Nothing
 *    exits the process (the reference stops in diagnostic.f90:317).
 *  - Host arrays are borrowed for the duration of the call; xyz vectors are
 *    interleaved (x0,y0,z0,x1,...), particle order is the caller's pg order.
 *  - Device memory is owned by the context.  Calls on one context are
 *    sequential and not thread-safe; separate contexts are independent.
 */
#ifndef SPHERA_HIP_H
#define SPHERA_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SPH_FP64 = 0, SPH_FP32_MIXED = 1 };

/* Scalars of Domain, Grid, Med(1) and input_any_t the hot path reads
 * (src/Modules/Hybrid_allocation_module.f90:159-168,1120-1131;
 *  src/Modules/Static_allocation_module.f90:226-239). */
typedef struct sph_params {
    int32_t precision;   /* SPH_FP64 | SPH_FP32_MIXED */
    int32_t capacity;    /* maximum particle count held by the context */
    int32_t nmaxpartj;   /* NMAXPARTJ: neighbour-list cap (lists truncate past it) */
    int32_t maxclose;    /* MAXCLOSEBOUNDFACES */
    int32_t ncd[3];      /* Grid%ncd */
    double extr_lo[3];   /* Grid%extr(:,1) */
    double dcd[3];       /* Grid%dcd */
    double dx, h;        /* Domain%dx, Domain%h */
    double coefke, coefkacl, eta, eta2;
    double grav[3], prif;                       /* Domain%grav, Domain%prif */
    double den0, eps, celerita, alfamon, kin_visc; /* Med(1) */
    double cfl, vsc_coeff, pesodt, tetap, tetav;   /* input_any_t */
    int32_t semi;        /* 1: SA-SPH frontiers active (Domain%tipo=="semi") */
    /* launch tunables of the particle kernels (0 = default; the autotuner in
     * sphera_amd/autotune.py picks them per case) */
    int32_t wg_size;     /* threads per workgroup: 64 (default), 128, 192 or 256 */
    int32_t chunk;       /* list entries gathered per step of the FP32 walks: 4 or 8 (default);
                            momentum pass, and the others where theirs is 0 */
    int32_t chunk_v, chunk_c, chunk_p; /* the same for velocity smoothing, continuity,
                                          pressure smoothing (0 = chunk) */
    int32_t nb_shape;    /* neighbour-search workgroup size x LDS candidate tile:
                            0 = 256 x 4096 in FP16 (default), 1 = 256 x 3072 in FP32,
                            2 = 128 x 1536 in FP32, 3 = 256 x 2816 in FP16 with 64
                            registers (8 waves per SIMD), 4 = 256 x 4096 in 16-bit
                            fixed point over the cell.  The lists are the same in
                            every shape (FP64 decides near the cut-off). */
    int32_t options;     /* SPH_OPT_* flags of the deck (below) */
} sph_params;

/* Deck options the hot path branches on (input_any_t of
 * src/Modules/Hybrid_allocation_module.f90:88-113,931; NumBEdges of
 * Static_allocation_module.f90:220; Domain%time_split and RKscheme).
 * SPH_OPT_DENSITY_THRESHOLDS is built: the 0.9/1.1 den0 clamp of
 * Leapfrog_continuity.f90:61-68; so are SPH_OPT_RK1 and
 * SPH_OPT_CONVEX_EDGES (3D contexts only).  The others are not
 * built on the device: sph_ctx_create / sph_ctx_create2d / sph_batch_create
 * refuse a context that sets them with status 95, rather than running the
 * case without them.  A frontier's slip_mode is 0, 1 or 2 (2: the SA-SPH
 * wall function of wall_function_for_SASPH.f90, built); other values are
 * refused with status 2. */
enum {
    SPH_OPT_DENSITY_THRESHOLDS = 1, /* density_thresholds == 1 */
    SPH_OPT_ALE3 = 2,               /* ALE3 (ALE3_term_momentum.f90, Leapfrog_continuity.f90:70-86) */
    SPH_OPT_C1 = 4,                 /* C1_BE or C1_divu (renormalisation, CalcVarLength.f90:285-324) */
    SPH_OPT_KTGF = 8,               /* KTGF_config > 0 (mixture viscosity, erosion) */
    SPH_OPT_CONVEX_EDGES = 16,      /* NumBEdges > 0 (AddElasticBoundaryReaction_3D.f90:106-152),
                                       built: set exactly when sph_faces.nedges > 0 */
    SPH_OPT_RK1 = 32,               /* time_split = 0, RKscheme = 1: explicit Euler, built
                                       (time_step_loop.f90:190-489, Euler.f90:61-241); such a
                                       context steps with sph_step / sph_step_graph only */
    SPH_OPT_RK2 = 64                /* time_split = 0, RKscheme = 2 (Heun.f90) */
};

/* SA-SPH frontier faces in BFaceList order and the grid-cell/face incidence
 * (BoundaryFace, Tratto, Partz, GCBFPointers, GCBFVector, kerneltab,
 * BoundIntegralTab).  T is column-major per face (T(r,c) at [9f+3c+r]);
 * LX: 4 nodes x (x,y); GX: 4 nodes x (x,y,z); btype: 0 fixe, 1 tapi, 2 open,
 * 3 sour, 4 velo, 5 flow.  A particle whose trajectory step crosses an open
 * face is removed at the next sort (CancelOutgoneParticles_3D.f90, then
 * OrdGrid1's swap-with-last).  gcbf_start is 1-based into gcbf_vec, which
 * holds 0-based face indices. */
typedef struct sph_faces {
    int32_t nf;
    const double *T, *LX, *GX, *area, *velocity, *tvel, *slip_coeff;
    const int32_t *nodes, *btype, *slip_mode, *lam_check;
    const int32_t *gcbf_count, *gcbf_start, *gcbf_vec;
    const double *kerneltab; /* 41 x 5 row-major */
    double ktdelta;
    const double *bitab;     /* 256 x 7 row-major */
    /* convex edges (BoundaryConvexEdge of FindBoundaryConvexEdges3D.f90, in
     * its order; NumBEdges of them): edge_face 2 per edge (0-based face
     * indices), edge_gx the first node (x,y,z), edge_comp the second node
     * minus the first, edge_len its length.  nedges = 0: none (the pointers
     * may be null). */
    int32_t nedges;
    const int32_t *edge_face;
    const double *edge_gx, *edge_comp, *edge_len;
} sph_faces;

typedef struct sph_ctx sph_ctx;

/* Context: allocates device memory for `capacity` particles and copies the
 * frontier tables.  Replaces the allocations of Gest_Trans.f90:130-211 and
 * the SA-SPH set-up of Gest_Trans.f90:692-701 (tables computed by the
 * caller). */
int sph_ctx_create(const sph_params* p, const sph_faces* f, sph_ctx** out);

/* 2D form: the reference's -DSPACE_2D build (x-z plane; particles at y = 0;
 * the grid has ncd[1] = 1 and every particle in row j = 1, CreaGrid.f90:69-71,
 * ParticleCellNumber.f90:59-61).  Frontiers are boundary sides in BoundarySide
 * order (DefineBoundarySideGeometry2D.f90:60-176, peri zones included): T and
 * RN (BoundaryMassForceMatrix2D) column-major 3x3 per side, v1 the first
 * vertex (x, 0, z), length, angle and previous (previous_side, 1-based, 0 =
 * none) of DefineBoundarySideRelativeAngles2D.f90, btype as for faces (6 =
 * peri), velocity = BoundarySide%velocity, tvel = Tratto%velocity; bit2d is
 * BoundIntegralTab2D (SA_SPH_module.f90:30-54, 22 x 6 row-major).  The phase
 * entry points below then run the 2D units: ComputeBoundaryDataTab_2D,
 * AddBoundaryContributions_to_ME2D, AddElasticBoundaryReaction_2D,
 * velocity_smoothing_SA_SPH_2D, AddBoundaryContribution_to_CE2D and
 * PressureSmoothing_2D.  Replaces Gest_Input.f90:292-293 (side geometry
 * computed by the caller).  v2 is the second vertex (x, 0, z) of each side:
 * read for open sides (btype 2), whose particles the next sort removes when
 * their step crosses the side (CancelOutgoneParticles_2D.f90:60-104,
 * OpenSide of Gest_Input.f90:600-613); it may be NULL when no side is open
 * (status 2 otherwise).  Open sides enter the boundary table as
 * SelectCloseBoundarySides2D.f90:99-107 selects them and add no wall term. */
typedef struct sph_sides2d {
    int32_t ns;
    const double *T, *RN, *v1, *length, *angle, *velocity, *tvel, *slip_coeff;
    const int32_t *btype, *previous, *slip_mode, *lam_check;
    const double *bit2d;
    const double *v2;
} sph_sides2d;
int sph_ctx_create2d(const sph_params* p, const sph_sides2d* s, sph_ctx** out);

/* DB-SPH walls (Domain%tipo = "bsph"): the wall elements ("semi-particles")
 * of the frontier's surface mesh, as the reference's pre-processing leaves
 * pg_w(1:n_w) (Import_ply_surface_meshes.f90, DBSPH_IC_surface_elements.f90,
 * semi_particle_volumes.f90; sphera_amd/dbsph.py computes them from a mesh):
 * centroid, unit normal into the fluid and velocity (xyz interleaved), area
 * ("weight"), semi-particle volume and mass, and the DBSPH section's
 * Gamma_limiter_flag, FS_allowed and negative_wall_p_allowed.  Replaces the
 * pg_w set-up of Gest_Input.f90:416-440,574-594 for the path.
 *
 * Called once on a context created with semi = 0, no faces, SPH_OPT_RK1
 * (time_split = 0, RKscheme = 1) and SPH_FP64, before sph_preloop.  The
 * context then runs the DB-SPH step: the fluid-wall lists and Shepard sums
 * of CalcVarLength.f90:462-540,842-877, the DB-SPH momentum terms of
 * inter_EqMoto.f90:229-235,286-322 and RHS_momentum_equation.f90:173-177,
 * Euler.f90:61-241's DB-SPH branch (Gamma += dShep dt, density from
 * rhoSPH_new / Gamma), PressureSmoothing_3D.f90:110-112, and the wall
 * elements' state (Gradients_to_MUSCL.f90, BC_wall_elements.f90).  Static
 * walls (no DBSPH kinematics), slip_ID = 0 (no VSL model), no MUSCL boundary
 * terms, no DB-SPH inlets or outlets, one medium.  Status 95 for a segmented
 * or FP32-mixed context, 2 for the other preconditions, 87 from a step when
 * a particle has more than 128 wall elements within 2h or a wall element
 * more than 128 particles. */
typedef struct sph_dbsph_walls_t {
    int32_t nw;
    const double *coord, *normal, *vel, *weight, *volume, *mass;
    int32_t gamma_limiter, fs_allowed, negative_wall_p_allowed;
    /* DBSPH%MUSCL_boundary_flag: the wall elements' terms in the particles'
     * MUSCL gradients (Gradients_to_MUSCL_boundary.f90) */
    int32_t muscl_boundary;
} sph_dbsph_walls_t;
int sph_dbsph_walls(sph_ctx* ctx, const sph_dbsph_walls_t* walls);
/* Moving DB-SPH walls (DBSPH_kinematics.f90, Euler.f90:152-159): nrec
 * records (t, vx, vy, vz, wx, wy, wz) of the mesh, the first at t <= 0, and
 * its rotation centre.  Each step the wall elements take the velocity
 * omega x (x - centre) + v at the step's time and move by dt times it; they
 * keep their cells, as the reference bins them once.  nrec = 0: static walls.
 * After sph_dbsph_walls, before sph_preloop.  Replaces the DBSPH%kinematics
 * and DBSPH%rotation_centre that ReadDBSPH.f90:115-141 reads. */
#define SPH_MAX_KIN 4096
int sph_dbsph_kinematics(sph_ctx* ctx, int32_t nrec, const double* records, const double* centre);
/* DB-SPH state at a time-step boundary, particles in pg order (any pointer
 * may be NULL): Gamma, sigma, rhoSPH_new, dShep, FS; the wall elements'
 * density, pressure and wet flags (pg_w%dens, %pres, %wet). */
int sph_dbsph_state(sph_ctx* ctx, double* gamma, double* sigma, double* rhosph, double* dshep, int32_t* fs,
                    double* wall_dens, double* wall_pres, int32_t* wall_wet);
int sph_ctx_destroy(sph_ctx* ctx);

/* Particle state in pg order: coord, vel (xyz interleaved), dens, pres, mass.
 * Initialises ids 0..n-1 and orders the particles on the grid (OrdGrid1). */
int sph_state_upload(sph_ctx* ctx, int32_t n, const double* coord, const double* vel,
                     const double* dens, const double* pres, const double* mass);
/* Current particle count (nag; particles leaving the grid are removed). */
int sph_particle_count(sph_ctx* ctx, int32_t* n);
/* Download in pg (id) order.  Any pointer may be NULL.  var equals vel at
 * every point where the reference reads it.  ids: the pg index of each
 * returned row (rows are sorted by it). */
int sph_state_download(sph_ctx* ctx, double* coord, double* vel, double* acc, double* dens,
                       double* pres, double* dden, double* vpres, int32_t* ids);
/* Laminar flags (pg%laminar_flag, set by the momentum pass,
 * inter_EqMoto.f90:269-283) in pg order, at a time-step boundary; the
 * reference writes them to its Paraview files. */
int sph_get_laminar(sph_ctx* ctx, int32_t* lam);
/* dt, dt_average, simulation_time, it, pair count of the current lists */
int sph_get_scalars(sph_ctx* ctx, double out[5]);
/* Fluid global quantities (fluid_global_quantities.f90:66-72): mass, linear
 * momentum (3), angular momentum about the origin (3), volume sum m/rho;
 * compensated FP64 sums (exact sum of the per-particle terms, rounded once). */
int sph_global_quantities(sph_ctx* ctx, double out[8]);
/* sph_global_quantities without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-workgroup partial sums live in one fixed device array
 * of the library (2048 workgroups x 16 doubles + 8: 262,208 bytes per GPU,
 * held until the process ends) instead of a per-call allocation; the
 * results equal sph_global_quantities' bit for bit.  Status 3 if the device
 * needs more than 2048 partial-sum workgroups.  Every call on one GPU uses
 * that one array, so calls that overlap on one GPU (from several host
 * threads) race; the call blocks until its results are copied. */
int sph_global_quantities_noalloc(sph_ctx* ctx, double out[8]);
/* Monitoring lines of the free-surface monitor, as ReadInputControlLines.f90
 * :97-124 stores them: icont[2*i], icont[2*i+1] = control_lines(i+1)%Icont
 * (1-based first and last point), points[3*j..] = control_points(j+1)%coord.
 * Copied to the device; nlines = 0 clears them.  Status 3 if a line's points
 * are outside 1..npoints. */
int sph_monitor_lines(sph_ctx* ctx, int32_t nlines, const int32_t* icont, int32_t npoints,
                      const double* points);
/* Free-surface monitor (calc_pelo.f90:82-131, cells as CalcVarp.f90:66-72):
 * per line, pelolib[3*i..] = mean of the two particles closest to the point
 * of the last filled cell along the line, + dx/2 on each coordinate;
 * zlft[3*i..] = the same at the last filled cell below the first empty cell
 * above fluid (z_lower_fluid_top).  0 where a line meets no fluid.  Taken at
 * a time-step boundary (status 121 otherwise); blocks until copied. */
int sph_free_surface(sph_ctx* ctx, double* pelolib, double* zlft);
/* Running total of pair interactions, sum over every list built so far of
 * sum_i (nPartIntorno(i) - 1) (the unit of the pair-interactions/s metric). */
int sph_pair_counter(sph_ctx* ctx, unsigned long long* total);
/* Debug: neighbour lists in pg order, entries are pg indices (0-based), in the
 * reference's list order.  stride >= nmaxpartj. */
int sph_get_neighbour_lists(sph_ctx* ctx, int32_t* count, int32_t* ids, int32_t stride);
/* Debug: the SA-SPH boundary table (BoundaryDataTab, ComputeBoundaryDataTab_3D
 * .f90:62-106) in pg order: ncbf[r] close faces of row r; for slot q < ncbf[r]
 * at [r*stride + q]: face (0-based index into sph_faces), loc (3), bi (8:
 * BoundaryVolumeIntegrals_P0 order), ig (9, column-major).  stride >= 8. */
int sph_get_boundary_table(sph_ctx* ctx, int32_t* ncbf, int32_t* face, double* loc, double* bi,
                           double* ig, int32_t stride);

/* ---- phases, in time_step_loop order ------------------------------------ */
int sph_dt(sph_ctx* ctx);              /* time_step_duration.f90:59-123 */
int sph_momentum(sph_ctx* ctx);        /* RHS_momentum_equation.f90:72-198 */
int sph_leapfrog_mom(sph_ctx* ctx);    /* Leapfrog_momentum.f90:66-111 */
int sph_velocity_smoothing(sph_ctx* ctx); /* velocity_smoothing.f90:74-133 + _2.f90:59-103 */
int sph_leapfrog_traj(sph_ctx* ctx);   /* Leapfrog_trajectories.f90:48-76 */
int sph_sort(sph_ctx* ctx);            /* OrdGrid1.f90:58-105 */
int sph_neighbours(sph_ctx* ctx);      /* CalcVarLength.f90:111-541 (fluid part) */
int sph_boundary_tab(sph_ctx* ctx);    /* ComputeBoundaryDataTab_3D.f90:62-106 (2D: _2D.f90:71-162) */
int sph_continuity(sph_ctx* ctx);      /* time_step_loop.f90:338-360 */
int sph_leapfrog_cont(sph_ctx* ctx);   /* Leapfrog_continuity.f90:51-88 */
int sph_eos(sph_ctx* ctx);             /* EoS_fluid_loop.f90:59-70 */
int sph_pressure_smoothing(sph_ctx* ctx); /* PressureSmoothing_3D.f90:80-161 (2D: _2D.f90:95-231) */

/* Pre-loop of time_step_loop.f90:82-173: lists, boundary table, first dt. */
int sph_preloop(sph_ctx* ctx);
/* n whole time steps (time_step_loop.f90:175-489, leapfrog, semi), all on the
 * device; checks the device status flags once at the end. */
int sph_step(sph_ctx* ctx, int32_t nsteps);
/* Same, enqueued on the context stream without the final synchronisation
 * and flag check (for overlap and graph capture by the caller). */
int sph_step_async(sph_ctx* ctx, int32_t nsteps);
/* Same, as one HIP graph of nsteps whole steps (captured at the first call
 * with this nsteps, replayed after): one launch instead of ~20 per step. */
int sph_step_graph(sph_ctx* ctx, int32_t nsteps);
int sph_sync(sph_ctx* ctx);

/* ---- serving ------------------------------------------------------------- *
 * No counterpart in the reference (one case per process there).  An ensemble
 * server keeps one context per request, each with its own stream, admits a
 * request when sph_ctx_bytes fits the free device memory, enqueues its steps
 * with sph_step_async and polls sph_query, so requests join and leave while
 * others run. */
/* Device bytes a context created with *p would hold. */
int sph_ctx_bytes(const sph_params* p, unsigned long long* bytes);
/* Free and total device memory (hipMemGetInfo). */
int sph_device_memory(unsigned long long* free_bytes, unsigned long long* total_bytes);
/* busy = 1 while work enqueued on the context's stream is running, else 0. */
int sph_query(sph_ctx* ctx, int32_t* busy);

/* ---- segmented batches --------------------------------------------------- *
 * No counterpart in the reference (one case per process there).  A batch
 * context holds nseg independent solves (dam breaks of uneven particle
 * counts, each with its own grid, h, tables and dt) concatenated in one set of
 * device arrays: every kernel of a step is launched once over all segments,
 * the workgroups finding their segment from per-block maps, so one step (or
 * one graph replay, sph_step_graph) advances every solve.  Each segment's
 * arithmetic is the one-solve context's, so its results equal those of the
 * same solve alone.  p[s] and f[s] describe segment s as for sph_ctx_create;
 * all share precision and wall treatment; launch tunables come from p[0].
 * The phase entry points, sph_preloop and the sph_step family act on every
 * segment.  Not available on a batch: sph_global_quantities, the free-surface
 * monitor and the debug dumps (status 2). */
int sph_batch_create(int32_t nseg, const sph_params* p, const sph_faces* const* f, sph_ctx** out);
/* Device bytes a batch context of these segments would hold. */
int sph_batch_bytes(int32_t nseg, const sph_params* p, unsigned long long* bytes);
/* Per-segment state: as sph_state_upload / sph_state_download /
 * sph_get_scalars / sph_particle_count for segment seg.  A batch is loaded by
 * one sph_segment_upload per segment; the call that completes the set orders
 * all segments on their grids (OrdGrid1) and leaves the context ready for
 * sph_preloop. */
int sph_segment_upload(sph_ctx* ctx, int32_t seg, int32_t n, const double* coord, const double* vel,
                       const double* dens, const double* pres, const double* mass);
int sph_segment_download(sph_ctx* ctx, int32_t seg, double* coord, double* vel, double* acc,
                         double* dens, double* pres, double* dden, double* vpres, int32_t* ids);
int sph_segment_scalars(sph_ctx* ctx, int32_t seg, double out[5]);
int sph_segment_count(sph_ctx* ctx, int32_t seg, int32_t* n);

/* Launch shape in use: wg_size, list entries prefetched per step of the
 * FP32 list walks, grid of the particle kernels. */
int sph_get_launch(sph_ctx* ctx, int32_t out[3]);
/* Per-kernel timing of the last sph_step: name/avg-ms pairs, for the bench. */
int sph_set_timing(sph_ctx* ctx, int32_t on);
int sph_get_timing(sph_ctx* ctx, char* names, int32_t names_len, double* ms, int32_t* calls,
                   int32_t max_kernels);

const char* sph_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
