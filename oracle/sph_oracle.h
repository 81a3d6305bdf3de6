/* CPU FP64 restatement of the SPHERA v10 particle-interaction hot path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker (or
 * as the timed CPU baseline).  The device library never links or calls it.
 *
 * Every routine follows one reference unit, in the reference's particle order
 * (pg order), with the reference's loop and summation order, so that on the
 * same inputs it reproduces the flang-built reference bit for bit (checked in
 * tests/test_oracle_reference.py against oracle/_ref/harness3d.x).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_params {
    int32_t n;              /* nag */
    int32_t nmaxpartj;      /* NMAXPARTJ */
    int32_t maxclose;       /* MAXCLOSEBOUNDFACES */
    int32_t ncd[3];
    double extr_lo[3], dcd[3];
    double dx, h, coefke, coefkacl, eta, eta2;
    double grav[3], prif;
    double den0, eps, celerita, alfamon, kin_visc, mumx;
    double cfl, vsc_coeff, pesodt, tetap, tetav;
    int32_t semi;           /* 1: SA-SPH walls active */
    int32_t density_thresholds; /* input_any_t%density_thresholds (1: 0.9/1.1 den0 clamp) */
    int32_t time_split;     /* Domain%time_split: 1 leapfrog, 0 explicit RK1 (RKscheme 1) */
} or_params;

/* nf faces: T[9*f] column-major (T(r,c) at [c*3+r]), LX[8*f] (4 nodes x 2),
 * GX[12*f] (4 nodes x 3), btype (0 fixe,1 tapi,2 open,3 sour,4 velo,5 flow),
 * velocity = BoundaryFace%velocity, tvel = Tratto%velocity. gcbf_start is
 * 1-based into gcbf_vec; gcbf_vec holds 0-based face indices. */
typedef struct or_faces {
    int32_t nf;
    const double *T, *LX, *GX, *area, *velocity, *tvel, *slip_coeff;
    const int32_t *nodes, *btype, *slip_mode, *lam_check;
    const int32_t *gcbf_count, *gcbf_start, *gcbf_vec;
    const double *kerneltab;  /* 41 x 5, row-major [row*5+col] */
    double ktdelta;
    const double *bitab;      /* 256 x 7, row-major */
    /* convex edges (FindBoundaryConvexEdges3D.f90 order): 2 face indices
     * (0-based) per edge, first node (x,y,z), second minus first, length */
    int32_t nedges;
    const int32_t *edge_face;
    const double *edge_gx, *edge_comp, *edge_len;
} or_faces;

void* or_create(const or_params* p, const or_faces* f, const double* coord, const double* vel,
                const double* dens, const double* pres, const double* mass);

/* 2D (SPACE_2D) boundary sides in BoundarySide order: T and RN column-major
 * 3x3 per side, v1 = first vertex (x, 0, z), btype codes as for faces
 * (6 = peri), previous = previous_side (1-based, 0 = none), velocity =
 * BoundarySide%velocity, tvel = Tratto%velocity, bit2d = BoundIntegralTab2D
 * (22 x 6, row-major).  The particles lie in the x-z plane (y = 0); the
 * grid has ncd[1] = 1; maxclose must be MAXCLOSEBOUNDSIDES (4).  v2 = second
 * vertex (x, 0, z); read for open sides (btype 2) only, may be NULL when no
 * side is open. */
typedef struct or_sides2d {
    int32_t ns;
    const double *T, *RN, *v1, *length, *angle, *velocity, *tvel, *slip_coeff;
    const int32_t *btype, *previous, *slip_mode, *lam_check;
    const double *bit2d;
    const double *v2;
} or_sides2d;

void* or_create2d(const or_params* p, const or_sides2d* s, const double* coord, const double* vel,
                  const double* dens, const double* pres, const double* mass);
void or_destroy(void* ctx);
int or_nag(void* ctx);

/* state exchange (pg order, xyz interleaved) */
void or_set_state(void* ctx, int n, const double* coord, const double* vel, const double* var,
                  const double* acc, const double* dens, const double* pres, const int32_t* laminar,
                  double dt, double dt_average, double time, int it);
void or_get_state(void* ctx, double* coord, double* vel, double* var, double* acc, double* dens,
                  double* pres, double* dden, double* vpres);
void or_get_scalars(void* ctx, double* out4); /* dt, dt_average, time, it */
int or_get_lists(void* ctx, int32_t* count, int32_t* ids, int32_t stride); /* ids 0-based */
/* first particle whose list (count, then entries in order) differs from the
 * oracle's, -1 if none (-2: lists not built); ids are 0-based, stride >= count */
int64_t or_compare_lists(void* ctx, const int32_t* count, const int32_t* ids, int32_t stride);
int or_get_btab(void* ctx, int32_t* ncbf, int32_t* face, double* loc, double* bi, double* ig);

/* phases (time_step_loop.f90 order); each returns 0 or a diagnostic code */
int or_ordgrid(void* ctx);          /* OrdGrid1 */
int or_calcvarlength(void* ctx);    /* CalcVarLength (fluid part) */
int or_boundary_tab(void* ctx);     /* ComputeBoundaryDataTab_3D (2D: _2D) */
int or_dt(void* ctx);               /* time_step_duration */
int or_momentum(void* ctx);         /* RHS_momentum_equation */
int or_leapfrog_mom(void* ctx, double dt_previous);
int or_vsmooth(void* ctx);          /* velocity_smoothing + velocity_smoothing_2 */
int or_leapfrog_traj(void* ctx);
int or_cancel_outgone(void* ctx);  /* CancelOutgoneParticles_3D: marks crossings of open faces */
long long or_outgone_count(void* ctx); /* particles marked by it since or_create */
int or_continuity(void* ctx);       /* SASPH_continuity + Continuity_Equation loop */
int or_leapfrog_cont(void* ctx);
int or_eos(void* ctx);
int or_psmooth(void* ctx);          /* PressureSmoothing_3D (2D: _2D) */
int or_preloop(void* ctx);          /* lists, boundary table, first dt */
int or_step(void* ctx, int nsteps); /* whole time steps (leapfrog, or RK1 when time_split = 0) */
int64_t or_pair_count(void* ctx);   /* sum_i (nPartIntorno(i)-1) of the current lists */
/* how often the rare branches ran since or_create: [0] elastic wall reactions
 * applied (AddElasticBoundaryReaction_3D/2D, vin < 0 below 0.25 dx), [1] zpmin
 * clamps (ComputeBoundaryVolumeIntegrals_P0.f90:84), [2] density_thresholds
 * clamps (Leapfrog_continuity.f90:61-68), [3] and [4] wall-function
 * interactions of the momentum equation below z0 e and on the logarithmic
 * branch (wall_function_for_SASPH.f90:69-78), [5] convex-edge reactions
 * applied (AddElasticBoundaryReaction_3D.f90:106-152) */
void or_branch_counts(void* ctx, int64_t* out6);
/* on != 0: or_calcvarlength builds the neighbour lists only (no rag or
 * PartKernel, about 1/15 of the memory), for list checks of large inputs;
 * the pair passes must not run afterwards */
void or_set_lists_only(void* ctx, int on);
int or_overflow(void* ctx);

/* DB-SPH walls (Domain%tipo = "bsph", 3D, explicit RK1): nw wall elements with
 * centroid, unit normal into the fluid, velocity (xyz interleaved), area
 * ("weight"), semi-particle volume and mass; flags = {Gamma_limiter_flag,
 * FS_allowed, negative_wall_p_allowed, MUSCL_boundary_flag}.  Called once after or_create (with
 * semi = 0 in or_params); or_preloop and or_step then run the DB-SPH branch.
 * Returns 2 if a wall element lies outside the grid. */
int or_set_dbsph(void* ctx, int nw, const double* coord, const double* normal, const double* vel,
                 const double* weight, const double* volume, const double* mass, const int32_t* flags);
/* Moving DB-SPH walls (DBSPH_kinematics.f90): nrec records (t, vx, vy, vz,
 * wx, wy, wz) of the mesh and its rotation centre; the first record's time
 * must not exceed 0.  Returns 2 otherwise. */
int or_set_dbsph_kinematics(void* ctx, int nrec, const double* records, const double* centre);
/* DB-SPH state in pg order (any pointer may be NULL): Gamma, sigma, rhoSPH_new,
 * dShep, drho (3), dvel (3x3 row-major), FS; wall elements' dens, pres, wet. */
void or_get_dbsph(void* ctx, double* gamma, double* sigma, double* rhosph, double* dshep, double* drho,
                  double* dvel, int32_t* fs, double* wall_dens, double* wall_pres, int32_t* wall_wet);
/* fluid-wall lists (0-based wall ids, the reference's order) */
int or_get_fw_lists(void* ctx, int32_t* count, int32_t* ids, int32_t stride);         /* particles whose list was truncated */

/* calc_pelo: free-surface position (pelolib) and lower-fluid-top height (zlft)
 * of each monitoring line, 3 values per line; icont: 2 per line, 1-based
 * first and last point as control_lines(:)%Icont; pts: 3 per point */
int or_free_surface(void* ctx, int nlines, const int32_t* icont, int npoints, const double* pts,
                    double* pelolib, double* zlft);

#ifdef __cplusplus
}
#endif
