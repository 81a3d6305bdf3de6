// Device-side structures shared by the kernels and the host orchestration.
#pragma once
#include <cmath>
#include <stdint.h>

#define SPH_MB 8  // boundary-table slots per particle (close SA-SPH faces)

namespace sph {

constexpr int NB_SHIFT = 27;
constexpr uint32_t NB_MASK = (1u << NB_SHIFT) - 1u;
// neighbour-list layout: entry k of particle p at ((k/NB_G)*N + p)*NB_G + k%NB_G,
// i.e. groups of NB_G consecutive entries per particle, the groups of one
// rank k/NB_G contiguous over p (32-byte vector stores and loads per lane)
#ifndef SPH_NB_GROUP
#define SPH_NB_GROUP 8
#endif
constexpr int NB_G = SPH_NB_GROUP;
__host__ __device__ inline size_t nb_index(size_t N, int p, int k) {
    const unsigned u = (unsigned)k;
    return ((size_t)(u / NB_G) * N + (size_t)p) * NB_G + (size_t)(u % NB_G);
}
// list capacity per particle: NMAXPARTJ rounded up to whole groups
__host__ __device__ inline int nb_stride(int nmaxpartj) { return (nmaxpartj + NB_G - 1) / NB_G * NB_G; }

// indices into the device scalar blocks
enum { SC_DT = 0, SC_DTPREV, SC_DTAVG, SC_TIME, SC_DTVEL, SC_TV1, SC_TP1, SC_COUNT };
enum {
    IS_N = 0, IS_IT, IS_ONGOING, IS_NB_OVERFLOW, IS_BT_OVERFLOW, IS_REMOVED, IS_UNSUPPORTED,
    IS_NPREV, IS_NRM, IS_RM_OVERFLOW, IS_DB_OVERFLOW, IS_COUNT
};

// most particles one sort may remove from the grid (renumbering list)
constexpr int RM_CAP = 4096;

// per-context constants, passed by value to every kernel
struct DP {
    int cap, nmaxpartj, nmax, semi, f32, nf;
    int ncd[3];
    double extr_lo[3], dcd[3];
    double h, dx, coefke, coefkacl, eta, eta2, squareh, square_doubleh, doubleh;
    double ke_coef, kacl_coef, ktdelta, GI;
    double grav[3], prif, den0, eps, celerita, alfamon, kin_visc;
    double cfl, vsc_coeff, pesodt, tetap, tetav;
    // FP32-mixed copies
    float f_h, f_unosuh, f_unosusquareh, f_eta, f_eta2, f_ke_coef, f_kacl_coef, f_coefke;
    float f_dcd[3], f_alfamon, f_celerita, f_kin_visc;
    // the density slot of the FP32 velocity records holds rho - den0; the
    // linear EoS p = eps/den0 (rho - den0) + prif gives the pressure from it
    float f_den0, f_eosk, f_prif;
    // every particle of the solve has the mass f_mass (one medium, one dx:
    // SetParticleParameters.f90:56-57): the pressure smoothing then reads one
    // (x, y, z, rho - den0) record per neighbour
    int umass;
    float f_mass;
    // density_thresholds == 1: Leapfrog_continuity.f90:61-68 clamps the
    // integrated density to [0.9, 1.1] den0
    int dthr;
    // time_split = 0, RKscheme = 1: the explicit RK1 (Euler) step order
    int rk1;
    // 3, or 2 for the x-z plane form (-DSPACE_2D build of the reference): one
    // cell row in y (every particle in row j = 1), boundary sides for faces
    int dim;
    // Placement of this solve in a segmented context (several independent
    // solves concatenated in one set of arrays; all zero for a context that
    // holds one solve).  cap is the capacity of the whole context, so list
    // and boundary-table strides are the same for every segment.
    int seg;   // segment index
    int base;  // first sorted slot of the segment (a multiple of 256)
    int scap;  // slots reserved for the segment
    int koff;  // first cell key: keys koff .. koff+nmax-1 are the cells, koff+nmax the
               // particles out of the grid (a multiple of 256)
    int roff;  // first cell row (fixed j, k) of the segment in the context's row list
    int foff;  // first frontier face of the segment in the context's face array
    int eoff;  // first convex edge of the segment in the context's edge array
    int ne;    // its convex edges (NumBEdges)
    int nopen; // its open faces (NumOpenFaces)
    // DB-SPH walls (Domain%tipo = "bsph"): nw wall elements; the DBSPH
    // section's Gamma_limiter_flag, FS_allowed and negative_wall_p_allowed
    int dbsph, nw, gamma_lim, fs_allowed, neg_wall_p, muscl_bnd, nkin;
    double kin_c[3];  // rotation centre of the wall kinematics
};

// Segment maps of a segmented context, passed by value to the kernels that
// take a segmented form.  Slots, keys and rows are partitioned in whole
// 256-blocks (slots, keys) or whole rows, so a workgroup of a particle or key
// kernel lies in one segment.
struct Seg {
    const DP* P;        // per-segment constants
    const int* pblk;    // segment of each 256-slot block
    const int* kblk;    // segment of each 256-key block
    const int* rseg;    // segment of each cell row
    int nseg, nrows;
};

// one convex frontier edge (BoundaryConvexEdge, FindBoundaryConvexEdges3D.f90):
// its faces (indices into the context's face array), first node, the vector
// to the second node and its length
struct DEdge {
    double gx[3], comp[3], len;
    int face[2];
};

// one SA-SPH frontier face (BoundaryFace + its Tratto/Partz data)
struct DFace {
    double T[9];    // column-major, T[6..8] = normal
    double LX[8];   // node local coordinates (x,y) x 4
    double GX[12];  // node global coordinates x 4
    double area, slip;
    double vel[3], tvel[3];
    // IsPointInternal's per-face constants (1/DueArea and the AA, BB, CC of
    // its two edges), evaluated once with the same expressions
    double pu, pA[2], pB[2], pC[2];
    // |grad| of the area coordinates csi0, csi1 and csi2 = 1 - csi0 - csi1
    // over the face plane (bounds of a coordinate over a circle)
    double pg[3];
    int nodes, btype, mode, lam;
    // 2D context: the array holds the boundary sides (BoundarySide of
    // DefineBoundarySideGeometry2D.f90) and these fields: RN(1,1), RN(1,3),
    // RN(3,1), RN(3,3) of BoundaryMassForceMatrix2D, the first vertex (x, z),
    // length, angle to the previous side and previous_side (1-based, 0 none),
    // the second vertex (x, z) of an open side (CancelOutgoneParticles_2D).
    // T, vel, tvel, slip, btype, mode and lam keep their meaning.
    double RN[4];
    double v1[2], v2[2];
    double len, angle;
    int prev;
};

// LocalNormalCoordinates.f90:55-87 constants of a face: node pairs (jj,kk) of
// the two area coordinates, DueArea = (3-fkod) area
inline void dface_point_coeffs(DFace& F) {
    const int fkod = F.nodes - 2;
    const double DueArea = (double)(3 - fkod) * F.area;
    F.pu = 1.0 / DueArea;
    const int a[2] = {fkod == 1 ? 1 : 2, fkod == 1 ? 2 : 3};
    const int b[2] = {fkod == 1 ? 2 : 3, 0};
    for (int e = 0; e < 2; ++e) {
        const double xj = F.LX[2 * a[e]], yj = F.LX[2 * a[e] + 1];
        const double xk = F.LX[2 * b[e]], yk = F.LX[2 * b[e] + 1];
        F.pA[e] = xj * yk - xk * yj;
        F.pB[e] = yj - yk;
        F.pC[e] = xk - xj;
    }
    for (int e = 0; e < 2; ++e) F.pg[e] = F.pu * std::sqrt(F.pB[e] * F.pB[e] + F.pC[e] * F.pC[e]);
    const double sb = F.pB[0] + F.pB[1], sc = F.pC[0] + F.pC[1];
    F.pg[2] = F.pu * std::sqrt(sb * sb + sc * sc);
}

// BoundIntegralTab (ComputeBoundaryIntegralTab.f90) as the FP32-mixed
// quadrature reads it.  The table holds BT_RINGS rings of BT_RDIRS
// directions; within a ring every direction has the same vertical offset
// b_z, the same weight da and the same |b_xy|.
constexpr int BT_RINGS = 8, BT_RDIRS = 32;

struct RingC {
    double bz, nr, rho;           // b_z, -1/b_z, |b_xy|
    float ck, da, R2;             // |b|/(-b_z h), weight, R_z
    float s0, s1, s00, s01, s11;  // sums over the ring of R0, R1, R0^2, R0 R1, R1^2
    int ok;                       // the ring's directions share b_z, da and |b_xy|
    int pad;
};
struct BitF64 {
    double b0, b1, b2, nr;  // offset, and -1/b2 (0 where b2 >= 0)
};
struct alignas(16) BtF32 {
    RingC rc[BT_RINGS];
    BitF64 bd[BT_RINGS * BT_RDIRS];
    float bf[BT_RINGS * BT_RDIRS][4];  // |b|/(-b_z h), da, R0, R1
    float br2[BT_RINGS * BT_RDIRS];    // R_z
    float kt[41 * 5];                  // kernel table, FP32
    float kt_inv_delta;
    int all_ok;
};

// bit: 256 x 7 (offset x3, da, R x3), kt: 41 x 5
inline void build_btf32(const double* bit, const double* kt, double ktdelta, double h, BtF32& T) {
    T.all_ok = 1;
    for (int r = 0; r < BT_RINGS; ++r) {
        const double* b = &bit[r * BT_RDIRS * 7];
        const double bz = b[2], da = b[3];
        const double rho = std::sqrt(b[0] * b[0] + b[1] * b[1]);
        int ok = bz < 0.0;
        double s0 = 0.0, s1 = 0.0, s00 = 0.0, s01 = 0.0, s11 = 0.0;
        for (int k = 0; k < BT_RDIRS; ++k) {
            const double* q = &b[k * 7];
            const double rk = std::sqrt(q[0] * q[0] + q[1] * q[1]);
            ok &= q[2] == bz && q[3] == da && std::fabs(rk - rho) <= 1.0e-12 * rho;
            s0 += q[4];
            s1 += q[5];
            s00 += q[4] * q[4];
            s01 += q[4] * q[5];
            s11 += q[5] * q[5];
        }
        const double nr = bz < 0.0 ? -1.0 / bz : 0.0;
        const double blen = std::sqrt(rho * rho + bz * bz);
        T.rc[r] = RingC{bz, nr, rho, (float)(blen * nr / h), (float)da, (float)b[6], (float)s0,
                        (float)s1, (float)s00, (float)s01, (float)s11, ok, 0};
        T.all_ok &= ok;
    }
    for (int t = 0; t < BT_RINGS * BT_RDIRS; ++t) {
        const double* b = &bit[t * 7];
        const double nr = b[2] < 0.0 ? -1.0 / b[2] : 0.0;
        const double blen = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        T.bd[t] = BitF64{b[0], b[1], b[2], nr};
        T.bf[t][0] = (float)(blen * nr / h);
        T.bf[t][1] = (float)b[3];
        T.bf[t][2] = (float)b[4];
        T.bf[t][3] = (float)b[5];
        T.br2[t] = (float)b[6];
    }
    for (int t = 0; t < 41 * 5; ++t) T.kt[t] = (float)kt[t];
    T.kt_inv_delta = (float)(1.0 / ktdelta);
}

// one BoundaryDataTab record (Hybrid_allocation_module.f90:796-805)
struct BTab {
    int face;
    int pad;
    double loc[3];
    double bi[8];
    double ig[9];
};

}  // namespace sph
