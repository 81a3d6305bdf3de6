// Device-side structures shared by the kernels and the host orchestration.
#pragma once
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
    IS_NPREV, IS_NRM, IS_RM_OVERFLOW, IS_COUNT
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
    int nodes, btype, mode, lam;
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
