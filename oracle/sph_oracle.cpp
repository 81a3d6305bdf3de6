// CPU FP64 restatement of the SPHERA v10 hot path (oracle; test infrastructure).
// Build: see oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off, no -march, so
// no FMA contraction, like the flang reference build).
//
// Particle order, loop order and the left-to-right evaluation of every FP
// expression follow the cited reference lines.  Indices are 0-based here; the
// reference's 1-based cell ids are kept as such (0 = outside the grid).
#include <atomic>
#include <cstdint>
#include "sph_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

const double GI = (double)9.80665f;  // Static_allocation_module.f90:75 (a default-real literal)
const double MinTotUnit = 0.95;      // PressureSmoothing_3D.f90:33

struct Ctx {
    or_params p;
    // faces
    int nf = 0;
    std::vector<double> T, LX, GX, area, fvel, tvel, slip_coeff;
    std::vector<int> nodes, btype, slip_mode, lam_check;
    // convex frontier edges (BoundaryConvexEdge)
    std::vector<int> edge_face;
    std::vector<double> edge_gx, edge_comp, edge_len;
    std::vector<int> gcbf_count, gcbf_start, gcbf_vec;
    double kt[41][5];
    double ktdelta = 0.0;
    double bit[256][7];
    // derived
    double squareh, doubleh, square_doubleh;
    int nmax;
    // state, pg order
    int n;
    std::vector<double> coord, vel, var, acc, dens, pres, dden, vpres, mass;
    std::vector<double> coord_old;  // CoordOld: the position before the last trajectory update
    std::vector<int> cella, kodvel, koddens, laminar;
    std::vector<int> id;  // not used by the physics; tracks swap-with-last removals
    // lists
    std::vector<int> icont, npartord, nnb, nbr;
    std::vector<double> rag, pk;
    // boundary table
    std::vector<int> bn, bface;
    std::vector<double> bloc, bbi, big;
    // 2D (x-z plane): boundary sides instead of faces
    int dim = 3;
    int ns = 0;
    std::vector<double> sT, sRN, sv1, sv2, slen, sang, stvel, sfvel, sslip;
    std::vector<int> sbtype, sprev, sslipmode, slam;
    double bit2d[22][6];
    double dt = 0.0, dt_average = 0.0, time = 0.0;
    int it = 0;
    int on_going = -999;
    int overflow = 0;
    bool lists_built = false;
    // how often the rare branches ran (checker counters, read by or_branch_counts):
    // elastic wall reactions applied, zpmin clamps of the volume integrals,
    // density_thresholds clamps, wall-function interactions below z0 e and above
    mutable std::atomic<long long> n_elastic{0}, n_zpmin{0}, n_dclamp{0}, n_wf_near{0}, n_wf_log{0},
        n_edge{0};
    long long n_outgone = 0;  // particles removed through open faces (OpCount)
    // lists only: CalcVarLength stores the neighbour lists without rag and
    // PartKernel (list checks of large clouds; the passes cannot run)
    bool lists_only = false;
    // DB-SPH walls (Domain%tipo = "bsph"): wall elements pg_w(1:n_w), their
    // cell lists (Icont_w, NPartOrd_w), the particles' DB-SPH state and the
    // fluid-wall lists (nPartIntorno_fw, PartIntorno_fw, rag_fw, kernel_fw)
    bool dbsph = false;
    int nw = 0;
    std::vector<double> wcoord, wnormal, wvel, wweight, wvolume, wmass, wdens, wpres;
    std::vector<int> wwet, icont_w, npartord_w;
    int gamma_limiter = 0, fs_allowed = 1, neg_wall_p = 0, muscl_boundary = 0;
    std::vector<double> kin;  // kinematics records (t, v, omega) of the one mesh
    double kin_centre[3] = {0.0, 0.0, 0.0};
    std::vector<double> gamma, sigma, rhosph, rhosph_old, dshep, drho, dvelg;
    std::vector<int> fs, nfw, fwl;
    std::vector<double> rag_fw, kfw;
};

inline double Tm(const Ctx& c, int f, int r, int col) { return c.T[9 * f + col * 3 + r]; }

// CellNumber.f90:53-59 (1-based i,j,k)
inline int cell_number(const Ctx& c, int i, int j, int k) {
    const int ni = c.p.ncd[0], nj = c.p.ncd[1], nk = c.p.ncd[2];
    if (i < 1 || i > ni || j < 1 || j > nj || k < 1 || k > nk) return 0;
    return ((k - 1) * nj + (j - 1)) * ni + i;
}

// ParticleCellNumber.f90:52-62
inline int particle_cell(const Ctx& c, const double* x) {
    const double xp = x[0] - c.p.extr_lo[0];
    const double yp = x[1] - c.p.extr_lo[1];
    const double zp = x[2] - c.p.extr_lo[2];
    const int i = (int)std::ceil(xp / c.p.dcd[0]);
    const int k = (int)std::ceil(zp / c.p.dcd[2]);
    const int j = c.dim == 2 ? 1 : (int)std::ceil(yp / c.p.dcd[1]);  // :59-61 (2D: j = 1)
    return cell_number(c, i, j, k);
}

// CellIndices.f90:51-63
inline void cell_indices(const Ctx& c, int nc, int& i, int& j, int& k) {
    const int ni = c.p.ncd[0], nj = c.p.ncd[1];
    const int nij = ni * nj;
    k = (nc - 1) / nij + 1;
    const int ncij = nc - nij * (k - 1);
    j = (ncij - 1) / ni + 1;
    i = ncij - ni * (j - 1);
}

// LocalNormalCoordinates.f90:55-87 (PX(1:2) used) + IsPointInternal.f90:54-72
inline bool point_internal(const Ctx& c, int f, double px0, double px1) {
    const int nodes = c.nodes[f];
    const int fkod = nodes - 2;
    static const int iseg[3] = {1, 2, 0};
    static const int mainod[2][3] = {{0, 1, 2}, {0, 2, 3}};
    const double DueArea = (double)(3 - fkod) * c.area[f];
    const double Usu = 1.0 / DueArea;
    double csi[3];
    for (int ii = 0; ii < 2; ++ii) {
        const int jj = iseg[ii];
        const int kk = iseg[jj];
        const double* lj = &c.LX[8 * f + 2 * mainod[fkod - 1][jj]];
        const double* lk = &c.LX[8 * f + 2 * mainod[fkod - 1][kk]];
        const double xj = lj[0], yj = lj[1], xk = lk[0], yk = lk[1];
        const double AA = xj * yk - xk * yj;
        const double BB = yj - yk;
        const double CC = xk - xj;
        csi[ii] = (AA + BB * px0 + CC * px1) * Usu;
    }
    csi[2] = 1.0 - (csi[0] + csi[1]);
    if (fkod == 1) {
        for (int ii = 0; ii < 3; ++ii)
            if (csi[ii] < 0.0 || csi[ii] > 1.0) return false;
        return true;
    }
    if (fkod == 2) {
        for (int ii = 0; ii < 2; ++ii)
            if (csi[ii] < 0.0 || csi[ii] > 1.0) return false;
        return true;
    }
    return false;
}

// InterpolateTable.f90:58-82 (columns 1..4)
inline void interpolate_table(const Ctx& c, double xval, double out[4]) {
    const double deltaX = c.ktdelta;
    const double xval0 = c.kt[0][0];
    const int nr0 = (int)((xval - xval0) / deltaX);
    if (nr0 <= 0) {
        for (int ic = 0; ic < 4; ++ic) out[ic] = c.kt[0][ic + 1];
    } else if (nr0 >= 40) {
        for (int ic = 0; ic < 4; ++ic) out[ic] = c.kt[40][ic + 1];
    } else {
        const int nr1 = nr0 + 1;
        const double csi = (xval - c.kt[nr0][0]) / deltaX;
        for (int ic = 0; ic < 4; ++ic)
            out[ic] = c.kt[nr0][ic + 1] + csi * (c.kt[nr1][ic + 1] - c.kt[nr0][ic + 1]);
    }
}

// ComputeBoundaryVolumeIntegrals_P0.f90:76-155
void volume_integrals(const Ctx& c, int f, const double locx[3], double bi[8], double ig[9]) {
    for (int q = 0; q < 8; ++q) bi[q] = 0.0;
    for (int q = 0; q < 9; ++q) ig[q] = 0.0;
    double IntWdV = 0.0, IntdWrm1dV = 0.0, IntGWZrm1dV = 0.0;
    double IntGWdV[3] = {0.0, 0.0, 0.0};
    double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double J0 = 0.0, J1 = 0.0, J2 = 0.0, J3 = 0.0;
    const double zpmin = 0.05 * c.p.h;
    double LocPi[3] = {locx[0], locx[1], locx[2]};
    if (LocPi[2] < 0.0) return;
    if (LocPi[2] < zpmin) {
        LocPi[2] = zpmin;
        c.n_zpmin.fetch_add(1, std::memory_order_relaxed);
    }
    double robpre = 2.1;
    for (int ipk = 0; ipk < 256; ++ipk) {
        double LocPj[3];
        for (int d = 0; d < 3; ++d) LocPj[d] = LocPi[d] + c.bit[ipk][d];
        if (LocPj[2] <= 0.0) {
            const double ts = LocPi[2] / (LocPi[2] - LocPj[2]);
            double PX[3];
            for (int d = 0; d < 3; ++d) PX[d] = LocPi[d] + (LocPj[d] - LocPi[d]) * ts;
            PX[2] = 0.0;
            if (point_internal(c, f, PX[0], PX[1])) {
                double dist = 0.0;
                for (int d = 0; d < 3; ++d) {
                    const double dd = PX[d] - LocPi[d];
                    dist = dist + dd * dd;
                }
                dist = std::sqrt(dist);
                const double da = c.bit[ipk][3];
                const double rob = dist / c.p.h;
                if (std::fabs(rob - robpre) > 1.0e-3) {
                    double iv[4];
                    interpolate_table(c, rob, iv);
                    J0 = iv[0] * da;
                    J1 = iv[1] * da;
                    J2 = iv[2] * da;
                    J3 = iv[3] * da;
                    robpre = rob;
                }
                const double R[3] = {c.bit[ipk][4], c.bit[ipk][5], c.bit[ipk][6]};
                const double RZ2 = R[2] * R[2];
                IntWdV = IntWdV + J0;
                IntdWrm1dV = IntdWrm1dV + J1;
                IntGWZrm1dV = IntGWZrm1dV + RZ2 * J1;
                for (int sd = 0; sd < 3; ++sd) {
                    IntGWdV[sd] = IntGWdV[sd] + R[sd] * J2;
                    for (int sj = sd; sj < 3; ++sj) G[sd][sj] = G[sd][sj] + R[sd] * R[sj] * J3;
                }
            }
        }
    }
    G[1][0] = G[0][1];
    G[2][0] = G[0][2];
    G[2][1] = G[1][2];
    if (LocPi[2] == zpmin) IntWdV = point_internal(c, f, LocPi[0], LocPi[1]) ? 0.5 : 0.0;
    bi[0] = 0.0;
    bi[1] = IntWdV;
    bi[2] = IntdWrm1dV;
    bi[3] = IntGWdV[0];
    bi[4] = IntGWdV[1];
    bi[5] = IntGWdV[2];
    bi[6] = IntGWZrm1dV;
    bi[7] = 0.0;
    // IntGiWrRdV(r,c) column-major, as the reference stores it
    for (int col = 0; col < 3; ++col)
        for (int r = 0; r < 3; ++r) ig[col * 3 + r] = G[r][col];
}

// FindCloseBoundaryFaces3D.f90:78-188 + ComputeBoundaryDataTab_3D.f90:62-106
void boundary_tab_particle(Ctx& c, int npi) {
    c.bn[npi] = 0;
    if (c.cella[npi] <= 0) return;
    const double* x = &c.coord[3 * npi];
    int nc = particle_cell(c, x);
    if (nc <= 0) return;
    int ic, jc, kc;
    cell_indices(c, nc, ic, jc, kc);
    const int maxc = c.p.maxclose;
    int* clob = &c.bface[(size_t)npi * maxc];
    double* locx = &c.bloc[(size_t)npi * maxc * 3];
    int ncbf = 0;
    for (int i = ic - 1; i <= ic + 1; ++i)
        for (int j = jc - 1; j <= jc + 1; ++j)
            for (int k = kc - 1; k <= kc + 1; ++k) {
                const int cn = cell_number(c, i, j, k);
                if (cn == 0) continue;
                const int nfc = c.gcbf_count[cn - 1];
                if (nfc <= 0) continue;
                const int ini = c.gcbf_start[cn - 1] - 1;
                for (int fl = ini; fl < ini + nfc; ++fl) {
                    const int f = c.gcbf_vec[fl];
                    double pinmin = 0.0;
                    const double pinmax = c.doubleh;
                    const int bt = c.btype[f];
                    if (bt == 3) pinmin = -c.doubleh;
                    bool there = false;
                    for (int q = 0; q < ncbf; ++q)
                        if (clob[q] == f) there = true;
                    if (there) continue;
                    const int nn = c.nodes[f];
                    const double* ref = &c.GX[12 * f + 3 * (nn - 1)];
                    double PX[3];
                    for (int sdi = 0; sdi < 3; ++sdi) {
                        double s = 0.0;
                        for (int sdj = 0; sdj < 3; ++sdj) s = s + Tm(c, f, sdj, sdi) * (x[sdj] - ref[sdj]);
                        PX[sdi] = s;
                    }
                    const double pin = PX[2];
                    if (pin >= pinmin && pin < pinmax) {
                        const bool needs_in = (bt == 3 || bt == 4 || bt == 5);
                        if (!needs_in || point_internal(c, f, PX[0], PX[1])) {
                            ncbf = ncbf + 1;
                            if (ncbf > maxc) { c.bn[npi] = -1; return; }
                            clob[ncbf - 1] = f;
                            locx[3 * (ncbf - 1) + 2] = pin;
                        }
                    }
                }
            }
    for (int q = 0; q < ncbf; ++q) {
        const int f = clob[q];
        const int nn = c.nodes[f];
        const double* ref = &c.GX[12 * f + 3 * (nn - 1)];
        for (int sdi = 0; sdi < 2; ++sdi) {
            double s = 0.0;
            for (int sdj = 0; sdj < 3; ++sdj) s = s + Tm(c, f, sdj, sdi) * (x[sdj] - ref[sdj]);
            locx[3 * q + sdi] = s;
        }
    }
    c.bn[npi] = ncbf;
    for (int q = 0; q < ncbf; ++q)
        volume_integrals(c, clob[q], &locx[3 * q], &c.bbi[((size_t)npi * maxc + q) * 8],
                         &c.big[((size_t)npi * maxc + q) * 9]);
}


// ---------------------------------------------------------------------------
// 2D (SPACE_2D build): boundary sides in the x-z plane.  Side s has T(r,c) at
// sT[9*s + 3*c + r] (column-major, 0-based r,c), first vertex sv1[3*s..].
// ---------------------------------------------------------------------------
const double PIGRECO = 3.14159265358979;         // Static_allocation_module.f90:50
const double KERNELCONST2D = 0.454728408833987;  // Static_allocation_module.f90:58
const double DELTAX_BIT = 0.1;                   // SA_SPH_module.f90:28

inline double ST(const Ctx& c, int s, int r, int col) { return c.sT[9 * s + col * 3 + r]; }
inline double SRN(const Ctx& c, int s, int r, int col) { return c.sRN[9 * s + col * 3 + r]; }

// wall_function_for_SASPH.f90:56-80 (no land-cover map): slip coefficient and
// mixing-length viscosity of one particle-frontier pair at distance r; d50 is
// ten times the frontier's roughness length (the callers pass
// BC_shear_stress_input * 10, AddBoundaryContributions_to_ME3D.f90:222-224).
// Nepero_number and k_v: Static_allocation_module.f90:51,53.
// Returns 0 outside the wall layer, 1 below z0 e, 2 on the logarithmic branch.
int wall_function(double u_t, double d50, double r, double dx, double& slip, double& nut) {
    const double e = 2.718281828, kv2 = 0.41 * 0.41;
    slip = 0.0;
    nut = 1.0e-15;
    if (r <= 0.5 * dx) {
        const double z0 = d50 / 10.0;
        if (r < z0 * e) {
            slip = 1.0;
            nut = kv2 * u_t * z0 * e;
            return 1;
        }
        slip = 1.0 / std::log(r / z0);
        nut = kv2 * u_t * r / std::log(r / z0);
        return 2;
    }
    return 0;
}

void count_wall_function(const Ctx& c, int branch) {
    if (branch == 1) c.n_wf_near.fetch_add(1, std::memory_order_relaxed);
    if (branch == 2) c.n_wf_log.fetch_add(1, std::memory_order_relaxed);
}

// w.f90:48-57
double w_fn(double r, double h, double coef) {
    const double sq = r / h;
    double q;
    if (sq <= 1.0) {
        q = 0.666666667 + sq * sq * (sq * 0.5 - 1.0);
    } else if (sq >= 2.0) {
        q = 0.0;
    } else {
        const double dms = 2.0 - sq;
        q = dms * dms * dms / 6.0;
    }
    return q * coef;
}

// WIntegr.f90:60-82
double wintegr(double ri, double hh) {
    const double ro = ri / hh;
    if (0.0 <= ro && ro < 1.0) {
        const double ro2 = ro * ro, ro3 = ro2 * ro;
        const double w1 = (0.5 + -0.375 * ro2 + 0.15 * ro3) * ro2;
        return KERNELCONST2D * (0.35 - w1);
    } else if (1.0 <= ro && ro < 2.0) {
        const double q1 = 2.0 - ro, q2 = q1 * q1, q4 = q2 * q2;
        const double w2 = (-0.125 + 0.05 * q1) * q4;
        return -KERNELCONST2D * w2;
    }
    return 0.0;
}

// J2Wro2.f90:55-70
double j2wro2(double ro) {
    const double a1 = 0.00833333333333333, a2 = 0.26666666666666667;
    const double ro2 = ro * ro, ro3 = ro2 * ro;
    if (0.0 <= ro && ro < 1.0) return KERNELCONST2D * (0.25 - (a1 * (40.0 - 36.0 * ro2 + 15.0 * ro3) * ro3));
    if (1.0 <= ro && ro < 2.0)
        return KERNELCONST2D * (a2 - (a1 * (80.0 - 90.0 * ro + 36.0 * ro2 - 5.0 * ro3) * ro3));
    return 0.0;
}

// InterpolateBoundaryIntegrals2D.f90:61-70 (cols: 1-based columns 1..5)
void interp2d(const Ctx& c, int ncols, const int* cols, double xx, double* func) {
    if (xx < c.bit2d[0][0]) xx = c.bit2d[0][0];
    const int ii = (int)((xx - c.bit2d[0][0]) / DELTAX_BIT) + 1;
    const double xi = c.bit2d[ii - 1][0];
    for (int nc = 0; nc < ncols; ++nc) {
        const int jj = cols[nc];
        const double fi = c.bit2d[ii - 1][jj], fip1 = c.bit2d[ii][jj];
        func[nc] = fi + (fip1 - fi) * (xx - xi) / DELTAX_BIT;
    }
}

// FindBoundaryIntersection2D.f90:59-86
void intersection2d(const Ctx& c, double xpi, double ypi, double sidelen, double& xpmin,
                    double& xpmax, double& interlen) {
    const double minlen = 0.01 * c.p.h;
    if (ypi < 0.0) ypi = 0.0;
    const double halfchord = std::sqrt(c.square_doubleh - ypi * ypi);
    const double XA = xpi - halfchord, XB = xpi + halfchord;
    xpmin = 0.0;
    if (XA > xpmin) xpmin = XA;
    xpmax = sidelen;
    if (XB < xpmax) xpmax = XB;
    interlen = xpmax - xpmin;
    if (interlen <= minlen) interlen = 0.0;
}

// ComputeSurfaceIntegral_WdS2D.f90:57-85
double surface_wds2d(const Ctx& c, double xpi, double ypi, double xpmin, double interlen) {
    double sum = 0.0;
    const double dsrif = c.p.h / 2.0;
    if (interlen > 0.0) {
        if (ypi < 0.0) ypi = 0.0;
        const double ypiq = ypi * ypi;
        int ndiv = (int)(interlen / dsrif + 0.5);
        if (ndiv < 2) ndiv = 2;
        const double deltas = interlen / ndiv;
        double xpip = xpmin - 0.5 * deltas;
        for (int ipt = 1; ipt <= ndiv; ++ipt) {
            xpip = xpip + deltas;
            const double dxpip = xpip - xpi;
            const double risq = dxpip * dxpip + ypiq;
            const double ris = std::sqrt(risq);
            sum = sum + w_fn(ris, c.p.h, c.p.coefke) * deltas;
        }
    }
    return sum;
}

// BoundaryVolumeIntegrals2D.f90:60-97
void volume_bvi2d(const Ctx& c, double xpi, double ypi, double xpmin, double xpmax, double interlen,
                  double& IntWdV, double* IntDpWdV) {
    const double dalfarif = PIGRECO / 10;
    IntWdV = 0.0;
    IntDpWdV[0] = IntDpWdV[1] = 0.0;
    if (interlen <= 0.0) return;
    const double yplimite = 0.05 * c.p.h;
    if (ypi < yplimite) ypi = yplimite;
    const double alfaA = std::atan2(xpmin - xpi, ypi), alfaB = std::atan2(xpmax - xpi, ypi);
    const double intalfa = alfaB - alfaA;
    int ndiv = (int)(intalfa / dalfarif + 0.5);
    if (ndiv < 1) ndiv = 1;
    const double dalfa = intalfa / ndiv;
    double alfa_k = alfaA - 0.5 * dalfa;
    for (int k = 1; k <= ndiv; ++k) {
        alfa_k = alfa_k + dalfa;
        const double sa = std::sin(alfa_k), ca = std::cos(alfa_k);
        const double rb = ypi / ca;
        const double rob = rb / c.p.h;
        const double mult = c.p.h * j2wro2(rob) * dalfa;
        IntDpWdV[0] = IntDpWdV[0] + sa * mult;
        IntDpWdV[1] = IntDpWdV[1] - ca * mult;
        IntWdV = IntWdV + wintegr(rb, c.p.h) * dalfa;
    }
    if (ypi == yplimite) IntWdV = 0.5;
}

// ComputeVolumeIntegral_WdV2D.f90:69-132 (called with Ncbslocal = 2)
int volume_wdv2d(const Ctx& c, int icbs, const int* side, const double* loc, double xpmin,
                 double xpmax, double interlen, double& V) {
    const double dalfarif = PIGRECO / 10;
    V = 0.0;
    if (interlen <= 0.0) return 0;
    const double yplimite = 0.05 * c.p.h;
    const double xpi = loc[2 * icbs], ypi = loc[2 * icbs + 1];
    if (ypi >= yplimite) {
        const double alfaA = std::atan2(xpmin - xpi, ypi), alfaB = std::atan2(xpmax - xpi, ypi);
        const double intalfa = alfaB - alfaA;
        int ndiv = (int)(intalfa / dalfarif + 0.5);
        if (ndiv < 2) ndiv = 2;
        const double dalfa = intalfa / ndiv;
        double alfa = alfaA - 0.5 * dalfa;
        for (int ipt = 1; ipt <= ndiv; ++ipt) {
            alfa = alfa + dalfa;
            const double ta = std::tan(alfa);
            const double ris = ypi * std::sqrt(1.0 + ta * ta);
            V = V + wintegr(ris, c.p.h) * dalfa;
        }
    } else {
        const int jcbs = 1 - icbs;
        const double ypj = loc[2 * jcbs + 1];
        if (ypj <= yplimite) {
            double angle = 0.0;
            if (c.sprev[side[0]] == side[1] + 1)
                angle = c.sang[side[0]];
            else if (c.sprev[side[1]] == side[0] + 1)
                angle = c.sang[side[1]];
            else
                return 84;  // diagnostic(8,4): sides not consecutive
            V = 0.5 * (angle + PIGRECO) / (2.0 * PIGRECO);
        } else {
            V = (xpmin >= xpi && xpi <= xpmax) ? 0.5 : 0.0;
        }
    }
    return 0;
}

// FindCloseBoundarySides2D.f90:68-170, SelectCloseBoundarySides2D.f90:66-120,
// ComputeBoundaryDataTab_2D.f90:71-162 for one particle.  bn = IntNcbs, or
// -code on error (88: more than MAXCLOSEBOUNDSIDES sides, 95: a side type
// other than fixe/tapi, 89: more than two selected sides, 84: sides not
// consecutive).
void boundary_tab_particle_2d(Ctx& c, int npi) {
    const size_t mc = c.p.maxclose;
    c.bn[npi] = 0;
    if (c.cella[npi] <= 0) return;
    const double* X = &c.coord[3 * npi];
    const double px = X[0], pz = X[2];
    const double dh = c.doubleh;
    int ncbs = 0;
    int clo[4];
    double loc[8];
    for (int isi = 0; isi < c.ns; ++isi) {
        if (c.sbtype[isi] == 6) continue;  // peri
        const double p1x = c.sv1[3 * isi], p1z = c.sv1[3 * isi + 2];
        const double P1P0 = px - p1x, P1P1 = pz - p1z;
        const double s0 = ST(c, isi, 0, 0), s1 = ST(c, isi, 2, 0);
        const double n0 = ST(c, isi, 0, 2), n1 = ST(c, isi, 2, 2);
        const double sidel = c.slen[isi];
        const double xp = P1P0 * s0 + P1P1 * s1;
        const double yp = P1P0 * n0 + P1P1 * n1;
        const double xpmin = -dh, xpmax = sidel + dh, ypmin = -dh, ypmax = dh;
        if ((xpmin < xp && xp < xpmax) && (ypmin < yp && yp < ypmax)) {
            double ypmx = 0.0, ypmn = 0.0;
            if (xp < 0.0) {
                ypmx = std::sqrt(c.square_doubleh - xp * xp);
                ypmn = -ypmx;
            } else if (xp <= sidel) {
                ypmx = ypmax;
                ypmn = ypmin;
            } else if (xp < xpmax) {
                const double Lmxpq = (sidel - xp) * (sidel - xp);
                ypmx = std::sqrt(c.square_doubleh - Lmxpq);
                ypmn = -ypmx;
            }
            if (ypmn < yp && yp < ypmx) {
                if (ncbs >= 4) { c.bn[npi] = -88; return; }
                clo[ncbs] = isi;
                loc[2 * ncbs] = xp;
                loc[2 * ncbs + 1] = yp;
                ncbs = ncbs + 1;
            }
        }
    }
    for (int q = 0; q < ncbs; ++q) {
        const int bt = c.sbtype[clo[q]];
        if (bt != 0 && bt != 1 && bt != 2) { c.bn[npi] = -95; return; }
    }
    if (ncbs == 2) {
        const int a = clo[0], b = clo[1];
        if (c.sprev[a] == b + 1)
            loc[2] = c.slen[b] - loc[2];
        else if (c.sprev[b] == a + 1)
            loc[0] = c.slen[a] - loc[0];
    }
    // SelectCloseBoundarySides2D.f90:66-120: fixe/tapi sides with y > 0; open
    // sides with y > 0 and 0 < x < side length
    int nint = 0;
    int iside[4];
    double iloc[8];
    for (int q = 0; q < ncbs; ++q) {
        bool keep = loc[2 * q + 1] > 0.0;
        if (c.sbtype[clo[q]] == 2) keep = keep && loc[2 * q] > 0.0 && loc[2 * q] < c.slen[clo[q]];
        if (keep) {
            iside[nint] = clo[q];
            iloc[2 * nint] = loc[2 * q];
            iloc[2 * nint + 1] = loc[2 * q + 1];
            nint = nint + 1;
        }
    }
    if (nint == 0) return;
    if (nint > 2) { c.bn[npi] = -89; return; }
    for (int icbs = 0; icbs < nint; ++icbs) {
        const size_t sl = (size_t)npi * mc + icbs;
        c.bface[sl] = iside[icbs];
        c.bloc[3 * sl] = iloc[2 * icbs];
        c.bloc[3 * sl + 1] = iloc[2 * icbs + 1];
        c.bloc[3 * sl + 2] = 0.0;
        double* BI = &c.bbi[8 * sl];
        const double ypi = iloc[2 * icbs + 1];
        const double deltai = ypi / c.p.h;
        double F[5];
        if (nint == 2) {
            double xpmin, xpmax, interlen;
            intersection2d(c, iloc[2 * icbs], ypi, c.slen[iside[icbs]], xpmin, xpmax, interlen);
            const double IntWdS = surface_wds2d(c, iloc[2 * icbs], ypi, xpmin, interlen);
            double IntWdV, IntDp[2];
            volume_bvi2d(c, iloc[2 * icbs], ypi, xpmin, xpmax, interlen, IntWdV, IntDp);
            double IntWdV1;
            const int rc = volume_wdv2d(c, icbs, iside, iloc, xpmin, xpmax, interlen, IntWdV1);
            if (rc) { c.bn[npi] = -rc; return; }
            const int cols[3] = {3, 4, 5};
            interp2d(c, 3, cols, deltai, F);
            BI[0] = IntWdS;
            BI[1] = IntWdV;
            BI[2] = IntWdV1;
            BI[3] = IntDp[0];
            BI[4] = IntDp[1];
            BI[5] = F[0] / c.squareh;
            BI[6] = F[1] / c.squareh;
            BI[7] = F[2] / c.squareh;
        } else {
            const int cols[5] = {1, 2, 3, 4, 5};
            interp2d(c, 5, cols, deltai, F);
            BI[0] = F[0] / c.p.h;
            BI[1] = F[1];
            BI[2] = F[1];
            BI[3] = 0.0;
            BI[4] = 0.0;
            BI[5] = F[2] / c.squareh;
            BI[6] = F[3] / c.squareh;
            BI[7] = F[4] / c.squareh;
        }
    }
    c.bn[npi] = nint;
}

}  // namespace

extern "C" {

void* or_create(const or_params* p, const or_faces* f, const double* coord, const double* vel,
                const double* dens, const double* pres, const double* mass) {
    Ctx* c = new Ctx();
    c->p = *p;
    c->n = p->n;
    const int n = p->n;
    c->squareh = p->h * p->h;
    c->doubleh = 2.0 * p->h;
    c->square_doubleh = c->doubleh * c->doubleh;
    c->nmax = p->ncd[0] * p->ncd[1] * p->ncd[2];
    if (f) {
        c->nf = f->nf;
        const int nf = f->nf;
        c->T.assign(f->T, f->T + 9 * nf);
        c->LX.assign(f->LX, f->LX + 8 * nf);
        c->GX.assign(f->GX, f->GX + 12 * nf);
        c->area.assign(f->area, f->area + nf);
        c->fvel.assign(f->velocity, f->velocity + 3 * nf);
        c->tvel.assign(f->tvel, f->tvel + 3 * nf);
        c->slip_coeff.assign(f->slip_coeff, f->slip_coeff + nf);
        c->nodes.assign(f->nodes, f->nodes + nf);
        c->btype.assign(f->btype, f->btype + nf);
        c->slip_mode.assign(f->slip_mode, f->slip_mode + nf);
        c->lam_check.assign(f->lam_check, f->lam_check + nf);
        c->gcbf_count.assign(f->gcbf_count, f->gcbf_count + c->nmax);
        c->gcbf_start.assign(f->gcbf_start, f->gcbf_start + c->nmax);
        int nv = 0;
        for (int i = 0; i < c->nmax; ++i) nv += c->gcbf_count[i];
        c->gcbf_vec.assign(f->gcbf_vec, f->gcbf_vec + nv);
        std::memcpy(c->kt, f->kerneltab, sizeof(c->kt));
        c->ktdelta = f->ktdelta;
        std::memcpy(c->bit, f->bitab, sizeof(c->bit));
        const int ne = f->nedges > 0 ? f->nedges : 0;
        c->edge_face.assign(f->edge_face, f->edge_face + 2 * ne);
        c->edge_gx.assign(f->edge_gx, f->edge_gx + 3 * ne);
        c->edge_comp.assign(f->edge_comp, f->edge_comp + 3 * ne);
        c->edge_len.assign(f->edge_len, f->edge_len + ne);
    } else {
        c->gcbf_count.assign(c->nmax, 0);
        c->gcbf_start.assign(c->nmax, 0);
    }
    c->coord.assign(coord, coord + 3 * n);
    c->vel.assign(vel, vel + 3 * n);
    c->var.assign(vel, vel + 3 * n);
    c->acc.assign(3 * n, 0.0);
    c->dens.assign(dens, dens + n);
    c->pres.assign(pres, pres + n);
    c->mass.assign(mass, mass + n);
    c->dden.assign(n, 0.0);
    c->vpres.assign(n, 0.0);
    c->cella.assign(n, 0);
    c->kodvel.assign(n, 0);
    c->koddens.assign(n, 0);
    c->laminar.assign(n, 0);
    c->id.resize(n);
    for (int i = 0; i < n; ++i) {
        c->id[i] = i;
        c->cella[i] = particle_cell(*c, &c->coord[3 * i]);  // SetParticleParameters.f90:129
    }
    or_ordgrid(c);
    return c;
}

void* or_create2d(const or_params* p, const or_sides2d* sd, const double* coord, const double* vel,
                  const double* dens, const double* pres, const double* mass) {
    Ctx* c = (Ctx*)or_create(p, nullptr, coord, vel, dens, pres, mass);
    c->dim = 2;
    const int ns = sd ? sd->ns : 0;
    c->ns = ns;
    if (ns > 0) {
        c->sT.assign(sd->T, sd->T + 9 * ns);
        c->sRN.assign(sd->RN, sd->RN + 9 * ns);
        c->sv1.assign(sd->v1, sd->v1 + 3 * ns);
        if (sd->v2) c->sv2.assign(sd->v2, sd->v2 + 3 * ns);
        c->slen.assign(sd->length, sd->length + ns);
        c->sang.assign(sd->angle, sd->angle + ns);
        c->sfvel.assign(sd->velocity, sd->velocity + 3 * ns);
        c->stvel.assign(sd->tvel, sd->tvel + 3 * ns);
        c->sslip.assign(sd->slip_coeff, sd->slip_coeff + ns);
        c->sbtype.assign(sd->btype, sd->btype + ns);
        c->sprev.assign(sd->previous, sd->previous + ns);
        c->sslipmode.assign(sd->slip_mode, sd->slip_mode + ns);
        c->slam.assign(sd->lam_check, sd->lam_check + ns);
        std::memcpy(c->bit2d, sd->bit2d, sizeof(c->bit2d));
        for (int k = 0; k < ns; ++k)
            if (sd->btype[k] == 2 && !sd->v2) {  // open sides need their second vertex
                or_destroy(c);
                return nullptr;
            }
    }
    // cells again with j = 1 (ParticleCellNumber.f90:59-61), then OrdGrid1
    for (int i = 0; i < c->n; ++i) c->cella[i] = particle_cell(*c, &c->coord[3 * i]);
    or_ordgrid(c);
    return c;
}

void or_destroy(void* ctx) { delete (Ctx*)ctx; }
int or_nag(void* ctx) { return ((Ctx*)ctx)->n; }

void or_set_state(void* ctx, int n, const double* coord, const double* vel, const double* var,
                  const double* acc, const double* dens, const double* pres, const int32_t* laminar,
                  double dt, double dt_average, double time, int it) {
    Ctx& c = *(Ctx*)ctx;
    c.n = n;
    c.coord.assign(coord, coord + 3 * n);
    c.vel.assign(vel, vel + 3 * n);
    c.var.assign(var, var + 3 * n);
    c.acc.assign(acc, acc + 3 * n);
    c.dens.assign(dens, dens + n);
    c.pres.assign(pres, pres + n);
    c.laminar.assign(laminar, laminar + n);
    c.dt = dt;
    c.dt_average = dt_average;
    c.time = time;
    c.it = it;
    c.on_going = it;
    c.cella.resize(n);
    for (int i = 0; i < n; ++i) c.cella[i] = particle_cell(c, &c.coord[3 * i]);
    or_ordgrid(ctx);
}

void or_get_state(void* ctx, double* coord, double* vel, double* var, double* acc, double* dens,
                  double* pres, double* dden, double* vpres) {
    Ctx& c = *(Ctx*)ctx;
    const size_t n = c.n;
    if (coord) std::memcpy(coord, c.coord.data(), 24 * n);
    if (vel) std::memcpy(vel, c.vel.data(), 24 * n);
    if (var) std::memcpy(var, c.var.data(), 24 * n);
    if (acc) std::memcpy(acc, c.acc.data(), 24 * n);
    if (dens) std::memcpy(dens, c.dens.data(), 8 * n);
    if (pres) std::memcpy(pres, c.pres.data(), 8 * n);
    if (dden) std::memcpy(dden, c.dden.data(), 8 * n);
    if (vpres) std::memcpy(vpres, c.vpres.data(), 8 * n);
}

void or_get_scalars(void* ctx, double* out) {
    Ctx& c = *(Ctx*)ctx;
    out[0] = c.dt;
    out[1] = c.dt_average;
    out[2] = c.time;
    out[3] = c.it;
}

int or_get_lists(void* ctx, int32_t* count, int32_t* ids, int32_t stride) {
    Ctx& c = *(Ctx*)ctx;
    if (!c.lists_built) return 1;
    const int mj = c.p.nmaxpartj;
    for (int i = 0; i < c.n; ++i) {
        count[i] = c.nnb[i];
        if (ids)
            for (int k = 0; k < c.nnb[i] && k < stride; ++k) ids[(size_t)i * stride + k] = c.nbr[(size_t)i * mj + k];
    }
    return 0;
}

int64_t or_compare_lists(void* ctx, const int32_t* count, const int32_t* ids, int32_t stride) {
    Ctx& c = *(Ctx*)ctx;
    if (!c.lists_built) return -2;
    const int mj = c.p.nmaxpartj;
    int64_t first = INT64_MAX;
#pragma omp parallel for schedule(static) reduction(min : first)
    for (int i = 0; i < c.n; ++i) {
        bool same = count[i] == c.nnb[i];
        for (int k = 0; same && k < c.nnb[i]; ++k)
            same = ids[(size_t)i * stride + k] == c.nbr[(size_t)i * mj + k];
        if (!same && i < first) first = i;
    }
    return first == INT64_MAX ? -1 : first;
}

int or_get_btab(void* ctx, int32_t* ncbf, int32_t* face, double* loc, double* bi, double* ig) {
    Ctx& c = *(Ctx*)ctx;
    const int mc = c.p.maxclose;
    for (int i = 0; i < c.n; ++i) {
        ncbf[i] = c.bn[i];
        for (int q = 0; q < c.bn[i]; ++q) {
            const size_t s = (size_t)i * mc + q;
            if (face) face[s] = c.bface[s];
            if (loc) std::memcpy(&loc[3 * s], &c.bloc[3 * s], 24);
            if (bi) std::memcpy(&bi[8 * s], &c.bbi[8 * s], 64);
            if (ig) std::memcpy(&ig[9 * s], &c.big[9 * s], 72);
        }
    }
    return 0;
}

// OrdGrid1.f90:58-105 (fluid part), with the swap-with-last removal of
// particles outside the grid (:62-83)
int or_ordgrid(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    const int nmax = c.nmax;
    c.icont.assign(nmax + 2, 0);
    std::vector<int> inc(nmax + 1, 0);
    int npi = 0;
    auto move_last = [&](int dst, int src) {
        for (int d = 0; d < 3; ++d) {
            c.coord[3 * dst + d] = c.coord[3 * src + d];
            c.vel[3 * dst + d] = c.vel[3 * src + d];
            c.var[3 * dst + d] = c.var[3 * src + d];
            c.acc[3 * dst + d] = c.acc[3 * src + d];
        }
        c.dens[dst] = c.dens[src];
        c.pres[dst] = c.pres[src];
        c.mass[dst] = c.mass[src];
        c.dden[dst] = c.dden[src];
        c.vpres[dst] = c.vpres[src];
        c.cella[dst] = c.cella[src];
        c.kodvel[dst] = c.kodvel[src];
        c.koddens[dst] = c.koddens[src];
        c.laminar[dst] = c.laminar[src];
        c.id[dst] = c.id[src];
        if (c.dbsph) {  // the whole pg record moves (OrdGrid1.f90:73-79)
            c.gamma[dst] = c.gamma[src];
            c.sigma[dst] = c.sigma[src];
            c.rhosph[dst] = c.rhosph[src];
            c.rhosph_old[dst] = c.rhosph_old[src];
            c.dshep[dst] = c.dshep[src];
            c.fs[dst] = c.fs[src];
            for (int d = 0; d < 3; ++d) c.drho[3 * dst + d] = c.drho[3 * src + d];
            for (int d = 0; d < 9; ++d) c.dvelg[9 * dst + d] = c.dvelg[9 * src + d];
        }
    };
    while (c.n > 0 && npi < c.n) {
        npi = npi + 1;
        const int i = npi - 1;
        const int ncel = particle_cell(c, &c.coord[3 * i]);
        if (c.cella[i] <= 0 || ncel <= 0 || ncel > nmax) {
            move_last(i, c.n - 1);
            c.n = c.n - 1;
            npi = npi - 1;
        } else {
            c.icont[ncel] += 1;
            c.cella[i] = ncel;
            c.icont[nmax + 1] += 1;
        }
    }
    c.coord.resize(3 * c.n);
    c.vel.resize(3 * c.n);
    c.var.resize(3 * c.n);
    c.acc.resize(3 * c.n);
    c.dens.resize(c.n);
    c.pres.resize(c.n);
    c.mass.resize(c.n);
    c.dden.resize(c.n);
    c.vpres.resize(c.n);
    c.cella.resize(c.n);
    c.kodvel.resize(c.n);
    c.koddens.resize(c.n);
    c.laminar.resize(c.n);
    c.id.resize(c.n);
    if (c.dbsph) {
        c.gamma.resize(c.n);
        c.sigma.resize(c.n);
        c.rhosph.resize(c.n);
        c.rhosph_old.resize(c.n);
        c.dshep.resize(c.n);
        c.fs.resize(c.n);
        c.drho.resize(3 * (size_t)c.n);
        c.dvelg.resize(9 * (size_t)c.n);
    }
    for (int i = nmax; i >= 1; --i) c.icont[i] = c.icont[i + 1] - c.icont[i];
    for (int i = 1; i <= nmax + 1; ++i) c.icont[i] = c.icont[i] + 1;
    c.npartord.assign(c.n + 1, 0);
    for (int i = 0; i < c.n; ++i) {
        const int ncel = c.cella[i];
        if (ncel == 0) continue;
        c.npartord[c.icont[ncel] + inc[ncel]] = i;  // 1-based slot, 0-based particle
        inc[ncel] += 1;
    }
    c.lists_built = false;
    return 0;
}

// CalcVarLength.f90:462-540: the wall elements of cell ncelj, visited after the
// cell's fluid particles (the fluid-wall list, kernel_fw, dShep, the wet
// flags, and the wall parts of sigma and rhoSPH_new).  A slot at exactly
// r = 2h keeps kernel values of zero here; the reference leaves that slot's
// previous contents.
static void dbsph_cell_walls(Ctx& c, int npi, int ncelj, bool& ov) {
    const int mj = c.p.nmaxpartj;
    const double ke_coef = c.p.coefke / c.p.h;
    const double* xi = &c.coord[3 * npi];
    const double* vi = &c.var[3 * npi];
    for (int fw = c.icont_w[ncelj]; fw <= c.icont_w[ncelj + 1] - 1; ++fw) {
        const int npj = c.npartord_w[fw];
        const double* xw = &c.wcoord[3 * npj];
        const double r0 = xi[0] - xw[0], r1 = xi[1] - xw[1], r2 = xi[2] - xw[2];
        const double rij = r0 * r0 + r1 * r1 + r2 * r2;
        if (rij > c.square_doubleh) continue;
        if (c.nfw[npi] >= mj) {  // the reference stops (diagnostic 10/9)
            ov = true;
            continue;
        }
        const size_t s = (size_t)npi * mj + c.nfw[npi];
        c.nfw[npi] = c.nfw[npi] + 1;
        c.fwl[s] = npj;
        const double rij2 = rij;
        const double q = std::sqrt(rij) / c.p.h;
        const double q2 = rij2 / c.squareh;
        const int iq = (int)q;
        c.rag_fw[3 * s] = r0;
        c.rag_fw[3 * s + 1] = r1;
        c.rag_fw[3 * s + 2] = r2;
        c.kfw[2 * s] = c.kfw[2 * s + 1] = 0.0;
        if (iq >= 2) continue;
        double wu = 0.666666667 + q2 * (q * 0.5 - 1.0);
        if (iq > 0) wu = (2.0 - q) * (2.0 - q) * (2.0 - q) * 0.166666666667;
        const double k1 = wu * c.p.coefke;
        c.kfw[2 * s] = k1;
        const double* nw = &c.wnormal[3 * npj];
        const double* vw = &c.wvel[3 * npj];
        c.dshep[npi] = c.dshep[npi] + k1 * c.wweight[npj] *
                                          (nw[0] * (vw[0] - vi[0]) + nw[1] * (vw[1] - vi[1]) +
                                           nw[2] * (vw[2] - vi[2]));
        if ((q * c.p.h) <= (1.3 * c.p.dx + c.wweight[npj] / 2.0))
            __atomic_store_n(&c.wwet[npj], 1, __ATOMIC_RELAXED);
        const double denom = 1.0 / (std::sqrt(rij) + c.p.eta);
        double gm = -2.0 * q + 1.5 * q2;
        if (iq > 0) gm = -gm + q2 - 2.0;
        c.kfw[2 * s + 1] = gm * ke_coef * denom;
        c.sigma[npi] = c.sigma[npi] + c.wvolume[npj] * k1;
        c.rhosph[npi] = c.rhosph[npi] + c.wmass[npj] * k1;  // NMedium = 1
    }
}

// CalcVarLength.f90:842-877: Gamma set from sigma on the pre-loop's first
// call (on_going_time_step = -2); afterwards the free-surface flag FS where
// sigma + 0.05 < Gamma, and the density from rhoSPH_new on a particle's
// first call
static void dbsph_after_lists(Ctx& c) {
    for (int npi = 0; npi < c.n; ++npi) {
        if (c.on_going == -2) {
            if (c.nfw[npi] == 0) {
                c.gamma[npi] = 1.0;
            } else {
                c.gamma[npi] = c.sigma[npi];
                if (c.gamma_limiter) c.gamma[npi] = std::min(c.gamma[npi], 1.0);
            }
        }
        if (c.on_going > -2) {
            const double msg = std::min(c.sigma[npi] + 0.05, c.gamma[npi]);
            if (c.fs_allowed && msg != c.gamma[npi]) c.fs[npi] = 1;
            if (c.rhosph_old[npi] == 0.0) {
                if (c.fs[npi] == 0) c.dens[npi] = c.rhosph[npi] / c.gamma[npi];
                else c.dens[npi] = c.rhosph[npi] / c.sigma[npi];
            }
        }
    }
}

// inter_EqMoto.f90:286-322 with slip_ID = 0: the wall elements' pressure
// terms; the wall shear term (DBSPH_BC_shear_viscosity_term.f90:46-49) and
// the semi-particle Morris term read the VSL velocity gradients and
// semi-particle viscosities, which are zero without the VSL model, so they
// add zeros
static void dbsph_wall_momentum(Ctx& c, int npi, double tp[3], double tv[3]) {
    const int mj = c.p.nmaxpartj;
    const double rhoi = c.dens[npi], pi = c.pres[npi];
    for (int k = 0; k < c.nfw[npi]; ++k) {
        const size_t s = (size_t)npi * mj + k;
        const int npj = c.fwl[s];
        const double wd = c.wdens[npj];
        const double A = pi / (rhoi * rhoi) + c.wpres[npj] / (wd * wd);
        const double* nw = &c.wnormal[3 * npj];
        const double* R = &c.rag_fw[3 * s];
        for (int d = 0; d < 3; ++d) {
            double ap = -wd * c.wweight[npj] * c.kfw[2 * s] * A * nw[d];
            ap = ap - c.wmass[npj] * A * R[d] * c.kfw[2 * s + 1];
            tp[d] = tp[d] + ap;
        }
    }
    for (int d = 0; d < 3; ++d) tv[d] = tv[d] + 0.0 + 0.0;
}

// Gradients_to_MUSCL.f90:47-91 (the density and velocity gradients of the
// particles with wall neighbours; others keep their previous values) and
// BC_wall_elements.f90:71-158 (each wall element's density and pressure:
// the Shepard average, over its fluid neighbours in pg order, of the
// partial-Riemann state rho* = rho_L + (u_L - u_R) rho_L / c_L, clamped at
// 10, and p* = EoS(rho*), clamped at -99 kPa)
static void dbsph_walls_state(Ctx& c) {
    const int mj = c.p.nmaxpartj;
    for (int npi = 0; npi < c.n; ++npi) {
        if (c.nfw[npi] <= 0) continue;
        double* g = &c.drho[3 * npi];
        double* V = &c.dvelg[9 * npi];  // V[3 r + k] = dvel(r+1, k+1)
        for (int d = 0; d < 3; ++d) g[d] = 0.0;
        for (int d = 0; d < 9; ++d) V[d] = 0.0;
        const double* vi = &c.var[3 * npi];
        for (int k = 0; k < c.nnb[npi]; ++k) {
            const size_t s = (size_t)npi * mj + k;
            const int npj = c.nbr[s];
            const double vol = c.mass[npj] / c.dens[npj];
            const double* R = &c.rag[3 * s];
            const double K0 = c.pk[4 * s];
            for (int d = 0; d < 3; ++d) g[d] = g[d] + R[d] * (c.dens[npj] - c.dens[npi]) * K0 * vol;
            for (int r = 0; r < 3; ++r)
                for (int d = 0; d < 3; ++d)
                    V[3 * r + d] = V[3 * r + d] + R[d] * (c.var[3 * npj + r] - vi[r]) * K0 * vol;
        }
        if (!c.muscl_boundary) continue;
        // Gradients_to_MUSCL_boundary.f90:44-86 (one medium): each wall
        // neighbour's Shepard term along its normal, then its gradient term
        const double sg = c.sigma[npi];
        for (int k = 0; k < c.nfw[npi]; ++k) {
            const size_t s = (size_t)npi * mj + k;
            const int w = c.fwl[s];
            const double* nw = &c.wnormal[3 * w];
            const double* vw = &c.wvel[3 * w];
            const double* R = &c.rag_fw[3 * s];
            const double dr = c.wdens[w] - c.dens[npi];
            const double Ww = c.wweight[w] * c.kfw[2 * s] / sg;
            for (int d = 0; d < 3; ++d) g[d] = g[d] + dr * nw[d] * Ww;
            for (int r = 0; r < 3; ++r)
                for (int d = 0; d < 3; ++d) V[3 * r + d] = V[3 * r + d] + (vw[r] - vi[r]) * nw[d] * Ww;
            const double vol = c.wmass[w] / c.wdens[w] / sg;
            for (int d = 0; d < 3; ++d) g[d] = g[d] + R[d] * dr * c.kfw[2 * s + 1] * vol;
            for (int r = 0; r < 3; ++r)
                for (int d = 0; d < 3; ++d)
                    V[3 * r + d] = V[3 * r + d] + R[d] * (vw[r] - vi[r]) * c.kfw[2 * s + 1] * vol;
        }
    }
    std::vector<double> den(c.nw, 0.0);
    std::vector<char> neigh(c.nw, 0);
    for (int w = 0; w < c.nw; ++w) {
        c.wpres[w] = 0.0;
        c.wdens[w] = c.p.den0;
    }
    for (int npi = 0; npi < c.n; ++npi) {
        const double* g = &c.drho[3 * npi];
        const double* V = &c.dvelg[9 * npi];
        const double* vi = &c.var[3 * npi];
        for (int k = 0; k < c.nfw[npi]; ++k) {
            const size_t s = (size_t)npi * mj + k;
            const int npj = c.fwl[s];
            if (!neigh[npj]) {
                c.wdens[npj] = 0.0;
                neigh[npj] = 1;
            }
            const double* R = &c.rag_fw[3 * s];
            const double* nw = &c.wnormal[3 * npj];
            const double* vw = &c.wvel[3 * npj];
            const double rhoL = c.dens[npi] - (g[0] * R[0] + g[1] * R[1] + g[2] * R[2]);
            double uC[3];
            for (int r = 0; r < 3; ++r) uC[r] = vi[r] - (V[3 * r] * R[0] + V[3 * r + 1] * R[1] + V[3 * r + 2] * R[2]);
            const double uL = uC[0] * nw[0] + uC[1] * nw[1] + uC[2] * nw[2];
            const double uR = vw[0] * nw[0] + vw[1] * nw[1] + vw[2] * nw[2];
            const double cL = std::sqrt(c.p.eps / rhoL);
            double rhostar = rhoL + (uL - uR) * rhoL / cL;
            if (rhostar < 10.0) rhostar = 10.0;
            double pstar = c.p.eps / c.p.den0 * (rhostar - c.p.den0) + c.p.prif;
            if (pstar < -99.0e3) pstar = -99.0e3;
            const double Ww = c.mass[npi] / c.dens[npi] * c.kfw[2 * s];
            den[npj] = den[npj] + Ww;
            c.wdens[npj] = c.wdens[npj] + rhostar * Ww;
            c.wpres[npj] = c.wpres[npj] + pstar * Ww;
        }
    }
    for (int w = 0; w < c.nw; ++w) {
        if (!neigh[w]) continue;
        c.wdens[w] = c.wdens[w] / den[w];
        c.wpres[w] = c.wpres[w] / den[w];
        if (!c.neg_wall_p && c.wpres[w] < 0.0) c.wpres[w] = 0.0;
    }
}

// CalcVarLength.f90:111-269 (fluid-fluid part: lists, rag, PartKernel)
int or_calcvarlength(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    const int n = c.n;
    const int mj = c.p.nmaxpartj;
    c.nnb.assign(n, 0);
    const bool lo = c.lists_only;
    if (c.nbr.size() < (size_t)n * mj) c.nbr.resize((size_t)n * mj);
    if (!lo && c.rag.size() < (size_t)n * mj * 3) {
        c.rag.resize((size_t)n * mj * 3);
        c.pk.resize((size_t)n * mj * 4);
    }
    const double ke_coef = c.p.coefke / c.p.h;
    const double kacl_coef = c.p.coefkacl / c.p.h;
    const bool db = c.dbsph && !lo;
    if (db) {
        c.nfw.assign(n, 0);
        if (c.fwl.size() < (size_t)n * mj) {
            c.fwl.resize((size_t)n * mj);
            c.rag_fw.resize((size_t)n * mj * 3);
            c.kfw.resize((size_t)n * mj * 2);
        }
    }
    int overflow = 0;
#pragma omp parallel for schedule(static) reduction(+ : overflow)
    for (int npi = 0; npi < n; ++npi) {
        if (db) {  // CalcVarLength.f90:134-140
            c.rhosph_old[npi] = c.rhosph[npi];
            c.rhosph[npi] = 0.0;
            c.dshep[npi] = 0.0;
            c.sigma[npi] = 0.0;
            c.fs[npi] = 0;
        }
        const int nceli = c.cella[npi];
        if (nceli == 0) continue;
        int ig, jg, kg;
        cell_indices(c, nceli, ig, jg, kg);
        const double* xi = &c.coord[3 * npi];
        int cnt = 0;
        bool ov = false;
        for (int jr = jg - 1; jr <= jg + 1; ++jr)
            for (int ir = ig - 1; ir <= ig + 1; ++ir)
                for (int kr = kg - 1; kr <= kg + 1; ++kr) {
                    const int ncelj = cell_number(c, ir, jr, kr);
                    if (ncelj == 0) continue;
                    for (int mm = c.icont[ncelj]; mm <= c.icont[ncelj + 1] - 1; ++mm) {
                        if (cnt >= mj) { ov = true; continue; }
                        if (db && c.nfw[npi] >= mj) { ov = true; continue; }
                        const int npj = c.npartord[mm];
                        const double* xj = &c.coord[3 * npj];
                        const double r0 = xi[0] - xj[0];
                        const double r1 = xi[1] - xj[1];
                        const double r2 = xi[2] - xj[2];
                        double rij = r0 * r0 + r1 * r1 + r2 * r2;
                        if (rij > c.square_doubleh) continue;
                        cnt = cnt + 1;
                        const size_t s = (size_t)npi * mj + (cnt - 1);
                        c.nbr[s] = npj;
                        if (lo) continue;
                        const double rij2 = rij;
                        rij = std::sqrt(rij);
                        const double q = rij / c.p.h;
                        const double q2 = rij2 / c.squareh;
                        const int iq = (int)q;
                        const double denom = 1.0 / (rij + c.p.eta);
                        c.rag[3 * s] = r0;
                        c.rag[3 * s + 1] = r1;
                        c.rag[3 * s + 2] = r2;
                        double* K = &c.pk[4 * s];
                        K[0] = K[1] = K[2] = K[3] = 0.0;
                        if (iq >= 2) continue;
                        double gm = -2.0 * q + 1.5 * q2;
                        double gw = -12.0 - 3.0 * q2 + 12.0 * q;
                        double wu = 0.666666667 + q2 * (q * 0.5 - 1.0);
                        if (iq > 0) {
                            gm = -gm + q2 - 2.0;
                            wu = (2.0 - q) * (2.0 - q) * (2.0 - q) * 0.166666666667;
                        }
                        gm = gm * ke_coef;
                        gw = gw * kacl_coef;
                        K[0] = gm * denom;
                        K[1] = K[0] / (rij2 + c.p.eta2);
                        K[2] = gw * denom;
                        K[3] = wu * c.p.coefke;
                        if (db) {  // CalcVarLength.f90:330-340 (one medium)
                            c.sigma[npi] = c.sigma[npi] + c.mass[npj] * K[3] / c.dens[npj];
                            c.rhosph[npi] = c.rhosph[npi] + c.mass[npj] * K[3];
                        }
                    }
                    if (db) dbsph_cell_walls(c, npi, ncelj, ov);
                }
        c.nnb[npi] = cnt;
        if (ov) overflow += 1;
    }
    c.overflow = overflow;
    c.lists_built = true;
    if (db) dbsph_after_lists(c);
    return 0;
}

int or_boundary_tab(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    const int n = c.n;
    const size_t mc = c.p.maxclose;
    c.bn.assign(n, 0);
    if (c.bface.size() < n * mc) {
        c.bface.resize(n * mc);
        c.bloc.resize(n * mc * 3);
        c.bbi.resize(n * mc * 8);
        c.big.resize(n * mc * 9);
    }
    if (c.dim == 2) {
        if (!c.p.semi || c.ns == 0) return 0;
        int code = 0;
        for (int npi = 0; npi < n; ++npi) {  // ComputeBoundaryDataTab_2D is a serial loop
            boundary_tab_particle_2d(c, npi);
            if (c.bn[npi] < 0 && code == 0) code = -c.bn[npi];
        }
        return code;
    }
    if (!c.p.semi || c.nf == 0) return 0;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
    for (int npi = 0; npi < n; ++npi) {
        boundary_tab_particle(c, npi);
        if (c.bn[npi] < 0) bad += 1;
    }
    return bad ? 86 : 0;  // diagnostic(8,6/7): too many close faces
}

// time_step_duration.f90:59-123 (liquid medium, dt_alfa_Mon=.false.)
int or_dt(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    double dtmin = 3.4e38, dt_vis = 3.4e38;
    if (c.n == 0) {
        c.dt = 0.0;
    } else {
        if (c.p.kin_visc > 1.0e-24) {
            dt_vis = c.p.vsc_coeff * c.squareh / (0.5 * c.p.kin_visc);
            dtmin = std::fmin(dtmin, dt_vis);
        }
        for (int i = 0; i < c.n; ++i) {
            const double* v = &c.vel[3 * i];
            const double U = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            const double dt_cfl = c.p.cfl * 2.0 * c.p.h / (c.p.celerita + U);
            dtmin = std::fmin(dtmin, std::fmin(dt_cfl, dt_vis));
        }
        if (c.on_going == 0)
            c.dt = dtmin;
        else
            c.dt = (1.0 - c.p.pesodt) * dtmin + c.p.pesodt * c.dt_average;
    }
    if (c.on_going == 0)
        c.dt_average = c.dt;
    else
        c.dt_average = (c.dt_average * (double)(c.on_going - 1) + c.dt) / (double)c.on_going;
    return 0;
}

// RHS_momentum_equation.f90:72-179 -> AddBoundaryContributions_to_ME3D.f90:
// 118-336, inter_EqMoto.f90:108-283, viscomon.f90:54-71, viscomorris.f90:
// 60-71, AddElasticBoundaryReaction_3D.f90:67-98
int or_momentum(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    const int n = c.n;
    const int mj = c.p.nmaxpartj;
    const size_t mc = c.p.maxclose;
    const double h = c.p.h;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int npi = 0; npi < n; ++npi) {
        double tp[3] = {0, 0, 0}, td[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
        c.kodvel[npi] = 0;
        const int ncbf = c.p.semi ? c.bn[npi] : 0;
        const double* vi = &c.var[3 * npi];
        // --- SA-SPH wall terms (called before the inner sums) ---
        if (ncbf > 0 && c.dim == 2) {
            // AddBoundaryContributions_to_ME2D.f90:121-405, fixe/tapi sides
            const double ro0 = c.p.den0, roi = c.dens[npi], pressi = c.pres[npi];
            const double distpimin = c.p.dx;
            double VMo[2] = {0.0, 0.0}, VSh[2] = {0.0, 0.0}, RGs[2] = {0.0, 0.0};
            double ALEt[3] = {0.0, 0.0, 0.0};
            const double g2[2] = {c.p.grav[0], c.p.grav[2]};
            for (int q = 0; q < ncbf; ++q) {
                const size_t sl = (size_t)npi * mc + q;
                const int sd = c.bface[sl];
                if (c.sbtype[sd] == 2) continue;  // AddBoundaryContributions_to_ME2D.f90:124
                const double xpi = c.bloc[3 * sl], ypi = c.bloc[3 * sl + 1];
                const double* BI = &c.bbi[8 * sl];
                const double IntWdS = BI[0], IntWdV = BI[2], IntWd1s0 = BI[5], IntWd3s0 = BI[6];
                const double ss[2] = {ST(c, sd, 0, 0), ST(c, sd, 2, 0)};
                const double nn[2] = {ST(c, sd, 0, 2), ST(c, sd, 2, 2)};
                double GravN = 0.0;
                GravN = GravN + c.p.grav[0] * nn[0];
                GravN = GravN + c.p.grav[2] * nn[1];
                const double sidevel[2] = {c.sbtype[sd] == 1 ? c.sfvel[3 * sd] : 0.0,
                                           c.sbtype[sd] == 1 ? c.sfvel[3 * sd + 2] : 0.0};
                const double Dvel[2] = {2.0 * (vi[0] - sidevel[0]), 2.0 * (vi[2] - sidevel[1])};
                const double DvelN = Dvel[0] * nn[0] + Dvel[1] * nn[1];
                const double distpi = std::fmax(ypi, distpimin);
                double pressib = pressi;
                const double pressibmin = 10.0 * ro0 * distpimin;
                if (DvelN <= 0.0) {
                    if (pressib < pressibmin) pressib = pressibmin;
                }
                const double PressB = pressib - ro0 * GravN * distpi;
                const double Qii = IntWdS * (pressib + PressB) / roi;
                double RG[2] = {0.0, 0.0};
                for (int i = 0; i < 2; ++i) {
                    const int ai = i == 0 ? 0 : 2;
                    for (int j = 0; j < 2; ++j) RG[i] = RG[i] + SRN(c, sd, ai, j == 0 ? 0 : 2) * g2[j];
                    RG[i] = RG[i] * IntWdV;
                }
                RGs[0] = RGs[0] + RG[0];
                RGs[1] = RGs[1] + RG[1];
                ALEt[0] = ALEt[0] - nn[0] * Qii;
                ALEt[2] = ALEt[2] - nn[1] * Qii;
                if (xpi >= 0.0 && xpi <= c.slen[sd]) {
                    double slip = 0.0, cinv = 0.0;
                    if (c.sslipmode[sd] == 1) {
                        slip = c.sslip[sd];
                        cinv = c.p.kin_visc;
                    } else if (c.sslipmode[sd] == 2) {
                        // AddBoundaryContributions_to_ME2D.f90:314-326
                        double ut[3];
                        for (int k = 0; k < 3; ++k)
                            ut[k] = (vi[k] - c.sfvel[3 * sd + k]) - (0.5 * DvelN * ST(c, sd, k, 2));
                        const double u = std::sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
                        count_wall_function(c, wall_function(u, c.sslip[sd] * 10.0, ypi, c.p.dx, slip, cinv));
                    }
                    const double Monvisc = c.p.alfamon * c.p.celerita * h;
                    const double TN = Monvisc * DvelN * IntWd3s0;
                    const double TT[2] = {TN * nn[0], TN * nn[1]};
                    double SVforce = 0.0;
                    if (c.laminar[npi] == 1 || c.slam[sd] == 0) SVforce = 2.0 * cinv * slip * IntWd1s0;
                    const double DvM[2] = {2.0 * vi[0], 2.0 * vi[2]};
                    for (int i = 0; i < 2; ++i) {
                        VMo[i] = VMo[i] + TT[i];
                        VSh[i] = VSh[i] + DvM[i] * SVforce;
                    }
                    (void)ss;
                }
            }
            tp[0] = tp[0] - RGs[0] - ALEt[0];
            tp[2] = tp[2] - RGs[1] - ALEt[2];
            td[0] = td[0] - VMo[0];
            td[2] = td[2] - VMo[1];
            tv[0] = tv[0] - VSh[0];
            tv[2] = tv[2] - VSh[1];
        } else if (ncbf > 0) {
            double ALEt[3] = {0, 0, 0}, gradpt[3] = {0, 0, 0}, VM[3] = {0, 0, 0}, VS[3] = {0, 0, 0};
            for (int q = 0; q < ncbf; ++q) {
                const size_t s = (size_t)npi * mc + q;
                const int f = c.bface[s];
                const int bt = c.btype[f];
                if (bt == 2) continue;
                if (bt != 0 && bt != 1) { bad += 1; break; }
                const double* L = &c.bloc[3 * s];
                if (!(L[2] > 0.0)) continue;
                const double* BI = &c.bbi[8 * s];
                const double* IG = &c.big[9 * s];
                double dvij[3], dvn = 0.0;
                for (int sd = 0; sd < 3; ++sd) {
                    dvij[sd] = 2.0 * (vi[sd] - c.fvel[3 * f + sd]);
                    dvn = dvn + Tm(c, f, sd, 2) * dvij[sd];
                }
                double gl[3];
                for (int sd = 0; sd < 3; ++sd) {
                    double a = 0.0;
                    for (int sj = 0; sj < 3; ++sj) a = a + Tm(c, f, sj, sd) * c.p.grav[sj];
                    gl[sd] = a;
                }
                double gloc[3], gglo[3];
                for (int r = 0; r < 3; ++r) {
                    double a = 0.0;
                    for (int k = 0; k < 3; ++k) a = a + IG[k * 3 + r] * gl[k];
                    gloc[r] = a;
                }
                for (int r = 0; r < 3; ++r) {
                    double a = 0.0;
                    for (int k = 0; k < 3; ++k) a = a + Tm(c, f, r, k) * gloc[k];
                    gglo[r] = a;
                }
                for (int d = 0; d < 3; ++d) gradpt[d] = gradpt[d] + gglo[d];
                double aloc[3], aglo[3];
                for (int d = 0; d < 3; ++d) aloc[d] = 2.0 * c.pres[npi] / c.dens[npi] * BI[3 + d];
                for (int r = 0; r < 3; ++r) {
                    double a = 0.0;
                    for (int k = 0; k < 3; ++k) a = a + Tm(c, f, r, k) * aloc[k];
                    aglo[r] = a;
                }
                for (int d = 0; d < 3; ++d) ALEt[d] = ALEt[d] + aglo[d];
                const double IntGWZ = BI[6];
                const double IntdW = BI[2];
                if (c.p.alfamon > 0.0) {
                    const double Monvisc = c.p.alfamon * c.p.celerita * h;
                    const double Mmult = -Monvisc * dvn * IntGWZ;
                    for (int d = 0; d < 3; ++d) VM[d] = VM[d] + Mmult * Tm(c, f, d, 2);
                }
                double slip = 0.0, cinv = 0.0;
                if (c.slip_mode[f] == 1) {
                    slip = c.slip_coeff[f];
                    cinv = c.p.kin_visc;
                } else if (c.slip_mode[f] == 2) {
                    // AddBoundaryContributions_to_ME3D.f90:207-227
                    double ut[3];
                    for (int d = 0; d < 3; ++d) ut[d] = 0.5 * (dvij[d] - dvn * Tm(c, f, d, 2));
                    const double u = std::sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
                    count_wall_function(c, wall_function(u, c.slip_coeff[f] * 10.0, L[2], c.p.dx, slip, cinv));
                }
                if (cinv > 0.0) {
                    if (c.laminar[npi] == 1 || c.lam_check[f] == 0) {
                        const double cm = cinv * IntdW * slip;
                        for (int d = 0; d < 3; ++d) {
                            const double dm = 2.0 * (vi[d] - c.fvel[3 * f + d]);
                            VS[d] = VS[d] + cm * dm;
                        }
                    }
                }
            }
            for (int d = 0; d < 3; ++d) {
                tp[d] = tp[d] + ALEt[d] + gradpt[d];
                td[d] = td[d] - VM[d];
                tv[d] = tv[d] + VS[d];
            }
        }
        // --- inner sums ---
        const int nn = c.lists_built ? c.nnb[npi] : 0;
        const double rhoi = c.dens[npi], pi = c.pres[npi];
        const double* vel_i = &c.vel[3 * npi];
        double tpa[3] = {0, 0, 0}, ale[3] = {0, 0, 0}, rvs[3] = {0, 0, 0};
        for (int k = 0; k < nn; ++k) {
            const size_t s = (size_t)npi * mj + k;
            const int npj = c.nbr[s];
            if (npj == npi) continue;
            const double rhoj = c.dens[npj], amassj = c.mass[npj], pj = c.pres[npj];
            const double* vj = &c.vel[3 * npj];
            const double dv[3] = {vj[0] - vel_i[0], vj[1] - vel_i[1], vj[2] - vel_i[2]};
            const double* R = &c.rag[3 * s];
            const double* K = &c.pk[4 * s];
            if (c.dbsph) {  // inter_EqMoto.f90:229-235
                const double alpha = pi / (rhoi * rhoi) + pj / (rhoj * rhoj);
                for (int d = 0; d < 3; ++d) tpa[d] = tpa[d] + -amassj * alpha * R[d] * K[0];
            } else {
                double alpha = (pj - pi) / (rhoi * rhoj);
                for (int d = 0; d < 3; ++d) {
                    const double ap = -amassj * alpha * R[d] * K[2];
                    tpa[d] = tpa[d] + ap;
                }
                alpha = (pi * (rhoj / rhoi + 1.0) + pj * (rhoi / rhoj - 1.0)) / (rhoi * rhoj);
                for (int d = 0; d < 3; ++d) ale[d] = ale[d] - amassj * alpha * R[d] * K[2];
            }
            // viscomon
            const double vrij = -dv[0] * R[0] - dv[1] * R[1] - dv[2] * R[2];
            const double rhot = rhoi + rhoj;
            const double celt = c.p.celerita + c.p.celerita;
            const double TermMon = c.p.alfamon * celt * h / rhot;
            for (int d = 0; d < 3; ++d) {
                const double ra = amassj * TermMon * vrij * R[d] * K[1];
                td[d] = td[d] + (ra + 0.0);
            }
            // viscomorris
            if (c.p.kin_visc > 1.0e-12 && c.p.kin_visc > 1.0e-12) {
                const double nui = c.p.kin_visc, nuj = c.p.kin_visc;
                const double rtil = nui * rhoi + nuj * rhoj + 1.0e-3;
                const double atil = 4.0 * (nui * nuj);
                const double fac = amassj * atil / rtil;
                const double rr = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
                for (int d = 0; d < 3; ++d) {
                    const double rv = fac * (-dv[d] * K[1] * rr);
                    tv[d] = tv[d] + rv;
                    rvs[d] = rvs[d] + rv;
                }
            }
        }
        for (int d = 0; d < 3; ++d) tp[d] = tp[d] + tpa[d] + ale[d];
        int lam = 0;
        if (c.p.kin_visc > 0.0) {
            const double apg = std::sqrt(tp[0] * tp[0] + tp[1] * tp[1] + tp[2] * tp[2]) + GI;
            const double amo = std::sqrt(rvs[0] * rvs[0] + rvs[1] * rvs[1] + rvs[2] * rvs[2]);
            double w = 0.0;
            if (apg != 0.0) w = amo / apg * 100.0;
            lam = (w > 5.0 || apg == 0.0) ? 1 : 0;
        }
        c.laminar[npi] = lam;
        double* a = &c.acc[3 * npi];
        if (c.dbsph) {
            dbsph_wall_momentum(c, npi, tp, tv);
            for (int d = 0; d < 3; ++d) a[d] = (tp[d] + td[d] + tv[d]) / c.gamma[npi] + c.p.grav[d];
            continue;
        }
        if (ncbf > 0 && c.dim == 2) {
            // AddElasticBoundaryReaction_2D.f90:66-99 over the IntNcbs sides
            double br[3] = {0, 0, 0};
            const double ypimin = 0.25 * c.p.dx;
            const double celer02 = c.p.eps / c.p.den0;
            for (int q = 0; q < ncbf; ++q) {
                const size_t sl = (size_t)npi * mc + q;
                const int sd = c.bface[sl];
                if (c.sbtype[sd] != 0 && c.sbtype[sd] != 1) continue;
                const double xpi = c.bloc[3 * sl], ypi = c.bloc[3 * sl + 1];
                if (ypi < ypimin && xpi > 0.0 && xpi < c.slen[sd]) {
                    double vin = 0.0;
                    for (int k = 0; k < 3; ++k) vin = vin + vi[k] * ST(c, sd, k, 2);
                    if (vin < 0.0) {
                        const double nr = -1.0 * celer02 * std::log((h + ypi - ypimin) / h) / h;
                        for (int d = 0; d < 3; ++d) br[d] = br[d] + nr * ST(c, sd, d, 2);
                        c.n_elastic.fetch_add(1, std::memory_order_relaxed);
                    }
                }
            }
            for (int d = 0; d < 3; ++d) a[d] = tp[d] + td[d] + tv[d] + c.p.grav[d] + br[d];
        } else if (ncbf > 0) {
            // AddElasticBoundaryReaction_3D.f90:72-104 (faces)
            double br[3] = {0, 0, 0};
            const double zimin = 0.25 * c.p.dx;
            const double celer02 = c.p.eps / c.p.den0;
            std::vector<char> rea(ncbf, 0);  // ReaFace
            for (int q = 0; q < ncbf; ++q) {
                const size_t s = (size_t)npi * mc + q;
                const int f = c.bface[s];
                const int bt = c.btype[f];
                if (bt != 0 && bt != 1) continue;
                const double* L = &c.bloc[3 * s];
                const double zi = L[2];
                if (zi < zimin && point_internal(c, f, L[0], L[1])) {
                    double vin = 0.0;
                    for (int sd = 0; sd < 3; ++sd) vin = vin + vi[sd] * Tm(c, f, sd, 2);
                    if (vin < 0.0) {
                        const double nr = -1.0 * celer02 * std::log((h + zi - zimin) / h) / h;
                        for (int d = 0; d < 3; ++d) br[d] = br[d] + nr * Tm(c, f, d, 2);
                        c.n_elastic.fetch_add(1, std::memory_order_relaxed);
                    }
                    rea[q] = 1;
                }
            }
            // AddElasticBoundaryReaction_3D.f90:106-152 (convex edges in
            // order; the subroutine returns at the first edge whose two faces
            // are not both among the particle's non-reacting close faces)
            const int nedge = (int)c.edge_len.size();
            for (int ne = 0; ne < nedge; ++ne) {
                int nclose = 0;
                for (int q = 0; q < ncbf; ++q) {
                    const int f = c.bface[(size_t)npi * mc + q];
                    if (f == c.edge_face[2 * ne] && !rea[q]) nclose += 1;
                    else if (f == c.edge_face[2 * ne + 1] && !rea[q]) nclose += 1;
                }
                if (nclose != 2) break;
                const double* X = &c.coord[3 * npi];
                const double* G = &c.edge_gx[3 * ne];
                const double* K = &c.edge_comp[3 * ne];
                double scaprod = 0.0;
                for (int sd = 0; sd < 3; ++sd) scaprod = scaprod + (X[sd] - G[sd]) * K[sd];
                const double edgelen2 = c.edge_len[ne] * c.edge_len[ne];
                const double tau = scaprod / edgelen2;
                if (tau >= 0.0 && tau <= 1.0) {
                    double XQ[3], QP[3], d2 = 0.0;
                    for (int sd = 0; sd < 3; ++sd) XQ[sd] = G[sd] + K[sd] * tau;
                    for (int sd = 0; sd < 3; ++sd) QP[sd] = X[sd] - XQ[sd];
                    for (int sd = 0; sd < 3; ++sd) d2 = d2 + QP[sd] * QP[sd];
                    const double zi = std::sqrt(d2);
                    if (zi < zimin) {
                        double cd[3], vin = 0.0;
                        for (int sd = 0; sd < 3; ++sd) cd[sd] = QP[sd] / zi;
                        for (int sd = 0; sd < 3; ++sd) vin = vin + vi[sd] * cd[sd];
                        if (vin < 0.0) {
                            const double nr = -1.0 * celer02 * std::log(zi / zimin) / zimin;
                            for (int d = 0; d < 3; ++d) br[d] = br[d] + nr * cd[d];
                            c.n_edge.fetch_add(1, std::memory_order_relaxed);
                        }
                    }
                }
            }
            for (int d = 0; d < 3; ++d) a[d] = tp[d] + td[d] + tv[d] + c.p.grav[d] + br[d];
        } else {
            for (int d = 0; d < 3; ++d) a[d] = tp[d] + td[d] + tv[d] + c.p.grav[d];
        }
    }
    return bad ? 93 : 0;
}

// Leapfrog_momentum.f90:66-86
int or_leapfrog_mom(void* ctx, double dt_previous) {
    Ctx& c = *(Ctx*)ctx;
    const double dtvel = 0.5 * (c.dt + dt_previous);
    for (int i = 0; i < c.n; ++i)
        for (int d = 0; d < 3; ++d) c.vel[3 * i + d] = c.vel[3 * i + d] + dtvel * c.acc[3 * i + d];
    return 0;
}

// velocity_smoothing.f90:74-133, velocity_smoothing_SA_SPH_3D.f90:64-122,
// velocity_smoothing_2.f90:59-103
int or_vsmooth(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    const int n = c.n;
    const int mj = c.p.nmaxpartj;
    const size_t mc = c.p.maxclose;
    std::vector<double> nv(3 * (size_t)n);
    if (c.p.tetav > 1.0e-9) {
#pragma omp parallel for schedule(static)
        for (int npi = 0; npi < n; ++npi) {
            double v[3] = {0.0, 0.0, 0.0};
            const double* vel_i = &c.vel[3 * npi];
            for (int k = 0; k < c.nnb[npi]; ++k) {
                const size_t s = (size_t)npi * mj + k;
                const int npj = c.nbr[s];
                const double pes = c.mass[npj] * c.pk[4 * s + 3] / c.dens[npj];
                for (int d = 0; d < 3; ++d) v[d] = v[d] + (c.vel[3 * npj + d] - vel_i[d]) * pes;
            }
            const int ncbf = c.p.semi ? c.bn[npi] : 0;
            for (int q = 0; q < ncbf && c.dim == 2; ++q) {
                // velocity_smoothing_SA_SPH_2D.f90:64-118 (fixe/tapi sides)
                const size_t sl = (size_t)npi * mc + q;
                const int sd = c.bface[sl];
                if (c.sbtype[sd] != 0 && c.sbtype[sd] != 1) continue;
                const double sss[2] = {ST(c, sd, 0, 0), ST(c, sd, 2, 0)};
                const double nnn[2] = {ST(c, sd, 0, 2), ST(c, sd, 2, 2)};
                const double DVG[2] = {2.0 * (c.stvel[3 * sd] - vel_i[0]),
                                       2.0 * (c.stvel[3 * sd + 2] - vel_i[2])};
                const double DVL0 = sss[0] * DVG[0] + sss[1] * DVG[1];
                const double DVL1 = nnn[0] * DVG[0] + nnn[1] * DVG[1];
                const double IntWdV = c.bbi[8 * sl + 2];
                double slip = 0.0;
                if (c.sslipmode[sd] == 1) {
                    slip = c.sslip[sd];
                } else if (c.sslipmode[sd] == 2) {
                    // velocity_smoothing_SA_SPH_2D.f90:93-106
                    double ut[3], nut;
                    for (int k = 0; k < 3; ++k)
                        ut[k] = (c.stvel[3 * sd + k] - vel_i[k]) - (0.5 * DVL1 * ST(c, sd, k, 2));
                    const double u = std::sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
                    wall_function(u, c.sslip[sd] * 10.0, c.bloc[3 * sl + 1], c.p.dx, slip, nut);
                }
                const double BCL0 = DVL0 * IntWdV * slip;
                const double BCL1 = DVL1 * IntWdV;
                v[0] = v[0] + (sss[0] * BCL0 + nnn[0] * BCL1);
                v[2] = v[2] + (sss[1] * BCL0 + nnn[1] * BCL1);
            }
            for (int q = 0; q < ncbf && c.dim == 3; ++q) {
                const size_t s = (size_t)npi * mc + q;
                const int f = c.bface[s];
                const int bt = c.btype[f];
                if (bt == 3 || bt == 4 || bt == 5) {
                    v[0] = v[1] = v[2] = 0.0;
                    break;
                }
                double DVG[3], DVL[3];
                for (int d = 0; d < 3; ++d) DVG[d] = 2.0 * (c.tvel[3 * f + d] - vel_i[d]);
                for (int i = 0; i < 3; ++i) {
                    double a = 0.0;
                    for (int j = 0; j < 3; ++j) a = a + Tm(c, f, j, i) * DVG[j];
                    DVL[i] = a;
                }
                const double IntWdV = c.bbi[8 * s + 1];
                if (bt == 0 || bt == 1) {
                    double slip = 0.0;
                    if (c.slip_mode[f] == 1) {
                        slip = c.slip_coeff[f];
                    } else if (c.slip_mode[f] == 2) {
                        // velocity_smoothing_SA_SPH_3D.f90:89-110
                        double ut[3], nut;
                        for (int d = 0; d < 3; ++d) ut[d] = 0.5 * (DVG[d] - DVL[2] * Tm(c, f, d, 2));
                        const double u = std::sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
                        wall_function(u, c.slip_coeff[f] * 10.0, c.bloc[3 * s + 2], c.p.dx, slip, nut);
                    }
                    double BCL[3];
                    BCL[0] = DVL[0] * IntWdV * slip;
                    BCL[1] = DVL[1] * IntWdV * slip;
                    BCL[2] = DVL[2] * IntWdV;
                    for (int i = 0; i < 3; ++i) {
                        double a = 0.0;
                        for (int j = 0; j < 3; ++j) a = a + Tm(c, f, i, j) * BCL[j];
                        v[i] = v[i] + a;
                    }
                }
            }
            for (int d = 0; d < 3; ++d) nv[3 * npi + d] = v[d];
        }
        const double tv1 = c.p.tetav * c.p.celerita * c.dt / c.p.h;
        for (int i = 0; i < n; ++i) {
            if (c.kodvel[i] == 0) {
                for (int d = 0; d < 3; ++d) {
                    c.var[3 * i + d] = c.vel[3 * i + d] + tv1 * nv[3 * i + d];
                    c.vel[3 * i + d] = c.var[3 * i + d];
                }
            } else {
                for (int d = 0; d < 3; ++d) c.var[3 * i + d] = c.vel[3 * i + d];
            }
        }
    } else {
        c.var = c.vel;
    }
    return 0;
}

// Leapfrog_trajectories.f90:48-76
int or_leapfrog_traj(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    c.coord_old.resize(3 * (size_t)c.n);
    for (int i = 0; i < c.n; ++i) {
        if (c.cella[i] <= 0) continue;
        for (int d = 0; d < 3; ++d) c.coord_old[3 * i + d] = c.coord[3 * i + d];
        for (int d = 0; d < 3; ++d) c.coord[3 * i + d] = c.coord[3 * i + d] + c.dt * c.vel[3 * i + d];
    }
    return 0;
}

// line_plane_intersection.f90:39-76 with three_plane_intersection.f90:42-66
// and Matrix_Inversion_3x3.f90:39-58: the point where the line through P1
// and P2 meets the plane through A, B and C, as the intersection of two
// planes through the line (each with a point 999 away from P1) and the
// plane; false when the 3x3 system's |determinant| is below 1e-36.
static bool line_plane(const double P1[3], const double P2[3], const double A[3], const double B[3],
                const double C[3], double X[3]) {
    const double Q1[3] = {P1[0] - 999.0, P1[1], P1[2]};
    const double Q2[3] = {P1[0], P1[1] - 999.0, P2[2] == P2[2] ? P1[2] - 999.0 : P1[2]};
    auto plane_line = [&](const double Q[3], double& a, double& b, double& cc, double& d) {
        a = (P2[1] - P1[1]) * (Q[2] - P1[2]) - (Q[1] - P1[1]) * (P2[2] - P1[2]);
        b = -((P2[0] - P1[0]) * (Q[2] - P1[2]) - (Q[0] - P1[0]) * (P2[2] - P1[2]));
        cc = (P2[0] - P1[0]) * (Q[1] - P1[1]) - (Q[0] - P1[0]) * (P2[1] - P1[1]);
        d = -(a * P1[0] + b * P1[1] + cc * P1[2]);
    };
    double m[3][3], dv[3];
    plane_line(Q1, m[0][0], m[0][1], m[0][2], dv[0]);
    plane_line(Q2, m[1][0], m[1][1], m[1][2], dv[1]);
    m[2][0] = (B[1] - A[1]) * (C[2] - A[2]) - (C[1] - A[1]) * (B[2] - A[2]);
    m[2][1] = -((B[0] - A[0]) * (C[2] - A[2]) - (C[0] - A[0]) * (B[2] - A[2]));
    m[2][2] = (B[0] - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (B[1] - A[1]);
    dv[2] = -(m[2][0] * A[0] + m[2][1] * A[1] + m[2][2] * A[2]);
    const double det = m[0][0] * m[1][1] * m[2][2] + m[1][0] * m[2][1] * m[0][2] +
                       m[2][0] * m[0][1] * m[1][2] - m[0][0] * m[2][1] * m[1][2] -
                       m[2][0] * m[1][1] * m[0][2] - m[1][0] * m[0][1] * m[2][2];
    if (!(std::fabs(det) >= 1.0e-36)) return false;
    double inv[3][3];
    inv[0][0] = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    inv[0][1] = m[0][2] * m[2][1] - m[0][1] * m[2][2];
    inv[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    inv[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    inv[1][1] = m[0][0] * m[2][2] - m[0][2] * m[2][0];
    inv[1][2] = m[0][2] * m[1][0] - m[0][0] * m[1][2];
    inv[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    inv[2][1] = m[0][1] * m[2][0] - m[0][0] * m[2][1];
    inv[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) inv[i][j] = inv[i][j] / det;
    const double b[3] = {-dv[0], -dv[1], -dv[2]};
    for (int i = 0; i < 3; ++i) X[i] = inv[i][0] * b[0] + inv[i][1] * b[1] + inv[i][2] * b[2];
    return true;
}

// CancelOutgoneParticles_3D.f90:67-160: a particle whose last trajectory step
// crossed an open face (from its inner to its outer side, through the face)
// gets cella = -1, and the next OrdGrid1 removes it.  Returns the number of
// particles marked.
int or_cancel_outgone(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    int marked = 0;
    if (c.coord_old.size() < 3 * (size_t)c.n) return 0;
    for (int f = 0; f < c.nf; ++f) {
        if (c.btype[f] != 2) continue;
        const int nodes = c.nodes[f];
        const double* G1 = &c.GX[12 * f];
        const double* G2 = &c.GX[12 * f + 3];
        const double* G3 = &c.GX[12 * f + 6];
        const double* Gn = &c.GX[12 * f + 3 * (nodes - 1)];
        const double dx21 = G2[0] - G1[0], dx31 = G3[0] - G1[0];
        const double dy21 = G2[1] - G1[1], dy31 = G3[1] - G1[1];
        const double dz21 = G2[2] - G1[2], dz31 = G3[2] - G1[2];
        double Pp[3][3];  // the face's nodes 1-3 in its local system (from its last node)
        const double* Gk[3] = {G1, G2, G3};
        for (int k = 0; k < 3; ++k)
            for (int sdi = 0; sdi < 3; ++sdi) {
                double a = 0.0;
                for (int sdj = 0; sdj < 3; ++sdj) a = a + Tm(c, f, sdj, sdi) * (Gk[k][sdj] - Gn[sdj]);
                Pp[k][sdi] = a;
            }
        auto det_of = [&](const double* X) {
            const double dx = X[0] - G1[0], dy = X[1] - G1[1], dz = X[2] - G1[2];
            return dx21 * dy31 * dz + dx31 * dy * dz21 + dx * dy21 * dz31 - dx21 * dy * dz31 -
                   dx31 * dy21 * dz - dx * dy31 * dz21;
        };
        for (int i = 0; i < c.n; ++i) {
            if (c.cella[i] == 0) continue;
            const double* Xn = &c.coord[3 * i];
            const double* Xo = &c.coord_old[3 * i];
            const double dn = det_of(Xn);
            if (!(dn <= 0.0)) continue;
            const double dpo = det_of(Xo);
            if (std::copysign(1.0, dn) == std::copysign(1.0, dpo)) continue;
            double Ln[3], Lo[3];
            for (int sdi = 0; sdi < 3; ++sdi) {
                double a = 0.0, b = 0.0;
                for (int sdj = 0; sdj < 3; ++sdj) {
                    a = a + Tm(c, f, sdj, sdi) * (Xn[sdj] - Gn[sdj]);
                    b = b + Tm(c, f, sdj, sdi) * (Xo[sdj] - Gn[sdj]);
                }
                Ln[sdi] = a;
                Lo[sdi] = b;
            }
            double LX[3];
            if (!line_plane(Ln, Lo, Pp[0], Pp[1], Pp[2], LX)) continue;
            if (point_internal(c, f, LX[0], LX[1])) {
                c.cella[i] = -1;
                marked += 1;
            }
        }
    }
    // CancelOutgoneParticles_2D.f90:60-104: the same over the open sides (x-z
    // plane): the step's end lies on the outer side of the side's line
    // (determinant <= 0) and its start did not, and the side's two vertices
    // lie on opposite sides of the step's line
    for (int s = 0; s < c.ns; ++s) {
        if (c.sbtype[s] != 2) continue;
        const double o1x = c.sv1[3 * s], o1z = c.sv1[3 * s + 2];
        const double o2x = c.sv2[3 * s], o2z = c.sv2[3 * s + 2];
        for (int i = 0; i < c.n; ++i) {
            if (c.cella[i] == 0) continue;
            const double nx = c.coord[3 * i], nz = c.coord[3 * i + 2];
            const double ox = c.coord_old[3 * i], oz = c.coord_old[3 * i + 2];
            const double dn = (nz - o1z) * (o2x - o1x) - (nx - o1x) * (o2z - o1z);
            if (!(dn <= 0.0)) continue;
            const double dold = (oz - o1z) * (o2x - o1x) - (ox - o1x) * (o2z - o1z);
            if (std::copysign(1.0, dn) == std::copysign(1.0, dold)) continue;
            const double d1 = (o1z - oz) * (nx - ox) - (o1x - ox) * (nz - oz);
            const double d2 = (o2z - oz) * (nx - ox) - (o2x - ox) * (nz - oz);
            if (std::copysign(1.0, d1) == std::copysign(1.0, d2)) continue;
            c.cella[i] = -1;
            marked += 1;
        }
    }
    c.n_outgone += marked;
    return marked;
}

// time_step_loop.f90:338-360: SASPH_continuity.f90:89-169 with
// AddBoundaryContribution_to_CE3D.f90:82-169, then Continuity_Equation.f90:97-249
int or_continuity(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    const int n = c.n;
    const int mj = c.p.nmaxpartj;
    const size_t mc = c.p.maxclose;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int npi = 0; npi < n; ++npi) {
        double dd = 0.0;
        c.koddens[npi] = 0;
        const double* vi = &c.var[3 * npi];
        const int ncbf = c.p.semi ? c.bn[npi] : 0;
        if (ncbf > 0 && c.dim == 2) {
            // SASPH_continuity.f90:126-169 with AddBoundaryContribution_to_CE2D.f90:
            // 74-150 (fixe/tapi sides): grad_u_SA(1), grad_w_SA(3)
            double gu[3] = {0.0, 0.0, 0.0}, gw[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < ncbf; ++q) {
                const size_t sl = (size_t)npi * mc + q;
                const int sd = c.bface[sl];
                if (c.sbtype[sd] != 0 && c.sbtype[sd] != 1) continue;
                const double IntWdS = c.bbi[8 * sl];
                double dvel[2];
                if (c.sbtype[sd] == 0) {
                    dvel[0] = 2.0 * (-vi[0]);
                    dvel[1] = 2.0 * (-vi[2]);
                } else {
                    dvel[0] = 2.0 * (c.sfvel[3 * sd] - vi[0]);
                    dvel[1] = 2.0 * (c.sfvel[3 * sd + 2] - vi[2]);
                }
                for (int ii = 0; ii < 2; ++ii) {
                    const int a = ii == 0 ? 0 : 2;
                    gu[a] = gu[a] - dvel[0] * ST(c, sd, a, 2) * IntWdS;
                    gw[a] = gw[a] - dvel[1] * ST(c, sd, a, 2) * IntWdS;
                }
            }
            if (c.koddens[npi] == 0) dd = dd - c.dens[npi] * (gu[0] + 0.0 + gw[2]);
        } else if (ncbf > 0) {
            double gu = 0.0, gv = 0.0, gw = 0.0;  // grad_u_SA(1), grad_v_SA(2), grad_w_SA(3)
            for (int q = 0; q < ncbf; ++q) {
                const size_t s = (size_t)npi * mc + q;
                const int f = c.bface[s];
                const int bt = c.btype[f];
                if (bt == 0 || bt == 1) {
                    if (c.bloc[3 * s + 2] > 0.0) {
                        double a2[3], nrm[3];
                        for (int d = 0; d < 3; ++d) {
                            a2[d] = 2.0 * (c.fvel[3 * f + d] - vi[d]);
                            nrm[d] = Tm(c, f, d, 2);
                        }
                        const double dp = a2[0] * nrm[0] + a2[1] * nrm[1] + a2[2] * nrm[2];
                        double dvel[3];
                        for (int d = 0; d < 3; ++d) dvel[d] = dp * nrm[d];
                        const double* BI = &c.bbi[8 * s];
                        double av[3];
                        for (int r = 0; r < 3; ++r) {
                            double a = 0.0;
                            for (int k = 0; k < 3; ++k) a = a + Tm(c, f, r, k) * BI[3 + k];
                            av[r] = a;
                        }
                        gu = gu - dvel[0] * av[0];
                        gv = gv - dvel[1] * av[1];
                        gw = gw - dvel[2] * av[2];
                    }
                } else if (bt == 4 || bt == 5 || bt == 3) {
                    bad += 1;
                    break;
                }
            }
            if (c.koddens[npi] == 0) dd = dd - c.dens[npi] * (gu + gv + gw);
        }
        const double rhoi = c.dens[npi];
        double a1 = 0.0, a2 = 0.0, a3 = 0.0;
        for (int k = 0; k < c.nnb[npi]; ++k) {
            const size_t s = (size_t)npi * mj + k;
            const int npj = c.nbr[s];
            const double rhoj = c.dens[npj], amassj = c.mass[npj];
            const double* vj = &c.var[3 * npj];
            const double dv0 = vj[0] - vi[0], dv1 = vj[1] - vi[1], dv2 = vj[2] - vi[2];
            const double* R = &c.rag[3 * s];
            const double K1 = c.pk[4 * s];
            a1 = a1 + rhoi * (amassj / rhoj) * K1 * dv0 * R[0];
            a2 = a2 + rhoi * (amassj / rhoj) * K1 * dv1 * R[1];
            a3 = a3 + rhoi * (amassj / rhoj) * K1 * dv2 * R[2];
        }
        c.dden[npi] = dd - (a1 + a2 + a3);
    }
    return bad ? 94 : 0;
}

// Leapfrog_continuity.f90:51-68 (semi: no division by the Shepard
// coefficient; the density_thresholds clamp when the deck sets it)
int or_leapfrog_cont(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    for (int i = 0; i < c.n; ++i) {
        c.dens[i] = c.dens[i] + c.dt * c.dden[i];
        if (c.p.density_thresholds == 1) {
            if (c.dens[i] < (0.9 * c.p.den0)) {
                c.dens[i] = 0.9 * c.p.den0;
                c.n_dclamp.fetch_add(1, std::memory_order_relaxed);
            }
            if (c.dens[i] > (1.1 * c.p.den0)) {
                c.dens[i] = 1.1 * c.p.den0;
                c.n_dclamp.fetch_add(1, std::memory_order_relaxed);
            }
        }
    }
    return 0;
}

void or_set_lists_only(void* ctx, int on) { ((Ctx*)ctx)->lists_only = on != 0; }

void or_branch_counts(void* ctx, int64_t* out6) {
    const Ctx& c = *(const Ctx*)ctx;
    out6[0] = c.n_elastic.load();
    out6[1] = c.n_zpmin.load();
    out6[2] = c.n_dclamp.load();
    out6[3] = c.n_wf_near.load();
    out6[4] = c.n_wf_log.load();
    out6[5] = c.n_edge.load();
}

long long or_outgone_count(void* ctx) { return ((Ctx*)ctx)->n_outgone; }

// EoS_fluid_loop.f90:59-70, EoS_barotropic_linear.f90:49-55
int or_eos(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    if (c.on_going <= 0) return 0;
    for (int i = 0; i < c.n; ++i) c.pres[i] = c.p.eps / c.p.den0 * (c.dens[i] - c.p.den0) + c.p.prif;
    return 0;
}

// PressureSmoothing_3D.f90:80-161
int or_psmooth(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    const int n = c.n;
    const int mj = c.p.nmaxpartj;
    const size_t mc = c.p.maxclose;
    if (!(c.p.tetap > 0.0)) return 0;
#pragma omp parallel for schedule(static)
    for (int npi = 0; npi < n; ++npi) {
        if (c.koddens[npi] != 0) continue;
        double diffp = 0.0;
        const int nsp = c.nnb[npi];
        if (nsp > 0) {
            const double p0i = c.p.prif, pi = c.pres[npi];
            double somp = 0.0, app = 0.0;
            for (int k = 0; k < nsp; ++k) {
                const size_t s = (size_t)npi * mj + k;
                const int npj = c.nbr[s];
                const double pes = c.mass[npj] * c.pk[4 * s + 3] / c.dens[npj];
                app = app + pes;
                somp = somp + (c.pres[npj] - pi) * pes;
            }
            const int ncbf = c.p.semi ? c.bn[npi] : 0;
            if (c.dim == 2) {
                // PressureSmoothing_2D.f90:124-214, SmoothVersion "b", fixe/tapi sides
                double VFT = 0.0;
                const double ro0i = c.p.den0;
                for (int q = 0; q < ncbf; ++q) {
                    const size_t sl = (size_t)npi * mc + q;
                    const int sd = c.bface[sl];
                    if (c.sbtype[sd] != 0 && c.sbtype[sd] != 1) continue;
                    const double sss[2] = {ST(c, sd, 0, 0), ST(c, sd, 2, 0)};
                    const double nnn[2] = {ST(c, sd, 0, 2), ST(c, sd, 2, 2)};
                    const double mf0 = sss[0] * c.p.grav[0] + sss[1] * c.p.grav[2];
                    const double mf1 = nnn[0] * c.p.grav[0] + nnn[1] * c.p.grav[2];
                    const double IntWdV = c.bbi[8 * sl + 1];
                    VFT = VFT + IntWdV;
                    somp = somp + ro0i * (mf0 * c.bbi[8 * sl + 3] + mf1 * c.bbi[8 * sl + 4]);
                }
                const double TetaP1 = c.p.tetap * c.p.celerita * c.dt / c.p.h;
                app = app + VFT + 0.0 + 0.0 + 0.0;
                c.vpres[npi] = pi + TetaP1 * (somp + (pi - pi) * 0.0 + (pi - pi) * 0.0 + (p0i - pi) * (1.0 - app));
                continue;
            }
            if (c.dbsph) {  // PressureSmoothing_3D.f90:110-112
                c.vpres[npi] = somp / app;
                continue;
            }
            for (int q = 0; q < ncbf; ++q) {
                const size_t s = (size_t)npi * mc + q;
                if (c.bloc[3 * s + 2] > 0.0) app = app + c.bbi[8 * s + 1];
            }
            if (app < MinTotUnit)
                diffp = somp + (p0i - pi) * (1.0 - app);
            else
                diffp = somp / app;
        }
        c.vpres[npi] = diffp;
    }
    const double tp1 = c.p.tetap * c.p.celerita * c.dt / c.p.h;
    for (int i = 0; i < n; ++i) {
        if (c.koddens[i] != 0) continue;
        if (c.dim == 2) {  // PressureSmoothing_2D.f90:217-231: p = vpres, rho = EoS^-1(p)
            c.pres[i] = c.vpres[i];
            c.dens[i] = c.p.den0 * (1.0 + (c.pres[i] - c.p.prif) / c.p.eps);
            continue;
        }
        c.pres[i] = c.pres[i] + tp1 * c.vpres[i];
        c.dens[i] = c.p.den0 * (1.0 + (c.pres[i] - c.p.prif) / c.p.eps);
    }
    return 0;
}

// time_step_loop.f90:82-173 (semi branch)
// time_step_loop.f90:82-135 for Domain%tipo = "bsph": CalcVarLength with
// on_going_time_step = -2 (Gamma from sigma), the reservoir removal (none
// here: no fictitious reservoirs), OrdGrid1, CalcVarLength with -1, then the
// wall elements' state (Gradients_to_MUSCL, BC_wall_elements)
static int preloop_dbsph(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    c.on_going = -2;
    or_calcvarlength(ctx);
    c.on_going = c.it;
    or_ordgrid(ctx);
    c.on_going = -1;
    or_calcvarlength(ctx);
    c.on_going = c.it;
    if (c.n > 0 && c.nw > 0) dbsph_walls_state(c);
    if (c.it == 0) or_dt(ctx);
    return c.overflow ? 89 : 0;
}

int or_preloop(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    if (c.dbsph) return preloop_dbsph(ctx);
    c.on_going = -999;
    int rc = or_calcvarlength(ctx);
    if (!rc) rc = or_boundary_tab(ctx);
    if (!rc && c.it == 0) rc = or_dt(ctx);
    return rc;
}

// time_step_loop.f90:175-489 with time_split=0, RKscheme=1 (one stage): lists
// and boundary table at the start of the stage (:192-212), RHS_momentum_equation,
// the continuity loop (:335-362), time_integration (time_integration.f90:52-65:
// Euler, then EoS_fluid_loop), PressureSmoothing, and OrdGrid1 at the end of
// the stage (:425-458).  Euler.f90:61-241 for semi: vel += dt acc
// (kodvel = 0), velocity_smoothing and var = vel + TetaV1 var, vel = var
// (:85-115), coord += dt var (:122-139), dens += dt dden (:231-241).  The
// pressure smoothing reads the kernel values and boundary table of the
// stage's start, as the reference's PartKernel(4) and BoundaryDataTab are.
static int step_rk1(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    c.it = c.it + 1;
    c.on_going = c.it;
    if (c.n > 0) or_dt(ctx);
    c.time = c.time + c.dt;
    or_calcvarlength(ctx);
    int rc = or_boundary_tab(ctx);
    if (rc) return rc;
    rc = or_momentum(ctx);
    if (rc) return rc;
    rc = or_continuity(ctx);
    if (rc) return rc;
    for (int i = 0; i < c.n; ++i)
        for (int d = 0; d < 3; ++d) c.vel[3 * i + d] = c.vel[3 * i + d] + c.dt * c.acc[3 * i + d];
    or_vsmooth(ctx);
    or_leapfrog_traj(ctx);
    for (int i = 0; i < c.n; ++i) c.dens[i] = c.dens[i] + c.dt * c.dden[i];
    or_eos(ctx);
    or_psmooth(ctx);
    // time_step_loop.f90:417-423: open faces, then OrdGrid1
    or_cancel_outgone(ctx);
    or_ordgrid(ctx);
    return 0;
}

// time_step_loop.f90:175-489 for Domain%tipo = "bsph" with RK1: no CalcVarLength
// at the stage's start (the lists are those of the previous step's end),
// RHS_momentum_equation, Continuity_Equation, then Euler.f90:61-241 for
// bsph: vel += dt acc, velocity_smoothing and var, coord += dt var,
// OrdGrid1, Gamma += dShep dt, CalcVarLength, dden / Gamma and
// dens = rhoSPH_new / Gamma (/ sigma at the free surface, not below
// 0.98 rho0); EoS_fluid_loop, PressureSmoothing_3D, and the wall elements'
// state (:470-473)
// DBSPH_kinematics.f90:52-107 and Euler.f90:152-159: the mesh's velocity and
// angular velocity at simulation_time from its records (the first record at or
// after it, linearly interpolated from the one before); each wall element's
// velocity omega x (x - centre) + v; then every element moves by dt vel.  The
// elements keep their cells (OrdGrid1.f90:108 bins them on the first step only).
static void dbsph_move_walls(Ctx& c) {
    const int nrec = (int)(c.kin.size() / 7);
    if (nrec > 0) {
        double va[3] = {0.0, 0.0, 0.0}, om[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < nrec; ++j) {
            const double* r = &c.kin[7 * j];
            if (r[0] >= c.time) {
                if (r[0] == c.time) {
                    for (int d = 0; d < 3; ++d) {
                        va[d] = r[1 + d];
                        om[d] = r[4 + d];
                    }
                } else {
                    const double* q = &c.kin[7 * (j - 1)];
                    for (int d = 0; d < 3; ++d) {
                        va[d] = q[1 + d] + (r[1 + d] - q[1 + d]) / (r[0] - q[0]) * (c.time - q[0]);
                        om[d] = q[4 + d] + (r[4 + d] - q[4 + d]) / (r[0] - q[0]) * (c.time - q[0]);
                    }
                }
                break;
            }
        }
        for (int w = 0; w < c.nw; ++w) {
            double rel[3], v[3];
            for (int d = 0; d < 3; ++d) rel[d] = c.wcoord[3 * w + d] - c.kin_centre[d];
            v[0] = om[1] * rel[2] - om[2] * rel[1];
            v[1] = om[2] * rel[0] - om[0] * rel[2];
            v[2] = om[0] * rel[1] - om[1] * rel[0];
            for (int d = 0; d < 3; ++d) c.wvel[3 * w + d] = v[d] + va[d];
        }
    }
    for (int w = 0; w < c.nw; ++w)
        for (int d = 0; d < 3; ++d) c.wcoord[3 * w + d] = c.wcoord[3 * w + d] + c.dt * c.wvel[3 * w + d];
}

int or_set_dbsph_kinematics(void* ctx, int nrec, const double* records, const double* centre) {
    Ctx& c = *(Ctx*)ctx;
    if (nrec < 0 || (nrec > 0 && (!records || !centre))) return 1;
    if (nrec > 0 && records[0] > 0.0) return 2;  // the reference would read before its first record
    c.kin.assign(records, records + 7 * (size_t)nrec);
    for (int d = 0; d < 3; ++d) c.kin_centre[d] = nrec > 0 ? centre[d] : 0.0;
    return 0;
}

static int step_dbsph(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    c.it = c.it + 1;
    c.on_going = c.it;
    if (c.n > 0) or_dt(ctx);
    c.time = c.time + c.dt;
    int rc = or_momentum(ctx);
    if (rc) return rc;
    rc = or_continuity(ctx);
    if (rc) return rc;
    for (int i = 0; i < c.n; ++i)
        for (int d = 0; d < 3; ++d) c.vel[3 * i + d] = c.vel[3 * i + d] + c.dt * c.acc[3 * i + d];
    or_vsmooth(ctx);
    or_leapfrog_traj(ctx);
    dbsph_move_walls(c);
    or_ordgrid(ctx);
    for (int i = 0; i < c.n; ++i) {
        c.gamma[i] = c.gamma[i] + c.dshep[i] * c.dt;
        if (c.gamma_limiter && c.gamma[i] >= 1.0) c.gamma[i] = 1.0;
    }
    or_calcvarlength(ctx);
    if (c.overflow) return 89;
    for (int i = 0; i < c.n; ++i) {
        if (c.gamma_limiter && c.nfw[i] == 0) c.gamma[i] = 1.0;
        c.dden[i] = c.dden[i] / c.gamma[i];
        if (c.fs[i] == 0) {
            c.dens[i] = c.rhosph[i] / c.gamma[i];
        } else {
            c.dens[i] = c.rhosph[i] / c.sigma[i];
            if (c.dens[i] < (0.98 * c.p.den0)) c.dens[i] = 0.98 * c.p.den0;
        }
    }
    or_eos(ctx);
    or_psmooth(ctx);
    if (c.n > 0 && c.nw > 0) dbsph_walls_state(c);
    return 0;
}

// time_step_loop.f90:175-489 (time_split=1, semi, no open/inlet/zmax zones)
int or_step(void* ctx, int nsteps) {
    Ctx& c = *(Ctx*)ctx;
    if (c.dbsph) {
        for (int s = 0; s < nsteps; ++s)
            if (int rc = step_dbsph(ctx)) return rc;
        return 0;
    }
    if (c.p.time_split == 0) {
        for (int s = 0; s < nsteps; ++s)
            if (int rc = step_rk1(ctx)) return rc;
        return 0;
    }
    for (int s = 0; s < nsteps; ++s) {
        c.it = c.it + 1;
        c.on_going = c.it;
        const double dt_prev = c.dt;
        if (c.n > 0) or_dt(ctx);
        c.time = c.time + c.dt;
        int rc = or_momentum(ctx);
        if (rc) return rc;
        or_leapfrog_mom(ctx, dt_prev);
        or_vsmooth(ctx);
        or_leapfrog_traj(ctx);
        or_cancel_outgone(ctx);  // time_step_loop.f90:265-269
        or_ordgrid(ctx);
        or_calcvarlength(ctx);
        rc = or_boundary_tab(ctx);
        if (rc) return rc;
        rc = or_continuity(ctx);
        if (rc) return rc;
        or_leapfrog_cont(ctx);
        or_eos(ctx);
        or_psmooth(ctx);
    }
    return 0;
}

int64_t or_pair_count(void* ctx) {
    Ctx& c = *(Ctx*)ctx;
    int64_t s = 0;
    for (int i = 0; i < c.n; ++i) s += c.nnb[i] - 1;
    return s;
}

int or_overflow(void* ctx) { return ((Ctx*)ctx)->overflow; }

// DB-SPH walls: the wall elements of DBSPH_IC_surface_elements.f90 and
// semi_particle_volumes.f90 (see sph_oracle.h); bins them into the grid
// cells in ascending id (OrdGrid1.f90:108-137) and starts the particles' and
// the walls' DB-SPH state as the reference's initial values do
int or_set_dbsph(void* ctx, int nw, const double* coord, const double* normal, const double* vel,
                 const double* weight, const double* volume, const double* mass, const int32_t* flags) {
    Ctx& c = *(Ctx*)ctx;
    c.dbsph = true;
    c.nw = nw;
    c.wcoord.assign(coord, coord + 3 * nw);
    c.wnormal.assign(normal, normal + 3 * nw);
    c.wvel.assign(vel, vel + 3 * nw);
    c.wweight.assign(weight, weight + nw);
    c.wvolume.assign(volume, volume + nw);
    c.wmass.assign(mass, mass + nw);
    c.wdens.assign(nw, c.p.den0);
    c.wpres.assign(nw, c.p.prif);
    c.wwet.assign(nw, 0);
    c.gamma_limiter = flags[0];
    c.fs_allowed = flags[1];
    c.neg_wall_p = flags[2];
    c.muscl_boundary = flags[3];
    const int n = c.n;
    c.gamma.assign(n, 0.0);
    c.sigma.assign(n, 0.0);
    c.rhosph.assign(n, 0.0);
    c.rhosph_old.assign(n, 0.0);
    c.dshep.assign(n, 0.0);
    c.fs.assign(n, 0);
    c.drho.assign(3 * (size_t)n, 0.0);
    c.dvelg.assign(9 * (size_t)n, 0.0);
    c.nfw.assign(n, 0);
    c.icont_w.assign(c.nmax + 2, 0);
    std::vector<int> cell(nw), inc(c.nmax + 1, 0);
    for (int w = 0; w < nw; ++w) {
        cell[w] = particle_cell(c, &c.wcoord[3 * w]);
        if (cell[w] <= 0) return 2;  // a wall element outside the grid
        c.icont_w[cell[w]] += 1;
        c.icont_w[c.nmax + 1] += 1;
    }
    for (int i = c.nmax; i >= 1; --i) c.icont_w[i] = c.icont_w[i + 1] - c.icont_w[i];
    for (int i = 1; i <= c.nmax + 1; ++i) c.icont_w[i] = c.icont_w[i] + 1;
    c.npartord_w.assign(nw + 1, 0);
    for (int w = 0; w < nw; ++w) {
        c.npartord_w[c.icont_w[cell[w]] + inc[cell[w]]] = w;
        inc[cell[w]] += 1;
    }
    return 0;
}

void or_get_dbsph(void* ctx, double* gamma, double* sigma, double* rhosph, double* dshep, double* drho,
                  double* dvel, int32_t* fs, double* wall_dens, double* wall_pres, int32_t* wall_wet) {
    Ctx& c = *(Ctx*)ctx;
    const int n = c.n;
    if (gamma) std::copy(c.gamma.begin(), c.gamma.begin() + n, gamma);
    if (sigma) std::copy(c.sigma.begin(), c.sigma.begin() + n, sigma);
    if (rhosph) std::copy(c.rhosph.begin(), c.rhosph.begin() + n, rhosph);
    if (dshep) std::copy(c.dshep.begin(), c.dshep.begin() + n, dshep);
    if (drho) std::copy(c.drho.begin(), c.drho.begin() + 3 * (size_t)n, drho);
    if (dvel) std::copy(c.dvelg.begin(), c.dvelg.begin() + 9 * (size_t)n, dvel);
    if (fs) std::copy(c.fs.begin(), c.fs.begin() + n, fs);
    if (wall_dens) std::copy(c.wdens.begin(), c.wdens.end(), wall_dens);
    if (wall_pres) std::copy(c.wpres.begin(), c.wpres.end(), wall_pres);
    if (wall_wet) std::copy(c.wwet.begin(), c.wwet.end(), wall_wet);
}

int or_get_fw_lists(void* ctx, int32_t* count, int32_t* ids, int32_t stride) {
    Ctx& c = *(Ctx*)ctx;
    const int mj = c.p.nmaxpartj;
    for (int i = 0; i < c.n; ++i) {
        const int k = i < (int)c.nfw.size() ? c.nfw[i] : 0;
        count[i] = k;
        for (int q = 0; q < k && q < stride; ++q) ids[(size_t)i * stride + q] = c.fwl[(size_t)i * mj + q];
    }
    return 0;
}

// Free-surface monitor lines: calc_pelo.f90:82-131, with the point cells of
// CalcVarp.f90:66-72 (ceiling of the offset over the cell size, CellNumber).
// icont holds control_lines(i)%Icont(1:2) (1-based, inclusive); the loop
// starts at Icont(1)+1 as the reference's does.  The two closest particles
// come from two minloc passes over an NMAXPARTJ array preset to 9e12
// (calc_pelo.f90:108-122).  Returns 3 if the cell lists are not built, 4 if
// icont is out of range.
int or_free_surface(void* ctx, int nlines, const int32_t* icont, int npoints, const double* pts,
                    double* pelolib, double* zlft) {
    Ctx& c = *(Ctx*)ctx;
    if ((int)c.icont.size() != c.nmax + 2) return 3;
    for (int i = 0; i < nlines; ++i)
        if (icont[2 * i] < 1 || icont[2 * i + 1] > npoints) return 4;
    const int nmpj = c.p.nmaxpartj;
    std::vector<int> particle_ID;
    std::vector<double> dis2;
    for (int i = 0; i < 3 * nlines; ++i) pelolib[i] = zlft[i] = 0.0;
    for (int i = 0; i < nlines; ++i) {
        bool zlft_flag = true, fluid_below = false;
        for (int j = icont[2 * i] + 1; j <= icont[2 * i + 1]; ++j) {
            const double* q = &pts[3 * (j - 1)];
            int ijk[3];
            bool inside = true;
            for (int d = 0; d < 3; ++d) {
                const double f = std::ceil((q[d] - c.p.extr_lo[d]) / c.p.dcd[d]);
                inside = inside && f >= 1.0 && f <= (double)c.p.ncd[d];
                ijk[d] = inside ? (int)f : 0;
            }
            const int ncel = inside ? cell_number(c, ijk[0], ijk[1], ijk[2]) : 0;
            if (ncel == 0) continue;
            if (c.icont[ncel + 1] <= c.icont[ncel]) {
                if (fluid_below) zlft_flag = false;
                continue;
            }
            const int cnt = c.icont[ncel + 1] - c.icont[ncel];
            const size_t len = (size_t)std::max(nmpj, cnt);
            particle_ID.assign(len, -1);
            dis2.assign(len, 9.0e12);
            int npartcel = 0;
            for (int mm = c.icont[ncel]; mm <= c.icont[ncel + 1] - 1; ++mm) {
                const int jj = c.npartord[mm];
                const double r0 = q[0] - c.coord[3 * jj];
                const double r1 = q[1] - c.coord[3 * jj + 1];
                const double r2 = q[2] - c.coord[3 * jj + 2];
                particle_ID[npartcel] = jj;
                dis2[npartcel] = r0 * r0 + r1 * r1 + r2 * r2;
                npartcel += 1;
            }
            auto minloc = [&]() {
                size_t m = 0;
                for (size_t t = 1; t < len; ++t)
                    if (dis2[t] < dis2[m]) m = t;
                return m;
            };
            const size_t m1 = minloc();
            dis2[m1] = 9.0e12;
            const size_t m2 = minloc();
            const int a = particle_ID[m1], b = particle_ID[m2];
            for (int d = 0; d < 3; ++d) {
                pelolib[3 * i + d] = (c.coord[3 * a + d] + c.coord[3 * b + d]) * 0.5 + c.p.dx / 2.0;
                if (zlft_flag) zlft[3 * i + d] = pelolib[3 * i + d];
            }
            fluid_below = true;
        }
    }
    return 0;
}

}  // extern "C"
