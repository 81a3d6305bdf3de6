// DB-SPH walls (Domain%tipo = "bsph", 3D, explicit RK1, FP64): the wall
// elements ("semi-particles") of a surface mesh take part in the fluid
// particles' Shepard sums, density and momentum, and carry a density and
// pressure of their own from a partial Riemann problem with the fluid.
//
//   k_dbsph_lists     CalcVarLength.f90:134-140,330-340,462-540,842-877 and, in
//                     the step, Euler.f90:185-241 around it: Gamma, the
//                     fluid-wall list, sigma, rhoSPH_new, dShep, FS, and the
//                     density from rhoSPH_new / Gamma with the EoS
//   k_momentum_dbsph  RHS_momentum_equation.f90:173-177 with inter_EqMoto.f90
//                     (DB-SPH pressure term :229-235, wall terms :286-322)
//   k_dbsph_muscl     Gradients_to_MUSCL.f90:47-91 (with the wall terms of
//                     Gradients_to_MUSCL_boundary.f90 when asked for)
//   k_dbsph_walls     BC_wall_elements.f90:71-158
//   k_dbsph_kinematics DBSPH_kinematics.f90 and the wall move of Euler.f90:152-159
//
// Wall elements do not move.  Their cell lists are built once on the host
// (OrdGrid1.f90:108-137): wstart[c-1] .. wstart[c]-1 index wids, the wall
// elements of cell c in ascending id, and wpos holds the same elements in
// that order as (x, y, z, id) so that a cell's candidates are one contiguous
// read.
#pragma once

namespace sph {

constexpr int SPH_MW = 128;   // fluid-wall list capacity per particle
constexpr int SPH_WF = 128;   // fluid neighbours a wall element gathers
constexpr int DB_G = 4;       // lanes per particle in k_dbsph_lists

// wall-element tables: w0 = (x, y, z, weight), w1 = (normal, volume),
// w2 = (velocity, mass)
struct DWall {
    const double4* w0;
    const double4* w1;
    const double4* w2;
    const int* wstart;
    const int* wids;
    const int* wcell;
    const double4* wpos;
};

// MODE 0: the pre-loop's first call (on_going_time_step = -2: Gamma = sigma,
// or 1 without wall neighbours); 1: its second call (Gamma kept, FS set);
// 2: inside Euler after OrdGrid1 (Gamma += dShep dt before the sums, then
// dden / Gamma and the density and pressure of the particle)
// The fluid sums walk the particle's neighbour list, whose entries carry the
// stencil code (dx+1) + 3(dy+1) + 9(dz+1) of their cell and come in the
// reference's cell order (y, x, z nest; ascending pg index in a cell): each
// cell's run of the list, then that cell's wall elements.
template <int MODE>
__global__ __launch_bounds__(256) void k_dbsph_lists(
    DP P, DWall W, const double4* __restrict__ xm, const double4* __restrict__ vel,
    const double2* __restrict__ rp, const int* __restrict__ cell, const uint32_t* __restrict__ nbr,
    const int* __restrict__ nnb,
    const double* __restrict__ gam_in, const double* __restrict__ dsh_in, double* __restrict__ gam_out,
    double* __restrict__ dsh_out, double* __restrict__ sig_out, double* __restrict__ rsph_out,
    int* __restrict__ fs_out, int* __restrict__ nfw, int* __restrict__ fwl, int* __restrict__ wwet,
    double* __restrict__ dden, double2* __restrict__ rp_out, const double* scal, int* iscal) {
    // DB_G lanes per particle: they take consecutive candidates of a cell
    // (fluid-list entries, then wall elements) side by side, and every lane
    // of the group adds the candidates' terms in candidate order, read from
    // their lanes with shuffles, so the sums keep the serial order
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / DB_G, g = t % DB_G;
    const int gb = (threadIdx.x & 63) & ~(DB_G - 1);  // the group's first lane
    const uint64_t gmask = (1ull << DB_G) - 1ull;
    if (i >= iscal[IS_N]) return;  // whole groups leave together
    const double4 xi = xm[i];
    const double4 vi = vel[i];  // var == vel at every CalcVarLength of the DB-SPH step
    double gamma = 0.0;
    if (MODE == 2) {  // Euler.f90:185-189
        gamma = gam_in[i] + dsh_in[i] * scal[SC_DT];
        if (P.gamma_lim && gamma >= 1.0) gamma = 1.0;
    } else if (MODE == 1) {
        gamma = gam_in[i];
    }
    double sigma = 0.0, rsph = 0.0, dshep = 0.0;
    int nf = 0;
    const int cap = SPH_MW < P.nmaxpartj ? SPH_MW : P.nmaxpartj;
    int ci, cj, ck;
    d_cell_indices(P, cell[i], ci, cj, ck);
    const size_t N = P.cap;
    const int nn = nnb[i];
    int kl = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
            for (int dz = -1; dz <= 1; ++dz) {
                // the cell's fluid neighbours, then its wall elements
                const uint32_t code = (uint32_t)((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1));
                for (;;) {
                    const int k = kl + g;
                    uint32_t e = NB_NONE;
                    if (k < nn) e = nbr[nb_index(N, i, k)];
                    const bool val = e != NB_NONE && (e >> NB_SHIFT) == code;
                    double cs = 0.0, cr = 0.0;
                    if (val) {
                        const double4 xj = xm[e & NB_MASK];
                        const double R0 = xi.x - xj.x, R1 = xi.y - xj.y, R2 = xi.z - xj.z;
                        K64 K;
                        kvals64(P, R0 * R0 + R1 * R1 + R2 * R2, false, false, true, K);
                        cs = xj.w * K.W / rp[e & NB_MASK].x;
                        cr = xj.w * K.W;
                    }
                    // the run of this code's entries is a prefix of the group
                    const int cnt = __popcll((__ballot(val) >> gb) & gmask);
#pragma unroll
                    for (int u = 0; u < DB_G; ++u) {
                        const double su = __shfl(cs, u, DB_G);
                        const double ru = __shfl(cr, u, DB_G);
                        if (u < cnt) {
                            sigma = sigma + su;
                            rsph = rsph + ru;
                        }
                    }
                    kl += cnt;
                    if (cnt < DB_G) break;
                }
                const int c = d_cell_number(P, ci + dx, cj + dy, ck + dz);
                if (c == 0) continue;
                const int wa = W.wstart[c - 1], wz = W.wstart[c];
                for (int wb = wa; wb < wz; wb += DB_G) {
                    const int w = wb + g;
                    bool inr = false, add = false;
                    int id = 0;
                    double dv = 0.0, sv = 0.0, rv = 0.0;
                    bool wet = false;
                    if (w < wz) {
                        const double4 a = W.wpos[w];
                        const double R0 = xi.x - a.x, R1 = xi.y - a.y, R2 = xi.z - a.z;
                        const double rij = R0 * R0 + R1 * R1 + R2 * R2;
                        if (rij <= P.square_doubleh) {
                            inr = true;
                            id = (int)a.w;
                            const double q = sqrt(rij) / P.h;
                            const double q2 = rij / P.squareh;
                            const int iq = (int)q;
                            if (iq < 2) {
                                add = true;
                                const double aw = W.w0[id].w;
                                const double4 b = W.w1[id];
                                const double4 gv = W.w2[id];
                                double wu = 0.666666667 + q2 * (q * 0.5 - 1.0);
                                if (iq > 0) wu = (2.0 - q) * (2.0 - q) * (2.0 - q) * 0.166666666667;
                                const double k1 = wu * P.coefke;
                                dv = k1 * aw * (b.x * (gv.x - vi.x) + b.y * (gv.y - vi.y) + b.z * (gv.z - vi.z));
                                wet = (q * P.h) <= (1.3 * P.dx + aw / 2.0);
                                sv = b.w * k1;
                                rv = gv.w * k1;
                            }
                        }
                    }
                    // in candidate order: the elements within 2h take the
                    // next list slots while there are any left
                    const uint64_t mk = (__ballot(inr) >> gb) & gmask;
                    const int slot = nf + __popcll(mk & ((1ull << g) - 1ull));
                    const bool listed = inr && slot < cap;
                    if (inr && !listed) atomicAdd(&iscal[IS_DB_OVERFLOW], 1);
                    if (listed) {
                        fwl[(size_t)slot * P.cap + i] = id;
                        if (add && wet) wwet[id] = 1;
                    }
                    nf = nf + __popcll(mk);
                    if (nf > cap) nf = cap;
                    const uint64_t ma = (__ballot(listed && add) >> gb) & gmask;
#pragma unroll
                    for (int u = 0; u < DB_G; ++u) {
                        const double du = __shfl(dv, u, DB_G);
                        const double su = __shfl(sv, u, DB_G);
                        const double ru = __shfl(rv, u, DB_G);
                        if ((ma >> u) & 1ull) {
                            dshep = dshep + du;
                            sigma = sigma + su;
                            rsph = rsph + ru;
                        }
                    }
                }
            }
    if (g != 0) return;
    nfw[i] = nf;
    sig_out[i] = sigma;
    rsph_out[i] = rsph;
    dsh_out[i] = dshep;
    int fs = 0;
    if (MODE == 0) {  // CalcVarLength.f90:845-853
        if (nf == 0) {
            gamma = 1.0;
        } else {
            gamma = sigma;
            if (P.gamma_lim) gamma = fmin(gamma, 1.0);
        }
    } else {  // CalcVarLength.f90:854-865
        const double msg = fmin(sigma + 0.05, gamma);
        if (P.fs_allowed && msg != gamma) fs = 1;
    }
    if (MODE == 2) {  // Euler.f90:225-239, then EoS_fluid_loop.f90:59-70
        if (P.gamma_lim && nf == 0) gamma = 1.0;
        dden[i] = dden[i] / gamma;
        double rho;
        if (fs == 0) {
            rho = rsph / gamma;
        } else {
            rho = rsph / sigma;
            if (rho < (0.98 * P.den0)) rho = 0.98 * P.den0;
        }
        rp_out[i] = make_double2(rho, P.eps / P.den0 * (rho - P.den0) + P.prif);
    }
    gam_out[i] = gamma;
    fs_out[i] = fs;
}

// RHS_momentum_equation.f90:173-177 and inter_EqMoto.f90 for bsph: the
// symmetric pressure term -m_j (p_i/rho_i^2 + p_j/rho_j^2) r PartKernel(1)
// (no ALE1 complement), the Monaghan and Morris terms as for SA-SPH, then the
// wall elements' pressure terms; the wall shear and semi-particle Morris
// terms read the VSL gradients and semi-particle viscosities, which are zero
// without the VSL model (slip_ID = 0), and add zeros.  acc = (tpres + tdiss
// + tvisc) / Gamma + g; Euler's vel += dt acc (kodvel = 0) is fused.
__global__ __launch_bounds__(256) void k_momentum_dbsph(
    DP P, DWall W, const double4* __restrict__ xm, const double4* __restrict__ vel,
    const double2* __restrict__ rp, const uint32_t* __restrict__ nbr, const int* __restrict__ nnb,
    const int* __restrict__ nfw, const int* __restrict__ fwl, const double* __restrict__ gam,
    const double* __restrict__ wdens, const double* __restrict__ wpres, int* __restrict__ lam_out,
    double4* __restrict__ acc_out, double4* __restrict__ vel_out, const double* scal, const int* iscal) {
    // DB_G lanes per particle take consecutive list entries (then wall
    // elements) side by side; every lane of the group adds the entries'
    // terms in list order, read from their lanes with shuffles
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / DB_G, g = t % DB_G;
    const int gb = (threadIdx.x & 63) & ~(DB_G - 1);
    const uint64_t gmask = (1ull << DB_G) - 1ull;
    if (i >= iscal[IS_N]) return;  // whole groups leave together
    const size_t N = P.cap;
    const double h = P.h;
    const double4 vi = vel[i];
    const double2 rpi = rp[i];
    const double4 xi = xm[i];
    const double rhoi = rpi.x, pi = rpi.y;
    double tp[3] = {0.0, 0.0, 0.0}, td[3] = {0.0, 0.0, 0.0}, tv[3] = {0.0, 0.0, 0.0};
    double tpa[3] = {0.0, 0.0, 0.0}, rvs[3] = {0.0, 0.0, 0.0};
    const double celt = P.celerita + P.celerita;
    const double nui = P.kin_visc, nuj = P.kin_visc;
    const bool morris = nui > 1.0e-12 && nuj > 1.0e-12;
    const int nn = nnb[i];
    for (int k0 = 0; k0 < nn; k0 += DB_G) {
        const int k = k0 + g;
        bool val = false;
        double cp[3] = {0.0, 0.0, 0.0}, cd[3] = {0.0, 0.0, 0.0}, cv[3] = {0.0, 0.0, 0.0};
        if (k < nn) {
            const int j = (int)(nbr[nb_index(N, i, k)] & NB_MASK);
            if (j != i) {
                val = true;
                const double4 xj = xm[j];
                const double4 vj = vel[j];
                const double2 rpj = rp[j];
                const double R0 = xi.x - xj.x, R1 = xi.y - xj.y, R2 = xi.z - xj.z;
                const double rij = R0 * R0 + R1 * R1 + R2 * R2;
                K64 K;
                kvals64(P, rij, true, false, false, K);
                const double rhoj = rpj.x, pj = rpj.y, amassj = xj.w;
                const double dv0 = vj.x - vi.x, dv1 = vj.y - vi.y, dv2 = vj.z - vi.z;
                const double alpha = pi / (rhoi * rhoi) + pj / (rhoj * rhoj);
                cp[0] = -amassj * alpha * R0 * K.K1;
                cp[1] = -amassj * alpha * R1 * K.K1;
                cp[2] = -amassj * alpha * R2 * K.K1;
                const double vrij = -dv0 * R0 - dv1 * R1 - dv2 * R2;
                const double TermMon = P.alfamon * celt * h / (rhoi + rhoj);
                cd[0] = amassj * TermMon * vrij * R0 * K.K2 + 0.0;
                cd[1] = amassj * TermMon * vrij * R1 * K.K2 + 0.0;
                cd[2] = amassj * TermMon * vrij * R2 * K.K2 + 0.0;
                if (morris) {
                    const double rtil = nui * rhoi + nuj * rhoj + 1.0e-3;
                    const double atil = 4.0 * (nui * nuj);
                    const double fac = amassj * atil / rtil;
                    const double rr = R0 * R0 + R1 * R1 + R2 * R2;
                    cv[0] = fac * (-dv0 * K.K2 * rr);
                    cv[1] = fac * (-dv1 * K.K2 * rr);
                    cv[2] = fac * (-dv2 * K.K2 * rr);
                }
            }
        }
        const uint64_t mv = (__ballot(val) >> gb) & gmask;
#pragma unroll
        for (int u = 0; u < DB_G; ++u) {
            double p_[3], d_[3], v_[3];
            for (int d = 0; d < 3; ++d) {
                p_[d] = __shfl(cp[d], u, DB_G);
                d_[d] = __shfl(cd[d], u, DB_G);
            }
            if (morris)
                for (int d = 0; d < 3; ++d) v_[d] = __shfl(cv[d], u, DB_G);
            if ((mv >> u) & 1ull) {
                for (int d = 0; d < 3; ++d) tpa[d] = tpa[d] + p_[d];
                for (int d = 0; d < 3; ++d) td[d] = td[d] + d_[d];
                if (morris) {
                    for (int d = 0; d < 3; ++d) tv[d] = tv[d] + v_[d];
                    for (int d = 0; d < 3; ++d) rvs[d] = rvs[d] + v_[d];
                }
            }
        }
    }
    for (int d = 0; d < 3; ++d) tp[d] = tp[d] + tpa[d] + 0.0;
    int lam = 0;
    if (P.kin_visc > 0.0) {
        const double apg = sqrt(tp[0] * tp[0] + tp[1] * tp[1] + tp[2] * tp[2]) + P.GI;
        const double amo = sqrt(rvs[0] * rvs[0] + rvs[1] * rvs[1] + rvs[2] * rvs[2]);
        double w = 0.0;
        if (apg != 0.0) w = amo / apg * 100.0;
        lam = (w > 5.0 || apg == 0.0) ? 1 : 0;
    }
    const int nf = nfw[i];
    for (int k0 = 0; k0 < nf; k0 += DB_G) {  // inter_EqMoto.f90:286-297
        const int k = k0 + g;
        double ap[3] = {0.0, 0.0, 0.0};
        if (k < nf) {
            const int id = fwl[(size_t)k * N + i];
            const double4 a = W.w0[id];
            const double4 b = W.w1[id];
            const double wm = W.w2[id].w;
            const double R[3] = {xi.x - a.x, xi.y - a.y, xi.z - a.z};
            K64 K;  // kernel_fw(1) = W, kernel_fw(2) = K1
            kvals64(P, R[0] * R[0] + R[1] * R[1] + R[2] * R[2], true, false, true, K);
            const double wd = wdens[id];
            const double A = pi / (rhoi * rhoi) + wpres[id] / (wd * wd);
            const double nw[3] = {b.x, b.y, b.z};
            for (int d = 0; d < 3; ++d) {
                ap[d] = -wd * a.w * K.W * A * nw[d];
                ap[d] = ap[d] - wm * A * R[d] * K.K1;
            }
        }
#pragma unroll
        for (int u = 0; u < DB_G; ++u) {
            double a_[3];
            for (int d = 0; d < 3; ++d) a_[d] = __shfl(ap[d], u, DB_G);
            if (k0 + u < nf)
                for (int d = 0; d < 3; ++d) tp[d] = tp[d] + a_[d];
        }
    }
    if (g != 0) return;
    lam_out[i] = lam;
    for (int d = 0; d < 3; ++d) tv[d] = tv[d] + 0.0 + 0.0;
    const double gm = gam[i];
    double a[3];
    for (int d = 0; d < 3; ++d) a[d] = (tp[d] + td[d] + tv[d]) / gm + P.grav[d];
    acc_out[i] = make_double4(a[0], a[1], a[2], 0.0);
    const double dt = scal[SC_DT];
    vel_out[i] = make_double4(vi.x + dt * a[0], vi.y + dt * a[1], vi.z + dt * a[2], vi.w);
}

// Gradients_to_MUSCL.f90:47-91 for the particles with wall neighbours:
// drho = sum r (rho_j - rho_i) PartKernel(1) m_j/rho_j and dvel(r, :) the
// same of var(r); drho[3 i + d], dvel[9 i + 3 r + d]
__global__ __launch_bounds__(256) void k_dbsph_muscl(DP P, DWall W, const double4* __restrict__ xm,
                                                     const double4* __restrict__ vel,
                                                     const double2* __restrict__ rp,
                                                     const uint32_t* __restrict__ nbr,
                                                     const int* __restrict__ nnb, const int* __restrict__ nfw,
                                                     const int* __restrict__ fwl, const double* __restrict__ sig,
                                                     const double* __restrict__ wdens,
                                                     double* __restrict__ drho, double* __restrict__ dvel,
                                                     const int* iscal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= iscal[IS_N] || nfw[i] <= 0) return;
    const double4 xi = xm[i];
    const double4 vi = vel[i];
    const double rhoi = rp[i].x;
    double g[3] = {0.0, 0.0, 0.0}, V[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    using RM = NRec<double4, double4, double2>;
    auto fetch = [&](uint32_t e) {
        const int j = (int)(e & NB_MASK);
        return RM{xm[j], vel[j], rp[j]};
    };
    walk_list<4, RM>(nbr, (size_t)P.cap, i, nnb[i], fetch, [&](uint32_t e, const RM& q) {
        const double4 xj = q.x;
        const double R[3] = {xi.x - xj.x, xi.y - xj.y, xi.z - xj.z};
        K64 K;
        kvals64(P, R[0] * R[0] + R[1] * R[1] + R[2] * R[2], true, false, false, K);
        const double vol = xj.w / q.r.x;
        const double dv[3] = {q.v.x - vi.x, q.v.y - vi.y, q.v.z - vi.z};
        for (int d = 0; d < 3; ++d) g[d] = g[d] + R[d] * (q.r.x - rhoi) * K.K1 * vol;
        for (int r = 0; r < 3; ++r)
            for (int d = 0; d < 3; ++d) V[3 * r + d] = V[3 * r + d] + R[d] * dv[r] * K.K1 * vol;
    });
    if (P.muscl_bnd) {  // Gradients_to_MUSCL_boundary.f90:44-86 (one medium)
        const double sg = sig[i];
        const int nf = nfw[i];
        const double v[3] = {vi.x, vi.y, vi.z};
        for (int k = 0; k < nf; ++k) {
            const int id = fwl[(size_t)k * P.cap + i];
            const double4 a = W.w0[id];
            const double4 b = W.w1[id];
            const double4 wv = W.w2[id];
            const double wd = wdens[id];
            const double R[3] = {xi.x - a.x, xi.y - a.y, xi.z - a.z};
            K64 K;  // kernel_fw(1) = W, kernel_fw(2) = K1
            kvals64(P, R[0] * R[0] + R[1] * R[1] + R[2] * R[2], true, false, true, K);
            const double nw[3] = {b.x, b.y, b.z};
            const double vw[3] = {wv.x, wv.y, wv.z};
            const double dr = wd - rhoi;
            const double Ww = a.w * K.W / sg;
            for (int d = 0; d < 3; ++d) g[d] = g[d] + dr * nw[d] * Ww;
            for (int r = 0; r < 3; ++r)
                for (int d = 0; d < 3; ++d) V[3 * r + d] = V[3 * r + d] + (vw[r] - v[r]) * nw[d] * Ww;
            const double vol = wv.w / wd / sg;
            for (int d = 0; d < 3; ++d) g[d] = g[d] + R[d] * dr * K.K1 * vol;
            for (int r = 0; r < 3; ++r)
                for (int d = 0; d < 3; ++d) V[3 * r + d] = V[3 * r + d] + R[d] * (vw[r] - v[r]) * K.K1 * vol;
        }
    }
    for (int d = 0; d < 3; ++d) drho[3 * (size_t)i + d] = g[d];
    for (int d = 0; d < 9; ++d) dvel[9 * (size_t)i + d] = V[d];
}

// DBSPH_kinematics.f90:52-107 and Euler.f90:152-159, one thread per sorted
// wall slot: with kinematics records (t, v, omega), the mesh's velocity and
// angular velocity at the step's time (the first record at or after it,
// linearly interpolated from the one before; zero past the last) give each
// element omega x (x - centre) + v; then every element moves by dt vel.
// The elements keep their cells: OrdGrid1.f90:108 bins them on the first
// step only.
__global__ __launch_bounds__(256) void k_dbsph_kinematics(DP P, const double* __restrict__ kin,
                                                          const int* __restrict__ wids, double4* w0,
                                                          double4* w2, double4* wpos, const double* scal) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.nw) return;
    const int id = wids[s];
    const double t = scal[SC_TIME], dt = scal[SC_DT];
    double4 a = w0[id];
    double4 g = w2[id];
    if (P.nkin > 0) {
        double va[3] = {0.0, 0.0, 0.0}, om[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < P.nkin; ++j) {
            const double* r = &kin[7 * j];
            if (r[0] >= t) {
                if (r[0] == t) {
                    for (int d = 0; d < 3; ++d) {
                        va[d] = r[1 + d];
                        om[d] = r[4 + d];
                    }
                } else {
                    const double* q = &kin[7 * (j - 1)];
                    for (int d = 0; d < 3; ++d) {
                        va[d] = q[1 + d] + (r[1 + d] - q[1 + d]) / (r[0] - q[0]) * (t - q[0]);
                        om[d] = q[4 + d] + (r[4 + d] - q[4 + d]) / (r[0] - q[0]) * (t - q[0]);
                    }
                }
                break;
            }
        }
        const double rel[3] = {a.x - P.kin_c[0], a.y - P.kin_c[1], a.z - P.kin_c[2]};
        g.x = (om[1] * rel[2] - om[2] * rel[1]) + va[0];
        g.y = (om[2] * rel[0] - om[0] * rel[2]) + va[1];
        g.z = (om[0] * rel[1] - om[1] * rel[0]) + va[2];
        w2[id] = g;
    }
    a.x = a.x + dt * g.x;
    a.y = a.y + dt * g.y;
    a.z = a.z + dt * g.z;
    w0[id] = a;
    wpos[s] = make_double4(a.x, a.y, a.z, (double)id);
}

// BC_wall_elements.f90:71-158, one wavefront per wall element: the fluid
// particles within 2h (the ones whose fluid-wall list holds the element),
// each giving the MUSCL left state at the element, rho* = rho_L + (u_L - u_R)
// rho_L / c_L (at least 10) and p* = EoS(rho*) (at least -99 kPa), averaged
// with the Shepard weights m/rho W in ascending pg index (the reference's
// serial scatter order); no neighbours: rho0 and 0; negative pressures set
// to 0 unless allowed.  The lanes share the candidates of the 27 cells and
// the contributions; each contribution goes to its rank by pg index in LDS,
// and one lane sums them in that order.
constexpr int DB_WPB = 4;  // wall elements (wavefronts) per 256-thread block
__global__ __launch_bounds__(256) void k_dbsph_walls(DP P, DWall W, const double4* __restrict__ xm,
                                                     const double4* __restrict__ vel,
                                                     const double2* __restrict__ rp,
                                                     const int* __restrict__ start,
                                                     const double* __restrict__ drho,
                                                     const double* __restrict__ dvel,
                                                     double* __restrict__ wdens, double* __restrict__ wpres,
                                                     int* iscal) {
    __shared__ int s_m[DB_WPB];
    __shared__ int s_pre[DB_WPB][28];  // candidates of the cells before cell t
    __shared__ int s_c0[DB_WPB][27];   // first sorted slot of cell t
    __shared__ int s_id[DB_WPB][SPH_WF];
    __shared__ double s_c[DB_WPB][3][SPH_WF];   // Ww, rho* Ww, p* Ww in arrival order
    __shared__ double s_r[DB_WPB][3][SPH_WF];   // the same by rank
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int w = blockIdx.x * DB_WPB + wv;
    const bool active = w < P.nw;
    if (lane == 0) s_m[wv] = 0;
    if (active && lane < 27) {  // the 27 cells around the element's cell
        int ci, cj, ck;
        d_cell_indices(P, W.wcell[w], ci, cj, ck);
        const int c = d_cell_number(P, ci + (lane / 3) % 3 - 1, cj + lane / 9 - 1, ck + lane % 3 - 1);
        const int s0 = c ? start[c - 1] : 0;
        s_c0[wv][lane] = s0;
        s_pre[wv][lane + 1] = c ? start[c] - s0 : 0;
    }
    __syncthreads();
    if (active && lane == 0) {
        s_pre[wv][0] = 0;
        for (int t = 1; t <= 27; ++t) s_pre[wv][t] += s_pre[wv][t - 1];
    }
    __syncthreads();
    if (active) {
        const double4 a = W.w0[w];
        const double4 b = W.w1[w];
        const double4 wg = W.w2[w];
        const double uR = wg.x * b.x + wg.y * b.y + wg.z * b.z;
        const int total = s_pre[wv][27];
        // the cells' candidates as one range: candidate k lies in the cell t
        // with s_pre[t] <= k < s_pre[t + 1]
        for (int k0 = lane; k0 < total; k0 += 64) {
            int lo = 0, hi = 27;  // s_pre[lo] <= k0 < s_pre[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_pre[wv][mid] <= k0) lo = mid; else hi = mid;
            }
            const int s = s_c0[wv][lo] + (k0 - s_pre[wv][lo]);
            const double4 xj = xm[s];
            const double R[3] = {xj.x - a.x, xj.y - a.y, xj.z - a.z};
            const double rij = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
            if (rij > P.square_doubleh) continue;
            const int k = atomicAdd(&s_m[wv], 1);
            if (k >= SPH_WF) continue;
            const double4 vj = vel[s];
            const double rhoj = rp[s].x;
            const double* g = &drho[3 * (size_t)s];
            const double* V = &dvel[9 * (size_t)s];
            const double rhoL = rhoj - (g[0] * R[0] + g[1] * R[1] + g[2] * R[2]);
            const double v[3] = {vj.x, vj.y, vj.z};
            double uC[3];
            for (int r = 0; r < 3; ++r)
                uC[r] = v[r] - (V[3 * r] * R[0] + V[3 * r + 1] * R[1] + V[3 * r + 2] * R[2]);
            const double uL = uC[0] * b.x + uC[1] * b.y + uC[2] * b.z;
            const double cL = sqrt(P.eps / rhoL);
            double rhostar = rhoL + (uL - uR) * rhoL / cL;
            if (rhostar < 10.0) rhostar = 10.0;
            double pstar = P.eps / P.den0 * (rhostar - P.den0) + P.prif;
            if (pstar < -99.0e3) pstar = -99.0e3;
            K64 K;
            kvals64(P, rij, false, false, true, K);
            const double Ww = xj.w / rhoj * K.W;
            s_id[wv][k] = (int)vj.w;
            s_c[wv][0][k] = Ww;
            s_c[wv][1][k] = rhostar * Ww;
            s_c[wv][2][k] = pstar * Ww;
        }
    }
    __syncthreads();
    const int m = s_m[wv] < SPH_WF ? s_m[wv] : SPH_WF;
    if (active) {
        if (lane == 0 && s_m[wv] > SPH_WF) atomicAdd(&iscal[IS_DB_OVERFLOW], 1);
        for (int k = lane; k < m; k += 64) {  // rank by pg index (ids are distinct)
            const int id = s_id[wv][k];
            int r = 0;
            for (int q = 0; q < m; ++q) r += s_id[wv][q] < id;
            s_r[wv][0][r] = s_c[wv][0][k];
            s_r[wv][1][r] = s_c[wv][1][k];
            s_r[wv][2][r] = s_c[wv][2][k];
        }
    }
    __syncthreads();
    if (active && lane == 0) {
        if (m == 0) {
            wdens[w] = P.den0;
            wpres[w] = 0.0;
        } else {
            double den = 0.0, dsum = 0.0, psum = 0.0;
            for (int r = 0; r < m; ++r) {
                den = den + s_r[wv][0][r];
                dsum = dsum + s_r[wv][1][r];
                psum = psum + s_r[wv][2][r];
            }
            double p = psum / den;
            if (!P.neg_wall_p && p < 0.0) p = 0.0;
            wdens[w] = dsum / den;
            wpres[w] = p;
        }
    }
}

}  // namespace sph
