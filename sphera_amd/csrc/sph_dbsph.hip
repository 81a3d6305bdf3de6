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
//   k_dbsph_muscl     Gradients_to_MUSCL.f90:47-91
//   k_dbsph_walls     BC_wall_elements.f90:71-158
//
// Wall elements do not move.  Their cell lists are built once on the host
// (OrdGrid1.f90:108-137): wstart[c-1] .. wstart[c]-1 index wids, the wall
// elements of cell c in ascending id.
#pragma once

namespace sph {

constexpr int SPH_MW = 128;   // fluid-wall list capacity per particle
constexpr int SPH_WF = 128;   // fluid neighbours a wall element gathers

// wall-element tables: w0 = (x, y, z, weight), w1 = (normal, volume),
// w2 = (velocity, mass)
struct DWall {
    const double4* w0;
    const double4* w1;
    const double4* w2;
    const int* wstart;
    const int* wids;
    const int* wcell;
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
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= iscal[IS_N]) return;
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
    bool ov = false;
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
                while (kl < nn) {
                    const uint32_t e = nbr[nb_index(N, i, kl)];
                    if ((e >> NB_SHIFT) != code) break;
                    const double4 xj = xm[e & NB_MASK];
                    const double R0 = xi.x - xj.x, R1 = xi.y - xj.y, R2 = xi.z - xj.z;
                    K64 K;
                    kvals64(P, R0 * R0 + R1 * R1 + R2 * R2, false, false, true, K);
                    sigma = sigma + xj.w * K.W / rp[e & NB_MASK].x;
                    rsph = rsph + xj.w * K.W;
                    ++kl;
                }
                const int c = d_cell_number(P, ci + dx, cj + dy, ck + dz);
                if (c == 0) continue;
                for (int w = W.wstart[c - 1]; w < W.wstart[c]; ++w) {
                    const int id = W.wids[w];
                    const double4 a = W.w0[id];
                    const double R0 = xi.x - a.x, R1 = xi.y - a.y, R2 = xi.z - a.z;
                    const double rij = R0 * R0 + R1 * R1 + R2 * R2;
                    if (rij > P.square_doubleh) continue;
                    if (nf >= SPH_MW || nf >= P.nmaxpartj) {
                        ov = true;
                        continue;
                    }
                    fwl[(size_t)nf * P.cap + i] = id;
                    nf = nf + 1;
                    const double q = sqrt(rij) / P.h;
                    const double q2 = rij / P.squareh;
                    const int iq = (int)q;
                    if (iq >= 2) continue;
                    double wu = 0.666666667 + q2 * (q * 0.5 - 1.0);
                    if (iq > 0) wu = (2.0 - q) * (2.0 - q) * (2.0 - q) * 0.166666666667;
                    const double k1 = wu * P.coefke;
                    const double4 b = W.w1[id];
                    const double4 g = W.w2[id];
                    dshep = dshep + k1 * a.w * (b.x * (g.x - vi.x) + b.y * (g.y - vi.y) + b.z * (g.z - vi.z));
                    if ((q * P.h) <= (1.3 * P.dx + a.w / 2.0)) wwet[id] = 1;
                    sigma = sigma + b.w * k1;
                    rsph = rsph + g.w * k1;
                }
            }
    if (ov) atomicAdd(&iscal[IS_DB_OVERFLOW], 1);
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
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= iscal[IS_N]) return;
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
    using RM = NRec<double4, double4, double2>;
    auto fetch = [&](uint32_t e) {
        const int j = (int)(e & NB_MASK);
        return RM{xm[j], vel[j], rp[j]};
    };
    walk_list<4, RM>(nbr, N, i, nnb[i], fetch, [&](uint32_t e, const RM& q) {
        if ((int)(e & NB_MASK) == i) return;
        const double4 xj = q.x;
        const double4 vj = q.v;
        const double2 rpj = q.r;
        const double R0 = xi.x - xj.x, R1 = xi.y - xj.y, R2 = xi.z - xj.z;
        const double rij = R0 * R0 + R1 * R1 + R2 * R2;
        K64 K;
        kvals64(P, rij, true, false, false, K);
        const double rhoj = rpj.x, pj = rpj.y, amassj = xj.w;
        const double dv0 = vj.x - vi.x, dv1 = vj.y - vi.y, dv2 = vj.z - vi.z;
        const double alpha = pi / (rhoi * rhoi) + pj / (rhoj * rhoj);
        tpa[0] = tpa[0] + -amassj * alpha * R0 * K.K1;
        tpa[1] = tpa[1] + -amassj * alpha * R1 * K.K1;
        tpa[2] = tpa[2] + -amassj * alpha * R2 * K.K1;
        const double vrij = -dv0 * R0 - dv1 * R1 - dv2 * R2;
        const double TermMon = P.alfamon * celt * h / (rhoi + rhoj);
        td[0] = td[0] + (amassj * TermMon * vrij * R0 * K.K2 + 0.0);
        td[1] = td[1] + (amassj * TermMon * vrij * R1 * K.K2 + 0.0);
        td[2] = td[2] + (amassj * TermMon * vrij * R2 * K.K2 + 0.0);
        if (morris) {
            const double rtil = nui * rhoi + nuj * rhoj + 1.0e-3;
            const double atil = 4.0 * (nui * nuj);
            const double fac = amassj * atil / rtil;
            const double rr = R0 * R0 + R1 * R1 + R2 * R2;
            const double w0 = fac * (-dv0 * K.K2 * rr);
            const double w1 = fac * (-dv1 * K.K2 * rr);
            const double w2 = fac * (-dv2 * K.K2 * rr);
            tv[0] = tv[0] + w0; tv[1] = tv[1] + w1; tv[2] = tv[2] + w2;
            rvs[0] = rvs[0] + w0; rvs[1] = rvs[1] + w1; rvs[2] = rvs[2] + w2;
        }
    });
    for (int d = 0; d < 3; ++d) tp[d] = tp[d] + tpa[d] + 0.0;
    int lam = 0;
    if (P.kin_visc > 0.0) {
        const double apg = sqrt(tp[0] * tp[0] + tp[1] * tp[1] + tp[2] * tp[2]) + P.GI;
        const double amo = sqrt(rvs[0] * rvs[0] + rvs[1] * rvs[1] + rvs[2] * rvs[2]);
        double w = 0.0;
        if (apg != 0.0) w = amo / apg * 100.0;
        lam = (w > 5.0 || apg == 0.0) ? 1 : 0;
    }
    lam_out[i] = lam;
    const int nf = nfw[i];
    for (int k = 0; k < nf; ++k) {  // inter_EqMoto.f90:286-297
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
            double ap = -wd * a.w * K.W * A * nw[d];
            ap = ap - wm * A * R[d] * K.K1;
            tp[d] = tp[d] + ap;
        }
    }
    for (int d = 0; d < 3; ++d) tv[d] = tv[d] + 0.0 + 0.0;
    const double g = gam[i];
    double a[3];
    for (int d = 0; d < 3; ++d) a[d] = (tp[d] + td[d] + tv[d]) / g + P.grav[d];
    acc_out[i] = make_double4(a[0], a[1], a[2], 0.0);
    const double dt = scal[SC_DT];
    vel_out[i] = make_double4(vi.x + dt * a[0], vi.y + dt * a[1], vi.z + dt * a[2], vi.w);
}

// Gradients_to_MUSCL.f90:47-91 for the particles with wall neighbours:
// drho = sum r (rho_j - rho_i) PartKernel(1) m_j/rho_j and dvel(r, :) the
// same of var(r); drho[3 i + d], dvel[9 i + 3 r + d]
__global__ __launch_bounds__(256) void k_dbsph_muscl(DP P, const double4* __restrict__ xm,
                                                     const double4* __restrict__ vel,
                                                     const double2* __restrict__ rp,
                                                     const uint32_t* __restrict__ nbr,
                                                     const int* __restrict__ nnb, const int* __restrict__ nfw,
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
    for (int d = 0; d < 3; ++d) drho[3 * (size_t)i + d] = g[d];
    for (int d = 0; d < 9; ++d) dvel[9 * (size_t)i + d] = V[d];
}

// BC_wall_elements.f90:71-158, one thread per wall element: the fluid
// particles within 2h (the ones whose fluid-wall list holds the element)
// in ascending pg index, each giving the MUSCL left state at the element,
// rho* = rho_L + (u_L - u_R) rho_L / c_L (at least 10) and p* = EoS(rho*)
// (at least -99 kPa), averaged with the Shepard weights m/rho W; no
// neighbours: rho0 and 0; negative pressures set to 0 unless allowed
__global__ __launch_bounds__(256) void k_dbsph_walls(DP P, DWall W, const double4* __restrict__ xm,
                                                     const double4* __restrict__ vel,
                                                     const double2* __restrict__ rp,
                                                     const int* __restrict__ start,
                                                     const double* __restrict__ drho,
                                                     const double* __restrict__ dvel,
                                                     double* __restrict__ wdens, double* __restrict__ wpres,
                                                     int* iscal) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= P.nw) return;
    const double4 a = W.w0[w];
    const double4 b = W.w1[w];
    const double4 wg = W.w2[w];
    int ids[SPH_WF], slots[SPH_WF];
    int m = 0;
    int ci, cj, ck;
    d_cell_indices(P, W.wcell[w], ci, cj, ck);
    for (int jr = cj - 1; jr <= cj + 1; ++jr)
        for (int ir = ci - 1; ir <= ci + 1; ++ir)
            for (int kr = ck - 1; kr <= ck + 1; ++kr) {
                const int c = d_cell_number(P, ir, jr, kr);
                if (c == 0) continue;
                for (int s = start[c - 1]; s < start[c]; ++s) {
                    const double4 xj = xm[s];
                    const double R0 = xj.x - a.x, R1 = xj.y - a.y, R2 = xj.z - a.z;
                    if (R0 * R0 + R1 * R1 + R2 * R2 > P.square_doubleh) continue;
                    if (m >= SPH_WF) {
                        atomicAdd(&iscal[IS_DB_OVERFLOW], 1);
                        continue;
                    }
                    // insertion in ascending pg index
                    const int id = (int)vel[s].w;
                    int k = m;
                    while (k > 0 && ids[k - 1] > id) {
                        ids[k] = ids[k - 1];
                        slots[k] = slots[k - 1];
                        --k;
                    }
                    ids[k] = id;
                    slots[k] = s;
                    ++m;
                }
            }
    if (m == 0) {
        wdens[w] = P.den0;
        wpres[w] = 0.0;
        return;
    }
    double den = 0.0, dsum = 0.0, psum = 0.0;
    const double uR = wg.x * b.x + wg.y * b.y + wg.z * b.z;
    for (int k = 0; k < m; ++k) {
        const int s = slots[k];
        const double4 xj = xm[s];
        const double4 vj = vel[s];
        const double rhoj = rp[s].x;
        const double R[3] = {xj.x - a.x, xj.y - a.y, xj.z - a.z};
        const double* g = &drho[3 * (size_t)s];
        const double* V = &dvel[9 * (size_t)s];
        const double rhoL = rhoj - (g[0] * R[0] + g[1] * R[1] + g[2] * R[2]);
        const double v[3] = {vj.x, vj.y, vj.z};
        double uC[3];
        for (int r = 0; r < 3; ++r) uC[r] = v[r] - (V[3 * r] * R[0] + V[3 * r + 1] * R[1] + V[3 * r + 2] * R[2]);
        const double uL = uC[0] * b.x + uC[1] * b.y + uC[2] * b.z;
        const double cL = sqrt(P.eps / rhoL);
        double rhostar = rhoL + (uL - uR) * rhoL / cL;
        if (rhostar < 10.0) rhostar = 10.0;
        double pstar = P.eps / P.den0 * (rhostar - P.den0) + P.prif;
        if (pstar < -99.0e3) pstar = -99.0e3;
        K64 K;
        kvals64(P, R[0] * R[0] + R[1] * R[1] + R[2] * R[2], false, false, true, K);
        const double Ww = xj.w / rhoj * K.W;
        den = den + Ww;
        dsum = dsum + rhostar * Ww;
        psum = psum + pstar * Ww;
    }
    double p = psum / den;
    if (!P.neg_wall_p && p < 0.0) p = 0.0;
    wdens[w] = dsum / den;
    wpres[w] = p;
}

}  // namespace sph
