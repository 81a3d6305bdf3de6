// MI355X (gfx950) kernels of the SPHERA v10 particle-interaction hot path.
//
// Layout in HBM (one context = one solve, capacity N particles), all arrays in
// cell-sorted order (cells ascending, particle id ascending inside a cell):
//   xm   double4 (x, y, z, mass)          x2 buffers (trajectory writes the other)
//   vel  double4 (vx, vy, vz, id)         x3 buffers (momentum -> smoothing -> sort)
//   rp   double2 (rho, p)                 x3 buffers (sort -> continuity -> smoothing)
//   lam  int32 laminar flag               x2 buffers
//   acc  double4, dden, vpres             outputs of the passes
//   FP32-mixed mirrors: xmf float4 (cell-relative x,y,z, mass), velf float4,
//   rpf float2, written by the same kernels that write the FP64 fields.
//   Neighbour list: entry k of particle i at nb_index(N, i, k) (groups of 8
//   entries per particle, a group rank contiguous over i), the sorted index of
//   the k-th neighbour in bits 0..26 and its stencil cell code
//   (dx+1)+3(dy+1)+9(dz+1) in bits 27..31; count in nnb[i].
//
// FP64 kernels are compiled with -ffp-contract=off and evaluate every
// expression in the order of the cited reference line, so sums over a list in
// the reference's order reproduce the reference's values.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "sph_device.h"

namespace sph {

// FP32 particle records of the passes that read velocities are interleaved:
// slot i of a record buffer holds the cell-relative position (x, y, z, mass)
// at float4 2i and (vx, vy, vz, rho) at 2i+1, so a neighbour's position and
// velocity come from one 32-byte run.  F4View is the strided view of one of
// the two halves.  The neighbour search and the pressure smoothing read the
// positions alone from the compact array xmf.
template <class T>
struct F4View {
    T* p;
    __host__ __device__ __forceinline__ T& operator[](size_t i) const { return p[2 * i]; }
};
using F4In = F4View<const float4>;
using F4Out = F4View<float4>;

// ----------------------------------------------------------------------------
// grid helpers: CellNumber.f90:53-59, ParticleCellNumber.f90:52-62,
// CellIndices.f90:51-63 (1-based cell ids, 0 = outside)
// ----------------------------------------------------------------------------
__device__ __forceinline__ int d_cell_number(const DP& P, int i, int j, int k) {
    if (i < 1 || i > P.ncd[0] || j < 1 || j > P.ncd[1] || k < 1 || k > P.ncd[2]) return 0;
    return ((k - 1) * P.ncd[1] + (j - 1)) * P.ncd[0] + i;
}
__device__ __forceinline__ int d_particle_cell(const DP& P, double x, double y, double z) {
    const double xp = x - P.extr_lo[0];
    const double yp = y - P.extr_lo[1];
    const double zp = z - P.extr_lo[2];
    const int i = (int)ceil(xp / P.dcd[0]);
    const int k = (int)ceil(zp / P.dcd[2]);
    const int j = P.dim == 2 ? 1 : (int)ceil(yp / P.dcd[1]);  // :59-61 (2D: j = 1)
    return d_cell_number(P, i, j, k);
}
__device__ __forceinline__ void d_cell_indices(const DP& P, int nc, int& i, int& j, int& k) {
    const int nij = P.ncd[0] * P.ncd[1];
    k = (nc - 1) / nij + 1;
    const int ncij = nc - nij * (k - 1);
    j = (ncij - 1) / P.ncd[0] + 1;
    i = ncij - P.ncd[0] * (j - 1);
}

// 8 XCDs each with their own L2: consecutive workgroups are dispatched round
// robin over the XCDs, so remap them to give every XCD one contiguous range
// of cell-sorted particles (their neighbour data then share that XCD's L2).
__device__ __forceinline__ int xcd_block(int nblocks) {
    const int b = blockIdx.x;
    const int per = (nblocks + 7) >> 3;
    const int xcd = b & 7;
    const int r = b >> 3;
    const int full = nblocks - (per - 1) * 8;  // XCDs that get `per` blocks
    int base;
    if (xcd < full) base = xcd * per;
    else base = full * per + (xcd - full) * (per - 1);
    return base + r;
}

// ----------------------------------------------------------------------------
// kernel function values: CalcVarLength.f90:243-265
// ----------------------------------------------------------------------------
struct K64 { double K1, K2, K3, W; };

__device__ __forceinline__ void kvals64(const DP& P, double rij, bool need12, bool need3,
                                        bool needW, K64& K) {
    const double rij2 = rij;
    const double r = sqrt(rij);
    const double q = r / P.h;
    const double q2 = rij2 / P.squareh;
    const int iq = (int)q;
    K.K1 = K.K2 = K.K3 = K.W = 0.0;
    if (iq >= 2) return;
    const double denom = 1.0 / (r + P.eta);
    if (need12) {
        double gm = -2.0 * q + 1.5 * q2;
        if (iq > 0) gm = -gm + q2 - 2.0;
        gm = gm * P.ke_coef;
        K.K1 = gm * denom;
        K.K2 = K.K1 / (rij2 + P.eta2);
    }
    if (need3) {
        double gw = -12.0 - 3.0 * q2 + 12.0 * q;
        gw = gw * P.kacl_coef;
        K.K3 = gw * denom;
    }
    if (needW) {
        double wu = 0.666666667 + q2 * (q * 0.5 - 1.0);
        if (iq > 0) wu = (2.0 - q) * (2.0 - q) * (2.0 - q) * 0.166666666667;
        K.W = wu * P.coefke;
    }
}

struct K32 { float K1, K2, K3, W; };

// FP32 form: same piecewise cubic spline, single-precision hardware sqrt/rcp
__device__ __forceinline__ void kvals32(const DP& P, float rij, bool need12, bool need3,
                                        bool needW, K32& K) {
    const float r = __builtin_amdgcn_sqrtf(rij);
    const float q = r * P.f_unosuh;
    const float q2 = rij * P.f_unosusquareh;
    // every listed pair has r <= 2h, where both branches of the spline and its
    // derivatives vanish at q = 2, so no cut-off test is needed
    K.K1 = K.K2 = K.K3 = K.W = 0.0f;
    const bool outer = q >= 1.0f;
    const float denom = __builtin_amdgcn_rcpf(r + P.f_eta);
    if (need12) {
        float gm = -2.0f * q + 1.5f * q2;
        if (outer) gm = -gm + q2 - 2.0f;
        K.K1 = gm * P.f_ke_coef * denom;
        K.K2 = K.K1 * __builtin_amdgcn_rcpf(rij + P.f_eta2);
    }
    if (need3) K.K3 = (-12.0f - 3.0f * q2 + 12.0f * q) * P.f_kacl_coef * denom;
    if (needW) {
        const float t = 2.0f - q;
        const float wu = outer ? t * t * t * 0.166666666667f : 0.666666667f + q2 * (q * 0.5f - 1.0f);
        K.W = wu * P.f_coefke;
    }
}

// The 27 stencil-cell offsets (dx,dy,dz) * dcd in FP32, indexed by the
// stencil code of a list entry; each workgroup keeps a copy in LDS.
__shared__ float4 g_soff[27];

__device__ __forceinline__ void stencil_offsets(const DP& P) {
    if (threadIdx.x < 27) {
        const int c = threadIdx.x;
        g_soff[c] = make_float4((float)(c % 3 - 1) * P.f_dcd[0], (float)((c / 3) % 3 - 1) * P.f_dcd[1],
                                (float)(c / 9 - 1) * P.f_dcd[2], 0.0f);
    }
    __syncthreads();
}

// cell-relative FP32 separation x_i - x_j from the stencil code of j
__device__ __forceinline__ void rag32(float4 xi, float4 xj, uint32_t code, float& r0, float& r1,
                                      float& r2) {
    const float4 o = g_soff[code >> NB_SHIFT];
    r0 = (xi.x - xj.x) - o.x;
    r1 = (xi.y - xj.y) - o.y;
    r2 = (xi.z - xj.z) - o.z;
}

// ----------------------------------------------------------------------------
// SA-SPH geometry: LocalNormalCoordinates.f90:55-87 + IsPointInternal.f90
// ----------------------------------------------------------------------------
// IsPointInternal.f90 on the face's precomputed constants (kept in registers
// by the callers that test many points of one face)
struct PIC {
    double pu, a0, b0, c0, a1, b1, c1;
    int fkod;
};
__device__ __forceinline__ PIC pic_of(const DFace& F) {
    return PIC{F.pu, F.pA[0], F.pB[0], F.pC[0], F.pA[1], F.pB[1], F.pC[1], F.nodes - 2};
}
__device__ __forceinline__ bool d_point_internal(const PIC& c, double px0, double px1) {
    const double csi0 = (c.a0 + c.b0 * px0 + c.c0 * px1) * c.pu;
    const double csi1 = (c.a1 + c.b1 * px0 + c.c1 * px1) * c.pu;
    const double csi2 = 1.0 - (csi0 + csi1);
    const bool in01 = !(csi0 < 0.0 || csi0 > 1.0) && !(csi1 < 0.0 || csi1 > 1.0);
    if (c.fkod == 1) return in01 && !(csi2 < 0.0 || csi2 > 1.0);
    if (c.fkod == 2) return in01;
    return false;
}
__device__ __forceinline__ bool d_point_internal(const DFace& F, double px0, double px1) {
    return d_point_internal(pic_of(F), px0, px1);
}

// ----------------------------------------------------------------------------
// sort: OrdGrid1.f90:58-105 as a device counting sort.  The particles arrive
// almost in cell order (they moved less than a cell since the last sort), so
// a workgroup's 256 particles fall in a narrow window of cells: it counts
// them in an LDS histogram and issues one global atomic per occupied cell.
// Inside the window a particle's rank among the workgroup's particles of its
// cell follows the thread order, so particles that stay in their cell keep
// their previous (ascending-id) order and k_cell_sort finds its runs sorted.
// ----------------------------------------------------------------------------
constexpr int HIST_W = 512;

// Segmented contexts (SG): the workgroup's segment comes from the slot-block
// map; slots, keys, ids and the removal list are the segment's.
template <bool SG>
__global__ __launch_bounds__(256) void k_cell_count(DP P0, Seg S, const double4* __restrict__ xm,
                                                    const double4* __restrict__ vel,
                                                    int* __restrict__ key, int* __restrict__ slot,
                                                    int* __restrict__ cnt, int* __restrict__ idslot,
                                                    int* __restrict__ rmids, int* iscal,
                                                    int* __restrict__ gone) {
    __shared__ int base[HIST_W];
    __shared__ int wcnt[4][HIST_W];  // particles of each window cell per wave
    __shared__ int cmin;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.pblk[(blockIdx.x * blockDim.x) >> 8]) : 0;
    const DP& P = SG ? S.P[sg] : P0;
    iscal += sg * IS_COUNT;
    rmids += sg * RM_CAP;
    const int n = iscal[IS_N];
    for (int b = threadIdx.x; b < HIST_W; b += blockDim.x)
        for (int w = 0; w < 4; ++w) wcnt[w][b] = 0;
    if (threadIdx.x == 0) cmin = 0x7fffffff;
    __syncthreads();
    int k = -1;
    if (i - P.base < n) {
        const double4 x = xm[i];
        const int c = d_particle_cell(P, x.x, x.y, x.z);
        const int id = (int)vel[i].w;
        idslot[P.base + id] = i;
        // the last trajectory step left through an open face: cella = -1 (read once)
        const bool out = gone[i] != 0;
        if (out) gone[i] = 0;
        if (c <= 0 || out) {  // out of the grid or cella <= 0: removed (OrdGrid1.f90:66-82)
            k = P.koff + P.nmax;
            atomicAdd(&iscal[IS_REMOVED], 1);
            const int q = atomicAdd(&iscal[IS_NRM], 1);
            if (q < RM_CAP) rmids[q] = id;
        } else {
            k = P.koff + c - 1;
            atomicMin(&cmin, k);
        }
        key[i] = k;
    }
    __syncthreads();
    const int c0 = cmin;
    int loc = -1, rank = 0;
    if (k >= 0) {
        const int d = k - c0;
        if (k < P.koff + P.nmax && d < HIST_W) loc = d;
        else slot[i] = atomicAdd(&cnt[k], 1);
    }
    // rank within the wave among the lanes of the same window cell, one
    // ballot per distinct cell of the wave
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long todo = __ballot(loc >= 0);
    while (todo) {
        const int lead = __ffsll((long long)todo) - 1;
        const int dk = __shfl(loc, lead, 64);
        const unsigned long long m = __ballot(loc == dk);
        if (loc == dk) rank = __popcll(m & below);
        if (lane == lead) wcnt[wid][dk] = __popcll(m);
        todo &= ~m;
    }
    __syncthreads();
    if (loc >= 0)
        for (int w = 0; w < wid; ++w) rank += wcnt[w][loc];
    for (int b = threadIdx.x; b < HIST_W; b += blockDim.x) {
        const int tot = wcnt[0][b] + wcnt[1][b] + wcnt[2][b] + wcnt[3][b];
        if (tot > 0) base[b] = atomicAdd(&cnt[c0 + b], tot);
    }
    __syncthreads();
    if (loc >= 0) slot[i] = base[loc] + rank;
}

// wave64 inclusive scan
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// block-level exclusive scan of 4*256 elements; writes per-block totals
__global__ __launch_bounds__(256) void k_scan_local(const int* __restrict__ in, int* __restrict__ out,
                                                    int* __restrict__ bsum, int m) {
    __shared__ int wsum[4];
    const int base = blockIdx.x * 1024 + threadIdx.x * 4;
    int v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = (base + q < m) ? in[base + q] : 0;
    const int tsum = v[0] + v[1] + v[2] + v[3];
    const int incl = wave_incl_scan(tsum);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wid; ++w) woff += wsum[w];
    int run = woff + incl - tsum;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (base + q < m) out[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 255) bsum[blockIdx.x] = woff + incl;
}

// Ids after the removals of this sort, as OrdGrid1.f90:62-83 leaves pg: the
// first loop walks npi upwards; a particle out of the grid is overwritten by
// pg(nag), nag decreases and the same npi is tested again.  So, for the removed
// ids r in ascending order while r < nag: the last particle moves to r unless
// it is itself removed (then it is dropped and the next last is tried).  One
// lane; removals are rare (particles leaving a closed tank).
// Segmented contexts: one lane per segment.
template <bool SG>
__global__ void k_renumber(DP P0, Seg S, int* __restrict__ rmids, const int* __restrict__ idslot,
                           double4* __restrict__ vel, int* iscal) {
    const int sg = SG ? (int)(blockIdx.x * blockDim.x + threadIdx.x) : 0;
    if (SG ? sg >= S.nseg : threadIdx.x != 0) return;
    const int pbase = SG ? S.P[sg].base : P0.base;
    iscal += sg * IS_COUNT;
    rmids += sg * RM_CAP;
    const int m = iscal[IS_NRM];
    if (m == 0) return;
    if (m > RM_CAP) {
        iscal[IS_RM_OVERFLOW] = 1;
        return;
    }
    for (int a = 1; a < m; ++a) {  // insertion sort, ascending
        const int v = rmids[a];
        int b = a - 1;
        while (b >= 0 && rmids[b] > v) {
            rmids[b + 1] = rmids[b];
            --b;
        }
        rmids[b + 1] = v;
    }
    int nag = iscal[IS_N];
    for (int q = 0; q < m; ++q) {
        const int r = rmids[q];
        if (r >= nag) break;
        for (;;) {
            const int last = --nag;
            if (last == r) break;
            int lo = 0, hi = m;  // is `last` removed too?
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (rmids[mid] < last) lo = mid + 1;
                else hi = mid;
            }
            if (lo < m && rmids[lo] == last) continue;
            vel[idslot[pbase + last]].w = (double)r;
            break;
        }
    }
}

// exclusive scan of the block totals by one workgroup
__global__ __launch_bounds__(1024) void k_scan_sums(int* __restrict__ bsum, int nb, int* total) {
    __shared__ int wsum[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int idx = base + threadIdx.x;
        const int v = idx < nb ? bsum[idx] : 0;
        const int incl = wave_incl_scan(v);
        const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int woff = 0;
        for (int w = 0; w < wid; ++w) woff += wsum[w];
        const int c0 = carry;
        if (idx < nb) bsum[idx] = c0 + woff + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c0 + woff + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void k_scan_add(int* __restrict__ out, const int* __restrict__ bsum, int m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] += bsum[i >> 10];
}

// Segmented contexts: the scan over the whole key space gives every key its
// start counted from slot 0; a segment's slots begin at its base instead.
// adj[s] = base_s - start[koff_s], added to every key of segment s.
__global__ void k_seg_adj(Seg S, const int* __restrict__ start, int* __restrict__ adj) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S.nseg) adj[s] = S.P[s].base - start[S.P[s].koff];
}
__global__ void k_seg_add(Seg S, int* __restrict__ start, const int* __restrict__ adj, int m) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) start[k] += adj[S.kblk[k >> 8]];
}

template <bool SG>
__global__ void k_scatter(DP P0, Seg S, const int* __restrict__ key, const int* __restrict__ slot,
                          const int* __restrict__ start, const double4* __restrict__ vel,
                          int* __restrict__ perm, int* __restrict__ pid, const int* iscal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.pblk[(blockIdx.x * blockDim.x) >> 8]) : 0;
    const int pbase = SG ? S.P[sg].base : P0.base;
    if (i - pbase >= iscal[sg * IS_COUNT + IS_N]) return;
    const int d = start[key[i]] + slot[i];
    perm[d] = i;
    pid[d] = (int)vel[i].w;
}

// ascending particle id inside every cell: the order OrdGrid1's 4th loop
// produces (NPartOrd), hence the reference's neighbour-list order.  A
// workgroup sorts the segments of 256 consecutive cells in LDS.
constexpr int CSORT_CAP = 6144;

template <bool SG>
__global__ __launch_bounds__(256) void k_cell_sort(DP P0, Seg S, const int* __restrict__ start,
                                                   int* __restrict__ perm, int* __restrict__ pid) {
    __shared__ int sid[CSORT_CAP];
    __shared__ int sperm[CSORT_CAP];
    const int c0 = blockIdx.x * 256;
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.kblk[blockIdx.x]) : 0;
    const DP& P = SG ? S.P[sg] : P0;
    const int c1 = min(c0 + 256, P.koff + P.nmax);
    if (c0 >= c1) return;  // a block of the out-of-grid key or of the padding
    const int g0 = start[c0], g1 = start[c1];
    const int len = g1 - g0;
    const int c = c0 + threadIdx.x;
    if (len <= CSORT_CAP) {
        for (int t = threadIdx.x; t < len; t += blockDim.x) {
            sid[t] = pid[g0 + t];
            sperm[t] = perm[g0 + t];
        }
        __syncthreads();
        if (c < c1) {
            const int b = start[c] - g0, e = start[c + 1] - g0;
            for (int a = b + 1; a < e; ++a) {
                const int kv = sid[a], pv = sperm[a];
                int q = a - 1;
                while (q >= b && sid[q] > kv) {
                    sid[q + 1] = sid[q];
                    sperm[q + 1] = sperm[q];
                    --q;
                }
                sid[q + 1] = kv;
                sperm[q + 1] = pv;
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < len; t += blockDim.x) perm[g0 + t] = sperm[t];
    } else if (c < c1) {  // dense window: sort in place in global memory
        const int b = start[c], e = start[c + 1];
        for (int a = b + 1; a < e; ++a) {
            const int kv = pid[a], pv = perm[a];
            int q = a - 1;
            while (q >= b && pid[q] > kv) {
                pid[q + 1] = pid[q];
                perm[q + 1] = perm[q];
                --q;
            }
            pid[q + 1] = kv;
            perm[q + 1] = pv;
        }
    }
}

// permutes the state into the new order and writes the FP32 cell-relative
// positions (used by the neighbour test in both precisions)
template <bool SG>
__global__ void k_gather(DP P0, Seg S, const int* __restrict__ perm, const int* __restrict__ key,
                         const double4* __restrict__ xm_s, const double4* __restrict__ vel_s,
                         const double2* __restrict__ rp_s, const int* __restrict__ lam_s,
                         const double4* __restrict__ acc_s, double4* __restrict__ xm_d,
                         double4* __restrict__ vel_d, double2* __restrict__ rp_d,
                         int* __restrict__ lam_d, double4* __restrict__ acc_d,
                         int* __restrict__ cell_d, float4* __restrict__ xmf, F4Out xmf0, F4Out xmf1,
                         F4Out velf,
                         float2* __restrict__ rpf, const int* __restrict__ start, int* iscal,
                         const double* __restrict__ dd_s, double* __restrict__ dd_d,
                         const double* __restrict__ vp_s, double* __restrict__ vp_d,
                         const double* __restrict__ g_s, double* __restrict__ g_d,
                         const double* __restrict__ h_s, double* __restrict__ h_d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.pblk[(blockIdx.x * blockDim.x) >> 8]) : 0;
    const DP& P = SG ? S.P[sg] : P0;
    iscal += sg * IS_COUNT;
    const int nin = start[P.koff + P.nmax] - P.base;  // particles still inside the grid
    if (p == P.base) {
        iscal[IS_NPREV] = iscal[IS_N];
        iscal[IS_N] = nin;
    }
    if (p - P.base >= nin) return;
    const int s = perm[p];
    const double4 x = xm_s[s];
    const double4 v = vel_s[s];
    const double2 r = rp_s[s];
    xm_d[p] = x;
    vel_d[p] = v;
    rp_d[p] = r;
    lam_d[p] = lam_s[s];
    acc_d[p] = acc_s[s];
    if (dd_s) {  // explicit RK1: dden and vpres of the step end in the new order
        dd_d[p] = dd_s[s];
        vp_d[p] = vp_s[s];
    }
    if (g_s) {  // DB-SPH: Gamma and dShep travel with the particle
        g_d[p] = g_s[s];
        h_d[p] = h_s[s];
    }
    const int c = key[s] - P.koff + 1;
    cell_d[p] = c;
    int ci, cj, ck;
    d_cell_indices(P, c, ci, cj, ck);
    const double ox = P.extr_lo[0] + P.dcd[0] * (double)(ci - 1);
    const double oy = P.extr_lo[1] + P.dcd[1] * (double)(cj - 1);
    const double oz = P.extr_lo[2] + P.dcd[2] * (double)(ck - 1);
    const float4 xr = make_float4((float)(x.x - ox), (float)(x.y - oy), (float)(x.z - oz), (float)x.w);
    xmf[p] = xr;
    if (P.f32) {  // the position half of both FP32 record buffers
        xmf0[p] = xr;
        xmf1[p] = xr;
        velf[p] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)(r.x - P.den0));
        rpf[p] = make_float2((float)r.x, (float)r.y);
    }
}

// ----------------------------------------------------------------------------
// neighbour lists: CalcVarLength.f90:133-242 — stencil y -> x -> z, ascending
// particle id in a cell, r^2 <= (2h)^2 in FP64, self included, truncation past
// NMAXPARTJ.  Candidates are tested on FP32 cell-relative positions;
// a candidate whose FP32 r^2 lies within a relative 4e-5 of (2h)^2 is decided
// by the reference's FP64 expression on the FP64 positions, so the lists are
// exactly the reference's.  (Each relative coordinate is rounded once from
// |x - corner| <= dcd, so the FP32 separation is within ~5e-7 dcd per axis and
// the FP32 r^2 within ~1.5e-6 relative of (2h)^2 near the cut-off.)
// ----------------------------------------------------------------------------
// List writer of one lane.  SPH_NB_BUF = 4 (default) or 8: the entries are
// collected in a shift register of that many registers and stored as one
// 16-byte (two for 8) vector store per full run, so the lane's list reaches
// L2 in whole 16/32-byte pieces instead of one 4-byte store per entry; flush()
// stores the last, partial run (its unused tail slots are past the count and
// never read).  A run is stored only while it lies inside the particle's
// list capacity (nb_stride(NMAXPARTJ)); the count keeps running, so the first
// NMAXPARTJ entries are the stored ones, as in the reference.  SPH_NB_BUF = 0:
// one 4-byte store per entry.
#ifndef SPH_NB_BUF
#define SPH_NB_BUF 4
#endif
#if SPH_NB_BUF == 0
struct NbOut {
    uint32_t* wp;
    int cnt;
    size_t jump;  // N*NB_G - (NB_G-1): from the last entry of a group to the first of the next
    __device__ __forceinline__ NbOut(uint32_t* nbr, size_t N, int p, int)
        : wp(nbr + nb_index(N, p, 0)), cnt(0), jump(N * NB_G - (NB_G - 1)) {}
    __device__ __forceinline__ void push(uint32_t e, int mj) {
        if (cnt < mj) *wp = e;
        ++cnt;
        wp += ((cnt & (NB_G - 1)) == 0) ? jump : (size_t)1;
    }
    __device__ __forceinline__ void flush() {}
};
#else
static_assert(SPH_NB_BUF == 4 || SPH_NB_BUF == 8, "SPH_NB_BUF is 0, 4 or 8");
static_assert(NB_G == 8, "the buffered list writer stores runs inside 8-entry groups");
struct NbOut {
    static constexpr int B = SPH_NB_BUF;
    uint32_t* wp;  // first slot of the current run
    int cnt;
    int lim;       // list capacity: runs ending past it are not stored
    size_t jump;   // from the last run of a group to the first slot of the particle's next group
    uint32_t b[B]; // b[B-1] is the newest entry
    __device__ __forceinline__ NbOut(uint32_t* nbr, size_t N, int p, int stride)
        : wp(nbr + nb_index(N, p, 0)), cnt(0), lim(stride), jump(N * NB_G - (NB_G - B)) {
#pragma unroll
        for (int u = 0; u < B; ++u) b[u] = 0u;
    }
    __device__ __forceinline__ void store_run() {
        uint4* q = reinterpret_cast<uint4*>(wp);
        q[0] = make_uint4(b[0], b[1], b[2], b[3]);
        if constexpr (B == 8) q[1] = make_uint4(b[4], b[5], b[6], b[7]);
    }
    __device__ __forceinline__ void push(uint32_t e, int) {
#pragma unroll
        for (int u = 0; u + 1 < B; ++u) b[u] = b[u + 1];
        b[B - 1] = e;
        ++cnt;
        if ((cnt & (B - 1)) == 0) {
            if (cnt <= lim) store_run();
            wp += ((cnt & (NB_G - 1)) == 0) ? jump : (size_t)B;
        }
    }
    // the r = cnt % B entries since the last run, moved to the run's first slots
    __device__ __forceinline__ void flush() {
        const int r = cnt & (B - 1);
        if (r == 0 || cnt > lim) return;
        uint32_t o[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            uint32_t v = 0u;
#pragma unroll
            for (int s = 1; s < B; ++s)  // slot u holds b[B - r + u] when u < r
                if (r == s && u < s) v = b[B - s + u];
            o[u] = v;
        }
#pragma unroll
        for (int u = 0; u < B; ++u) b[u] = o[u];
        store_run();
    }
};
#endif

// The neighbour test of one candidate and the list append: FP32 on
// cell-relative positions, and the reference's FP64 expression on the FP64
// positions inside the guard band.  xo: FP32 position of particle p relative
// to the candidate's cell corner (its own cell-relative position minus the
// stencil offset), xj: the candidate's cell-relative position, m: its sorted
// index.  The FP32 r^2 (fused multiply-adds) is within ~3e-7 relative of the
// exact one, far inside the 4e-5 band.  Every accepted candidate is counted;
// the first NMAXPARTJ are stored (the reference keeps the first NMAXPARTJ of
// its list, CalcVarLength.f90:200-207), so the loops have no early exit.
// Q: xj holds the candidate's 16-bit fixed-point coordinates (0..65535 over
// the cell), decoded with the per-axis step sc in one fused multiply-add.
template <bool Q = false>
__device__ __forceinline__ void nb_test(const DP& P, const double4* __restrict__ xm, double4 xi64,
                                        float4 xo, float4 xj, int m, uint32_t ctag, float lo, float hi,
                                        NbOut& w, int mj, float4 sc = float4{}) {
    float r0, r1, r2;
    if constexpr (Q) {
        r0 = __builtin_fmaf(xj.x, -sc.x, xo.x);
        r1 = __builtin_fmaf(xj.y, -sc.y, xo.y);
        r2 = __builtin_fmaf(xj.z, -sc.z, xo.z);
    } else {
        r0 = xo.x - xj.x;
        r1 = xo.y - xj.y;
        r2 = xo.z - xj.z;
    }
    const float r2f = __builtin_fmaf(r0, r0, __builtin_fmaf(r1, r1, r2 * r2));
    if (r2f <= hi) {
        bool acc = r2f < lo;
        if (!acc) {  // guard band: the reference's FP64 test
            const double4 xj64 = xm[m];
            const double d0 = xi64.x - xj64.x;
            const double d1 = xi64.y - xj64.y;
            const double d2 = xi64.z - xj64.z;
            const double rij = d0 * d0 + d1 * d1 + d2 * d2;
            acc = !(rij > P.square_doubleh);
        }
        if (acc) w.push((uint32_t)m | ctag, mj);
    }
}

// Candidates [b,e) of one stencil cell, four loads ahead: load(x) gives the
// FP32 position of slot x, gl + x its sorted index.
template <bool Q = false, class Load>
__device__ __forceinline__ void nb_cell(const DP& P, const double4* __restrict__ xm, double4 xi64,
                                        float4 xi, float4 o, uint32_t ctag, int b, int e, int gl,
                                        Load&& load, float lo, float hi, NbOut& w, int mj,
                                        float4 sc = float4{}) {
    const float4 xo = make_float4(xi.x - o.x, xi.y - o.y, xi.z - o.z, 0.0f);
    int x = b;
    for (; x + 4 <= e; x += 4) {
        const float4 q0 = load(x), q1 = load(x + 1), q2 = load(x + 2), q3 = load(x + 3);
        nb_test<Q>(P, xm, xi64, xo, q0, gl + x, ctag, lo, hi, w, mj, sc);
        nb_test<Q>(P, xm, xi64, xo, q1, gl + x + 1, ctag, lo, hi, w, mj, sc);
        nb_test<Q>(P, xm, xi64, xo, q2, gl + x + 2, ctag, lo, hi, w, mj, sc);
        nb_test<Q>(P, xm, xi64, xo, q3, gl + x + 3, ctag, lo, hi, w, mj, sc);
    }
    for (; x < e; ++x) nb_test<Q>(P, xm, xi64, xo, load(x), gl + x, ctag, lo, hi, w, mj, sc);
}

// List of particle p from global memory (candidates gathered through L1/L2).
__device__ int nb_particle_global(const DP& P, int p, const double4* __restrict__ xm,
                                  const float4* __restrict__ xmf, const int* __restrict__ cell,
                                  const int* __restrict__ start, uint32_t* __restrict__ nbr,
                                  float lo, float hi, bool& ov) {
    const int ni = P.ncd[0], nj = P.ncd[1], nk = P.ncd[2];
    const int mj = P.nmaxpartj;
    const size_t N = P.cap;
    int ic, jc, kc;
    d_cell_indices(P, cell[p], ic, jc, kc);
    const float4 xi = xmf[p];
    const double4 xi64 = xm[p];
    NbOut w(nbr, N, p, nb_stride(mj));
    auto load = [&](int x) { return xmf[x]; };
    for (int dy = -1; dy <= 1; ++dy) {
        const int jj = jc + dy;
        if (jj < 1 || jj > nj) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int ii = ic + dx;
            if (ii < 1 || ii > ni) continue;
            for (int dz = -1; dz <= 1; ++dz) {
                const int kk = kc + dz;
                if (kk < 1 || kk > nk) continue;
                const int cn = ((kk - 1) * nj + (jj - 1)) * ni + ii;
                const int code = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                nb_cell(P, xm, xi64, xi, g_soff[code], (uint32_t)code << NB_SHIFT, start[P.koff + cn - 1],
                        start[P.koff + cn], 0, load, lo, hi, w, mj);
            }
        }
    }
    w.flush();
    ov = w.cnt > mj;
    return min(w.cnt, mj);
}

// Tiled form.  The particles of one row of cells (fixed j,k; cells i are
// contiguous in the sorted order, hence in memory) are cut into blocks of at
// most NB_WG.  For a block over cells ia..ib the FP32 positions of the 9
// neighbouring row segments (rows j+-1, k+-1 over cells ia-1..ib+1) are staged
// in LDS once, and every lane walks its 27 stencil cells in the reference's
// order out of LDS.  A block over more than NB_MAXC cells or with more than
// NB_TILE candidates takes the global path above.  Persistent workgroups take
// the blocks of one contiguous range per XCD.
// Launch shapes of the tiled search (workgroup size, candidate tile, FP16
// tile, register bound): the default 256 x 4096 FP16 (26 KB of LDS, 6 waves
// per SIMD, registers held to 80 so that 6 fit) and three the autotuner may
// pick instead: 256 x 3072 FP32 (36 KB,
// 4 waves), 128 x 1536 FP32, and 256 x 2816 FP16 held to 64 registers for 8
// waves per SIMD.
constexpr int NB_WG = 256;
constexpr int NB_MAXC = 48;
constexpr int NB_TILE = 3072;
// tk: tile coordinates in FP32 (NB_T32), FP16 (NB_T16) or 16-bit fixed point
// over the cell (NB_TQ16); minb: workgroups per CU the registers must allow
enum { NB_T32 = 0, NB_T16 = 1, NB_TQ16 = 2 };
struct NbShape { int wg, tile; int tk; int minb; };
constexpr int NB_NSHAPES = 5;
constexpr NbShape NB_SHAPES[NB_NSHAPES] = {{256, 4096, NB_T16, 6},  {256, 3072, NB_T32, 1},
                                           {128, 1536, NB_T32, 1},  {256, 2816, NB_T16, 8},
                                           {256, 4096, NB_TQ16, 6}};

// rowblk[r] = first block of row r (exclusive scan of ceil(count_r / bw), bw
// the search's workgroup size),
// rowblk[R] = number of blocks.  One workgroup.  Segmented contexts: the rows
// of all segments, each row read with its segment's grid.
template <bool SG>
__global__ __launch_bounds__(1024) void k_nb_rowblocks(DP P0, Seg S, const int* __restrict__ start,
                                                       int* __restrict__ rowblk, int bw) {
    __shared__ int wsum[16];
    __shared__ int carry;
    const int R = SG ? S.nrows : P0.ncd[1] * P0.ncd[2];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < R; base += 1024) {
        const int r = base + threadIdx.x;
        int v = 0;
        if (r < R) {
            const DP& P = SG ? S.P[S.rseg[r]] : P0;
            const int ni = P.ncd[0];
            const int rl = r - P.roff;
            const int k0 = P.koff + rl * ni;
            v = (start[k0 + ni] - start[k0] + bw - 1) / bw;
        }
        const int incl = wave_incl_scan(v);
        const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int woff = 0;
        for (int w = 0; w < wid; ++w) woff += wsum[w];
        const int c0 = carry;
        if (r < R) rowblk[r] = c0 + woff + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c0 + woff + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) rowblk[R] = carry;
}

// TK = NB_T16: the tile holds the candidates' cell-relative positions in FP16
// (6 bytes each: more workgroups per CU).  A coordinate in [0, dcd] rounds by
// at most 2^-11 dcd, so the FP32 r^2 from it is within ~2e-3 relative of the
// exact one near (2h)^2; the FP64 guard band is widened to 5e-3 to keep the
// lists exactly the reference's.  TK = NB_TQ16: 16-bit fixed point over the
// cell instead (q = rint(x 65535 / dcd), also 6 bytes): a coordinate rounds
// by at most 2^-17 dcd, the FP32 r^2 is within ~2.8e-5 (dcd / 2h) relative of
// the exact one near (2h)^2 (dcd / 2h < 1.5 on the reference's grid), and the
// guard band is 2e-4, so the FP64 test runs 25 times less often.
template <bool SG, int WG = NB_WG, int TILE = NB_TILE, int TK = NB_T32, int MINB = 1>
__global__ __launch_bounds__(WG, MINB) void k_neighbours(DP P0, Seg S, const double4* __restrict__ xm,
                                                      const float4* __restrict__ xmf,
                                                      const int* __restrict__ cell,
                                                      const int* __restrict__ start,
                                                      const int* __restrict__ rowblk,
                                                      uint32_t* __restrict__ nbr, int* __restrict__ nnb,
                                                      int* iscal, unsigned long long* pairs) {
    using TXY = typename std::conditional<TK == NB_T16, __half2,
                                          typename std::conditional<TK == NB_TQ16, uint32_t, float2>::type>::type;
    using TZ = typename std::conditional<TK == NB_T16, __half,
                                         typename std::conditional<TK == NB_TQ16, uint16_t, float>::type>::type;
    constexpr bool Q = TK == NB_TQ16;
    __shared__ TXY tile_xy[TILE];  // staged candidates: (x, y) and z, 12 (FP16: 6) bytes each
    __shared__ TZ tile_z[TILE];
    __shared__ int bnd[9][NB_MAXC + 4];  // LDS offset of the start of cell ia-1+u of row t
    __shared__ int seg_g[9];             // sorted index of a segment's first particle
    __shared__ int seg_l[9];             // its LDS offset
    __shared__ int seg_n[9];
    __shared__ int s_row, s_ia, s_ib, s_p0, s_p1, s_global, s_seg;
    if (!SG) stencil_offsets(P0);
    int cur_seg = -1;
    const int R = SG ? S.nrows : P0.ncd[1] * P0.ncd[2];
    const int nblocks = rowblk[R];
    const int per_xcd = gridDim.x >> 3;  // gridDim.x is a multiple of 8
    const int xcd = blockIdx.x & 7;
    const int b_lo = (int)((long long)nblocks * xcd / 8), b_hi = (int)((long long)nblocks * (xcd + 1) / 8);
    unsigned long long mypairs = 0;
    for (int blk = b_lo + (blockIdx.x >> 3); blk < b_hi; blk += per_xcd) {
        if (threadIdx.x == 0) {
            // row of the block: last r with rowblk[r] <= blk
            int a = 0, b = R;  // rowblk[a] <= blk < rowblk[b]
            while (b - a > 1) {
                const int m = (a + b) >> 1;
                if (rowblk[m] <= blk) a = m;
                else b = m;
            }
            const int r = a;
            const int sg0 = SG ? S.rseg[r] : 0;
            const DP& Q = SG ? S.P[sg0] : P0;
            const int rl = r - Q.roff;  // the row in its segment's grid
            const int qi = Q.ncd[0];
            const int p0 = start[Q.koff + rl * qi] + (blk - rowblk[r]) * WG;
            const int p1 = min(p0 + WG, start[Q.koff + (rl + 1) * qi]);
            const int ia = cell[p0] - rl * qi, ib = cell[p1 - 1] - rl * qi;
            s_seg = sg0;
            s_row = rl;
            s_ia = ia;
            s_ib = ib;
            s_p0 = p0;
            s_p1 = p1;
            s_global = ib - ia + 1 > NB_MAXC;
        }
        __syncthreads();
        const int sg = SG ? __builtin_amdgcn_readfirstlane(s_seg) : 0;
        const DP& P = SG ? S.P[sg] : P0;
        if (SG && sg != cur_seg) {  // the stencil offsets of this segment's grid
            stencil_offsets(P);
            cur_seg = sg;
        }
        const int ni = P.ncd[0], nj = P.ncd[1], nk = P.ncd[2];
        const float sq = (float)P.square_doubleh;
        const float lo = sq * (1.0f - 4.0e-5f), hi = sq * (1.0f + 4.0e-5f);
        const float band = TK == NB_T16 ? 5.0e-3f : (Q ? 2.0e-4f : 4.0e-5f);  // guard band of the tiled test
        // fixed-point step per axis (NB_TQ16) and its inverse for the staging
        const float4 qs = make_float4(P.f_dcd[0] * (1.0f / 65535.0f), P.f_dcd[1] * (1.0f / 65535.0f),
                                      P.f_dcd[2] * (1.0f / 65535.0f), 0.0f);
        const float4 qi = make_float4(65535.0f / P.f_dcd[0], 65535.0f / P.f_dcd[1], 65535.0f / P.f_dcd[2], 0.0f);
        const float tlo = sq * (1.0f - band), thi = sq * (1.0f + band);
        const int mj = P.nmaxpartj;
        const size_t N = P.cap;
        const int r = s_row, ia = s_ia, ib = s_ib, p0 = s_p0, p1 = s_p1;
        const int jc = r % nj + 1, kc = r / nj + 1;
        const int ca = max(ia - 1, 1), cb = min(ib + 1, ni);
        if (!s_global) {
            if (threadIdx.x < 9) {
                const int t = threadIdx.x;
                const int jj = jc + t / 3 - 1, kk = kc + t % 3 - 1;
                int g0 = 0, g1 = 0;
                if (jj >= 1 && jj <= nj && kk >= 1 && kk <= nk) {
                    const int rb = P.koff + ((kk - 1) * nj + (jj - 1)) * ni;
                    g0 = start[rb + ca - 1];
                    g1 = start[rb + cb];
                }
                seg_g[t] = g0;
                seg_n[t] = g1 - g0;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int off = 0;
                for (int t = 0; t < 9; ++t) {
                    seg_l[t] = off;
                    off += seg_n[t];
                }
                s_global = off > TILE;
            }
            __syncthreads();
        }
        const int p = p0 + threadIdx.x;
        if (!s_global) {
            // cell boundaries inside the tile: u = 0 .. ib-ia+3 is the start of cell
            // ia-1+u (cells ia-1 .. ib+1, and the end of ib+1)
            const int nu = ib - ia + 4;
            for (int x = threadIdx.x; x < 9 * nu; x += WG) {
                const int t = x / nu, u = x % nu;
                const int jj = jc + t / 3 - 1, kk = kc + t % 3 - 1;
                const int ii = ia - 1 + u;
                int v;
                if (jj < 1 || jj > nj || kk < 1 || kk > nk || ii < ca) v = seg_l[t];
                else if (ii > cb) v = seg_l[t] + seg_n[t];
                else v = seg_l[t] + start[P.koff + ((kk - 1) * nj + (jj - 1)) * ni + ii - 1] - seg_g[t];
                bnd[t][u] = v;
            }
            // stage the candidates' FP32 positions, four loads in flight per thread
            {
                const int total = seg_l[8] + seg_n[8];
                for (int e0 = threadIdx.x; e0 < total; e0 += 4 * WG) {
                    float4 v[4];
                    int dst[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int e = e0 + u * WG;
                        dst[u] = -1;
                        if (e < total) {
                            int t = 0;
#pragma unroll
                            for (int q = 1; q < 9; ++q) t += e >= seg_l[q];
                            v[u] = xmf[seg_g[t] + (e - seg_l[t])];
                            dst[u] = e;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (dst[u] >= 0) {
                            if constexpr (TK == NB_T16) {
                                tile_xy[dst[u]] = __floats2half2_rn(v[u].x, v[u].y);
                                tile_z[dst[u]] = __float2half_rn(v[u].z);
                            } else if constexpr (Q) {
                                auto q16 = [](float a) {  // 0 <= a <= 65535 (+ rounding)
                                    return (uint32_t)__float2uint_rn(fminf(fmaxf(a, 0.0f), 65535.0f));
                                };
                                tile_xy[dst[u]] = q16(v[u].x * qi.x) | (q16(v[u].y * qi.y) << 16);
                                tile_z[dst[u]] = (uint16_t)q16(v[u].z * qi.z);
                            } else {
                                tile_xy[dst[u]] = make_float2(v[u].x, v[u].y);
                                tile_z[dst[u]] = v[u].z;
                            }
                        }
                }
            }
            __syncthreads();
            if (p < p1) {
                const int ucell = cell[p] - r * ni - ia + 1;  // u index of the own cell
                const float4 xi = xmf[p];
                const double4 xi64 = xm[p];
                NbOut w(nbr, N, p, nb_stride(mj));
                auto load = [&](int x) {
                    if constexpr (TK == NB_T16) {
                        const float2 a = __half22float2(tile_xy[x]);
                        return make_float4(a.x, a.y, __half2float(tile_z[x]), 0.0f);
                    } else if constexpr (Q) {
                        const uint32_t a = tile_xy[x];
                        return make_float4((float)(a & 0xFFFFu), (float)(a >> 16), (float)tile_z[x], 0.0f);
                    } else {
                        const float2 a = tile_xy[x];
                        return make_float4(a.x, a.y, tile_z[x], 0.0f);
                    }
                };
                for (int dy = 0; dy < 3; ++dy)
                    for (int dx = 0; dx < 3; ++dx)
                        for (int dz = 0; dz < 3; ++dz) {
                            const int t = 3 * dy + dz;
                            const int code = dx + 3 * dy + 9 * dz;
                            nb_cell<Q>(P, xm, xi64, xi, g_soff[code], (uint32_t)code << NB_SHIFT,
                                       bnd[t][ucell + dx - 1], bnd[t][ucell + dx], seg_g[t] - seg_l[t], load,
                                       tlo, thi, w, mj, qs);
                        }
                w.flush();
                const bool ov = w.cnt > mj;
                const int cnt = min(w.cnt, mj);
                nnb[p] = cnt;
                mypairs += (unsigned long long)(cnt - 1);
                if (ov) atomicAdd(&iscal[sg * IS_COUNT + IS_NB_OVERFLOW], 1);
            }
        } else if (p < p1) {
            bool ov;
            const int cnt = nb_particle_global(P, p, xm, xmf, cell, start, nbr, lo, hi, ov);
            nnb[p] = cnt;
            mypairs += (unsigned long long)(cnt - 1);
            if (ov) atomicAdd(&iscal[sg * IS_COUNT + IS_NB_OVERFLOW], 1);
        }
        __syncthreads();  // the tile and the block descriptors are reused
    }
    // pair-interaction counter: sum_i (nPartIntorno(i) - 1), self excluded
    unsigned long long v = mypairs;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(pairs, v);
}

// ----------------------------------------------------------------------------
// SA-SPH volume integrals (FP64)
// ----------------------------------------------------------------------------
__device__ void d_volume_integrals(const DP& P, const DFace& F, const double* __restrict__ kt,
                                   const double* __restrict__ bit, const double locx[3],
                                   double bi[8], double ig[9]) {
    double IntWdV = 0.0, IntdWrm1dV = 0.0, IntGWZrm1dV = 0.0;
    double G0 = 0.0, G1 = 0.0, G2 = 0.0;
    double M00 = 0.0, M01 = 0.0, M02 = 0.0, M11 = 0.0, M12 = 0.0, M22 = 0.0;
    double J0 = 0.0, J1 = 0.0, J2 = 0.0, J3 = 0.0;
    for (int q = 0; q < 8; ++q) bi[q] = 0.0;
    for (int q = 0; q < 9; ++q) ig[q] = 0.0;
    const double zpmin = 0.05 * P.h;
    double L0 = locx[0], L1 = locx[1], L2 = locx[2];
    if (L2 < 0.0) return;
    if (L2 < zpmin) L2 = zpmin;
    double robpre = 2.1;
    const PIC pic = pic_of(F);
    for (int ipk = 0; ipk < 256; ++ipk) {
        const double* b = &bit[ipk * 7];
        const double Pj0 = L0 + b[0], Pj1 = L1 + b[1], Pj2 = L2 + b[2];
        if (!(Pj2 <= 0.0)) continue;
        const double ts = L2 / (L2 - Pj2);
        const double X0 = L0 + (Pj0 - L0) * ts;
        const double X1 = L1 + (Pj1 - L1) * ts;
        if (!d_point_internal(pic, X0, X1)) continue;
        double dist = 0.0;
        {
            const double d0 = X0 - L0, d1 = X1 - L1, d2 = 0.0 - L2;
            dist = dist + d0 * d0;
            dist = dist + d1 * d1;
            dist = dist + d2 * d2;
        }
        dist = sqrt(dist);
        const double da = b[3];
        const double rob = dist / P.h;
        if (fabs(rob - robpre) > 1.0e-3) {
            const double deltaX = P.ktdelta;
            const int nr0 = (int)((rob - kt[0]) / deltaX);
            double iv0, iv1, iv2, iv3;
            if (nr0 <= 0) {
                iv0 = kt[1]; iv1 = kt[2]; iv2 = kt[3]; iv3 = kt[4];
            } else if (nr0 >= 40) {
                iv0 = kt[40 * 5 + 1]; iv1 = kt[40 * 5 + 2]; iv2 = kt[40 * 5 + 3]; iv3 = kt[40 * 5 + 4];
            } else {
                const double* r0 = &kt[nr0 * 5];
                const double* r1 = &kt[(nr0 + 1) * 5];
                const double csi = (rob - r0[0]) / deltaX;
                iv0 = r0[1] + csi * (r1[1] - r0[1]);
                iv1 = r0[2] + csi * (r1[2] - r0[2]);
                iv2 = r0[3] + csi * (r1[3] - r0[3]);
                iv3 = r0[4] + csi * (r1[4] - r0[4]);
            }
            J0 = iv0 * da;
            J1 = iv1 * da;
            J2 = iv2 * da;
            J3 = iv3 * da;
            robpre = rob;
        }
        const double R0 = b[4], R1 = b[5], R2 = b[6];
        const double RZ2 = R2 * R2;
        IntWdV = IntWdV + J0;
        IntdWrm1dV = IntdWrm1dV + J1;
        IntGWZrm1dV = IntGWZrm1dV + RZ2 * J1;
        G0 = G0 + R0 * J2;
        M00 = M00 + R0 * R0 * J3;
        M01 = M01 + R0 * R1 * J3;
        M02 = M02 + R0 * R2 * J3;
        G1 = G1 + R1 * J2;
        M11 = M11 + R1 * R1 * J3;
        M12 = M12 + R1 * R2 * J3;
        G2 = G2 + R2 * J2;
        M22 = M22 + R2 * R2 * J3;
    }
    if (L2 == zpmin) IntWdV = d_point_internal(F, L0, L1) ? 0.5 : 0.0;
    bi[1] = IntWdV;
    bi[2] = IntdWrm1dV;
    bi[3] = G0;
    bi[4] = G1;
    bi[5] = G2;
    bi[6] = IntGWZrm1dV;
    // IntGiWrRdV column-major
    ig[0] = M00; ig[1] = M01; ig[2] = M02;
    ig[3] = M01; ig[4] = M11; ig[5] = M12;
    ig[6] = M02; ig[7] = M12; ig[8] = M22;
}

// ----------------------------------------------------------------------------
// boundary table: ComputeBoundaryDataTab_3D.f90:62-106 in two kernels.
// k_boundary_find: FindCloseBoundaryFaces3D.f90:78-188 per particle, appends
// every (particle, close face) pair to a compact work list; k_boundary_int:
// ComputeBoundaryVolumeIntegrals_P0.f90:76-155 for one pair per thread, so the
// 256-direction quadrature keeps every lane busy.
// ----------------------------------------------------------------------------
template <bool SG>
__global__ __launch_bounds__(256) void k_boundary_find(DP P0, Seg S, const DFace* __restrict__ faces,
                                                       const int* __restrict__ ncount,
                                                       const int* __restrict__ nstart,
                                                       const int* __restrict__ nvec,
                                                       const double4* __restrict__ xm,
                                                       int* __restrict__ bn, BTab* __restrict__ bt,
                                                       int* __restrict__ work, int* nwork, int* iscal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.pblk[(blockIdx.x * blockDim.x) >> 8]) : 0;
    const DP& P = SG ? S.P[sg] : P0;
    iscal += sg * IS_COUNT;
    if (i - P.base >= iscal[IS_N]) return;
    const double4 x = xm[i];
    const double xc[3] = {x.x, x.y, x.z};
    const int nc = d_particle_cell(P, x.x, x.y, x.z);
    // the faces listed in the 27 cells around the particle's cell, each once
    // and in the order the reference's cell walk meets them (a face accepted
    // once is skipped when met again; one rejected is rejected again)
    const int nf = nc > 0 ? ncount[P.koff + nc - 1] : 0;
    if (nf == 0) {  // no frontier face listed within reach
        bn[i] = 0;
        return;
    }
    int ncbf = 0;
    int clob[SPH_MB];
    double lz[SPH_MB];
    const int ini = nstart[P.koff + nc - 1];
    for (int fl = ini; fl < ini + nf; ++fl) {
        const int f = nvec[fl];
        const DFace& F = faces[f];
        const double pinmin = (F.btype == 3) ? -P.doubleh : 0.0;
        const double* ref = &F.GX[3 * (F.nodes - 1)];
        double PX[3];
        for (int sdi = 0; sdi < 3; ++sdi) {
            double s = 0.0;
            for (int sdj = 0; sdj < 3; ++sdj) s = s + F.T[3 * sdi + sdj] * (xc[sdj] - ref[sdj]);
            PX[sdi] = s;
        }
        const double pin = PX[2];
        if (pin >= pinmin && pin < P.doubleh) {
            const bool need_in = (F.btype == 3 || F.btype == 4 || F.btype == 5);
            if (!need_in || d_point_internal(F, PX[0], PX[1])) {
                if (ncbf >= SPH_MB) {
                    atomicAdd(&iscal[IS_BT_OVERFLOW], 1);
                    continue;
                }
                clob[ncbf] = f;
                lz[ncbf] = pin;
                ++ncbf;
            }
        }
    }
    bn[i] = ncbf;
    if (ncbf == 0) return;
    const int w0 = atomicAdd(nwork, ncbf);
    for (int q = 0; q < ncbf; ++q) {
        const DFace& F = faces[clob[q]];
        const double* ref = &F.GX[3 * (F.nodes - 1)];
        BTab& rec = bt[(size_t)q * P.cap + i];
        for (int sdi = 0; sdi < 2; ++sdi) {
            double s = 0.0;
            for (int sdj = 0; sdj < 3; ++sdj) s = s + F.T[3 * sdi + sdj] * (xc[sdj] - ref[sdj]);
            rec.loc[sdi] = s;
        }
        rec.loc[2] = lz[q];
        rec.face = clob[q];
        work[w0 + q] = i | (q << NB_SHIFT);
    }
}

// Segmented contexts: a workgroup's pairs may belong to several segments, so
// each lane reads its segment's constants and tables (kernel and direction
// tables of every segment, consecutive) from global memory.
template <bool SG>
__global__ __launch_bounds__(256) void k_boundary_int(DP P0, Seg S, const DFace* __restrict__ faces,
                                                      const double* __restrict__ kt_g,
                                                      const double* __restrict__ bit_g,
                                                      BTab* __restrict__ bt,
                                                      const int* __restrict__ work,
                                                      const int* nwork) {
    __shared__ double kt_s[SG ? 1 : 41 * 5];
    __shared__ double bit_s[SG ? 1 : 256 * 7];
    if (!SG) {
        for (int t = threadIdx.x; t < 41 * 5; t += blockDim.x) kt_s[t] = kt_g[t];
        for (int t = threadIdx.x; t < 256 * 7; t += blockDim.x) bit_s[t] = bit_g[t];
        __syncthreads();
    }
    const int nw = *nwork;
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += gridDim.x * blockDim.x) {
        const int e = work[w];
        const int i = e & (int)NB_MASK, q = (int)((uint32_t)e >> NB_SHIFT);
        const int sg = SG ? S.pblk[i >> 8] : 0;
        const DP& P = SG ? S.P[sg] : P0;
        const double* kt = SG ? kt_g + (size_t)sg * (41 * 5) : kt_s;
        const double* bit = SG ? bit_g + (size_t)sg * (256 * 7) : bit_s;
        BTab& rec = bt[(size_t)q * P.cap + i];
        const double loc[3] = {rec.loc[0], rec.loc[1], rec.loc[2]};
        double bi[8], ig[9];
        d_volume_integrals(P, faces[rec.face], kt, bit, loc, bi, ig);
        for (int t = 0; t < 8; ++t) rec.bi[t] = bi[t];
        for (int t = 0; t < 9; ++t) rec.ig[t] = ig[t];
    }
}

// ----------------------------------------------------------------------------
// 2D form (x-z plane): boundary table over the boundary sides.
// FindCloseBoundarySides2D.f90:68-170, SelectCloseBoundarySides2D.f90:66-120,
// ComputeBoundaryDataTab_2D.f90:71-162 with FindBoundaryIntersection2D.f90:
// 59-86, ComputeSurfaceIntegral_WdS2D.f90:57-85, BoundaryVolumeIntegrals2D
// .f90:60-97, ComputeVolumeIntegral_WdV2D.f90:69-132 and
// InterpolateBoundaryIntegrals2D.f90:61-70.  One lane per particle over the
// sides (a tank has a handful); every expression in the reference's order.
// The integrals use the device's sin, cos, tan and atan2 (the host reference
// calls the C library's).
// ----------------------------------------------------------------------------
constexpr double PIGRECO_D = 3.14159265358979;
constexpr double KERNELCONST2D_D = 0.454728408833987;
constexpr double DELTAX_BIT_D = 0.1;

// w.f90:48-57
__device__ __forceinline__ double d_w2(double r, double h, double coef) {
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
__device__ __forceinline__ double d_wintegr(double ri, double hh) {
    const double ro = ri / hh;
    if (0.0 <= ro && ro < 1.0) {
        const double ro2 = ro * ro, ro3 = ro2 * ro;
        const double w1 = (0.5 + -0.375 * ro2 + 0.15 * ro3) * ro2;
        return KERNELCONST2D_D * (0.35 - w1);
    } else if (1.0 <= ro && ro < 2.0) {
        const double q1 = 2.0 - ro, q2 = q1 * q1, q4 = q2 * q2;
        return -KERNELCONST2D_D * ((-0.125 + 0.05 * q1) * q4);
    }
    return 0.0;
}
// J2Wro2.f90:55-70
__device__ __forceinline__ double d_j2wro2(double ro) {
    const double a1 = 0.00833333333333333, a2 = 0.26666666666666667;
    const double ro2 = ro * ro, ro3 = ro2 * ro;
    if (0.0 <= ro && ro < 1.0) return KERNELCONST2D_D * (0.25 - (a1 * (40.0 - 36.0 * ro2 + 15.0 * ro3) * ro3));
    if (1.0 <= ro && ro < 2.0)
        return KERNELCONST2D_D * (a2 - (a1 * (80.0 - 90.0 * ro + 36.0 * ro2 - 5.0 * ro3) * ro3));
    return 0.0;
}
// BoundIntegralTab2D row r (0-based), column c; 22 x 6 row-major
__device__ __forceinline__ void d_interp2d(const double* __restrict__ tab, int ncols, const int* cols,
                                           double xx, double* func) {
    if (xx < tab[0]) xx = tab[0];
    // rows 1..21 hold ii (0 <= delta < 2 for a selected side); the clamp only
    // keeps a non-finite position inside the table
    const int ii = min(max((int)((xx - tab[0]) / DELTAX_BIT_D) + 1, 1), 21);
    const double xi = tab[(ii - 1) * 6];
    for (int nc = 0; nc < ncols; ++nc) {
        const int jj = cols[nc];
        const double fi = tab[(ii - 1) * 6 + jj], fip1 = tab[ii * 6 + jj];
        func[nc] = fi + (fip1 - fi) * (xx - xi) / DELTAX_BIT_D;
    }
}

__global__ __launch_bounds__(256) void k_boundary_2d(DP P, const DFace* __restrict__ sides,
                                                     const double* __restrict__ tab,
                                                     const double4* __restrict__ xm,
                                                     const int* __restrict__ cell,
                                                     int* __restrict__ bn, BTab* __restrict__ bt,
                                                     int* iscal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.cap) return;
    if (i >= iscal[IS_N]) {
        bn[i] = 0;
        return;
    }
    int nint = 0;
    const double4 X = xm[i];
    const double px = X.x, pz = X.z;
    const double dh = P.doubleh;
    int ncbs = 0;
    int clo[4];
    double loc[8];
    bool bad = false;
    for (int isi = 0; isi < P.nf && !bad; ++isi) {
        const DFace& S = sides[isi];
        if (S.btype == 6) continue;  // peri
        const double P1P0 = px - S.v1[0], P1P1 = pz - S.v1[1];
        const double s0 = S.T[0], s1 = S.T[2], n0 = S.T[6], n1 = S.T[8];
        const double sidel = S.len;
        const double xp = P1P0 * s0 + P1P1 * s1;
        const double yp = P1P0 * n0 + P1P1 * n1;
        const double xpmin = -dh, xpmax = sidel + dh, ypmin = -dh, ypmax = dh;
        if ((xpmin < xp && xp < xpmax) && (ypmin < yp && yp < ypmax)) {
            double ypmx = 0.0, ypmn = 0.0;
            if (xp < 0.0) {
                ypmx = sqrt(P.square_doubleh - xp * xp);
                ypmn = -ypmx;
            } else if (xp <= sidel) {
                ypmx = ypmax;
                ypmn = ypmin;
            } else if (xp < xpmax) {
                const double Lmxpq = (sidel - xp) * (sidel - xp);
                ypmx = sqrt(P.square_doubleh - Lmxpq);
                ypmn = -ypmx;
            }
            if (ypmn < yp && yp < ypmx) {
                if (ncbs >= 4) {
                    bad = true;
                    break;
                }
                clo[ncbs] = isi;
                loc[2 * ncbs] = xp;
                loc[2 * ncbs + 1] = yp;
                ncbs = ncbs + 1;
            }
        }
    }
    for (int q = 0; q < ncbs && !bad; ++q) {
        const int b = sides[clo[q]].btype;
        if (b != 0 && b != 1 && b != 2) {
            atomicAdd(&iscal[IS_UNSUPPORTED], 1);
            bad = true;
        }
    }
    if (bad) {
        if (ncbs >= 4) atomicAdd(&iscal[IS_BT_OVERFLOW], 1);
        bn[i] = 0;
        return;
    }
    if (ncbs == 2) {
        const int a = clo[0], b = clo[1];
        if (sides[a].prev == b + 1)
            loc[2] = sides[b].len - loc[2];
        else if (sides[b].prev == a + 1)
            loc[0] = sides[a].len - loc[0];
    }
    // SelectCloseBoundarySides2D.f90:66-120: fixe/tapi sides with y > 0; open
    // sides with y > 0 and 0 < x < side length
    int iside[4];
    double iloc[8];
    for (int q = 0; q < ncbs; ++q) {
        bool keep = loc[2 * q + 1] > 0.0;
        if (sides[clo[q]].btype == 2) keep = keep && loc[2 * q] > 0.0 && loc[2 * q] < sides[clo[q]].len;
        if (keep) {
            iside[nint] = clo[q];
            iloc[2 * nint] = loc[2 * q];
            iloc[2 * nint + 1] = loc[2 * q + 1];
            nint = nint + 1;
        }
    }
    if (nint > 2) {
        atomicAdd(&iscal[IS_BT_OVERFLOW], 1);
        bn[i] = 0;
        return;
    }
    const double h = P.h;
    const double yplimite = 0.05 * h;
    const double dalfarif = PIGRECO_D / 10;
    for (int icbs = 0; icbs < nint; ++icbs) {
        BTab& B = bt[(size_t)icbs * P.cap + i];
        const double xpi = iloc[2 * icbs];
        double ypi = iloc[2 * icbs + 1];
        B.face = iside[icbs];
        B.loc[0] = xpi;
        B.loc[1] = ypi;
        B.loc[2] = 0.0;
        const double deltai = ypi / h;
        double F[5];
        if (nint == 2) {
            // FindBoundaryIntersection2D (ypi > 0 here)
            const double halfchord = sqrt(P.square_doubleh - ypi * ypi);
            const double XA = xpi - halfchord, XB = xpi + halfchord;
            double xpmin = 0.0;
            if (XA > xpmin) xpmin = XA;
            double xpmax = sides[iside[icbs]].len;
            if (XB < xpmax) xpmax = XB;
            double interlen = xpmax - xpmin;
            if (interlen <= 0.01 * h) interlen = 0.0;
            // ComputeSurfaceIntegral_WdS2D
            double IntWdS = 0.0;
            if (interlen > 0.0) {
                const double dsrif = h / 2.0;
                const double ypiq = ypi * ypi;
                int ndiv = (int)(interlen / dsrif + 0.5);
                if (ndiv < 2) ndiv = 2;
                const double deltas = interlen / ndiv;
                double xpip = xpmin - 0.5 * deltas;
                for (int ipt = 1; ipt <= ndiv; ++ipt) {
                    xpip = xpip + deltas;
                    const double dxpip = xpip - xpi;
                    const double ris = sqrt(dxpip * dxpip + ypiq);
                    IntWdS = IntWdS + d_w2(ris, h, P.coefke) * deltas;
                }
            }
            // BoundaryVolumeIntegrals2D
            double IntWdV = 0.0, Dp0 = 0.0, Dp1 = 0.0;
            if (interlen > 0.0) {
                double yb = ypi;
                if (yb < yplimite) yb = yplimite;
                const double alfaA = atan2(xpmin - xpi, yb), alfaB = atan2(xpmax - xpi, yb);
                const double intalfa = alfaB - alfaA;
                int ndiv = (int)(intalfa / dalfarif + 0.5);
                if (ndiv < 1) ndiv = 1;
                const double dalfa = intalfa / ndiv;
                double alfa_k = alfaA - 0.5 * dalfa;
                for (int k = 1; k <= ndiv; ++k) {
                    alfa_k = alfa_k + dalfa;
                    const double sa = sin(alfa_k), ca = cos(alfa_k);
                    const double rb = yb / ca;
                    const double mult = h * d_j2wro2(rb / h) * dalfa;
                    Dp0 = Dp0 + sa * mult;
                    Dp1 = Dp1 - ca * mult;
                    IntWdV = IntWdV + d_wintegr(rb, h) * dalfa;
                }
                if (yb == yplimite) IntWdV = 0.5;
            }
            // ComputeVolumeIntegral_WdV2D (Ncbslocal = 2)
            double IntWdV1 = 0.0;
            if (interlen > 0.0) {
                if (ypi >= yplimite) {
                    const double alfaA = atan2(xpmin - xpi, ypi), alfaB = atan2(xpmax - xpi, ypi);
                    const double intalfa = alfaB - alfaA;
                    int ndiv = (int)(intalfa / dalfarif + 0.5);
                    if (ndiv < 2) ndiv = 2;
                    const double dalfa = intalfa / ndiv;
                    double alfa = alfaA - 0.5 * dalfa;
                    for (int ipt = 1; ipt <= ndiv; ++ipt) {
                        alfa = alfa + dalfa;
                        const double ta = tan(alfa);
                        const double ris = ypi * sqrt(1.0 + ta * ta);
                        IntWdV1 = IntWdV1 + d_wintegr(ris, h) * dalfa;
                    }
                } else {
                    const double ypj = iloc[2 * (1 - icbs) + 1];
                    if (ypj <= yplimite) {
                        double angle = 0.0;
                        if (sides[iside[0]].prev == iside[1] + 1)
                            angle = sides[iside[0]].angle;
                        else if (sides[iside[1]].prev == iside[0] + 1)
                            angle = sides[iside[1]].angle;
                        else
                            atomicAdd(&iscal[IS_UNSUPPORTED], 1);  // sides not consecutive
                        IntWdV1 = 0.5 * (angle + PIGRECO_D) / (2.0 * PIGRECO_D);
                    } else {
                        IntWdV1 = (xpmin >= xpi && xpi <= xpmax) ? 0.5 : 0.0;
                    }
                }
            }
            const int cols[3] = {3, 4, 5};
            d_interp2d(tab, 3, cols, deltai, F);
            B.bi[0] = IntWdS;
            B.bi[1] = IntWdV;
            B.bi[2] = IntWdV1;
            B.bi[3] = Dp0;
            B.bi[4] = Dp1;
            B.bi[5] = F[0] / P.squareh;
            B.bi[6] = F[1] / P.squareh;
            B.bi[7] = F[2] / P.squareh;
        } else {
            const int cols[5] = {1, 2, 3, 4, 5};
            d_interp2d(tab, 5, cols, deltai, F);
            B.bi[0] = F[0] / h;
            B.bi[1] = F[1];
            B.bi[2] = F[1];
            B.bi[3] = 0.0;
            B.bi[4] = 0.0;
            B.bi[5] = F[2] / P.squareh;
            B.bi[6] = F[3] / P.squareh;
            B.bi[7] = F[4] / P.squareh;
        }
    }
    bn[i] = nint;
}

// ----------------------------------------------------------------------------
// FP32-mixed form of the same quadrature, eight lanes per (particle, face)
// pair, lane r taking ring r of BoundIntegralTab.  The rays of a ring meet
// the face plane on one circle of radius ts |b_xy| around (L0, L1), all at
// the same distance rob h = L2 |b| / (-b_z).  IsPointInternal's area
// coordinates are affine over the plane, so over that circle each stays
// within its centre value +- radius |grad|: a circle that clears every bound
// (or lies beyond one) by BT_MARGIN has every direction inside (outside) the
// face, as the FP64 per-direction test finds it; other rings run that test
// per direction (FP64, with -1/b_z precomputed).  The reference's reuse of
// the kernel-table values (refresh when rob moves by more than 1e-3) changes
// state only at the first kept direction of a ring, so the lanes replay it
// over the rings' (count, rob) in ring order; interpolation and sums are
// FP32 and the eight ring sums are reduced across the lanes.
// ----------------------------------------------------------------------------
constexpr double BT_MARGIN = 1.0e-10;

// 1: the circle (centre c0, c1, radius rad) lies inside the face, -1: outside,
// 0: undecided
__device__ __forceinline__ int ring_side(const DFace& F, const PIC& c, double c0, double c1, double rad) {
    if (c.fkod != 1 && c.fkod != 2) return -1;
    const double u0 = (c.a0 + c.b0 * c0 + c.c0 * c1) * c.pu;
    const double u1 = (c.a1 + c.b1 * c0 + c.c1 * c1) * c.pu;
    const double m0 = rad * F.pg[0] + BT_MARGIN, m1 = rad * F.pg[1] + BT_MARGIN;
    bool in = u0 - m0 >= 0.0 && u0 + m0 <= 1.0 && u1 - m1 >= 0.0 && u1 + m1 <= 1.0;
    bool out = u0 + m0 < 0.0 || u0 - m0 > 1.0 || u1 + m1 < 0.0 || u1 - m1 > 1.0;
    if (c.fkod == 1) {
        const double u2 = 1.0 - (u0 + u1), m2 = rad * F.pg[2] + BT_MARGIN;
        in = in && u2 - m2 >= 0.0 && u2 + m2 <= 1.0;
        out = out || u2 + m2 < 0.0 || u2 - m2 > 1.0;
    }
    return in ? 1 : (out ? -1 : 0);
}

// kernel-table values at rob times the weight da (InterpolateTable.f90:58-82)
__device__ __forceinline__ void kt_values_f32(const BtF32& T, float rob, float da, float J[4]) {
#pragma clang fp contract(fast)
    const float* kt = T.kt;
    const int nr0 = (int)((rob - kt[0]) * T.kt_inv_delta);
    if (nr0 <= 0) {
        for (int c = 0; c < 4; ++c) J[c] = kt[1 + c] * da;
    } else if (nr0 >= 40) {
        for (int c = 0; c < 4; ++c) J[c] = kt[40 * 5 + 1 + c] * da;
    } else {
        const float* r0 = &kt[nr0 * 5];
        const float csi = (rob - r0[0]) * T.kt_inv_delta;
        for (int c = 0; c < 4; ++c) J[c] = (r0[1 + c] + csi * (r0[6 + c] - r0[1 + c])) * da;
    }
}

// accumulators: IntWdV, IntdWrm1dV, IntGWZrm1dV, G0..2, M00 M01 M02 M11 M12 M22;
// n directions with common R_z and unit-vector sums s = (R0, R1, R0^2, R0R1, R1^2)
__device__ __forceinline__ void bt_add(float a[12], const float J[4], float n, float R2, const float s[5]) {
#pragma clang fp contract(fast)
    a[0] += n * J[0];
    a[1] += n * J[1];
    a[2] += R2 * R2 * n * J[1];
    a[3] += J[2] * s[0];
    a[4] += J[2] * s[1];
    a[5] += J[2] * R2 * n;
    a[6] += J[3] * s[2];
    a[7] += J[3] * s[3];
    a[8] += J[3] * R2 * s[0];
    a[9] += J[3] * s[4];
    a[10] += J[3] * R2 * s[1];
    a[11] += J[3] * R2 * R2 * n;
}

// a table without the ring structure: one lane, direction by direction
__device__ void bt_pair_f32_dirs(const BtF32& T, const DFace& F, double L0, double L1, double L2,
                                 float a[12]) {
    const PIC pic = pic_of(F);
    const float L2f = (float)L2;
    float robpre = 2.1f, J[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < BT_RINGS * BT_RDIRS; ++k) {
        const BitF64 g = T.bd[k];
        if (!(L2 + g.b2 <= 0.0)) continue;
        const double ts = L2 * g.nr;
        if (!d_point_internal(pic, L0 + g.b0 * ts, L1 + g.b1 * ts)) continue;
        const float* f = T.bf[k];
        const float rob = L2f * f[0];
        if (fabsf(rob - robpre) > 1.0e-3f) {
            kt_values_f32(T, rob, f[1], J);
            robpre = rob;
        }
        const float s[5] = {f[2], f[3], f[2] * f[2], f[2] * f[3], f[3] * f[3]};
        bt_add(a, J, 1.f, T.br2[k], s);
    }
}

// Segmented contexts: the lanes of a pair read their segment's table (one
// per segment, consecutive) from global memory.
template <bool SG>
__global__ __launch_bounds__(256) void k_boundary_int_f32(DP P0, Seg S, const DFace* __restrict__ faces,
                                                          const BtF32* __restrict__ tab,
                                                          BTab* __restrict__ bt,
                                                          const int* __restrict__ work,
                                                          const int* nwork) {
    __shared__ BtF32 T_s;
    const int nw = *nwork;
    const int per_block = blockDim.x / BT_RINGS;
    if (blockIdx.x * per_block >= nw) return;  // no pair for this workgroup
    if (!SG) {
        const int4* src = reinterpret_cast<const int4*>(tab);
        int4* dst = reinterpret_cast<int4*>(&T_s);
        for (int t = threadIdx.x; t < (int)(sizeof(BtF32) / 16); t += blockDim.x) dst[t] = src[t];
        __syncthreads();
    }
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = gt & (BT_RINGS - 1);
    const int stride = (gridDim.x * blockDim.x) / BT_RINGS;
    // the eight lanes of a pair stay together: same w, same trip count
    for (int w = gt / BT_RINGS; w < nw; w += stride) {
        const int e = work[w];
        const int i = e & (int)NB_MASK, q = (int)((uint32_t)e >> NB_SHIFT);
        const int sg = SG ? S.pblk[i >> 8] : 0;
        const DP& P = SG ? S.P[sg] : P0;
        const BtF32& T = SG ? tab[sg] : T_s;
        BTab& rec = bt[(size_t)q * P.cap + i];
        const double L0 = rec.loc[0], L1 = rec.loc[1];
        double L2 = rec.loc[2];
        const DFace& F = faces[rec.face];
        const double zpmin = 0.05 * P.h;
        float a[12];
        for (int t = 0; t < 12; ++t) a[t] = 0.f;
        const bool neg = L2 < 0.0;
        if (L2 < zpmin) L2 = zpmin;
        if (!neg && !T.all_ok) {
            if (r == 0) bt_pair_f32_dirs(T, F, L0, L1, L2, a);
        } else if (!neg) {
            const RingC R = T.rc[r];
            const PIC pic = pic_of(F);
            float n = 0.f, s[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
            if (L2 + R.bz <= 0.0) {  // the plane test, common to the ring
                const double ts = L2 * R.nr;
                const int side = ring_side(F, pic, L0, L1, ts * R.rho);
                if (side > 0) {
                    n = (float)BT_RDIRS;
                    s[0] = R.s0; s[1] = R.s1; s[2] = R.s00; s[3] = R.s01; s[4] = R.s11;
                } else if (side == 0) {
                    for (int k = r * BT_RDIRS; k < (r + 1) * BT_RDIRS; ++k) {
                        const BitF64 g = T.bd[k];
                        if (!d_point_internal(pic, L0 + g.b0 * ts, L1 + g.b1 * ts)) continue;
                        const float* f = T.bf[k];
                        n += 1.f;
                        s[0] += f[2];
                        s[1] += f[3];
                        s[2] += f[2] * f[2];
                        s[3] += f[2] * f[3];
                        s[4] += f[3] * f[3];
                    }
                }
            }
            const float L2f = (float)L2;
            const float rob = L2f * R.ck;
            // the reuse rule over the rings in order: the ring whose values ring r uses
            float robpre = 2.1f;
            int src = -1;
            for (int k = 0; k < BT_RINGS; ++k) {
                const float nk = __shfl(n, k, BT_RINGS), rk = __shfl(rob, k, BT_RINGS);
                if (nk > 0.f && fabsf(rk - robpre) > 1.0e-3f) {
                    robpre = rk;
                    if (k <= r) src = k;
                }
            }
            if (n > 0.f && src >= 0) {
                float J[4];
                kt_values_f32(T, L2f * T.rc[src].ck, T.rc[src].da, J);
                bt_add(a, J, n, R.R2, s);
            }
        }
        for (int o = BT_RINGS / 2; o > 0; o >>= 1)
            for (int t = 0; t < 12; ++t) a[t] += __shfl_xor(a[t], o, BT_RINGS);
        // bi(1:8), then IntGiWrRdV column-major: 17 doubles, spread over the lanes
        double v[17];
        v[0] = 0.0;
        v[1] = neg ? 0.0 : (L2 == zpmin ? (d_point_internal(F, L0, L1) ? 0.5 : 0.0) : (double)a[0]);
        v[2] = a[1]; v[3] = a[3]; v[4] = a[4]; v[5] = a[5]; v[6] = a[2]; v[7] = 0.0;
        v[8] = a[6]; v[9] = a[7]; v[10] = a[8];
        v[11] = a[7]; v[12] = a[9]; v[13] = a[10];
        v[14] = a[8]; v[15] = a[10]; v[16] = a[11];
        double lo = 0.0, hi = 0.0;
#pragma unroll
        for (int t = 0; t < BT_RINGS; ++t)
            if (t == r) {
                lo = v[t];
                hi = v[t + 8];
            }
        double* out = rec.bi;  // bi[8] and ig[9] are contiguous
        out[r] = lo;
        out[r + 8] = hi;
        if (r == 0) out[16] = v[16];
    }
}

// ----------------------------------------------------------------------------
// Global quantities of the fluid (fluid_global_quantities.f90:66-72: mass,
// linear momentum, angular momentum about the origin, volume) as compensated
// FP64 sums: every partial sum is a double-double (s + c, Knuth's TwoSum on
// each addition), so the result is the exact sum of the terms rounded once,
// whatever the reduction order.
// ----------------------------------------------------------------------------
struct DD {
    double s, c;
};
__device__ __forceinline__ DD dd_add(DD x, DD y) {
    const double s = x.s + y.s;
    const double bp = s - x.s;
    const double e = (x.s - (s - bp)) + (y.s - bp);
    const double c = e + (x.c + y.c);
    const double s2 = s + c;  // renormalise (|s| >= |c|)
    return DD{s2, c - (s2 - s)};
}
__device__ __forceinline__ DD dd_shfl_xor(DD v, int o) {
    return DD{__shfl_xor(v.s, o, 64), __shfl_xor(v.c, o, 64)};
}

constexpr int GQ_N = 8;  // mass, lin mom x3, ang mom x3, volume

__global__ __launch_bounds__(256) void k_global_partial(const double4* __restrict__ xm,
                                                        const double4* __restrict__ vel,
                                                        const double2* __restrict__ rp,
                                                        double* __restrict__ part, const int* iscal) {
    __shared__ double red[4][GQ_N][2];
    const int n = iscal[IS_N];
    DD acc[GQ_N];
    for (int q = 0; q < GQ_N; ++q) acc[q] = DD{0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double4 x = xm[i];
        const double4 v = vel[i];
        const double m = x.w;
        // Vector_Product(coord, vel)
        const double a0 = x.y * v.z - x.z * v.y;
        const double a1 = x.z * v.x - x.x * v.z;
        const double a2 = x.x * v.y - x.y * v.x;
        const double t[GQ_N] = {m, m * v.x, m * v.y, m * v.z, m * a0, m * a1, m * a2, m / rp[i].x};
#pragma unroll
        for (int q = 0; q < GQ_N; ++q) acc[q] = dd_add(acc[q], DD{t[q], 0.0});
    }
#pragma unroll
    for (int q = 0; q < GQ_N; ++q)
        for (int o = 32; o > 0; o >>= 1) acc[q] = dd_add(acc[q], dd_shfl_xor(acc[q], o));
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < GQ_N; ++q) {
            red[wid][q][0] = acc[q].s;
            red[wid][q][1] = acc[q].c;
        }
    __syncthreads();
    if (threadIdx.x < GQ_N) {
        const int q = threadIdx.x;
        DD r = DD{red[0][q][0], red[0][q][1]};
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = dd_add(r, DD{red[w][q][0], red[w][q][1]});
        part[(size_t)blockIdx.x * GQ_N * 2 + 2 * q] = r.s;
        part[(size_t)blockIdx.x * GQ_N * 2 + 2 * q + 1] = r.c;
    }
}

__global__ void k_global_final(const double* __restrict__ part, int nblocks, double* __restrict__ out) {
    const int q = threadIdx.x;
    if (q >= GQ_N) return;
    DD r = DD{0.0, 0.0};
    for (int b = 0; b < nblocks; ++b)
        r = dd_add(r, DD{part[(size_t)b * GQ_N * 2 + 2 * q], part[(size_t)b * GQ_N * 2 + 2 * q + 1]});
    out[q] = r.s + r.c;
}

// ----------------------------------------------------------------------------
// Free-surface monitor lines: calc_pelo.f90:82-131.  One workgroup per line;
// each thread takes one point of a FS_WG-point chunk: its cell (the ceiling
// form of CalcVarp.f90:66-72, with out-of-range indices giving cell 0 before
// any integer conversion), then the two particles of that cell closest to
// the point.  The reference takes them with two minloc passes over an array
// preset to 9e12; a running (first, second) minimum with strict comparisons
// in cell order (ascending particle index, as NPartOrd) picks the same pair,
// because a point and a particle of the same cell are closer than 9e12.  A
// cell holding one particle gives that particle twice.  Thread 0 then runs
// the reference's sequential flags over the chunk: pelolib takes the last
// filled cell, z_lower_fluid_top stops at the first empty cell above fluid;
// cells outside the grid are skipped.
// icont: control_lines(:)%Icont, 1-based (first, last); the reference starts
// at Icont(1)+1, which is 0-based index Icont(1).
// out[3*line + d] = pelolib, out[3*(nlines + line) + d] = z_lower_fluid_top.
// ----------------------------------------------------------------------------
constexpr int FS_WG = 256;

__device__ __forceinline__ int d_point_cell(const DP& P, double x, double y, double z) {
    const double q[3] = {x, y, z};
    int ijk[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double f = ceil((q[d] - P.extr_lo[d]) / P.dcd[d]);
        if (!(f >= 1.0 && f <= (double)P.ncd[d])) return 0;  // also NaN
        ijk[d] = (int)f;
    }
    return d_cell_number(P, ijk[0], ijk[1], ijk[2]);
}

__global__ __launch_bounds__(FS_WG) void k_free_surface(DP P, const double4* __restrict__ xm,
                                                       const int* __restrict__ start,
                                                       const double* __restrict__ pts,
                                                       const int* __restrict__ icont, int nlines,
                                                       double* __restrict__ out) {
    __shared__ double val[FS_WG][3];
    __shared__ int st[FS_WG];  // 0 outside the grid, 1 empty cell, 2 filled cell
    const int line = blockIdx.x;
    const int j0 = icont[2 * line], j1 = icont[2 * line + 1];
    bool zflag = true, below = false;
    double pel[3] = {0.0, 0.0, 0.0}, zl[3] = {0.0, 0.0, 0.0};
    const double half_dx = P.dx / 2.0;
    for (int base = j0; base < j1; base += FS_WG) {
        const int j = base + (int)threadIdx.x;
        int s = 0;
        if (j < j1) {
            const double qx = pts[3 * j], qy = pts[3 * j + 1], qz = pts[3 * j + 2];
            const int c = d_point_cell(P, qx, qy, qz);
            if (c > 0) {
                const int b = start[c - 1], e = start[c];
                s = e > b ? 2 : 1;
                if (s == 2) {
                    double m1 = 0.0, m2 = 0.0;
                    int i1 = -1, i2 = -1;
                    for (int m = b; m < e; ++m) {
                        const double4 x = xm[m];
                        const double r0 = qx - x.x, r1 = qy - x.y, r2 = qz - x.z;
                        const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
                        if (i1 < 0 || d2 < m1) {
                            m2 = m1;
                            i2 = i1;
                            m1 = d2;
                            i1 = m;
                        } else if (i2 < 0 || d2 < m2) {
                            m2 = d2;
                            i2 = m;
                        }
                    }
                    if (i2 < 0) i2 = i1;
                    const double4 a = xm[i1], bb = xm[i2];
                    val[threadIdx.x][0] = (a.x + bb.x) * 0.5 + half_dx;
                    val[threadIdx.x][1] = (a.y + bb.y) * 0.5 + half_dx;
                    val[threadIdx.x][2] = (a.z + bb.z) * 0.5 + half_dx;
                }
            }
        }
        st[threadIdx.x] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int nt = min(FS_WG, j1 - base);
            for (int t = 0; t < nt; ++t) {
                if (st[t] == 0) continue;
                if (st[t] == 1) {
                    if (below) zflag = false;
                    continue;
                }
                for (int d = 0; d < 3; ++d) {
                    pel[d] = val[t][d];
                    if (zflag) zl[d] = val[t][d];
                }
                below = true;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) {
            out[3 * line + d] = pel[d];
            out[3 * (nlines + line) + d] = zl[d];
        }
}

// ----------------------------------------------------------------------------
// dt: time_step_duration.f90:74-116.  Positive doubles order like their bit
// patterns, so the minimum over particles is an exact atomicMin on uint64.
// ----------------------------------------------------------------------------
__device__ __forceinline__ double d_dt_cfl(const DP& P, double vx, double vy, double vz) {
    const double U = sqrt(vx * vx + vy * vy + vz * vz);
    return P.cfl * 2.0 * P.h / (P.celerita + U);
}

__device__ __forceinline__ void block_min_to(unsigned long long v, unsigned long long* dst) {
    __shared__ unsigned long long red[16];
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(v, o, 64);
        v = t < v ? t : v;
    }
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = red[w] < m ? red[w] : m;
        atomicMin(dst, m);
    }
}

template <bool SG>
__global__ __launch_bounds__(256) void k_dtmin(DP P0, Seg S, const double4* __restrict__ vel,
                                               unsigned long long* dtmin, const int* iscal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.pblk[(blockIdx.x * blockDim.x) >> 8]) : 0;
    const DP& P = SG ? S.P[sg] : P0;
    dtmin += sg;
    unsigned long long v = 0x7FF0000000000000ull;
    if (i - P.base < iscal[sg * IS_COUNT + IS_N]) {
        const double4 u = vel[i];
        v = (unsigned long long)__double_as_longlong(d_dt_cfl(P, u.x, u.y, u.z));
    }
    block_min_to(v, dtmin);
}

// step prologue (time_step_loop.f90:180-186): dt_previous_step, dt, time
// Segmented contexts: one lane per segment.
template <bool SG>
__global__ void k_dt_begin(DP P0, Seg S, double* scal, unsigned long long* dtmin, int* iscal, int advance) {
    const int sg = SG ? (int)(blockIdx.x * blockDim.x + threadIdx.x) : 0;
    if (SG ? sg >= S.nseg : (threadIdx.x != 0 || blockIdx.x != 0)) return;
    const DP& P = SG ? S.P[sg] : P0;
    scal += sg * SC_COUNT;
    iscal += sg * IS_COUNT;
    dtmin += sg;
    if (advance) {
        iscal[IS_IT] += 1;
        iscal[IS_ONGOING] = iscal[IS_IT];
        scal[SC_DTPREV] = scal[SC_DT];
    }
    const int og = iscal[IS_ONGOING];
    double dtmin_v = 3.4e38;
    double dt_vis = 3.4e38;
    if (P.kin_visc > 1.0e-24) {
        dt_vis = P.vsc_coeff * P.squareh / (0.5 * P.kin_visc);
        dtmin_v = fmin(dtmin_v, dt_vis);
    }
    const double mparts = __longlong_as_double((long long)*dtmin);
    if (iscal[IS_N] == 0) {
        scal[SC_DT] = 0.0;
    } else {
        dtmin_v = fmin(dtmin_v, fmin(mparts, dt_vis));
        if (og == 0) scal[SC_DT] = dtmin_v;
        else scal[SC_DT] = (1.0 - P.pesodt) * dtmin_v + P.pesodt * scal[SC_DTAVG];
    }
    const double dt = scal[SC_DT];
    if (og == 0) scal[SC_DTAVG] = dt;
    else scal[SC_DTAVG] = (scal[SC_DTAVG] * (double)(og - 1) + dt) / (double)og;
    if (advance) scal[SC_TIME] = scal[SC_TIME] + dt;
    scal[SC_DTVEL] = 0.5 * (dt + scal[SC_DTPREV]);
    scal[SC_TV1] = P.tetav * P.celerita * dt / P.h;
    scal[SC_TP1] = P.tetap * P.celerita * dt / P.h;
    *dtmin = 0x7FF0000000000000ull;
}

// ----------------------------------------------------------------------------
// The four pair passes.  A persistent workgroup takes a cell block, stages the
// block's 9 neighbour rows into LDS (coalesced, contiguous in sorted order)
// and gives each home particle one lane that walks its neighbour list in the
// reference's order, reading neighbour records from LDS.  Blocks whose rows
// exceed the LDS tile read their neighbours from global memory instead.
// FP64: the reference's expressions, operation for operation.  FP32-mixed:
// FP32 pair math on cell-relative separations, FP32 sums per particle, FP64
// wall terms, FP64 integration.
// ----------------------------------------------------------------------------
template <bool F32>
struct Rec;
template <>
struct Rec<false> {
    using X = double4;
    using V = double4;
    using RP = double2;
};
template <>
struct Rec<true> {
    using X = float4;
    using V = float4;
    using RP = float2;
};


constexpr uint32_t NB_NONE = 0xFFFFFFFFu;

// Walks particle i's neighbour list in the reference's order, CH entries at a
// time: the CH list entries and then the CH neighbour records are all loaded
// before any is used, so their latencies overlap; body() then consumes them
// strictly in list order (the FP64 sums keep the reference's order).
template <int CH, class R, class Fetch, class Body>
__device__ __forceinline__ void walk_list(const uint32_t* __restrict__ nbr, size_t N, int i, int nn,
                                          Fetch&& fetch, Body&& body) {
    static_assert(CH % 4 == 0, "list walks read whole 16-byte quarters of a group");
    for (int k0 = 0; k0 < nn; k0 += CH) {
        uint32_t ev[CH];
        if constexpr (NB_G >= 4) {
#pragma unroll
            for (int q = 0; q < CH / 4; ++q) {
                const uint4 v = *reinterpret_cast<const uint4*>(nbr + nb_index(N, i, k0 + 4 * q));
                ev[4 * q + 0] = v.x;
                ev[4 * q + 1] = v.y;
                ev[4 * q + 2] = v.z;
                ev[4 * q + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int u = 0; u < CH; ++u) ev[u] = nbr[nb_index(N, i, k0 + u)];
        }
#pragma unroll
        for (int u = 0; u < CH; ++u)
            if (k0 + u >= nn) ev[u] = NB_NONE;
        R rq[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u)
            if (ev[u] != NB_NONE) rq[u] = fetch(ev[u]);
#pragma unroll
        for (int u = 0; u < CH; ++u)
            if (ev[u] != NB_NONE) body(ev[u], rq[u]);
    }
}

template <class X, class V, class RP>
struct NRec {
    X x;
    V v;
    RP r;
};
template <class X, class RP>
struct NRec2 {
    X x;
    RP r;
};

// wall_function_for_SASPH.f90:56-80 (no land-cover map), FP64: slip
// coefficient and mixing-length viscosity of one particle-frontier pair at
// distance r; d50 is ten times the frontier's roughness length.  Nepero_number
// and k_v: Static_allocation_module.f90:51,53.
__device__ __forceinline__ void wall_function(double u_t, double d50, double r, double dx,
                                              double& slip, double& nut) {
    const double e = 2.718281828, kv2 = 0.41 * 0.41;
    slip = 0.0;
    nut = 1.0e-15;
    if (r <= 0.5 * dx) {
        const double z0 = d50 / 10.0;
        if (r < z0 * e) {
            slip = 1.0;
            nut = kv2 * u_t * z0 * e;
        } else {
            const double L = log(r / z0);
            slip = 1.0 / L;
            nut = kv2 * u_t * r / L;
        }
    }
}

// line_plane_intersection.f90:39-76 with three_plane_intersection.f90:42-66
// and Matrix_Inversion_3x3.f90:39-58, FP64: the point where the line through
// P1 and P2 meets the plane through A, B and C (two planes through the line,
// each with a point 999 away from P1, and the plane); false when the 3x3
// system's |determinant| is below 1e-36.
__device__ bool d_line_plane(const double P1[3], const double P2[3], const double A[3], const double B[3],
                             const double C[3], double X[3]) {
    const double Q1[3] = {P1[0] - 999.0, P1[1], P1[2]};
    const double Q2[3] = {P1[0], P1[1] - 999.0, P2[2] == P2[2] ? P1[2] - 999.0 : P1[2]};
    double m[3][3], dv[3];
    for (int r = 0; r < 2; ++r) {
        const double* Q = r == 0 ? Q1 : Q2;
        m[r][0] = (P2[1] - P1[1]) * (Q[2] - P1[2]) - (Q[1] - P1[1]) * (P2[2] - P1[2]);
        m[r][1] = -((P2[0] - P1[0]) * (Q[2] - P1[2]) - (Q[0] - P1[0]) * (P2[2] - P1[2]));
        m[r][2] = (P2[0] - P1[0]) * (Q[1] - P1[1]) - (Q[0] - P1[0]) * (P2[1] - P1[1]);
        dv[r] = -(m[r][0] * P1[0] + m[r][1] * P1[1] + m[r][2] * P1[2]);
    }
    m[2][0] = (B[1] - A[1]) * (C[2] - A[2]) - (C[1] - A[1]) * (B[2] - A[2]);
    m[2][1] = -((B[0] - A[0]) * (C[2] - A[2]) - (C[0] - A[0]) * (B[2] - A[2]));
    m[2][2] = (B[0] - A[0]) * (C[1] - A[1]) - (C[0] - A[0]) * (B[1] - A[1]);
    dv[2] = -(m[2][0] * A[0] + m[2][1] * A[1] + m[2][2] * A[2]);
    const double det = m[0][0] * m[1][1] * m[2][2] + m[1][0] * m[2][1] * m[0][2] +
                       m[2][0] * m[0][1] * m[1][2] - m[0][0] * m[2][1] * m[1][2] -
                       m[2][0] * m[1][1] * m[0][2] - m[1][0] * m[0][1] * m[2][2];
    if (!(fabs(det) >= 1.0e-36)) return false;
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

// CancelOutgoneParticles_3D.f90:67-158 for one open face, FP64: the step from
// xo to xn ends on the face's outer side, started on its inner side, and
// the segment crosses the face's plane inside the face.
__device__ bool open_face_crossed(const DFace& F, const double xo[3], const double xn[3]) {
    const int nodes = F.nodes;
    const double* G1 = &F.GX[0];
    const double* G2 = &F.GX[3];
    const double* G3 = &F.GX[6];
    const double* Gn = &F.GX[3 * (nodes - 1)];
    const double dx21 = G2[0] - G1[0], dx31 = G3[0] - G1[0];
    const double dy21 = G2[1] - G1[1], dy31 = G3[1] - G1[1];
    const double dz21 = G2[2] - G1[2], dz31 = G3[2] - G1[2];
    auto det_of = [&](const double* X) {
        const double dx = X[0] - G1[0], dy = X[1] - G1[1], dz = X[2] - G1[2];
        return dx21 * dy31 * dz + dx31 * dy * dz21 + dx * dy21 * dz31 - dx21 * dy * dz31 - dx31 * dy21 * dz -
               dx * dy31 * dz21;
    };
    const double dn = det_of(xn);
    if (!(dn <= 0.0)) return false;
    const double dpo = det_of(xo);
    if (copysign(1.0, dn) == copysign(1.0, dpo)) return false;
    // local coordinates (T(:,i) . (X - node(nodes))) of the face's nodes 1-3 and of both ends
    double Pp[3][3], Ln[3], Lo[3];
    const double* Gk[3] = {G1, G2, G3};
    for (int sdi = 0; sdi < 3; ++sdi) {
        for (int k = 0; k < 3; ++k) {
            double a = 0.0;
            for (int sdj = 0; sdj < 3; ++sdj) a = a + F.T[3 * sdi + sdj] * (Gk[k][sdj] - Gn[sdj]);
            Pp[k][sdi] = a;
        }
        double a = 0.0, b = 0.0;
        for (int sdj = 0; sdj < 3; ++sdj) {
            a = a + F.T[3 * sdi + sdj] * (xn[sdj] - Gn[sdj]);
            b = b + F.T[3 * sdi + sdj] * (xo[sdj] - Gn[sdj]);
        }
        Ln[sdi] = a;
        Lo[sdi] = b;
    }
    double LX[3];
    if (!d_line_plane(Ln, Lo, Pp[0], Pp[1], Pp[2], LX)) return false;
    return d_point_internal(pic_of(F), LX[0], LX[1]);
}

// CancelOutgoneParticles_2D.f90:80-101 for one open side (x-z plane), FP64:
// the step from xo to xn ends on the outer side of the side's line
// (determinant <= 0), started on the other, and the side's two vertices lie
// on opposite sides of the step's line.
__device__ bool open_side_crossed(const DFace& S, const double xo[3], const double xn[3]) {
    const double o1x = S.v1[0], o1z = S.v1[1], o2x = S.v2[0], o2z = S.v2[1];
    const double dn = (xn[2] - o1z) * (o2x - o1x) - (xn[0] - o1x) * (o2z - o1z);
    if (!(dn <= 0.0)) return false;
    const double dold = (xo[2] - o1z) * (o2x - o1x) - (xo[0] - o1x) * (o2z - o1z);
    if (copysign(1.0, dn) == copysign(1.0, dold)) return false;
    const double d1 = (o1z - xo[2]) * (xn[0] - xo[0]) - (o1x - xo[0]) * (xn[2] - xo[2]);
    const double d2 = (o2z - xo[2]) * (xn[0] - xo[0]) - (o2x - xo[0]) * (xn[2] - xo[2]);
    return copysign(1.0, d1) != copysign(1.0, d2);
}

// SA-SPH terms of the momentum equation (AddBoundaryContributions_to_ME3D.f90:
// 118-336), FP64; tp/td/tv start at zero
__device__ __forceinline__ void sa_momentum(const DP& P, const DFace* __restrict__ faces,
                                            const BTab* __restrict__ bt, int i, int ncbf,
                                            const double v_i[3], double2 rpi, int lam_prev,
                                            double tp[3], double td[3], double tv[3], int* iscal) {
    const double h = P.h;
    double ALEt[3] = {0, 0, 0}, gradpt[3] = {0, 0, 0}, VM[3] = {0, 0, 0}, VS[3] = {0, 0, 0};
    for (int q = 0; q < ncbf; ++q) {
        const BTab& B = bt[(size_t)q * P.cap + i];
        const DFace& F = faces[B.face];
        if (F.btype == 2) continue;
        if (F.btype != 0 && F.btype != 1) {
            atomicAdd(&iscal[IS_UNSUPPORTED], 1);
            break;
        }
        if (!(B.loc[2] > 0.0)) continue;
        double dvij[3], dvn = 0.0;
        for (int sd = 0; sd < 3; ++sd) {
            dvij[sd] = 2.0 * (v_i[sd] - F.vel[sd]);
            dvn = dvn + F.T[6 + sd] * dvij[sd];
        }
        double gl[3];
        for (int sd = 0; sd < 3; ++sd) {
            double a = 0.0;
            for (int sj = 0; sj < 3; ++sj) a = a + F.T[3 * sd + sj] * P.grav[sj];
            gl[sd] = a;
        }
        double gloc[3];
        for (int r = 0; r < 3; ++r) {
            double a = 0.0;
            for (int k = 0; k < 3; ++k) a = a + B.ig[3 * k + r] * gl[k];
            gloc[r] = a;
        }
        for (int r = 0; r < 3; ++r) {
            double a = 0.0;
            for (int k = 0; k < 3; ++k) a = a + F.T[3 * k + r] * gloc[k];
            gradpt[r] = gradpt[r] + a;
        }
        double aloc[3];
        for (int d = 0; d < 3; ++d) aloc[d] = 2.0 * rpi.y / rpi.x * B.bi[3 + d];
        for (int r = 0; r < 3; ++r) {
            double a = 0.0;
            for (int k = 0; k < 3; ++k) a = a + F.T[3 * k + r] * aloc[k];
            ALEt[r] = ALEt[r] + a;
        }
        if (P.alfamon > 0.0) {
            const double Monvisc = P.alfamon * P.celerita * h;
            const double Mmult = -Monvisc * dvn * B.bi[6];
            for (int d = 0; d < 3; ++d) VM[d] = VM[d] + Mmult * F.T[6 + d];
        }
        double slip = 0.0, cinv = 0.0;
        if (F.mode == 1) {
            slip = F.slip;
            cinv = P.kin_visc;
        } else if (F.mode == 2) {
            // AddBoundaryContributions_to_ME3D.f90:207-227
            double ut[3];
            for (int d = 0; d < 3; ++d) ut[d] = 0.5 * (dvij[d] - dvn * F.T[6 + d]);
            wall_function(sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]), F.slip * 10.0,
                          B.loc[2], P.dx, slip, cinv);
        }
        if (cinv > 0.0 && (lam_prev == 1 || F.lam == 0)) {
            const double cm = cinv * B.bi[2] * slip;
            for (int d = 0; d < 3; ++d) VS[d] = VS[d] + cm * (2.0 * (v_i[d] - F.vel[d]));
        }
    }
    for (int d = 0; d < 3; ++d) {
        tp[d] = tp[d] + ALEt[d] + gradpt[d];
        td[d] = td[d] - VM[d];
        tv[d] = tv[d] + VS[d];
    }
}

// AddBoundaryContributions_to_ME2D.f90:121-405 for fixe/tapi sides (2D form);
// tp/td/tv start at zero, as the reference's tpres/tdiss/tvisc do
__device__ __forceinline__ void sa_momentum_2d(const DP& P, const DFace* __restrict__ sides,
                                               const BTab* __restrict__ bt, int i, int ncbf,
                                               const double v_i[3], double2 rpi, int lam_prev,
                                               double tp[3], double td[3], double tv[3]) {
    const double ro0 = P.den0, roi = rpi.x, pressi = rpi.y;
    const double distpimin = P.dx;
    double VMo0 = 0.0, VMo1 = 0.0, VSh0 = 0.0, VSh1 = 0.0, RG0s = 0.0, RG1s = 0.0;
    double ALE0 = 0.0, ALE2 = 0.0;
    const double g0 = P.grav[0], g1 = P.grav[2];
    for (int q = 0; q < ncbf; ++q) {
        const BTab& B = bt[(size_t)q * P.cap + i];
        const DFace& S = sides[B.face];
        if (S.btype == 2) continue;  // AddBoundaryContributions_to_ME2D.f90:124
        const double xpi = B.loc[0], ypi = B.loc[1];
        const double IntWdS = B.bi[0], IntWdV = B.bi[2], IntWd1s0 = B.bi[5], IntWd3s0 = B.bi[6];
        const double n0 = S.T[6], n1 = S.T[8];
        double GravN = 0.0;
        GravN = GravN + P.grav[0] * n0;
        GravN = GravN + P.grav[2] * n1;
        const double sv0 = S.btype == 1 ? S.vel[0] : 0.0, sv1 = S.btype == 1 ? S.vel[2] : 0.0;
        const double Dv0 = 2.0 * (v_i[0] - sv0), Dv1 = 2.0 * (v_i[2] - sv1);
        const double DvelN = Dv0 * n0 + Dv1 * n1;
        const double distpi = fmax(ypi, distpimin);
        double pressib = pressi;
        const double pressibmin = 10.0 * ro0 * distpimin;
        if (DvelN <= 0.0) {
            if (pressib < pressibmin) pressib = pressibmin;
        }
        const double PressB = pressib - ro0 * GravN * distpi;
        const double Qii = IntWdS * (pressib + PressB) / roi;
        double RG0 = 0.0, RG1 = 0.0;
        RG0 = RG0 + S.RN[0] * g0;
        RG0 = RG0 + S.RN[1] * g1;
        RG0 = RG0 * IntWdV;
        RG1 = RG1 + S.RN[2] * g0;
        RG1 = RG1 + S.RN[3] * g1;
        RG1 = RG1 * IntWdV;
        RG0s = RG0s + RG0;
        RG1s = RG1s + RG1;
        ALE0 = ALE0 - n0 * Qii;
        ALE2 = ALE2 - n1 * Qii;
        if (xpi >= 0.0 && xpi <= S.len) {
            double slip = 0.0, cinv = 0.0;
            if (S.mode == 1) {
                slip = S.slip;
                cinv = P.kin_visc;
            } else if (S.mode == 2) {
                // AddBoundaryContributions_to_ME2D.f90:314-326
                double ut[3];
                for (int k = 0; k < 3; ++k) ut[k] = (v_i[k] - S.vel[k]) - (0.5 * DvelN * S.T[6 + k]);
                wall_function(sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]), S.slip * 10.0, ypi,
                              P.dx, slip, cinv);
            }
            const double Monvisc = P.alfamon * P.celerita * P.h;
            const double TN = Monvisc * DvelN * IntWd3s0;
            double SVforce = 0.0;
            if (lam_prev == 1 || S.lam == 0) SVforce = 2.0 * cinv * slip * IntWd1s0;
            VMo0 = VMo0 + TN * n0;
            VMo1 = VMo1 + TN * n1;
            VSh0 = VSh0 + (2.0 * v_i[0]) * SVforce;
            VSh1 = VSh1 + (2.0 * v_i[2]) * SVforce;
        }
    }
    tp[0] = tp[0] - RG0s - ALE0;
    tp[2] = tp[2] - RG1s - ALE2;
    td[0] = td[0] - VMo0;
    td[2] = td[2] - VMo1;
    tv[0] = tv[0] - VSh0;
    tv[2] = tv[2] - VSh1;
}

// AddElasticBoundaryReaction_2D.f90:66-99 over the selected sides
__device__ __forceinline__ void elastic_2d(const DP& P, const DFace* __restrict__ sides,
                                           const BTab* __restrict__ bt, int i, int ncbf,
                                           const double v_i[3], double br[3]) {
    const double h = P.h;
    const double ypimin = 0.25 * P.dx;
    const double celer02 = P.eps / P.den0;
    for (int q = 0; q < ncbf; ++q) {
        const BTab& B = bt[(size_t)q * P.cap + i];
        const DFace& S = sides[B.face];
        if (S.btype != 0 && S.btype != 1) continue;  // AddElasticBoundaryReaction_2D.f90:75
        const double xpi = B.loc[0], ypi = B.loc[1];
        if (ypi < ypimin && xpi > 0.0 && xpi < S.len) {
            double vin = 0.0;
            for (int k = 0; k < 3; ++k) vin = vin + v_i[k] * S.T[6 + k];
            if (vin < 0.0) {
                const double nr = -1.0 * celer02 * log((h + ypi - ypimin) / h) / h;
                for (int d = 0; d < 3; ++d) br[d] = br[d] + nr * S.T[6 + d];
            }
        }
    }
}

// ----------------------------------------------------------------------------
// momentum: RHS_momentum_equation.f90:72-179, inter_EqMoto.f90:108-283,
// viscomon.f90:54-71, viscomorris.f90:60-71, AddElasticBoundaryReaction_3D.f90:
// 67-98; fused Leapfrog_momentum.f90:66-86
// ----------------------------------------------------------------------------
template <bool F32, int CH, bool SG>
__global__ __launch_bounds__(256) void k_momentum(
    DP P0, Seg S, const DFace* __restrict__ faces, const DEdge* __restrict__ edges,
    const int* __restrict__ bn, const BTab* __restrict__ bt, const double4* __restrict__ xm,
    const double4* __restrict__ vel, const double2* __restrict__ rp, F4In xmf,
    F4In velf, const float2* __restrict__ rpf, const int* __restrict__ lam_in,
    int* __restrict__ lam_out, const uint32_t* __restrict__ nbr, const int* __restrict__ nnb,
    double4* __restrict__ acc_out, double4* __restrict__ vel_out, F4Out velf_out,
    const double* scal, int* iscal) {
    const int vb = xcd_block(gridDim.x);
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.pblk[(vb * (int)blockDim.x) >> 8]) : 0;
    const DP& P = SG ? S.P[sg] : P0;
    scal += sg * SC_COUNT;
    iscal += sg * IS_COUNT;
    stencil_offsets(P);
    const double h = P.h;
    // Leapfrog_momentum.f90:66-86 integrates with dtvel; the explicit RK1
    // branch (Euler.f90:85-93) with dt
    const double dtvel = P.rk1 ? scal[SC_DT] : scal[SC_DTVEL];
    const size_t N = P.cap;
    {
        const int i = vb * blockDim.x + threadIdx.x;
        if (i - P.base < iscal[IS_N]) {
            const double4 vi = vel[i];  // var == vel here
            const double2 rpi = rp[i];
            double tp[3] = {0.0, 0.0, 0.0}, td[3] = {0.0, 0.0, 0.0}, tv[3] = {0.0, 0.0, 0.0};
            const int ncbf = P.semi ? bn[i] : 0;
            const double v_i[3] = {vi.x, vi.y, vi.z};
            if (ncbf > 0) {
                if (P.dim == 2)
                    sa_momentum_2d(P, faces, bt, i, ncbf, v_i, rpi, lam_in[i], tp, td, tv);
                else
                    sa_momentum(P, faces, bt, i, ncbf, v_i, rpi, lam_in[i], tp, td, tv, iscal);
            }
            const int nn = nnb[i];
            double rvs[3];
            if constexpr (!F32) {
                const double4 xi = xm[i];
                const double rhoi = rpi.x, pi = rpi.y;
                double tpa[3] = {0, 0, 0}, ale[3] = {0, 0, 0};
                rvs[0] = rvs[1] = rvs[2] = 0.0;
                const double celt = P.celerita + P.celerita;
                const double nui = P.kin_visc, nuj = P.kin_visc;
                const bool morris = nui > 1.0e-12 && nuj > 1.0e-12;
                using RM = NRec<double4, double4, double2>;
                auto fetch = [&](uint32_t e) {
                    const int j = (int)(e & NB_MASK);
                    return RM{xm[j], vel[j], rp[j]};
                };
                walk_list<4, RM>(nbr, N, i, nn, fetch, [&](uint32_t e, const RM& q) {
                    if ((int)(e & NB_MASK) == i) return;
                    const double4 xj = q.x;
                    const double4 vj = q.v;
                    const double2 rpj = q.r;
                    const double R0 = xi.x - xj.x, R1 = xi.y - xj.y, R2 = xi.z - xj.z;
                    const double rij = R0 * R0 + R1 * R1 + R2 * R2;
                    K64 K;
                    kvals64(P, rij, true, true, false, K);
                    const double rhoj = rpj.x, pj = rpj.y, amassj = xj.w;
                    const double dv0 = vj.x - vi.x, dv1 = vj.y - vi.y, dv2 = vj.z - vi.z;
                    double alpha = (pj - pi) / (rhoi * rhoj);
                    tpa[0] = tpa[0] + -amassj * alpha * R0 * K.K3;
                    tpa[1] = tpa[1] + -amassj * alpha * R1 * K.K3;
                    tpa[2] = tpa[2] + -amassj * alpha * R2 * K.K3;
                    alpha = (pi * (rhoj / rhoi + 1.0) + pj * (rhoi / rhoj - 1.0)) / (rhoi * rhoj);
                    ale[0] = ale[0] - amassj * alpha * R0 * K.K3;
                    ale[1] = ale[1] - amassj * alpha * R1 * K.K3;
                    ale[2] = ale[2] - amassj * alpha * R2 * K.K3;
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
                for (int d = 0; d < 3; ++d) tp[d] = tp[d] + tpa[d] + ale[d];
            } else {
                                // symmetric form -m_j (p_i/rho_i^2 + p_j/rho_j^2) r K3 of the
                // reference's difference term plus its ALE1 complement
                const float4 xf = xmf[i];
                const float4 vf = velf[i];
                const float2 rpfi = rpf[i];
                const float rhoi = rpfi.x, pi = rpfi.y;
                const float pri = pi * __builtin_amdgcn_rcpf(rhoi * rhoi);
                float s_p0 = 0.f, s_p1 = 0.f, s_p2 = 0.f, s_d0 = 0.f, s_d1 = 0.f, s_d2 = 0.f;
                float s_v0 = 0.f, s_v1 = 0.f, s_v2 = 0.f;
                const float mon = P.f_alfamon * (P.f_celerita + P.f_celerita) * P.f_h;
                const float nu = P.f_kin_visc;
                const float atil = 4.0f * (nu * nu);
                const bool morris = P.kin_visc > 1.0e-12;
                // neighbour records: (x, m) and (v, rho - den0); the pressure
                // follows from the linear EoS, so one 32-byte run per neighbour
                using RM = NRec2<float4, float4>;
                auto fetch = [&](uint32_t e) {
                    const int j = (int)(e & NB_MASK);
                    return RM{xmf[j], velf[j]};
                };
                walk_list<CH, RM>(nbr, N, i, nn, fetch, [&](uint32_t e, const RM& q) {
                    if ((int)(e & NB_MASK) == i) return;
                    const float4 xj = q.x;
                    const float4 vj = q.r;
                    float R0, R1, R2;
                    rag32(xf, xj, e, R0, R1, R2);
                    const float rij = R0 * R0 + R1 * R1 + R2 * R2;
                    K32 K;
                    kvals32(P, rij, true, true, false, K);
                    const float rhoj = P.f_den0 + vj.w, pj = __builtin_fmaf(P.f_eosk, vj.w, P.f_prif);
                    const float mj = xj.w;
                    const float dv0 = vj.x - vf.x, dv1 = vj.y - vf.y, dv2 = vj.z - vf.z;
                    const float cp = -mj * (pri + pj * __builtin_amdgcn_rcpf(rhoj * rhoj)) * K.K3;
                    s_p0 += cp * R0; s_p1 += cp * R1; s_p2 += cp * R2;
                    const float vrij = -(dv0 * R0 + dv1 * R1 + dv2 * R2);
                    const float cd = mj * mon * __builtin_amdgcn_rcpf(rhoi + rhoj) * vrij * K.K2;
                    s_d0 += cd * R0; s_d1 += cd * R1; s_d2 += cd * R2;
                    if (morris) {
                        const float fac = mj * atil * __builtin_amdgcn_rcpf(nu * rhoi + nu * rhoj + 1.0e-3f);
                        const float cv = -fac * K.K2 * rij;
                        s_v0 += cv * dv0; s_v1 += cv * dv1; s_v2 += cv * dv2;
                    }
                });
                tp[0] = tp[0] + (double)s_p0; tp[1] = tp[1] + (double)s_p1; tp[2] = tp[2] + (double)s_p2;
                td[0] = td[0] + (double)s_d0; td[1] = td[1] + (double)s_d1; td[2] = td[2] + (double)s_d2;
                tv[0] = tv[0] + (double)s_v0; tv[1] = tv[1] + (double)s_v1; tv[2] = tv[2] + (double)s_v2;
                rvs[0] = s_v0; rvs[1] = s_v1; rvs[2] = s_v2;
            }
            int lam = 0;
            if (P.kin_visc > 0.0) {
                const double apg = sqrt(tp[0] * tp[0] + tp[1] * tp[1] + tp[2] * tp[2]) + P.GI;
                const double amo = sqrt(rvs[0] * rvs[0] + rvs[1] * rvs[1] + rvs[2] * rvs[2]);
                double w = 0.0;
                if (apg != 0.0) w = amo / apg * 100.0;
                lam = (w > 5.0 || apg == 0.0) ? 1 : 0;
            }
            lam_out[i] = lam;
            double a[3];
            if (ncbf > 0 && P.dim == 2) {
                double br[3] = {0.0, 0.0, 0.0};
                elastic_2d(P, faces, bt, i, ncbf, v_i, br);
                for (int d = 0; d < 3; ++d) a[d] = tp[d] + td[d] + tv[d] + P.grav[d] + br[d];
            } else if (ncbf > 0) {
                double br[3] = {0.0, 0.0, 0.0};
                const double zimin = 0.25 * P.dx;
                const double celer02 = P.eps / P.den0;
                // AddElasticBoundaryReaction_3D.f90:72-104 (faces); rea: the
                // slots whose face reacted (ReaFace), for the edges below
                static_assert(SPH_MB <= 32, "one ReaFace bit per boundary-table slot");
                unsigned rea = 0u;
                for (int q = 0; q < ncbf; ++q) {
                    const BTab& B = bt[(size_t)q * P.cap + i];
                    const DFace& F = faces[B.face];
                    if (F.btype != 0 && F.btype != 1) continue;
                    const double zi = B.loc[2];
                    if (zi < zimin && d_point_internal(F, B.loc[0], B.loc[1])) {
                        double vin = 0.0;
                        for (int sd = 0; sd < 3; ++sd) vin = vin + v_i[sd] * F.T[6 + sd];
                        if (vin < 0.0) {
                            const double nr = -1.0 * celer02 * log((h + zi - zimin) / h) / h;
                            for (int d = 0; d < 3; ++d) br[d] = br[d] + nr * F.T[6 + d];
                        }
                        rea |= 1u << q;
                    }
                }
                // AddElasticBoundaryReaction_3D.f90:106-152 (convex edges, in
                // order; the reference returns at the first edge whose two
                // faces are not both among the particle's non-reacting faces)
                for (int ne = 0; ne < P.ne; ++ne) {
                    const DEdge& E = edges[P.eoff + ne];
                    int nclose = 0;
                    for (int q = 0; q < ncbf; ++q) {
                        const int f = bt[(size_t)q * P.cap + i].face;
                        const bool rq = (rea >> q) & 1u;
                        if (f == E.face[0] && !rq) ++nclose;
                        else if (f == E.face[1] && !rq) ++nclose;
                    }
                    if (nclose != 2) break;
                    const double4 xi = xm[i];
                    const double x[3] = {xi.x, xi.y, xi.z};
                    double scaprod = 0.0;
                    for (int sd = 0; sd < 3; ++sd) scaprod = scaprod + (x[sd] - E.gx[sd]) * E.comp[sd];
                    const double tau = scaprod / (E.len * E.len);
                    if (tau >= 0.0 && tau <= 1.0) {
                        double QP[3], d2 = 0.0;
                        for (int sd = 0; sd < 3; ++sd) QP[sd] = x[sd] - (E.gx[sd] + E.comp[sd] * tau);
                        for (int sd = 0; sd < 3; ++sd) d2 = d2 + QP[sd] * QP[sd];
                        const double zi = sqrt(d2);
                        if (zi < zimin) {
                            double cd[3], vin = 0.0;
                            for (int sd = 0; sd < 3; ++sd) cd[sd] = QP[sd] / zi;
                            for (int sd = 0; sd < 3; ++sd) vin = vin + v_i[sd] * cd[sd];
                            if (vin < 0.0) {
                                const double nr = -1.0 * celer02 * log(zi / zimin) / zimin;
                                for (int d = 0; d < 3; ++d) br[d] = br[d] + nr * cd[d];
                            }
                        }
                    }
                }
                for (int d = 0; d < 3; ++d) a[d] = tp[d] + td[d] + tv[d] + P.grav[d] + br[d];
            } else {
                for (int d = 0; d < 3; ++d) a[d] = tp[d] + td[d] + tv[d] + P.grav[d];
            }
            acc_out[i] = make_double4(a[0], a[1], a[2], 0.0);
            // Leapfrog_momentum (RK1: Euler): vel = vel + dtvel*acc (kodvel == 0)
            const double4 vn =
                make_double4(vi.x + dtvel * a[0], vi.y + dtvel * a[1], vi.z + dtvel * a[2], vi.w);
            vel_out[i] = vn;
            if (F32) velf_out[i] = make_float4((float)vn.x, (float)vn.y, (float)vn.z, (float)(rpi.x - P.den0));
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// velocity smoothing: velocity_smoothing.f90:74-133, velocity_smoothing_SA_SPH
// _3D.f90:64-122 (FP64), velocity_smoothing_2.f90:59-103; fused
// Leapfrog_trajectories.f90:48-76 and the particle part of the next
// time_step_duration.  Neighbour density travels in velf.w (FP32).  Explicit
// RK1: the same smoothing, update and positions (Euler.f90:94-139), then the
// density update and EoS of Euler.f90:231-241 and time_integration.f90:52-65.
// ----------------------------------------------------------------------------
template <bool F32, int CH, bool SG>
__global__ __launch_bounds__(256) void k_vsmooth(
    DP P0, Seg S, const DFace* __restrict__ faces,
    const int* __restrict__ bn, const BTab* __restrict__ bt, const double4* __restrict__ xm,
    const double4* __restrict__ vel1, const double2* __restrict__ rp, F4In xmf,
    F4In velf1, const uint32_t* __restrict__ nbr, const int* __restrict__ nnb,
    double4* __restrict__ vel_out, double4* __restrict__ xm_out,
    const double* scal, unsigned long long* dtmin, int* iscal, const double* __restrict__ dden_in,
    double2* __restrict__ rp_out, float2* __restrict__ rpf_out, float4* __restrict__ xrf_out,
    int* __restrict__ gone) {
    using X = typename Rec<F32>::X;
    using V = typename Rec<F32>::V;
    using RT = typename std::conditional<F32, float, double>::type;
    const int vb = xcd_block(gridDim.x);
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.pblk[(vb * (int)blockDim.x) >> 8]) : 0;
    const DP& P = SG ? S.P[sg] : P0;
    scal += sg * SC_COUNT;
    iscal += sg * IS_COUNT;
    const bool eos = iscal[IS_ONGOING] > 0;
    dtmin += sg;
    stencil_offsets(P);
    const size_t N = P.cap;
    const bool smooth = P.tetav > 1.0e-9;
    const double tv1 = scal[SC_TV1];
    const double dt = scal[SC_DT];
    unsigned long long dtbits = 0x7FF0000000000000ull;
    {
        const int i = vb * blockDim.x + threadIdx.x;
        if (i - P.base < iscal[IS_N]) {
            const double4 xi = xm[i];
            const double4 vi = vel1[i];
            double v[3] = {0.0, 0.0, 0.0};
            if (smooth) {
                const int nn = nnb[i];
                if constexpr (!F32) {
                    using RV = NRec<double4, double4, double>;
                    auto fetch = [&](uint32_t e) {
                        const int j = (int)(e & NB_MASK);
                            return RV{xm[j], vel1[j], rp[j].x};
                    };
                    walk_list<4, RV>(nbr, N, i, nn, fetch, [&](uint32_t e, const RV& q) {
                        const double4 xj = q.x;
                        const double4 vj = q.v;
                        const double rhoj = q.r;
                        const double R0 = xi.x - xj.x, R1 = xi.y - xj.y, R2 = xi.z - xj.z;
                        const double rij = R0 * R0 + R1 * R1 + R2 * R2;
                        K64 K;
                        kvals64(P, rij, false, false, true, K);
                        const double pes = xj.w * K.W / rhoj;
                        v[0] = v[0] + (vj.x - vi.x) * pes;
                        v[1] = v[1] + (vj.y - vi.y) * pes;
                        v[2] = v[2] + (vj.z - vi.z) * pes;
                                        });
                } else {
                                        const float4 xf = xmf[i];
                    const float4 vf = velf1[i];
                    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
                    using RV = NRec2<float4, float4>;
                    auto fetch = [&](uint32_t e) {
                        const int j = (int)(e & NB_MASK);
                            return RV{xmf[j], velf1[j]};
                    };
                    walk_list<CH, RV>(nbr, N, i, nn, fetch, [&](uint32_t e, const RV& q) {
                        const float4 xj = q.x;
                        const float4 vj = q.r;
                        float R0, R1, R2;
                        rag32(xf, xj, e, R0, R1, R2);
                        K32 K;
                        kvals32(P, R0 * R0 + R1 * R1 + R2 * R2, false, false, true, K);
                        const float pes = xj.w * K.W * __builtin_amdgcn_rcpf(P.f_den0 + vj.w);
                        s0 += (vj.x - vf.x) * pes;
                        s1 += (vj.y - vf.y) * pes;
                        s2 += (vj.z - vf.z) * pes;
                                        });
                    v[0] = s0; v[1] = s1; v[2] = s2;
                }
                const int ncbf = P.semi ? bn[i] : 0;
                // velocity_smoothing_SA_SPH_2D.f90:64-118 (fixe/tapi sides)
                for (int q = 0; q < ncbf && P.dim == 2; ++q) {
                    const BTab& B = bt[(size_t)q * P.cap + i];
                    const DFace& S = faces[B.face];
                    if (S.btype != 0 && S.btype != 1) continue;
                    const double s0 = S.T[0], s1 = S.T[2], n0 = S.T[6], n1 = S.T[8];
                    const double DVG0 = 2.0 * (S.tvel[0] - vi.x), DVG1 = 2.0 * (S.tvel[2] - vi.z);
                    const double DVL0 = s0 * DVG0 + s1 * DVG1;
                    const double DVL1 = n0 * DVG0 + n1 * DVG1;
                    const double IntWdV = B.bi[2];
                    double slip = S.mode == 1 ? S.slip : 0.0;
                    if (S.mode == 2) {
                        // velocity_smoothing_SA_SPH_2D.f90:93-106
                        const double u0 = (S.tvel[0] - vi.x) - (0.5 * DVL1 * S.T[6]);
                        const double u1 = (S.tvel[1] - vi.y) - (0.5 * DVL1 * S.T[7]);
                        const double u2 = (S.tvel[2] - vi.z) - (0.5 * DVL1 * S.T[8]);
                        double nut;
                        wall_function(sqrt(u0 * u0 + u1 * u1 + u2 * u2), S.slip * 10.0, B.loc[1], P.dx,
                                      slip, nut);
                    }
                    const double BCL0 = DVL0 * IntWdV * slip;
                    const double BCL1 = DVL1 * IntWdV;
                    v[0] = v[0] + (s0 * BCL0 + n0 * BCL1);
                    v[2] = v[2] + (s1 * BCL0 + n1 * BCL1);
                }
                for (int q = 0; q < ncbf && P.dim == 3; ++q) {
                    const BTab& B = bt[(size_t)q * P.cap + i];
                    const DFace& F = faces[B.face];
                    if (F.btype == 3 || F.btype == 4 || F.btype == 5) {
                        v[0] = v[1] = v[2] = 0.0;
                        break;
                    }
                    if (F.btype != 0 && F.btype != 1) continue;
                    const double DVG[3] = {2.0 * (F.tvel[0] - vi.x), 2.0 * (F.tvel[1] - vi.y),
                                           2.0 * (F.tvel[2] - vi.z)};
                    double DVL[3];
                    for (int ii = 0; ii < 3; ++ii) {
                        double a = 0.0;
                        for (int jj = 0; jj < 3; ++jj) a = a + F.T[3 * ii + jj] * DVG[jj];
                        DVL[ii] = a;
                    }
                    const double IntWdV = B.bi[1];
                    double slip = F.mode == 1 ? F.slip : 0.0;
                    if (F.mode == 2) {
                        // velocity_smoothing_SA_SPH_3D.f90:89-110
                        double ut[3], nut;
                        for (int d = 0; d < 3; ++d) ut[d] = 0.5 * (DVG[d] - DVL[2] * F.T[6 + d]);
                        wall_function(sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]), F.slip * 10.0,
                                      B.loc[2], P.dx, slip, nut);
                    }
                    const double BCL[3] = {DVL[0] * IntWdV * slip, DVL[1] * IntWdV * slip,
                                           DVL[2] * IntWdV};
                    for (int ii = 0; ii < 3; ++ii) {
                        double a = 0.0;
                        for (int jj = 0; jj < 3; ++jj) a = a + F.T[3 * jj + ii] * BCL[jj];
                        v[ii] = v[ii] + a;
                    }
                }
            }
            double4 vn = vi;
            if (smooth) vn = make_double4(vi.x + tv1 * v[0], vi.y + tv1 * v[1], vi.z + tv1 * v[2], vi.w);
            vel_out[i] = vn;

            const double xn[3] = {xi.x + dt * vn.x, xi.y + dt * vn.y, xi.z + dt * vn.z};
            xm_out[i] = make_double4(xn[0], xn[1], xn[2], xi.w);
            // the next time_step_duration runs over the particles the coming
            // OrdGrid1 keeps: not those out of the grid, nor those that left
            // through an open face
            bool stays = d_particle_cell(P, xn[0], xn[1], xn[2]) > 0;
            if (P.nopen > 0) {  // CancelOutgoneParticles_3D / _2D on the step just made
                const double xo[3] = {xi.x, xi.y, xi.z};
                int g = 0;
                for (int f = P.foff; f < P.foff + P.nf; ++f)
                    if (faces[f].btype == 2 &&
                        (P.dim == 2 ? open_side_crossed(faces[f], xo, xn) : open_face_crossed(faces[f], xo, xn)))
                        g = 1;
                gone[i] = g;
                stays = stays && !g;
            }
            if (P.rk1 && !P.dbsph) {
                // explicit RK1: Euler.f90:231-241 (dens = dens + dt dden) and
                // EoS_fluid_loop.f90:59-70 (time_integration.f90:52-65); the
                // pressure smoothing that follows reads these with the
                // stage's kernel values, hence the old cell-relative position
                const double2 rpi = rp[i];
                const double rho1 = rpi.x + dt * dden_in[i];
                const double p1 = eos ? P.eps / P.den0 * (rho1 - P.den0) + P.prif : rpi.y;
                rp_out[i] = make_double2(rho1, p1);
                if (F32) {
                    rpf_out[i] = make_float2((float)rho1, (float)p1);
                    const float4 xo = xmf[i];
                    xrf_out[i] = make_float4(xo.x, xo.y, xo.z, (float)(rho1 - P.den0));
                }
            }
            if (stays) {
                const unsigned long long b =
                    (unsigned long long)__double_as_longlong(d_dt_cfl(P, vn.x, vn.y, vn.z));
                dtbits = b < dtbits ? b : dtbits;
            }
        }
        __syncthreads();
    }
    block_min_to(dtbits, dtmin);
}

// ----------------------------------------------------------------------------
// continuity: SASPH_continuity.f90:89-169 with AddBoundaryContribution_to_CE3D
// .f90:82-169 (FP64), Continuity_Equation.f90:97-235; fused
// Leapfrog_continuity.f90:51-68 (with the density_thresholds clamp) and EoS_fluid_loop.f90:59-70
// ----------------------------------------------------------------------------
template <bool F32, int CH, bool SG>
__global__ __launch_bounds__(256) void k_continuity(
    DP P0, Seg S, const DFace* __restrict__ faces,
    const int* __restrict__ bn, const BTab* __restrict__ bt, const double4* __restrict__ xm,
    const double4* __restrict__ vel, const double2* __restrict__ rp, F4In xmf,
    F4In velf, const uint32_t* __restrict__ nbr, const int* __restrict__ nnb,
    double* __restrict__ dden_out, double2* __restrict__ rp_out, float2* __restrict__ rpf_out,
    float4* __restrict__ xrf_out,
    const double* scal, int* iscal) {
    using X = typename Rec<F32>::X;
    using V = typename Rec<F32>::V;
    using RT = typename std::conditional<F32, float, double>::type;
    const int vb = xcd_block(gridDim.x);
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.pblk[(vb * (int)blockDim.x) >> 8]) : 0;
    const DP& P = SG ? S.P[sg] : P0;
    scal += sg * SC_COUNT;
    iscal += sg * IS_COUNT;
    stencil_offsets(P);
    const size_t N = P.cap;
    const double dt = scal[SC_DT];
    const bool eos = iscal[IS_ONGOING] > 0;
    {
        const int i = vb * blockDim.x + threadIdx.x;
        if (i - P.base < iscal[IS_N]) {
            const double4 vi = vel[i];
            const double2 rpi = rp[i];
            const double rhoi = rpi.x;
            double dd = 0.0;
            const int ncbf = P.semi ? bn[i] : 0;
            if (ncbf > 0 && P.dim == 2) {
                // SASPH_continuity.f90:126-169 with AddBoundaryContribution_to_CE2D
                // .f90:74-150 (fixe/tapi sides): grad_u_SA(1) and grad_w_SA(3)
                double gu0 = 0.0, gw2 = 0.0;
                for (int q = 0; q < ncbf; ++q) {
                    const BTab& B = bt[(size_t)q * P.cap + i];
                    const DFace& S = faces[B.face];
                    if (S.btype != 0 && S.btype != 1) continue;
                    const double IntWdS = B.bi[0];
                    double dv0, dv1;
                    if (S.btype == 0) {
                        dv0 = 2.0 * (-vi.x);
                        dv1 = 2.0 * (-vi.z);
                    } else {
                        dv0 = 2.0 * (S.vel[0] - vi.x);
                        dv1 = 2.0 * (S.vel[2] - vi.z);
                    }
                    gu0 = gu0 - dv0 * S.T[6] * IntWdS;
                    gw2 = gw2 - dv1 * S.T[8] * IntWdS;
                }
                dd = dd - rhoi * (gu0 + 0.0 + gw2);
            } else if (ncbf > 0) {
                double gu = 0.0, gv = 0.0, gw = 0.0;
                const double v_i[3] = {vi.x, vi.y, vi.z};
                for (int q = 0; q < ncbf; ++q) {
                    const BTab& B = bt[(size_t)q * P.cap + i];
                    const DFace& F = faces[B.face];
                    if (F.btype == 0 || F.btype == 1) {
                        if (B.loc[2] > 0.0) {
                            double a2[3];
                            for (int d = 0; d < 3; ++d) a2[d] = 2.0 * (F.vel[d] - v_i[d]);
                            const double dp = a2[0] * F.T[6] + a2[1] * F.T[7] + a2[2] * F.T[8];
                            double av[3];
                            for (int r = 0; r < 3; ++r) {
                                double a = 0.0;
                                for (int k = 0; k < 3; ++k) a = a + F.T[3 * k + r] * B.bi[3 + k];
                                av[r] = a;
                            }
                            gu = gu - dp * F.T[6] * av[0];
                            gv = gv - dp * F.T[7] * av[1];
                            gw = gw - dp * F.T[8] * av[2];
                        }
                    } else if (F.btype >= 3) {
                        atomicAdd(&iscal[IS_UNSUPPORTED], 1);
                        break;
                    }
                }
                dd = dd - rhoi * (gu + gv + gw);
            }
            const int nn = nnb[i];
            double inner;
            if constexpr (!F32) {
                const double4 xi = xm[i];
                double a1 = 0.0, a2 = 0.0, a3 = 0.0;
                using RC = NRec<double4, double4, double>;
                auto fetch = [&](uint32_t e) {
                    const int j = (int)(e & NB_MASK);
                    return RC{xm[j], vel[j], rp[j].x};
                };
                walk_list<4, RC>(nbr, N, i, nn, fetch, [&](uint32_t e, const RC& q) {
                    const double4 xj = q.x;
                    const double4 vj = q.v;
                    const double rhoj = q.r;
                    const double R0 = xi.x - xj.x, R1 = xi.y - xj.y, R2 = xi.z - xj.z;
                    const double rij = R0 * R0 + R1 * R1 + R2 * R2;
                    K64 K;
                    kvals64(P, rij, true, false, false, K);
                    const double amassj = xj.w;
                    a1 = a1 + rhoi * (amassj / rhoj) * K.K1 * (vj.x - vi.x) * R0;
                    a2 = a2 + rhoi * (amassj / rhoj) * K.K1 * (vj.y - vi.y) * R1;
                    a3 = a3 + rhoi * (amassj / rhoj) * K.K1 * (vj.z - vi.z) * R2;
                                });
                inner = a1 + a2 + a3;
            } else {
                                const float4 xf = xmf[i];
                const float4 vf = velf[i];
                float s = 0.f;
                using RC = NRec2<float4, float4>;
                auto fetch = [&](uint32_t e) {
                    const int j = (int)(e & NB_MASK);
                    return RC{xmf[j], velf[j]};
                };
                walk_list<CH, RC>(nbr, N, i, nn, fetch, [&](uint32_t e, const RC& q) {
                    const float4 xj = q.x;
                    const float4 vj = q.r;
                    float R0, R1, R2;
                    rag32(xf, xj, e, R0, R1, R2);
                    K32 K;
                    kvals32(P, R0 * R0 + R1 * R1 + R2 * R2, true, false, false, K);
                    s += xj.w * __builtin_amdgcn_rcpf(P.f_den0 + vj.w) * K.K1 *
                         ((vj.x - vf.x) * R0 + (vj.y - vf.y) * R1 + (vj.z - vf.z) * R2);
                                });
                inner = rhoi * (double)s;
            }
            const double dden = dd - inner;
            dden_out[i] = dden;
            // RK1: the density is integrated after the velocities and positions
            // (Euler.f90:231-241, in k_vsmooth)
            if (!P.rk1) {
                double rho1 = rhoi + dt * dden;
                if (P.dthr) {  // Leapfrog_continuity.f90:61-68
                    if (rho1 < 0.9 * P.den0) rho1 = 0.9 * P.den0;
                    if (rho1 > 1.1 * P.den0) rho1 = 1.1 * P.den0;
                }
                const double p1 = eos ? P.eps / P.den0 * (rho1 - P.den0) + P.prif : rpi.y;
                rp_out[i] = make_double2(rho1, p1);
                if (F32) {
                    rpf_out[i] = make_float2((float)rho1, (float)p1);
                    // (cell-relative position, rho - den0) for the pressure smoothing
                    const float4 xo = xmf[i];
                    xrf_out[i] = make_float4(xo.x, xo.y, xo.z, (float)(rho1 - P.den0));
                }
            }
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// pressure smoothing: PressureSmoothing_3D.f90:80-161 (both loops)
// ----------------------------------------------------------------------------
template <bool F32, int CH, bool SG>
__global__ __launch_bounds__(256) void k_psmooth(
    DP P0, Seg S, const DFace* __restrict__ faces, const int* __restrict__ bn, const BTab* __restrict__ bt,
    const double4* __restrict__ xm, const double2* __restrict__ rp, const float4* __restrict__ xmf,
    const float2* __restrict__ rpf, const float4* __restrict__ xrf, const uint32_t* __restrict__ nbr,
    const int* __restrict__ nnb,
    double* __restrict__ vpres_out, double2* __restrict__ rp_out, float2* __restrict__ rpf_out,
    float* __restrict__ velw_out,
    const double* scal, const int* iscal) {
    using X = typename Rec<F32>::X;
    using RPT = typename Rec<F32>::RP;
    const int vb = xcd_block(gridDim.x);
    const int sg = SG ? __builtin_amdgcn_readfirstlane(S.pblk[(vb * (int)blockDim.x) >> 8]) : 0;
    const DP& P = SG ? S.P[sg] : P0;
    scal += sg * SC_COUNT;
    iscal += sg * IS_COUNT;
    stencil_offsets(P);
    const size_t N = P.cap;
    const double tp1 = scal[SC_TP1];
    {
        const int i = vb * blockDim.x + threadIdx.x;
        if (i - P.base < iscal[IS_N]) {
            const double2 rpi = rp[i];
            const int nsp = nnb[i];
            double diffp = 0.0;
            if (nsp > 0) {
                const double p0i = P.prif, pi = rpi.y;
                double somp, app;
                if constexpr (!F32) {
                    const double4 xi = xm[i];
                    somp = 0.0;
                    app = 0.0;
                    using RP2 = NRec2<double4, double2>;
                    auto fetch = [&](uint32_t e) {
                        const int j = (int)(e & NB_MASK);
                            return RP2{xm[j], rp[j]};
                    };
                    walk_list<4, RP2>(nbr, N, i, nsp, fetch, [&](uint32_t e, const RP2& q) {
                        const double4 xj = q.x;
                        const double2 rpj = q.r;
                        const double R0 = xi.x - xj.x, R1 = xi.y - xj.y, R2 = xi.z - xj.z;
                        K64 K;
                        kvals64(P, R0 * R0 + R1 * R1 + R2 * R2, false, false, true, K);
                        const double pes = xj.w * K.W / rpj.x;
                        app = app + pes;
                        somp = somp + (rpj.y - pi) * pes;
                                        });
                } else {
                    const float4 xf = xmf[i];
                    const float pif = rpf[i].y;
                    float s_a = 0.f, s_p = 0.f;
                    if (P.umass) {
                        // one (x, y, z, rho - den0) record per neighbour: m_j is
                        // the solve's mass, p_j - p_i the linear EoS of the
                        // density difference
                        const float wi = xrf[i].w;
                        auto fetch = [&](uint32_t e) { return xrf[(int)(e & NB_MASK)]; };
                        walk_list<CH, float4>(nbr, N, i, nsp, fetch, [&](uint32_t e, const float4& xj) {
                            float R0, R1, R2;
                            rag32(xf, xj, e, R0, R1, R2);
                            K32 K;
                            kvals32(P, R0 * R0 + R1 * R1 + R2 * R2, false, false, true, K);
                            const float pes = P.f_mass * K.W * __builtin_amdgcn_rcpf(P.f_den0 + xj.w);
                            s_a += pes;
                            s_p += P.f_eosk * (xj.w - wi) * pes;
                        });
                    } else {
                        using RP2 = NRec2<float4, float2>;
                        auto fetch = [&](uint32_t e) {
                            const int j = (int)(e & NB_MASK);
                            return RP2{xmf[j], rpf[j]};
                        };
                        walk_list<CH, RP2>(nbr, N, i, nsp, fetch, [&](uint32_t e, const RP2& q) {
                            const float4 xj = q.x;
                            const float2 rpj = q.r;
                            float R0, R1, R2;
                            rag32(xf, xj, e, R0, R1, R2);
                            K32 K;
                            kvals32(P, R0 * R0 + R1 * R1 + R2 * R2, false, false, true, K);
                            const float pes = xj.w * K.W * __builtin_amdgcn_rcpf(rpj.x);
                            s_a += pes;
                            s_p += (rpj.y - pif) * pes;
                        });
                    }
                    somp = s_p;
                    app = s_a;
                }
                const int ncbf = P.semi ? bn[i] : 0;
                if (P.dim == 2) {
                    // PressureSmoothing_2D.f90:124-214, SmoothVersion "b", fixe/tapi
                    // sides; vpres is the smoothed pressure itself
                    double VFT = 0.0;
                    for (int q = 0; q < ncbf; ++q) {
                        const BTab& B = bt[(size_t)q * P.cap + i];
                        const DFace& S = faces[B.face];
                        if (S.btype != 0 && S.btype != 1) continue;
                        const double mf0 = S.T[0] * P.grav[0] + S.T[2] * P.grav[2];
                        const double mf1 = S.T[6] * P.grav[0] + S.T[8] * P.grav[2];
                        VFT = VFT + B.bi[1];
                        somp = somp + P.den0 * (mf0 * B.bi[3] + mf1 * B.bi[4]);
                    }
                    app = app + VFT + 0.0 + 0.0 + 0.0;
                    const double pz = (pi - pi) * 0.0;
                    diffp = pi + tp1 * (somp + pz + pz + (p0i - pi) * (1.0 - app));
                } else {
                    for (int q = 0; q < ncbf; ++q) {
                        const BTab& B = bt[(size_t)q * P.cap + i];
                        if (B.loc[2] > 0.0) app = app + B.bi[1];
                    }
                    if (P.dbsph) diffp = somp / app;  // PressureSmoothing_3D.f90:110-112
                    else if (app < 0.95) diffp = somp + (p0i - pi) * (1.0 - app);
                    else diffp = somp / app;
                }
            } else if (P.dim == 2) {
                diffp = rpi.y;  // no list (never: self is listed): pressure unchanged
            }
            vpres_out[i] = diffp;
            double p2 = rpi.y, rho2 = rpi.x;
            if (P.dim == 2) {  // PressureSmoothing_2D.f90:217-231: p = vpres, rho = EoS^-1(p)
                if (P.tetap > 0.0) {
                    p2 = diffp;
                    rho2 = P.den0 * (1.0 + (p2 - P.prif) / P.eps);
                }
            } else if (P.tetap > 0.0) {
                p2 = rpi.y + tp1 * diffp;
                rho2 = P.den0 * (1.0 + (p2 - P.prif) / P.eps);
            }
            rp_out[i] = make_double2(rho2, p2);
            if (F32) {
                rpf_out[i] = make_float2((float)rho2, (float)p2);
                // the density slot of the velocity record the next momentum
                // pass reads (element 8 i: 32-byte records, w of the second half)
                velw_out[8 * (size_t)i] = (float)(rho2 - P.den0);
            }
        }
        __syncthreads();
    }
}

}  // namespace sph
