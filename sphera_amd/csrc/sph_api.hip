// Host orchestration and C ABI (include/sphera_hip.h) of the device hot path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/sphera_hip.h"
#include "sph_kernels.hip"
#include "sph_dbsph.hip"

using namespace sph;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPC(expr)                                                                  \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return fail(100 + (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

enum Phase {
    PH_EMPTY = 0, PH_SORTED,  // state uploaded and ordered
    PH_READY,                 // lists, boundary table, dt available (step boundary)
    PH_DT, PH_MOM, PH_VS, PH_SORT, PH_NB, PH_BT, PH_CONT, PH_PS
};

struct Timer {
    std::string name;
    double ms = 0.0;
    int calls = 0;
};

}  // namespace

struct sph_ctx {
    DP P{};  // a one-solve context's constants; segment 0's (with the context capacity) otherwise
    // segmented context: nseg solves in one set of arrays (see Seg)
    int nseg = 1;
    std::vector<DP> segs;
    Seg S{};
    DP* d_P = nullptr;
    int *d_pblk = nullptr, *d_kblk = nullptr, *d_rseg = nullptr, *d_adj = nullptr;
    int nkeys = 1, nrows = 1;  // cell keys (cells + out-of-grid bucket per segment), cell rows
    int nb_stride_max = 1;     // largest NMAXPARTJ of the segments (list capacity)
    std::vector<int> staged;   // segments uploaded since the last commit
    int precision = SPH_FP64;
    hipStream_t stream = nullptr;
    // faces / tables
    DFace* d_faces = nullptr;
    DEdge* d_edges = nullptr;  // convex frontier edges of every segment
    // per cell key: the frontier faces listed in its 27-cell neighbourhood,
    // each once (count, start into d_gvec)
    int *d_gcount = nullptr, *d_gstart = nullptr, *d_gvec = nullptr;
    double *d_kt = nullptr, *d_bit = nullptr;
    BtF32* d_btf = nullptr;  // the quadrature table in the FP32-mixed layout
    // state
    double4 *xm[2] = {}, *vel[3] = {}, *acc[2] = {};
    double2* rp[3] = {};
    int* lam[2] = {};
    float4* xmf = nullptr;  // FP32 cell-relative positions (x, y, z, mass)
    float4* xrf = nullptr;  // FP32 (x, y, z, rho - den0) after continuity, for the pressure smoothing
    // FP32 record buffers, (position, velocity) float4 pairs per particle:
    // [0] read by momentum and continuity, [1] by velocity smoothing
    float4* rec[2] = {};
    float2* rpf[3] = {};
    double *dden = nullptr, *vpres = nullptr;
    // explicit RK1: dden and vpres in the order after the step's closing sort
    double *dden2 = nullptr, *vpres2 = nullptr;
    int* gone = nullptr;  // 1: the last trajectory step left through an open face
    int *cell = nullptr, *key = nullptr, *slot = nullptr, *cnt = nullptr, *start = nullptr,
        *perm = nullptr, *pid = nullptr, *bsum = nullptr;
    int *work = nullptr, *nwork = nullptr;  // (particle, face) pairs of the boundary table
    int* rowblk = nullptr;                  // first neighbour-search block of each cell row
    int *idslot = nullptr, *rmids = nullptr;  // id -> slot, ids removed by a sort
    int nb_grid = 8;                        // persistent workgroups of the neighbour search
    int nb_shape = 0;                       // its launch shape (index into NB_SHAPES)
    int wg = 256, grid = 1, bt_grid = 1024;  // launch shape of the particle kernels
    int chunk = 8;                            // list entries gathered per step of the FP32 walks
    int chunk_v = 8, chunk_c = 8, chunk_p = 8;  // the same for smoothing, continuity, pressure smoothing
    std::vector<std::pair<int, hipGraphExec_t>> graphs;  // captured n-step sequences
    uint32_t* nbr = nullptr;
    int* nnb = nullptr;
    int* bn = nullptr;
    BTab* bt = nullptr;
    double* scal = nullptr;
    int* iscal = nullptr;
    unsigned long long* dtmin = nullptr;
    unsigned long long* pairs = nullptr;  // running pair-interaction count
    // DB-SPH walls: the wall-element tables (w0, w1, w2, cell lists), their
    // density, pressure and wet flags; per particle Gamma and dShep (two
    // buffers each: the sort gathers [0] into [1]), sigma, rhoSPH_new, FS,
    // the fluid-wall lists and the MUSCL gradients
    double4 *d_w0 = nullptr, *d_w1 = nullptr, *d_w2 = nullptr;
    int *d_wstart = nullptr, *d_wids = nullptr, *d_wcell = nullptr;
    double4* d_wpos = nullptr;
    double* d_kin = nullptr;  // wall kinematics records (t, v, omega)
    double *d_wdens = nullptr, *d_wpres = nullptr;
    int* d_wwet = nullptr;
    double *gam[2] = {}, *dsh[2] = {}, *sig = nullptr, *rsph = nullptr, *drho = nullptr, *dvelg = nullptr;
    int *fsf = nullptr, *nfw = nullptr, *fwl = nullptr;
    // monitoring lines of the free-surface monitor
    int mon_nlines = 0, mon_npoints = 0;
    int* mon_icont = nullptr;
    double *mon_pts = nullptr, *mon_out = nullptr;
    Phase phase = PH_EMPTY;
    bool timing = false;
    std::vector<Timer> timers;
    std::vector<hipEvent_t> ev;
    int n_upload = 0;
    int device = 0;  // the GPU current at creation (its stream's device)

    ~sph_ctx();
    int alloc();
    // device bytes held by the context; with dry set, alloc() only counts them
    size_t bytes = 0;
    bool dry = false;
    template <class T>
    hipError_t da(T** p, size_t n) {
        bytes += sizeof(T) * (n ? n : 1);
        return dry ? hipSuccess : hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
    }
};

namespace {

template <class T>
hipError_t dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * (n ? n : 1));
}

__global__ void k_fill_int(int* p, int v, int m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = v;
}

inline int nblk(int n, int b = 256) { return (n + b - 1) / b; }

struct KTime {
    sph_ctx* c;
    int idx;
    KTime(sph_ctx* cc, const char* name) : c(cc), idx(-1) {
        if (!c->timing) return;
        for (size_t t = 0; t < c->timers.size(); ++t)
            if (c->timers[t].name == name) idx = (int)t;
        if (idx < 0) {
            c->timers.push_back(Timer{name});
            idx = (int)c->timers.size() - 1;
        }
        hipEvent_t a, b;
        hipEventCreate(&a);
        hipEventCreate(&b);
        c->ev.push_back(a);
        c->ev.push_back(b);
        hipEventRecord(a, c->stream);
    }
    ~KTime() {
        if (!c->timing || idx < 0) return;
        hipEventRecord(c->ev.back(), c->stream);
        evidx.push_back(idx);
    }
    static thread_local std::vector<int> evidx;
};
thread_local std::vector<int> KTime::evidx;

int collect_timing(sph_ctx* c) {
    if (!c->timing) return 0;
    HIPC(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < KTime::evidx.size(); ++k) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]);
        c->timers[KTime::evidx[k]].ms += ms;
        c->timers[KTime::evidx[k]].calls += 1;
    }
    for (auto e : c->ev) hipEventDestroy(e);
    c->ev.clear();
    KTime::evidx.clear();
    return 0;
}

// ---- device steps --------------------------------------------------------
// position / velocity halves of FP32 record buffer k
inline F4In XF(const sph_ctx* c, int k) { return F4In{c->rec[k]}; }
inline F4In VF(const sph_ctx* c, int k) { return F4In{c->rec[k] + 1}; }
inline F4Out XFo(const sph_ctx* c, int k) { return F4Out{c->rec[k]}; }
inline F4Out VFo(const sph_ctx* c, int k) { return F4Out{c->rec[k] + 1}; }
void scan_ints(sph_ctx* c, const int* in, int* out, int* bs, int m, int* total) {
    const int nb = (m + 1023) / 1024;
    k_scan_local<<<nb, 256, 0, c->stream>>>(in, out, bs, m);
    k_scan_sums<<<1, 1024, 0, c->stream>>>(bs, nb, total);
    k_scan_add<<<nblk(m), 256, 0, c->stream>>>(out, bs, m);
}

// Launches kernel K in its one-solve (<false>) or segmented (<true>) form.
#define SEGL(K, G, B, ...)                                                                 \
    do {                                                                                   \
        if (c->nseg > 1) K<true><<<(G), (B), 0, c->stream>>>(c->P, c->S, __VA_ARGS__);      \
        else K<false><<<(G), (B), 0, c->stream>>>(c->P, c->S, __VA_ARGS__);                 \
    } while (0)

__global__ void k_fill_strided(int* p, int stride, int v, int m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[(size_t)i * stride] = v;
}

// Buffer of the current (rho, p) at a step boundary: the pressure smoothing
// writes rp[0]; the leapfrog step's sort comes before it, the RK1 step's sort
// after it (into rp[1]).
inline int cur_rp(const sph_ctx* c) { return c->P.dbsph ? 0 : (c->P.rk1 ? 1 : 0); }

// cell sort + block descriptors (OrdGrid1)
int run_sort(sph_ctx* c) {
    const int m = c->nkeys;
    const int cap = c->P.cap;
    const int ns = c->nseg;
    {
        KTime t(c, "sort_count");
        k_fill_int<<<nblk(m), 256, 0, c->stream>>>(c->cnt, 0, m);
        k_fill_strided<<<nblk(ns, 64), 64, 0, c->stream>>>(c->iscal + IS_NRM, IS_COUNT, 0, ns);
        SEGL(k_cell_count, nblk(cap), 256, c->xm[1], c->vel[2], c->key, c->slot, c->cnt, c->idslot, c->rmids,
             c->iscal, c->gone);
        SEGL(k_renumber, nblk(ns, 64), 64, c->rmids, c->idslot, c->vel[2], c->iscal);
    }
    {
        KTime t(c, "sort_scan");
        scan_ints(c, c->cnt, c->start, c->bsum, m, nullptr);
        if (ns > 1) {  // each segment's slots start at its base
            k_seg_adj<<<nblk(ns, 64), 64, 0, c->stream>>>(c->S, c->start, c->d_adj);
            k_seg_add<<<nblk(m), 256, 0, c->stream>>>(c->S, c->start, c->d_adj, m);
        }
    }
    {
        KTime t(c, "sort_scatter");
        SEGL(k_scatter, nblk(cap), 256, c->key, c->slot, c->start, c->vel[2], c->perm, c->pid, c->iscal);
        SEGL(k_cell_sort, ns > 1 ? m / 256 : (c->P.nmax + 255) / 256, 256, c->start, c->perm, c->pid);
    }
    {
        KTime t(c, "sort_gather");
        SEGL(k_gather, nblk(cap), 256, c->perm, c->key, c->xm[1], c->vel[2], c->rp[0], c->lam[0], c->acc[1],
             c->xm[0], c->vel[0], c->rp[1], c->lam[1], c->acc[0], c->cell, c->xmf, XFo(c, 0), XFo(c, 1),
             VFo(c, 0), c->rpf[1], c->start, c->iscal, c->P.rk1 ? c->dden : nullptr, c->dden2,
             c->P.rk1 ? c->vpres : nullptr, c->vpres2, c->P.dbsph ? c->gam[0] : nullptr, c->gam[1],
             c->dsh[0], c->dsh[1]);
    }
    return 0;
}

// the tiled search in launch shape `shape` of NB_SHAPES
#define NB_KERNEL(SG, k) \
    k_neighbours<SG, NB_SHAPES[k].wg, NB_SHAPES[k].tile, NB_SHAPES[k].tk, NB_SHAPES[k].minb>

template <bool SG>
void launch_nb(sph_ctx* c) {
    const int wg = NB_SHAPES[c->nb_shape].wg;
    auto go = [&](auto kern) {
        kern<<<c->nb_grid, wg, 0, c->stream>>>(c->P, c->S, c->xm[0], c->xmf, c->cell, c->start, c->rowblk,
                                               c->nbr, c->nnb, c->iscal, c->pairs);
    };
    switch (c->nb_shape) {
        case 1: go(NB_KERNEL(SG, 1)); break;
        case 2: go(NB_KERNEL(SG, 2)); break;
        case 3: go(NB_KERNEL(SG, 3)); break;
        case 4: go(NB_KERNEL(SG, 4)); break;
        default: go(NB_KERNEL(SG, 0)); break;
    }
}

int run_neighbours(sph_ctx* c) {
    KTime t(c, "neighbours");
    SEGL(k_nb_rowblocks, 1, 1024, c->start, c->rowblk, NB_SHAPES[c->nb_shape].wg);
    if (c->nseg > 1) launch_nb<true>(c);
    else launch_nb<false>(c);
    return 0;
}

int run_btab(sph_ctx* c) {
    if (!c->P.semi || c->P.nf == 0) {
        k_fill_int<<<nblk(c->P.cap), 256, 0, c->stream>>>(c->bn, 0, c->P.cap);
        return 0;
    }
    if (c->P.dim == 2) {  // ComputeBoundaryDataTab_2D
        KTime t(c, "boundary_2d");
        k_boundary_2d<<<nblk(c->P.cap), 256, 0, c->stream>>>(c->P, c->d_faces, c->d_bit, c->xm[0], c->cell,
                                                            c->bn, c->bt, c->iscal);
        return 0;
    }
    {
        KTime t(c, "boundary_find");
        k_fill_int<<<1, 64, 0, c->stream>>>(c->nwork, 0, 1);
        SEGL(k_boundary_find, nblk(c->P.cap), 256, c->d_faces, c->d_gcount, c->d_gstart, c->d_gvec,
             c->xm[0], c->bn, c->bt, c->work, c->nwork, c->iscal);
    }
    KTime t(c, "boundary_int");
    if (c->P.f32)  // eight lanes per pair
        SEGL(k_boundary_int_f32, 4 * c->bt_grid, 256, c->d_faces, c->d_btf, c->bt, c->work, c->nwork);
    else
        SEGL(k_boundary_int, c->bt_grid, 256, c->d_faces, c->d_kt, c->d_bit, c->bt, c->work, c->nwork);
    return 0;
}

int run_dt(sph_ctx* c, int advance, bool reduce) {
    KTime t(c, "dt");
    if (reduce) SEGL(k_dtmin, nblk(c->P.cap), 256, c->vel[0], c->dtmin, c->iscal);
    SEGL(k_dt_begin, nblk(c->nseg, 64), 64, c->scal, c->dtmin, c->iscal, advance);
    return 0;
}

// The pair passes: FP32-mixed with the list-walk chunk of the pass (4 or 8;
// segmented contexts use 8), or FP64; one-solve or segmented form.
#define PAIRL(K, CHV, ...)                                                                          \
    do {                                                                                            \
        if (c->P.f32) {                                                                             \
            if (c->nseg > 1) K<true, 8, true><<<c->grid, c->wg, 0, c->stream>>>(c->P, c->S, __VA_ARGS__); \
            else if ((CHV) == 4) K<true, 4, false><<<c->grid, c->wg, 0, c->stream>>>(c->P, c->S, __VA_ARGS__); \
            else K<true, 8, false><<<c->grid, c->wg, 0, c->stream>>>(c->P, c->S, __VA_ARGS__);        \
        } else {                                                                                    \
            if (c->nseg > 1) K<false, 4, true><<<c->grid, c->wg, 0, c->stream>>>(c->P, c->S, __VA_ARGS__); \
            else K<false, 4, false><<<c->grid, c->wg, 0, c->stream>>>(c->P, c->S, __VA_ARGS__);      \
        }                                                                                           \
    } while (0)

int run_momentum(sph_ctx* c) {
    KTime t(c, "momentum");
    PAIRL(k_momentum, c->chunk, c->d_faces, c->d_edges, c->bn, c->bt, c->xm[0], c->vel[0], c->rp[cur_rp(c)], XF(c, 0),
          VF(c, 0), c->rpf[cur_rp(c)], c->lam[1], c->lam[0], c->nbr, c->nnb, c->acc[1], c->vel[1], VFo(c, 1), c->scal, c->iscal);
    return 0;
}

int run_vsmooth(sph_ctx* c) {
    KTime t(c, "velocity_smoothing");
    PAIRL(k_vsmooth, c->chunk_v, c->d_faces, c->bn, c->bt, c->xm[0], c->vel[1], c->rp[cur_rp(c)], XF(c, 1),
          VF(c, 1), c->nbr, c->nnb, c->vel[2], c->xm[1], c->scal, c->dtmin, c->iscal, c->dden, c->rp[2],
          c->rpf[2], c->xrf, c->gone);
    return 0;
}

int run_continuity(sph_ctx* c) {
    KTime t(c, "continuity");
    PAIRL(k_continuity, c->chunk_c, c->d_faces, c->bn, c->bt, c->xm[0], c->vel[0],
          c->P.dbsph ? c->rp[0] : c->rp[1], XF(c, 0),
          VF(c, 0), c->nbr, c->nnb, c->dden, c->rp[2], c->rpf[2], c->xrf, c->scal, c->iscal);
    return 0;
}

int run_psmooth(sph_ctx* c) {
    KTime t(c, "pressure_smoothing");
    // w of the velocity half of FP32 record buffer 0 (read by the next momentum pass)
    float* velw = c->P.f32 ? reinterpret_cast<float*>(c->rec[0] + 1) + 3 : nullptr;
    PAIRL(k_psmooth, c->chunk_p, c->d_faces, c->bn, c->bt, c->xm[0], c->rp[2], c->xmf, c->rpf[2], c->xrf,
          c->nbr, c->nnb, c->vpres, c->rp[0], c->rpf[0], velw, c->scal, c->iscal);
    return 0;
}

// DB-SPH walls (one FP64 solve): CalcVarLength's DB-SPH sums in mode 0/1
// (pre-loop) or 2 (Euler, after the sort: the state in rp[1], the new density
// and pressure into rp[2])
DWall dwall(const sph_ctx* c) { return DWall{c->d_w0, c->d_w1, c->d_w2, c->d_wstart, c->d_wids, c->d_wcell, c->d_wpos}; }

int run_dbsph_lists(sph_ctx* c, int mode) {
    KTime t(c, "dbsph_lists");
    const int cap = c->P.cap;
    if (mode == 0)
        k_dbsph_lists<0><<<nblk(DB_G * cap), 256, 0, c->stream>>>(c->P, dwall(c), c->xm[0], c->vel[0], c->rp[0], c->cell,
                                                           c->nbr, c->nnb, c->gam[0], c->dsh[0], c->gam[0], c->dsh[0],
                                                           c->sig, c->rsph, c->fsf, c->nfw, c->fwl, c->d_wwet,
                                                           c->dden2, c->rp[2], c->scal, c->iscal);
    else if (mode == 1)
        k_dbsph_lists<1><<<nblk(DB_G * cap), 256, 0, c->stream>>>(c->P, dwall(c), c->xm[0], c->vel[0], c->rp[0], c->cell,
                                                           c->nbr, c->nnb, c->gam[0], c->dsh[0], c->gam[0], c->dsh[0],
                                                           c->sig, c->rsph, c->fsf, c->nfw, c->fwl, c->d_wwet,
                                                           c->dden2, c->rp[2], c->scal, c->iscal);
    else
        k_dbsph_lists<2><<<nblk(DB_G * cap), 256, 0, c->stream>>>(c->P, dwall(c), c->xm[0], c->vel[0], c->rp[1], c->cell,
                                                           c->nbr, c->nnb, c->gam[1], c->dsh[1], c->gam[0], c->dsh[0],
                                                           c->sig, c->rsph, c->fsf, c->nfw, c->fwl, c->d_wwet,
                                                           c->dden2, c->rp[2], c->scal, c->iscal);
    return 0;
}

int run_momentum_dbsph(sph_ctx* c) {
    KTime t(c, "momentum");
    k_momentum_dbsph<<<nblk(DB_G * c->P.cap), 256, 0, c->stream>>>(c->P, dwall(c), c->xm[0], c->vel[0], c->rp[0], c->nbr,
                                                            c->nnb, c->nfw, c->fwl, c->gam[0], c->d_wdens,
                                                            c->d_wpres, c->lam[0], c->acc[1], c->vel[1], c->scal,
                                                            c->iscal);
    return 0;
}

// PressureSmoothing_3D.f90 for bsph on the new lists: rp[2] -> rp[0]
int run_psmooth_dbsph(sph_ctx* c) {
    KTime t(c, "pressure_smoothing");
    PAIRL(k_psmooth, c->chunk_p, c->d_faces, c->bn, c->bt, c->xm[0], c->rp[2], c->xmf, c->rpf[2], c->xrf,
          c->nbr, c->nnb, c->vpres2, c->rp[0], c->rpf[0], nullptr, c->scal, c->iscal);
    return 0;
}

// the wall elements' velocity and position in Euler, after the particles'
int run_dbsph_kinematics(sph_ctx* c) {
    KTime t(c, "dbsph_kinematics");
    k_dbsph_kinematics<<<nblk(c->P.nw), 256, 0, c->stream>>>(c->P, c->d_kin, c->d_wids, c->d_w0, c->d_w2,
                                                             c->d_wpos, c->scal);
    return 0;
}

// Gradients_to_MUSCL and BC_wall_elements on the step's final state
int run_dbsph_walls(sph_ctx* c) {
    KTime t(c, "dbsph_walls");
    k_dbsph_muscl<<<nblk(c->P.cap), 256, 0, c->stream>>>(c->P, dwall(c), c->xm[0], c->vel[0], c->rp[0], c->nbr,
                                                         c->nnb, c->nfw, c->fwl, c->sig, c->d_wdens, c->drho,
                                                         c->dvelg, c->iscal);
    k_dbsph_walls<<<(c->P.nw + DB_WPB - 1) / DB_WPB, 256, 0, c->stream>>>(c->P, dwall(c), c->xm[0], c->vel[0], c->rp[0], c->start,
                                                        c->drho, c->dvelg, c->d_wdens, c->d_wpres, c->iscal);
    return 0;
}

int check_flags(sph_ctx* c) {
    std::vector<int> all((size_t)IS_COUNT * c->nseg);
    HIPC(hipMemcpyAsync(all.data(), c->iscal, sizeof(int) * all.size(), hipMemcpyDeviceToHost, c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    int is[IS_COUNT] = {0};
    for (int s = 0; s < c->nseg; ++s)
        for (int k : {(int)IS_UNSUPPORTED, (int)IS_RM_OVERFLOW, (int)IS_BT_OVERFLOW, (int)IS_DB_OVERFLOW})
            is[k] |= all[(size_t)s * IS_COUNT + k];
    if (is[IS_UNSUPPORTED])
        return fail(95, "particle close to a boundary type the device path does not implement "
                        "(sour/velo/flow)");
    if (is[IS_RM_OVERFLOW])
        return fail(102, "more than " + std::to_string(RM_CAP) +
                             " particles left the grid in one step (renumbering list)");
    if (is[IS_BT_OVERFLOW])
        return fail(86, "a particle has more close boundary faces than the device table holds (" +
                            std::to_string(SPH_MB) + ")");
    if (is[IS_DB_OVERFLOW])
        return fail(87, "a DB-SPH list is full: more than " + std::to_string(SPH_MW) +
                            " wall elements within 2h of a particle, " + std::to_string(SPH_WF) +
                            " particles within 2h of a wall element, or NMAXPARTJ fluid neighbours");
    return 0;
}

int expect(sph_ctx* c, Phase need, const char* what) {
    if (c->P.rk1)
        return fail(2, std::string(what) + ": the phase entry points follow the leapfrog order; an "
                                           "explicit RK1 context steps with sph_step / sph_step_graph");
    if (c->phase != need)
        return fail(120, std::string(what) + " called out of the time_step_loop order");
    return 0;
}

}  // namespace

sph_ctx::~sph_ctx() {
    auto F = [](void* p) {
        if (p) hipFree(p);
    };
    F(d_faces); F(d_edges); F(d_gcount); F(d_gstart); F(d_gvec); F(d_kt); F(d_bit); F(d_btf);
    F(d_P); F(d_pblk); F(d_kblk); F(d_rseg); F(d_adj);
    for (auto p : xm) F(p);
    for (auto p : vel) F(p);
    for (auto p : acc) F(p);
    for (auto p : rp) F(p);
    for (auto p : lam) F(p);
    F(xmf); F(xrf);
    for (auto p : rec) F(p);
    for (auto p : rpf) F(p);
    F(dden); F(vpres); F(dden2); F(vpres2); F(gone); F(rowblk); F(idslot); F(rmids); F(cell); F(key); F(slot); F(cnt); F(start); F(perm); F(pid); F(bsum);
    F(work); F(nwork);
    F(nbr); F(nnb); F(bn); F(bt); F(scal); F(iscal); F(dtmin); F(pairs);
    F(mon_icont); F(mon_pts); F(mon_out);
    F(d_w0); F(d_w1); F(d_w2); F(d_wstart); F(d_wids); F(d_wcell); F(d_wpos); F(d_kin); F(d_wdens); F(d_wpres); F(d_wwet);
    for (auto p : gam) F(p);
    for (auto p : dsh) F(p);
    F(sig); F(rsph); F(drho); F(dvelg); F(fsf); F(nfw); F(fwl);
    for (auto e : ev) hipEventDestroy(e);
    for (auto& g : graphs) hipGraphExecDestroy(g.second);
    if (stream) hipStreamDestroy(stream);
}

int sph_ctx::alloc() {
    const size_t N = P.cap;
    const size_t m = nkeys;
    const size_t ns = nseg;
    for (auto& p : xm) HIPC(da(&p, N));
    for (auto& p : vel) HIPC(da(&p, N));
    for (auto& p : acc) HIPC(da(&p, N));
    for (auto& p : rp) HIPC(da(&p, N));
    for (auto& p : lam) HIPC(da(&p, N));
    HIPC(da(&xmf, N));
    HIPC(da(&xrf, P.f32 ? N : 1));
    for (auto& r : rec) HIPC(da(&r, P.f32 ? 2 * N : 1));
    for (auto& p : rpf) HIPC(da(&p, P.f32 ? N : 1));
    HIPC(da(&dden, N));
    HIPC(da(&vpres, N));
    HIPC(da(&dden2, P.rk1 ? N : 1));
    HIPC(da(&vpres2, P.rk1 ? N : 1));
    HIPC(da(&gone, N));
    HIPC(da(&cell, N));
    HIPC(da(&key, N));
    HIPC(da(&slot, N));
    HIPC(da(&perm, N));
    HIPC(da(&pid, N));
    HIPC(da(&work, P.semi ? N * SPH_MB : 1));
    HIPC(da(&nwork, 1));
    HIPC(da(&cnt, m));
    HIPC(da(&start, m + 1));
    HIPC(da(&bsum, (m + 1023) / 1024 + 1));
    HIPC(da(&rowblk, (size_t)nrows + 1));
    HIPC(da(&idslot, N));
    HIPC(da(&rmids, RM_CAP * ns));
    HIPC(da(&nbr, N * (size_t)nb_stride(nb_stride_max)));
    HIPC(da(&nnb, N));
    HIPC(da(&bn, N));
    HIPC(da(&bt, P.semi ? N * SPH_MB : 1));
    HIPC(da(&scal, SC_COUNT * ns));
    HIPC(da(&iscal, IS_COUNT * ns));
    HIPC(da(&dtmin, ns));
    HIPC(da(&pairs, 1));
    if (dry) return 0;
    HIPC(hipMemset(pairs, 0, sizeof(unsigned long long)));
    HIPC(hipMemset(acc[0], 0, sizeof(double4) * N));
    HIPC(hipMemset(acc[1], 0, sizeof(double4) * N));
    HIPC(hipMemset(lam[0], 0, sizeof(int) * N));
    HIPC(hipMemset(lam[1], 0, sizeof(int) * N));
    HIPC(hipMemset(dden, 0, sizeof(double) * N));
    HIPC(hipMemset(vpres, 0, sizeof(double) * N));
    if (P.rk1) {
        HIPC(hipMemset(dden2, 0, sizeof(double) * N));
        HIPC(hipMemset(vpres2, 0, sizeof(double) * N));
    }
    HIPC(hipMemset(nnb, 0, sizeof(int) * N));
    HIPC(hipMemset(gone, 0, sizeof(int) * N));
    HIPC(hipMemset(bn, 0, sizeof(int) * N));
    HIPC(hipMemset(scal, 0, sizeof(double) * SC_COUNT * ns));
    HIPC(hipMemset(iscal, 0, sizeof(int) * IS_COUNT * ns));
    return 0;
}

namespace {

// Deck options the device does not build: status 95 names the first one set
// (the reference would run them; a device run without them would differ).
int check_options(int32_t opt) {
    static const struct { int32_t bit; const char* what; } NOT_BUILT[] = {
        {SPH_OPT_ALE3, "ALE3 (ALE3_term_momentum.f90, Leapfrog_continuity.f90:70-86)"},
        {SPH_OPT_C1, "C1_BE / C1_divu renormalisation (CalcVarLength.f90:285-324)"},
        {SPH_OPT_KTGF, "KTGF_config > 0 (mixture viscosity and erosion)"},
        {SPH_OPT_RK2, "the RK2 (Heun) time integration (Heun.f90)"},
    };
    const int32_t known = SPH_OPT_DENSITY_THRESHOLDS | SPH_OPT_ALE3 | SPH_OPT_C1 | SPH_OPT_KTGF |
                          SPH_OPT_CONVEX_EDGES | SPH_OPT_RK1 | SPH_OPT_RK2;
    if ((opt & SPH_OPT_RK1) && (opt & SPH_OPT_RK2)) return fail(2, "RK1 and RK2 are exclusive");
    if (opt & ~known) return fail(2, "unknown option bits " + std::to_string(opt & ~known));
    for (const auto& o : NOT_BUILT)
        if (opt & o.bit) return fail(95, std::string("not built on the device: ") + o.what);
    return 0;
}

// slip_coefficient_mode of a fixed or tapis frontier: 0 (no wall shear), 1
// (slip coefficient from input) or 2 (SA-SPH wall function,
// AddBoundaryContributions_to_ME3D.f90:207-230); ReadInputBoundaries.f90:
// 186-211 reads no other value.
int check_slip_mode(int mode, int btype) {
    if ((btype == 0 || btype == 1) && (mode < 0 || mode > 2))
        return fail(2, "slip_coefficient_mode must be 0, 1 or 2, got " + std::to_string(mode));
    return 0;
}

// Validates one solve's parameters and fills its constants (placement fields
// zero).  The reference's values: ReadInput.f90:198-205 (h, 2h, eta, eta2),
// Gest_Input.f90:266-289 (kernel constants).
int fill_dp(const sph_params* p, const sph_faces* f, DP& P, bool dry = false) {
    if (p->capacity <= 0 || p->capacity > (int)NB_MASK) return fail(2, "capacity out of range");
    if (p->nmaxpartj <= 0) return fail(2, "nmaxpartj must be positive");
    for (int d = 0; d < 3; ++d)
        if (p->ncd[d] <= 0 || !(p->dcd[d] > 0.0)) return fail(2, "invalid grid");
    if (p->precision != SPH_FP64 && p->precision != SPH_FP32_MIXED) return fail(2, "invalid precision");
    if (p->semi && !dry && (!f || f->nf <= 0)) return fail(2, "SA-SPH requested without frontier faces");
    const long long nmax = (long long)p->ncd[0] * p->ncd[1] * p->ncd[2];
    if (nmax >= (1LL << 30)) return fail(2, "grid too large");
    if (int rc = check_options(p->options)) return rc;
    if (f && f->nf > 0 && (!f->slip_mode || !f->btype)) return fail(1, "null frontier table");
    for (int q = 0; f && q < f->nf; ++q)
        if (int rc = check_slip_mode(f->slip_mode[q], f->btype[q])) return rc;
    // convex edges: SPH_OPT_CONVEX_EDGES says NumBEdges > 0, the faces carry them
    const int ne = (f && p->semi) ? f->nedges : 0;
    if (f && f->nedges < 0) return fail(2, "negative number of convex edges");
    if (((p->options & SPH_OPT_CONVEX_EDGES) != 0) != (ne > 0))
        return fail(2, "SPH_OPT_CONVEX_EDGES is set exactly when the frontier faces carry convex edges");
    if (ne > 0 && (!f->edge_face || !f->edge_gx || !f->edge_comp || !f->edge_len))
        return fail(1, "null convex-edge table");
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 2; ++k)
            if (f->edge_face[2 * e + k] < 0 || f->edge_face[2 * e + k] >= f->nf)
                return fail(2, "convex edge face index out of range");
    P = DP{};
    P.cap = p->capacity;
    P.scap = p->capacity;
    P.nmaxpartj = p->nmaxpartj;
    P.nmax = (int)nmax;
    P.semi = p->semi ? 1 : 0;
    P.f32 = p->precision == SPH_FP32_MIXED;
    P.dthr = (p->options & SPH_OPT_DENSITY_THRESHOLDS) ? 1 : 0;
    P.rk1 = (p->options & SPH_OPT_RK1) ? 1 : 0;
    P.nf = f ? f->nf : 0;
    P.ne = ne;
    P.nopen = 0;  // open faces (CancelOutgoneParticles_3D)
    for (int q = 0; p->semi && f && q < f->nf; ++q) P.nopen += f->btype[q] == 2;
    P.dim = 3;
    for (int d = 0; d < 3; ++d) {
        P.ncd[d] = p->ncd[d];
        P.extr_lo[d] = p->extr_lo[d];
        P.dcd[d] = p->dcd[d];
        P.grav[d] = p->grav[d];
    }
    P.h = p->h;
    P.dx = p->dx;
    P.coefke = p->coefke;
    P.coefkacl = p->coefkacl;
    P.eta = p->eta;
    P.eta2 = p->eta2;
    P.squareh = p->h * p->h;
    P.doubleh = 2.0 * p->h;
    P.square_doubleh = P.doubleh * P.doubleh;
    P.ke_coef = p->coefke / p->h;
    P.kacl_coef = p->coefkacl / p->h;
    P.ktdelta = f ? f->ktdelta : 0.05;
    P.GI = (double)9.80665f;  // Static_allocation_module.f90:75, a default-real literal
    P.prif = p->prif;
    P.den0 = p->den0;
    P.eps = p->eps;
    P.celerita = p->celerita;
    P.alfamon = p->alfamon;
    P.kin_visc = p->kin_visc;
    P.cfl = p->cfl;
    P.vsc_coeff = p->vsc_coeff;
    P.pesodt = p->pesodt;
    P.tetap = p->tetap;
    P.tetav = p->tetav;
    P.f_h = (float)P.h;
    P.f_unosuh = (float)(1.0 / P.h);
    P.f_unosusquareh = (float)(1.0 / P.squareh);
    P.f_eta = (float)P.eta;
    P.f_eta2 = (float)P.eta2;
    P.f_ke_coef = (float)P.ke_coef;
    P.f_kacl_coef = (float)P.kacl_coef;
    P.f_coefke = (float)P.coefke;
    for (int d = 0; d < 3; ++d) P.f_dcd[d] = (float)P.dcd[d];
    P.f_alfamon = (float)P.alfamon;
    P.f_celerita = (float)P.celerita;
    P.f_kin_visc = (float)P.kin_visc;
    P.f_den0 = (float)P.den0;
    P.f_eosk = (float)(P.eps / P.den0);
    P.f_prif = (float)P.prif;
    return 0;
}

inline int round256(long long v) { return (int)((v + 255) / 256 * 256); }

// Places nseg solves in one context: slots, keys and rows of segment s follow
// those of segment s-1 (slots and keys in whole 256-blocks).  One solve keeps
// its own sizes.  Fills c->P (segment 0 with the context capacity), c->segs,
// nkeys, nrows, nseg and the list stride; allocates nothing.
int layout_segments(sph_ctx* c, int nseg, const sph_params* p, const sph_faces* const* f = nullptr,
                    bool dry = false) {
    c->nseg = nseg;
    c->segs.assign(nseg, DP{});
    long long slots = 0, keys = 0, rows = 0, faces = 0, edges = 0;
    int nmaxpartj = 0;
    for (int s = 0; s < nseg; ++s) {
        const sph_faces* fs = f ? f[s] : nullptr;
        if (int rc = fill_dp(&p[s], fs, c->segs[s], dry)) return rc;
        DP& P = c->segs[s];
        if (p[s].precision != p[0].precision || (p[s].semi != 0) != (p[0].semi != 0) ||
            ((p[s].options ^ p[0].options) & SPH_OPT_RK1))
            return fail(2, "the segments of a batch share precision, wall treatment and time integration");
        P.seg = s;
        P.base = (int)slots;
        P.koff = (int)keys;
        P.roff = (int)rows;
        P.foff = (int)faces;
        P.eoff = (int)edges;
        if (nseg == 1) {
            slots = P.cap;
            keys = (long long)P.nmax + 1;
        } else {
            slots += round256(P.scap);
            keys += round256((long long)P.nmax + 1);
        }
        rows += (long long)P.ncd[1] * P.ncd[2];
        faces += P.nf;
        edges += P.ne;
        nmaxpartj = std::max(nmaxpartj, P.nmaxpartj);
        if (slots > (long long)NB_MASK || keys >= (1LL << 30))
            return fail(2, "batch too large for one context");
    }
    for (auto& P : c->segs) P.cap = (int)slots;
    c->P = c->segs[0];
    c->nkeys = (int)keys;
    c->nrows = (int)rows;
    c->nb_stride_max = nmaxpartj;
    return 0;
}

// Launch shape of the particle kernels and the persistent grids.
int setup_launch(sph_ctx* c, const sph_params* p) {
    if (hipStreamCreate(&c->stream) != hipSuccess) return fail(101, "hipStreamCreate failed");
    int dev = 0, ncu = 256;
    hipGetDevice(&dev);
    c->device = dev;
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    c->wg = p->wg_size > 0 ? p->wg_size : 64;
    if (c->wg < 64 || c->wg > 256 || c->wg % 64) return fail(2, "wg_size must be 64, 128, 192 or 256");
    c->grid = (c->P.cap + c->wg - 1) / c->wg;
    c->chunk = p->chunk > 0 ? p->chunk : 8;
    c->chunk_v = p->chunk_v > 0 ? p->chunk_v : c->chunk;
    c->chunk_c = p->chunk_c > 0 ? p->chunk_c : c->chunk;
    c->chunk_p = p->chunk_p > 0 ? p->chunk_p : c->chunk;
    auto bad = [](int v) { return v != 4 && v != 8; };
    if (bad(c->chunk) || bad(c->chunk_v) || bad(c->chunk_c) || bad(c->chunk_p))
        return fail(2, "chunk must be 4 or 8");
    if (c->nseg > 1)  // the segmented pair passes are built for whole-group walks
        c->chunk = c->chunk_v = c->chunk_c = c->chunk_p = 8;
    c->bt_grid = 4 * ncu;
    // persistent neighbour-search workgroups: as many as fit a CU (LDS
    // tile), times the CUs, a multiple of the 8 XCDs
    c->nb_shape = p->nb_shape;
    if (c->nb_shape < 0 || c->nb_shape >= NB_NSHAPES) return fail(2, "nb_shape must be 0 to 4");
    if (NB_SHAPES[c->nb_shape].tk == NB_T16)  // FP16 keeps its relative precision for cells of 1e-4..1e4 m
        for (const auto& Q : c->segs)
            for (int d = 0; d < 3; ++d)
                if (!(Q.dcd[d] >= 1.0e-4 && Q.dcd[d] <= 1.0e4)) c->nb_shape = 1;
    int per_cu = 0;
    hipError_t oe;
    const int nwg = NB_SHAPES[c->nb_shape].wg;
    switch (c->nb_shape) {
        case 1: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, NB_KERNEL(false, 1), nwg, 0); break;
        case 2: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, NB_KERNEL(false, 2), nwg, 0); break;
        case 3: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, NB_KERNEL(false, 3), nwg, 0); break;
        case 4: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, NB_KERNEL(false, 4), nwg, 0); break;
        default: oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, NB_KERNEL(false, 0), nwg, 0); break;
    }
    if (oe != hipSuccess || per_cu <= 0) per_cu = 3;
    c->nb_grid = (per_cu * ncu + 7) / 8 * 8;
    return 0;
}

// The frontier tables of every segment, concatenated: faces (indices offset
// by foff), per-key incidence (count, 1-based start into the face-index
// vector, faces in the 27-cell neighbourhood), kernel and direction tables
// and the FP32 quadrature table, one per segment.
int upload_tables(sph_ctx* c, const sph_faces* const* f) {
    const int nseg = c->nseg;
    long long nf = 0, nv = 0;
    for (int s = 0; s < nseg; ++s) {
        const DP& P = c->segs[s];
        if (!f || !f[s] || f[s]->nf <= 0) return 0;  // no frontier tables (no SA-SPH walls)
        (void)P;
        nf += f[s]->nf;
        for (int i = 0; i < P.nmax; ++i) nv += f[s]->gcbf_count[i];
    }
    std::vector<DFace> hf;
    hf.reserve(nf);
    std::vector<DEdge> he;
    // per key: count and 0-based start of the cell's near-face list in gvec
    std::vector<int> gcount(c->nkeys, 0), gstart(c->nkeys, 0), gvec;
    gvec.reserve(nv);
    std::vector<double> kt, bit;
    std::vector<BtF32> btf(nseg);
    for (int s = 0; s < nseg; ++s) {
        const sph_faces* fs = f[s];
        const DP& P = c->segs[s];
        for (int q = 0; q < fs->nf; ++q) {
            DFace F;
            std::memset(&F, 0, sizeof(F));
            std::memcpy(F.T, fs->T + 9 * q, sizeof(F.T));
            std::memcpy(F.LX, fs->LX + 8 * q, sizeof(F.LX));
            std::memcpy(F.GX, fs->GX + 12 * q, sizeof(F.GX));
            F.area = fs->area[q];
            F.slip = fs->slip_coeff[q];
            for (int d = 0; d < 3; ++d) {
                F.vel[d] = fs->velocity[3 * q + d];
                F.tvel[d] = fs->tvel[3 * q + d];
            }
            F.nodes = fs->nodes[q];
            F.btype = fs->btype[q];
            F.mode = fs->slip_mode[q];
            F.lam = fs->lam_check[q];
            if (F.nodes != 3 && F.nodes != 4) return fail(2, "SA-SPH faces must be triangles or quadrilaterals");
            dface_point_coeffs(F);
            hf.push_back(F);
        }
        long long sv = 0;
        for (int i = 0; i < P.nmax; ++i) sv += fs->gcbf_count[i];
        for (long long q = 0; q < sv; ++q)
            if (fs->gcbf_vec[q] < 0 || fs->gcbf_vec[q] >= fs->nf) return fail(2, "gcbf_vec entry out of range");
        // per cell: the frontier faces listed in its 27-cell neighbourhood, each
        // once, in the order FindCloseBoundaryFaces3D.f90 meets them (cells by
        // di, dj, dk, faces in GCBFVector order), so the particle's table keeps
        // the reference's face order
        const int ni = P.ncd[0], nj = P.ncd[1], nk = P.ncd[2];
        std::vector<int> seen;
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < nj; ++j)
                for (int i = 0; i < ni; ++i) {
                    seen.clear();
                    for (int di = -1; di <= 1; ++di)
                        for (int dj = -1; dj <= 1; ++dj)
                            for (int dk = -1; dk <= 1; ++dk) {
                                const int ii = i + di, jj = j + dj, kk = k + dk;
                                if (ii < 0 || ii >= ni || jj < 0 || jj >= nj || kk < 0 || kk >= nk) continue;
                                const size_t cn = ((size_t)kk * nj + jj) * ni + ii;
                                const int ini = fs->gcbf_start[cn] - 1;
                                for (int q = ini; q < ini + fs->gcbf_count[cn]; ++q) {
                                    const int fq = fs->gcbf_vec[q] + P.foff;
                                    if (std::find(seen.begin(), seen.end(), fq) == seen.end()) seen.push_back(fq);
                                }
                            }
                    const size_t key = P.koff + ((size_t)k * nj + j) * ni + i;
                    gcount[key] = (int)seen.size();
                    gstart[key] = (int)gvec.size();
                    gvec.insert(gvec.end(), seen.begin(), seen.end());
                }
        for (int e = 0; e < P.ne; ++e) {
            DEdge E;
            std::memset(&E, 0, sizeof(E));
            for (int d = 0; d < 3; ++d) {
                E.gx[d] = fs->edge_gx[3 * e + d];
                E.comp[d] = fs->edge_comp[3 * e + d];
            }
            E.len = fs->edge_len[e];
            E.face[0] = fs->edge_face[2 * e] + P.foff;
            E.face[1] = fs->edge_face[2 * e + 1] + P.foff;
            he.push_back(E);
        }
        kt.insert(kt.end(), fs->kerneltab, fs->kerneltab + 41 * 5);
        bit.insert(bit.end(), fs->bitab, fs->bitab + 256 * 7);
        build_btf32(fs->bitab, fs->kerneltab, fs->ktdelta, P.h, btf[s]);
    }
    if (gvec.empty()) gvec.push_back(0);
    if (he.empty()) he.push_back(DEdge{});
    auto up = [&](auto** d, const auto& h) {
        using T = typename std::remove_reference<decltype(h[0])>::type;
        return hipMalloc((void**)d, sizeof(T) * h.size()) == hipSuccess &&
               hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) == hipSuccess;
    };
    bool ok = up(&c->d_faces, hf) && up(&c->d_gcount, gcount) && up(&c->d_gstart, gstart) &&
              up(&c->d_gvec, gvec) && up(&c->d_kt, kt) && up(&c->d_bit, bit) &&
              up(&c->d_btf, btf) && up(&c->d_edges, he);
    if (!ok) return fail(102, "device allocation/copy of frontier tables failed");
    return 0;
}

// Segment maps of a segmented context: per-segment constants, segment of
// every 256-slot block, 256-key block and cell row.
int upload_segment_maps(sph_ctx* c) {
    if (c->nseg == 1) return 0;
    std::vector<int> pblk(c->P.cap / 256), kblk(c->nkeys / 256), rseg(c->nrows);
    for (int s = 0; s < c->nseg; ++s) {
        const DP& P = c->segs[s];
        const int sb = (s + 1 < c->nseg ? c->segs[s + 1].base : c->P.cap);
        for (int b = P.base / 256; b < sb / 256; ++b) pblk[b] = s;
        const int ke = (s + 1 < c->nseg ? c->segs[s + 1].koff : c->nkeys);
        for (int b = P.koff / 256; b < ke / 256; ++b) kblk[b] = s;
        for (int r = 0; r < P.ncd[1] * P.ncd[2]; ++r) rseg[P.roff + r] = s;
    }
    bool ok = hipMalloc(&c->d_P, sizeof(DP) * c->nseg) == hipSuccess &&
              hipMemcpy(c->d_P, c->segs.data(), sizeof(DP) * c->nseg, hipMemcpyHostToDevice) == hipSuccess &&
              hipMalloc(&c->d_pblk, sizeof(int) * pblk.size()) == hipSuccess &&
              hipMemcpy(c->d_pblk, pblk.data(), sizeof(int) * pblk.size(), hipMemcpyHostToDevice) == hipSuccess &&
              hipMalloc(&c->d_kblk, sizeof(int) * kblk.size()) == hipSuccess &&
              hipMemcpy(c->d_kblk, kblk.data(), sizeof(int) * kblk.size(), hipMemcpyHostToDevice) == hipSuccess &&
              hipMalloc(&c->d_rseg, sizeof(int) * rseg.size()) == hipSuccess &&
              hipMemcpy(c->d_rseg, rseg.data(), sizeof(int) * rseg.size(), hipMemcpyHostToDevice) == hipSuccess &&
              hipMalloc(&c->d_adj, sizeof(int) * c->nseg) == hipSuccess;
    if (!ok) return fail(102, "device allocation/copy of the segment maps failed");
    c->S = Seg{c->d_P, c->d_pblk, c->d_kblk, c->d_rseg, c->nseg, c->nrows};
    return 0;
}

int create_ctx(int nseg, const sph_params* p, const sph_faces* const* f, sph_ctx** out) {
    sph_ctx* c = new sph_ctx();
    int rc = layout_segments(c, nseg, p, f);
    if (!rc) c->precision = p[0].precision;
    if (!rc) rc = setup_launch(c, &p[0]);
    if (!rc) rc = c->alloc();
    if (!rc) rc = upload_tables(c, f);
    if (!rc) rc = upload_segment_maps(c);
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

}  // namespace

// Partial sums of sph_global_quantities_noalloc: GQ_SCRATCH_BLOCKS
// workgroups' double-double sums (2 x GQ_N doubles each), then the GQ_N
// results.  One copy per GPU, held until the process ends (262,208 bytes).
// Host code never passes this name to a kernel or a copy: it uses the
// device address scratch_on() looked up for the device.
constexpr int GQ_SCRATCH_BLOCKS = 2048;
static __device__ __align__(256) double scratch[2 * GQ_N * GQ_SCRATCH_BLOCKS + GQ_N];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on GPU dev, which the caller has made current:
// looked up once per device with hipGetSymbolAddress, then reused.
int scratch_on(int dev, double** out) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(2, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(100 + (int)e, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return 0;
}

}  // namespace

extern "C" {

const char* sph_last_error(void) { return g_err.c_str(); }

int sph_ctx_bytes(const sph_params* p, unsigned long long* bytes) {
    if (!p || !bytes) return fail(1, "null argument");
    if (p->capacity <= 0 || p->nmaxpartj <= 0) return fail(2, "invalid capacity or nmaxpartj");
    sph_ctx c;
    c.dry = true;
    if (int rc = layout_segments(&c, 1, p, nullptr, true)) return rc;
    c.alloc();
    *bytes = c.bytes;
    return 0;
}

int sph_device_memory(unsigned long long* free_bytes, unsigned long long* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(1, "null argument");
    size_t f = 0, t = 0;
    HIPC(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
    return 0;
}

int sph_query(sph_ctx* c, int32_t* busy) {
    if (!c || !busy) return fail(1, "null argument");
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipErrorNotReady) {
        *busy = 1;
        return 0;
    }
    if (e != hipSuccess) return fail(100, std::string("hipStreamQuery: ") + hipGetErrorString(e));
    *busy = 0;
    return 0;
}

int sph_ctx_create(const sph_params* p, const sph_faces* f, sph_ctx** out) {
    if (!p || !out) return fail(1, "null argument");
    return create_ctx(1, p, f ? &f : nullptr, out);
}

int sph_batch_create(int32_t nseg, const sph_params* p, const sph_faces* const* f, sph_ctx** out) {
    if (!p || !out) return fail(1, "null argument");
    if (nseg < 1) return fail(2, "a batch holds at least one segment");
    return create_ctx(nseg, p, f, out);
}

int sph_batch_bytes(int32_t nseg, const sph_params* p, unsigned long long* bytes) {
    if (!p || !bytes) return fail(1, "null argument");
    if (nseg < 1) return fail(2, "a batch holds at least one segment");
    sph_ctx c;
    c.dry = true;
    if (int rc = layout_segments(&c, nseg, p, nullptr, true)) return rc;
    c.alloc();
    *bytes = c.bytes;
    return 0;
}

int sph_ctx_create2d(const sph_params* p, const sph_sides2d* sd, sph_ctx** out) {
    if (!p || !out) return fail(1, "null argument");
    if (p->ncd[1] != 1) return fail(2, "a 2D grid has one cell row in y (ncd[1] = 1)");
    if (p->semi && (!sd || sd->ns <= 0)) return fail(2, "SA-SPH requested without boundary sides");
    if (int rc = check_options(p->options)) return rc;
    if (p->options & SPH_OPT_CONVEX_EDGES) return fail(2, "convex edges are a 3D frontier feature");
    int nopen = 0;
    for (int k = 0; sd && k < sd->ns; ++k) {
        if (int rc = check_slip_mode(sd->slip_mode[k], sd->btype[k])) return rc;
        if (sd->btype[k] == 2) nopen += 1;
    }
    if (nopen > 0 && !sd->v2) return fail(2, "open sides need their second vertex (sides.v2)");
    sph_params q = *p;
    q.semi = 0;
    int rc = sph_ctx_create(&q, nullptr, out);
    if (rc) return rc;
    sph_ctx* c = *out;
    DP& P = c->P;
    P.dim = 2;
    P.semi = p->semi ? 1 : 0;
    P.nf = sd ? sd->ns : 0;
    P.nopen = P.semi ? nopen : 0;  // open sides (CancelOutgoneParticles_2D)
    if (P.nf > 0) {
        std::vector<DFace> hs(P.nf);
        for (int k = 0; k < P.nf; ++k) {
            DFace& S = hs[k];
            std::memset(&S, 0, sizeof(S));
            std::memcpy(S.T, sd->T + 9 * k, sizeof(S.T));
            const double* RN = sd->RN + 9 * k;  // column-major: RN(r,c) at [3c+r]
            S.RN[0] = RN[0];
            S.RN[1] = RN[6];
            S.RN[2] = RN[2];
            S.RN[3] = RN[8];
            S.v1[0] = sd->v1[3 * k];
            S.v1[1] = sd->v1[3 * k + 2];
            if (sd->btype[k] == 2) {
                S.v2[0] = sd->v2[3 * k];
                S.v2[1] = sd->v2[3 * k + 2];
            }
            S.len = sd->length[k];
            S.angle = sd->angle[k];
            S.prev = sd->previous[k];
            S.slip = sd->slip_coeff[k];
            for (int d = 0; d < 3; ++d) {
                S.vel[d] = sd->velocity[3 * k + d];
                S.tvel[d] = sd->tvel[3 * k + d];
            }
            S.btype = sd->btype[k];
            S.mode = sd->slip_mode[k];
            S.lam = sd->lam_check[k];
            S.nodes = 2;
            if (S.prev < 0 || S.prev > P.nf) {
                sph_ctx_destroy(c);
                *out = nullptr;
                return fail(2, "previous side out of range");
            }

        }
        // the context was created without SA-SPH faces, so its boundary
        // table holds one record: give it the SPH_MB slots per particle that
        // k_boundary_2d writes
        bool ok = true;
        if (P.semi) {
            hipFree(c->bt);
            c->bt = nullptr;
            ok = hipMalloc(&c->bt, sizeof(BTab) * (size_t)P.cap * SPH_MB) == hipSuccess;
        }
        ok = ok && hipMalloc(&c->d_faces, sizeof(DFace) * P.nf) == hipSuccess &&
             hipMalloc(&c->d_bit, sizeof(double) * 22 * 6) == hipSuccess;
        ok = ok && hipMemcpy(c->d_faces, hs.data(), sizeof(DFace) * P.nf, hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemcpy(c->d_bit, sd->bit2d, sizeof(double) * 22 * 6, hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) {
            sph_ctx_destroy(c);
            *out = nullptr;
            return fail(102, "device allocation/copy of boundary sides failed");
        }
    }
    return 0;
}

int sph_ctx_destroy(sph_ctx* c) {
    if (c) {
        hipStreamSynchronize(c->stream);
        delete c;
    }
    return 0;
}

int sph_dbsph_walls(sph_ctx* c, const sph_dbsph_walls_t* w) {
    if (!c || !w) return fail(1, "null argument");
    if (c->P.dbsph) return fail(2, "the context already has DB-SPH walls");
    if (c->phase != PH_EMPTY && c->phase != PH_SORTED)
        return fail(120, "sph_dbsph_walls is called before sph_preloop");
    if (c->nseg != 1) return fail(95, "DB-SPH walls run in a one-solve context");
    if (c->P.f32) return fail(95, "DB-SPH walls run in FP64 (precision SPH_FP64)");
    if (!c->P.rk1) return fail(2, "DB-SPH walls need the explicit RK1 step (SPH_OPT_RK1: time_split = 0)");
    if (c->P.dim != 3) return fail(2, "DB-SPH walls are a 3D frontier");
    if (c->P.semi || c->P.nf) return fail(2, "a DB-SPH context has no SA-SPH frontier (semi = 0, no faces)");
    const int nw = w->nw;
    if (nw <= 0) return fail(2, "no wall elements");
    if (!w->coord || !w->normal || !w->vel || !w->weight || !w->volume || !w->mass)
        return fail(1, "null wall-element array");
    DP& P = c->P;
    // wall elements into the grid cells, ascending id per cell (OrdGrid1.f90:108-137)
    std::vector<int> wcell(nw), count(P.nmax + 1, 0);
    for (int k = 0; k < nw; ++k) {
        const double* x = w->coord + 3 * k;
        const int i = (int)std::ceil((x[0] - P.extr_lo[0]) / P.dcd[0]);
        const int j = (int)std::ceil((x[1] - P.extr_lo[1]) / P.dcd[1]);
        const int kk = (int)std::ceil((x[2] - P.extr_lo[2]) / P.dcd[2]);
        if (i < 1 || i > P.ncd[0] || j < 1 || j > P.ncd[1] || kk < 1 || kk > P.ncd[2])
            return fail(2, "a wall element lies outside the grid");
        wcell[k] = ((kk - 1) * P.ncd[1] + (j - 1)) * P.ncd[0] + i;
        count[wcell[k]] += 1;
    }
    std::vector<int> wstart(P.nmax + 1, 0), fill(P.nmax + 1, 0), wids(nw);
    for (int cc = 1; cc <= P.nmax; ++cc) wstart[cc] = wstart[cc - 1] + count[cc];
    for (int k = 0; k < nw; ++k) {
        const int cc = wcell[k];
        wids[wstart[cc - 1] + fill[cc]] = k;
        fill[cc] += 1;
    }
    std::vector<double4> hpos(nw);
    for (int w2 = 0; w2 < nw; ++w2) {
        const double* x = w->coord + 3 * wids[w2];
        hpos[w2] = make_double4(x[0], x[1], x[2], (double)wids[w2]);
    }
    std::vector<double4> h0(nw), h1(nw), h2(nw);
    for (int k = 0; k < nw; ++k) {
        h0[k] = make_double4(w->coord[3 * k], w->coord[3 * k + 1], w->coord[3 * k + 2], w->weight[k]);
        h1[k] = make_double4(w->normal[3 * k], w->normal[3 * k + 1], w->normal[3 * k + 2], w->volume[k]);
        h2[k] = make_double4(w->vel[3 * k], w->vel[3 * k + 1], w->vel[3 * k + 2], w->mass[k]);
    }
    const size_t N = P.cap;
    bool ok = dalloc(&c->d_w0, nw) == hipSuccess && dalloc(&c->d_w1, nw) == hipSuccess &&
              dalloc(&c->d_w2, nw) == hipSuccess && dalloc(&c->d_wstart, P.nmax + 1) == hipSuccess &&
              dalloc(&c->d_wids, nw) == hipSuccess && dalloc(&c->d_wcell, nw) == hipSuccess && dalloc(&c->d_wpos, nw) == hipSuccess &&
              dalloc(&c->d_wdens, nw) == hipSuccess && dalloc(&c->d_wpres, nw) == hipSuccess &&
              dalloc(&c->d_wwet, nw) == hipSuccess;
    for (auto& p : c->gam) ok = ok && dalloc(&p, N) == hipSuccess;
    for (auto& p : c->dsh) ok = ok && dalloc(&p, N) == hipSuccess;
    ok = ok && dalloc(&c->sig, N) == hipSuccess && dalloc(&c->rsph, N) == hipSuccess &&
         dalloc(&c->drho, 3 * N) == hipSuccess && dalloc(&c->dvelg, 9 * N) == hipSuccess &&
         dalloc(&c->fsf, N) == hipSuccess && dalloc(&c->nfw, N) == hipSuccess &&
         dalloc(&c->fwl, N * SPH_MW) == hipSuccess;
    if (!ok) return fail(102, "device allocation of the DB-SPH arrays failed");
    // DBSPH_IC_surface_elements.f90: dens = den0, pres = prif, wet = 0; the
    // particles' DB-SPH sums start at zero
    std::vector<double> d0(nw, P.den0), p0(nw, P.prif);
    HIPC(hipMemcpy(c->d_w0, h0.data(), sizeof(double4) * nw, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_w1, h1.data(), sizeof(double4) * nw, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_w2, h2.data(), sizeof(double4) * nw, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_wstart, wstart.data(), sizeof(int) * (P.nmax + 1), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_wids, wids.data(), sizeof(int) * nw, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_wcell, wcell.data(), sizeof(int) * nw, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_wpos, hpos.data(), sizeof(double4) * nw, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_wdens, d0.data(), sizeof(double) * nw, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_wpres, p0.data(), sizeof(double) * nw, hipMemcpyHostToDevice));
    HIPC(hipMemset(c->d_wwet, 0, sizeof(int) * nw));
    for (auto p : c->gam) HIPC(hipMemset(p, 0, sizeof(double) * N));
    for (auto p : c->dsh) HIPC(hipMemset(p, 0, sizeof(double) * N));
    HIPC(hipMemset(c->nfw, 0, sizeof(int) * N));
    HIPC(hipMemset(c->fsf, 0, sizeof(int) * N));
    P.dbsph = 1;
    P.nw = nw;
    P.gamma_lim = w->gamma_limiter ? 1 : 0;
    P.fs_allowed = w->fs_allowed ? 1 : 0;
    P.neg_wall_p = w->negative_wall_p_allowed ? 1 : 0;
    P.muscl_bnd = w->muscl_boundary ? 1 : 0;
    return 0;
}

int sph_dbsph_kinematics(sph_ctx* c, int32_t nrec, const double* records, const double* centre) {
    if (!c) return fail(1, "null argument");
    if (!c->P.dbsph) return fail(2, "sph_dbsph_kinematics needs DB-SPH walls (sph_dbsph_walls first)");
    if (c->phase != PH_EMPTY && c->phase != PH_SORTED)
        return fail(120, "sph_dbsph_kinematics is called before sph_preloop");
    if (nrec < 0 || nrec > SPH_MAX_KIN) return fail(2, "kinematics records: 0 to 4096");
    if (nrec > 0 && (!records || !centre)) return fail(1, "null kinematics array");
    if (nrec > 0 && !(records[0] <= 0.0))
        return fail(2, "the first kinematics record's time must not exceed 0 (DBSPH_kinematics.f90 reads before it)");
    if (c->d_kin) {
        HIPC(hipFree(c->d_kin));
        c->d_kin = nullptr;
    }
    if (nrec > 0) {
        if (dalloc(&c->d_kin, 7 * (size_t)nrec) != hipSuccess) return fail(102, "device allocation failed");
        HIPC(hipMemcpy(c->d_kin, records, sizeof(double) * 7 * nrec, hipMemcpyHostToDevice));
    }
    c->P.nkin = nrec;
    for (int d = 0; d < 3; ++d) c->P.kin_c[d] = nrec > 0 ? centre[d] : 0.0;
    return 0;
}

int sph_dbsph_state(sph_ctx* c, double* gamma, double* sigma, double* rhosph, double* dshep, int32_t* fs,
                    double* wall_dens, double* wall_pres, int32_t* wall_wet) {
    if (!c) return fail(1, "null argument");
    if (!c->P.dbsph) return fail(2, "the context has no DB-SPH walls");
    if (c->phase != PH_READY) return fail(121, "the DB-SPH state is defined at a time-step boundary");
    HIPC(hipStreamSynchronize(c->stream));
    int n;
    HIPC(hipMemcpy(&n, c->iscal + IS_N, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<double4> hv(n);
    std::vector<double> hg(n), hs(n), hr(n), hd(n);
    std::vector<int> hf(n);
    if (n) {
        HIPC(hipMemcpy(hv.data(), c->vel[0], sizeof(double4) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hg.data(), c->gam[0], sizeof(double) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hs.data(), c->sig, sizeof(double) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hr.data(), c->rsph, sizeof(double) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hd.data(), c->dsh[0], sizeof(double) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hf.data(), c->fsf, sizeof(int) * n, hipMemcpyDeviceToHost));
    }
    for (int s = 0; s < n; ++s) {  // rows in pg (id) order
        const int r = (int)hv[s].w;
        if (r < 0 || r >= n) return fail(103, "particle ids are not 0..n-1");
        if (gamma) gamma[r] = hg[s];
        if (sigma) sigma[r] = hs[s];
        if (rhosph) rhosph[r] = hr[s];
        if (dshep) dshep[r] = hd[s];
        if (fs) fs[r] = hf[s];
    }
    const int nw = c->P.nw;
    if (wall_dens) HIPC(hipMemcpy(wall_dens, c->d_wdens, sizeof(double) * nw, hipMemcpyDeviceToHost));
    if (wall_pres) HIPC(hipMemcpy(wall_pres, c->d_wpres, sizeof(double) * nw, hipMemcpyDeviceToHost));
    if (wall_wet) HIPC(hipMemcpy(wall_wet, c->d_wwet, sizeof(int) * nw, hipMemcpyDeviceToHost));
    return 0;
}

int sph_state_upload(sph_ctx* c, int32_t n, const double* coord, const double* vel,
                     const double* dens, const double* pres, const double* mass) {
    if (!c) return fail(1, "null argument");
    if (c->nseg > 1) return fail(2, "a segmented context is loaded with sph_segment_upload");
    return sph_segment_upload(c, 0, n, coord, vel, dens, pres, mass);
}

int sph_segment_upload(sph_ctx* c, int32_t seg, int32_t n, const double* coord, const double* vel,
                       const double* dens, const double* pres, const double* mass) {
    if (!c || !coord || !vel || !dens || !pres || !mass) return fail(1, "null argument");
    if (seg < 0 || seg >= c->nseg) return fail(3, "segment index out of range");
    const DP& P = c->nseg > 1 ? c->segs[seg] : c->P;
    if (n < 0 || n > P.scap) return fail(3, "particle count exceeds the segment capacity");
    std::vector<double4> hx(n), hv(n);
    std::vector<double2> hr(n);
    for (int i = 0; i < n; ++i) {
        hx[i] = make_double4(coord[3 * i], coord[3 * i + 1], coord[3 * i + 2], mass[i]);
        hv[i] = make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], (double)i);
        hr[i] = make_double2(dens[i], pres[i]);
    }
    int is[IS_COUNT] = {0};
    is[IS_N] = n;
    is[IS_ONGOING] = -999;
    double sc[SC_COUNT] = {0};
    const size_t b = P.base;
    // one mass for the whole solve: the pressure smoothing's one-record form
    bool um = n > 0;
    for (int i = 1; i < n && um; ++i) um = mass[i] == mass[0];
    const int umass = um ? 1 : 0;
    const float f_mass = um ? (float)mass[0] : 0.0f;
    if (umass != P.umass || f_mass != P.f_mass) {
        DP& Q = c->nseg > 1 ? c->segs[seg] : c->P;
        Q.umass = umass;
        Q.f_mass = f_mass;
        if (c->nseg > 1) {
            if (seg == 0) {
                c->P.umass = umass;
                c->P.f_mass = f_mass;
            }
            HIPC(hipMemcpy(c->d_P + seg, &Q, sizeof(DP), hipMemcpyHostToDevice));
        }
        // captured graphs hold the constants they were captured with
        for (auto& g : c->graphs) hipGraphExecDestroy(g.second);
        c->graphs.clear();
    }
    if (c->staged.empty()) {  // first segment of a new upload: clear the whole context
        HIPC(hipMemsetAsync(c->lam[0], 0, sizeof(int) * c->P.cap, c->stream));
        HIPC(hipMemsetAsync(c->acc[1], 0, sizeof(double4) * c->P.cap, c->stream));
        HIPC(hipMemsetAsync(c->iscal, 0, sizeof(int) * IS_COUNT * c->nseg, c->stream));
        HIPC(hipMemsetAsync(c->scal, 0, sizeof(double) * SC_COUNT * c->nseg, c->stream));
        c->staged.assign(c->nseg, 0);
        c->phase = PH_EMPTY;
    }
    if (n) {
        HIPC(hipMemcpyAsync(c->xm[1] + b, hx.data(), sizeof(double4) * n, hipMemcpyHostToDevice, c->stream));
        HIPC(hipMemcpyAsync(c->vel[2] + b, hv.data(), sizeof(double4) * n, hipMemcpyHostToDevice, c->stream));
        HIPC(hipMemcpyAsync(c->rp[0] + b, hr.data(), sizeof(double2) * n, hipMemcpyHostToDevice, c->stream));
    }
    HIPC(hipMemcpyAsync(c->iscal + (size_t)seg * IS_COUNT, is, sizeof(is), hipMemcpyHostToDevice, c->stream));
    HIPC(hipMemcpyAsync(c->scal + (size_t)seg * SC_COUNT, sc, sizeof(sc), hipMemcpyHostToDevice, c->stream));
    // the staging buffers go out of scope: wait for the copies
    HIPC(hipStreamSynchronize(c->stream));
    c->staged[seg] = 1;
    for (int s = 0; s < c->nseg; ++s)
        if (!c->staged[s]) return 0;  // the other segments come in later calls
    c->staged.clear();
    int rc = run_sort(c);
    if (rc) return rc;
    // the sort leaves (rho,p) in the buffer the continuity pass reads; the
    // state at rest keeps it in the buffer the momentum pass reads
    HIPC(hipMemcpyAsync(c->rp[0], c->rp[1], sizeof(double2) * c->P.cap, hipMemcpyDeviceToDevice, c->stream));
    if (c->P.f32)
        HIPC(hipMemcpyAsync(c->rpf[0], c->rpf[1], sizeof(float2) * c->P.cap, hipMemcpyDeviceToDevice,
                            c->stream));
    HIPC(hipStreamSynchronize(c->stream));
    collect_timing(c);
    c->n_upload = n;
    c->phase = PH_SORTED;
    return 0;
}

int sph_particle_count(sph_ctx* c, int32_t* n) { return sph_segment_count(c, 0, n); }

int sph_segment_count(sph_ctx* c, int32_t seg, int32_t* n) {
    if (!c || !n) return fail(1, "null argument");
    if (seg < 0 || seg >= c->nseg) return fail(3, "segment index out of range");
    HIPC(hipStreamSynchronize(c->stream));
    HIPC(hipMemcpy(n, c->iscal + (size_t)seg * IS_COUNT + IS_N, sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

int sph_state_download(sph_ctx* c, double* coord, double* vel, double* acc, double* dens,
                       double* pres, double* dden, double* vpres, int32_t* ids) {
    return sph_segment_download(c, 0, coord, vel, acc, dens, pres, dden, vpres, ids);
}

int sph_segment_download(sph_ctx* c, int32_t seg, double* coord, double* vel, double* acc, double* dens,
                         double* pres, double* dden, double* vpres, int32_t* ids) {
    if (!c) return fail(1, "null argument");
    if (seg < 0 || seg >= c->nseg) return fail(3, "segment index out of range");
    if (c->phase != PH_SORTED && c->phase != PH_READY && c->phase != PH_PS)
        return fail(121, "state download is only defined at a time-step boundary");
    HIPC(hipStreamSynchronize(c->stream));
    int n;
    HIPC(hipMemcpy(&n, c->iscal + (size_t)seg * IS_COUNT + IS_N, sizeof(int), hipMemcpyDeviceToHost));
    const size_t b = c->nseg > 1 ? c->segs[seg].base : 0;
    std::vector<double4> hx(n), hv(n), ha(n);
    std::vector<double2> hr(n);
    std::vector<double> hd(n), hp(n);
    if (n) {
        HIPC(hipMemcpy(hx.data(), c->xm[0] + b, sizeof(double4) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hv.data(), c->vel[0] + b, sizeof(double4) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(ha.data(), c->acc[0] + b, sizeof(double4) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hr.data(), c->rp[cur_rp(c)] + b, sizeof(double2) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hd.data(), (c->P.rk1 ? c->dden2 : c->dden) + b, sizeof(double) * n,
                       hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hp.data(), (c->P.rk1 ? c->vpres2 : c->vpres) + b, sizeof(double) * n,
                       hipMemcpyDeviceToHost));
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return hv[a].w < hv[b].w; });
    for (int r = 0; r < n; ++r) {
        const int s = order[r];
        if (coord) { coord[3 * r] = hx[s].x; coord[3 * r + 1] = hx[s].y; coord[3 * r + 2] = hx[s].z; }
        if (vel) { vel[3 * r] = hv[s].x; vel[3 * r + 1] = hv[s].y; vel[3 * r + 2] = hv[s].z; }
        if (acc) { acc[3 * r] = ha[s].x; acc[3 * r + 1] = ha[s].y; acc[3 * r + 2] = ha[s].z; }
        if (dens) dens[r] = hr[s].x;
        if (pres) pres[r] = hr[s].y;
        if (dden) dden[r] = hd[s];
        if (vpres) vpres[r] = hp[s];
        if (ids) ids[r] = (int32_t)hv[s].w;
    }
    return 0;
}

int sph_get_laminar(sph_ctx* c, int32_t* lam) {
    if (!c || !lam) return fail(1, "null argument");
    if (c->nseg > 1) return fail(2, "not available on a segmented context");
    if (c->phase != PH_SORTED && c->phase != PH_READY && c->phase != PH_PS)
        return fail(121, "the laminar flags are taken at a time-step boundary");
    HIPC(hipStreamSynchronize(c->stream));
    int n;
    HIPC(hipMemcpy(&n, c->iscal + IS_N, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> hl(n);
    std::vector<double4> hv(n);
    if (n) {
        // the momentum pass writes lam[0]; every sort gathers it into lam[1]
        // (in the leapfrog step before the next momentum pass reads it, in
        // the RK1 step at its end), so lam[1] holds the current flags
        HIPC(hipMemcpy(hl.data(), c->lam[1], sizeof(int) * n, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(hv.data(), c->vel[0], sizeof(double4) * n, hipMemcpyDeviceToHost));
    }
    for (int s = 0; s < n; ++s) lam[(int)hv[s].w] = hl[s];
    return 0;
}

int sph_get_scalars(sph_ctx* c, double out[5]) { return sph_segment_scalars(c, 0, out); }

int sph_segment_scalars(sph_ctx* c, int32_t seg, double out[5]) {
    if (!c || !out) return fail(1, "null argument");
    if (seg < 0 || seg >= c->nseg) return fail(3, "segment index out of range");
    HIPC(hipStreamSynchronize(c->stream));
    double sc[SC_COUNT];
    int is[IS_COUNT];
    HIPC(hipMemcpy(sc, c->scal + (size_t)seg * SC_COUNT, sizeof(sc), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(is, c->iscal + (size_t)seg * IS_COUNT, sizeof(is), hipMemcpyDeviceToHost));
    out[0] = sc[SC_DT];
    out[1] = sc[SC_DTAVG];
    out[2] = sc[SC_TIME];
    out[3] = is[IS_IT];
    const int n = is[IS_N];
    const size_t b = c->nseg > 1 ? c->segs[seg].base : 0;
    std::vector<int> nn(n);
    if (n) HIPC(hipMemcpy(nn.data(), c->nnb + b, sizeof(int) * n, hipMemcpyDeviceToHost));
    long long pc = 0;
    for (int i = 0; i < n; ++i) pc += nn[i] - 1;
    out[4] = (double)pc;
    return 0;
}

int sph_global_quantities(sph_ctx* c, double out[8]) {
    if (!c || !out) return fail(1, "null argument");
    if (c->nseg > 1) return fail(2, "not available on a segmented context");
    if (c->phase != PH_SORTED && c->phase != PH_READY && c->phase != PH_PS)
        return fail(121, "global quantities are taken at a time-step boundary");
    const int nb = 2 * c->bt_grid / 4;  // two workgroups per CU
    double* part = nullptr;
    double* dout = nullptr;
    HIPC(hipMalloc(&part, sizeof(double) * 2 * GQ_N * nb));
    HIPC(hipMalloc(&dout, sizeof(double) * GQ_N));
    k_global_partial<<<nb, 256, 0, c->stream>>>(c->xm[0], c->vel[0], c->rp[cur_rp(c)], part, c->iscal);
    k_global_final<<<1, 64, 0, c->stream>>>(part, nb, dout);
    hipError_t e = hipMemcpyAsync(out, dout, sizeof(double) * GQ_N, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(part);
    hipFree(dout);
    if (e != hipSuccess) return fail(100, std::string("global quantities: ") + hipGetErrorString(e));
    return 0;
}

int sph_global_quantities_noalloc(sph_ctx* c, double out[8]) {
    if (!c || !out) return fail(1, "null argument");
    if (c->nseg > 1) return fail(2, "not available on a segmented context");
    if (c->phase != PH_SORTED && c->phase != PH_READY && c->phase != PH_PS)
        return fail(121, "global quantities are taken at a time-step boundary");
    // the workgroup partition of sph_global_quantities, so the double-double
    // partial sums and the rounded results are the same bit for bit
    const int nb = 2 * c->bt_grid / 4;
    if (nb > GQ_SCRATCH_BLOCKS)
        return fail(3, "the global sums need " + std::to_string(nb) + " partial-sum workgroups; the "
                       "library's scratch holds " + std::to_string(GQ_SCRATCH_BLOCKS));
    int prev = 0;
    HIPC(hipGetDevice(&prev));
    if (prev != c->device) HIPC(hipSetDevice(c->device));
    double* base = nullptr;
    int rc = scratch_on(c->device, &base);
    if (!rc) {
        // every partial slot the final pass reads (b < nb) is written by the
        // first pass of this call, so nothing has to start at zero
        double* part = base;
        double* dout = base + 2 * GQ_N * GQ_SCRATCH_BLOCKS;
        k_global_partial<<<nb, 256, 0, c->stream>>>(c->xm[0], c->vel[0], c->rp[cur_rp(c)], part, c->iscal);
        k_global_final<<<1, 64, 0, c->stream>>>(part, nb, dout);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, dout, sizeof(double) * GQ_N, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(100 + (int)e, std::string("global quantities: ") + hipGetErrorString(e));
    }
    if (prev != c->device) hipSetDevice(prev);
    return rc;
}

int sph_monitor_lines(sph_ctx* c, int32_t nlines, const int32_t* icont, int32_t npoints,
                      const double* points) {
    if (!c || nlines < 0 || npoints < 0 || (nlines && !icont) || (npoints && !points))
        return fail(1, "null or negative argument");
    for (int i = 0; i < nlines; ++i)
        if (icont[2 * i] < 1 || icont[2 * i + 1] > npoints)
            return fail(3, "monitoring line " + std::to_string(i + 1) + " points outside 1.." +
                               std::to_string(npoints));
    HIPC(hipStreamSynchronize(c->stream));
    auto F = [](auto*& p) {
        if (p) hipFree(p);
        p = nullptr;
    };
    F(c->mon_icont);
    F(c->mon_pts);
    F(c->mon_out);
    c->mon_nlines = c->mon_npoints = 0;
    if (nlines == 0) return 0;
    HIPC(dalloc(&c->mon_icont, 2 * (size_t)nlines));
    HIPC(dalloc(&c->mon_pts, 3 * (size_t)npoints));
    HIPC(dalloc(&c->mon_out, 6 * (size_t)nlines));
    HIPC(hipMemcpy(c->mon_icont, icont, sizeof(int) * 2 * nlines, hipMemcpyHostToDevice));
    if (npoints) HIPC(hipMemcpy(c->mon_pts, points, sizeof(double) * 3 * npoints, hipMemcpyHostToDevice));
    c->mon_nlines = nlines;
    c->mon_npoints = npoints;
    return 0;
}

int sph_free_surface(sph_ctx* c, double* pelolib, double* zlft) {
    if (!c || ((!pelolib || !zlft) && c->mon_nlines)) return fail(1, "null argument");
    if (c->nseg > 1) return fail(2, "not available on a segmented context");
    if (c->phase != PH_SORTED && c->phase != PH_READY && c->phase != PH_PS)
        return fail(121, "the free surface is taken at a time-step boundary");
    const int nl = c->mon_nlines;
    if (nl == 0) return 0;
    k_free_surface<<<nl, FS_WG, 0, c->stream>>>(c->P, c->xm[0], c->start, c->mon_pts, c->mon_icont, nl,
                                                c->mon_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(pelolib, c->mon_out, sizeof(double) * 3 * nl, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(zlft, c->mon_out + 3 * nl, sizeof(double) * 3 * nl, hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(100, std::string("free surface: ") + hipGetErrorString(e));
    return 0;
}

int sph_pair_counter(sph_ctx* c, unsigned long long* total) {
    if (!c || !total) return fail(1, "null argument");
    HIPC(hipStreamSynchronize(c->stream));
    HIPC(hipMemcpy(total, c->pairs, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return 0;
}

int sph_get_neighbour_lists(sph_ctx* c, int32_t* count, int32_t* ids, int32_t stride) {
    if (!c || !count) return fail(1, "null argument");
    if (c->nseg > 1) return fail(2, "not available on a segmented context");
    if (stride < c->P.nmaxpartj && ids) return fail(3, "stride smaller than nmaxpartj");
    HIPC(hipStreamSynchronize(c->stream));
    int n;
    HIPC(hipMemcpy(&n, c->iscal + IS_N, sizeof(int), hipMemcpyDeviceToHost));
    const size_t N = c->P.cap;
    std::vector<int> nn(n);
    std::vector<double4> hv(n);
    HIPC(hipMemcpy(nn.data(), c->nnb, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(hv.data(), c->vel[0], sizeof(double4) * n, hipMemcpyDeviceToHost));
    int kmax = 0;
    for (int i = 0; i < n; ++i) kmax = std::max(kmax, nn[i]);
    const size_t ngroups = (size_t)(kmax + NB_G - 1) / NB_G;
    std::vector<uint32_t> nb(ngroups * N * NB_G);
    if (kmax) HIPC(hipMemcpy(nb.data(), c->nbr, sizeof(uint32_t) * nb.size(), hipMemcpyDeviceToHost));
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return hv[a].w < hv[b].w; });
    for (int r = 0; r < n; ++r) {
        const int s = order[r];
        count[r] = nn[s];
        if (ids)
            for (int k = 0; k < nn[s]; ++k)
                ids[(size_t)r * stride + k] = (int32_t)hv[nb[nb_index(N, s, k)] & NB_MASK].w;
    }
    return 0;
}

int sph_get_boundary_table(sph_ctx* c, int32_t* ncbf, int32_t* face, double* loc, double* bi,
                           double* ig, int32_t stride) {
    if (!c || !ncbf || !face || !loc || !bi || !ig) return fail(1, "null argument");
    if (c->nseg > 1) return fail(2, "not available on a segmented context");
    if (stride < SPH_MB) return fail(3, "stride smaller than the boundary-table slots");
    HIPC(hipStreamSynchronize(c->stream));
    int n;
    HIPC(hipMemcpy(&n, c->iscal + IS_N, sizeof(int), hipMemcpyDeviceToHost));
    const size_t N = c->P.cap;
    std::vector<int> nc(n);
    std::vector<double4> hv(n);
    HIPC(hipMemcpy(nc.data(), c->bn, sizeof(int) * n, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(hv.data(), c->vel[0], sizeof(double4) * n, hipMemcpyDeviceToHost));
    int qmax = 0;
    for (int i = 0; i < n; ++i) qmax = std::max(qmax, nc[i]);
    std::vector<BTab> hb((size_t)qmax * N);
    if (qmax) HIPC(hipMemcpy(hb.data(), c->bt, sizeof(BTab) * hb.size(), hipMemcpyDeviceToHost));
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return hv[a].w < hv[b].w; });
    for (int r = 0; r < n; ++r) {
        const int s = order[r];
        ncbf[r] = nc[s];
        for (int q = 0; q < nc[s]; ++q) {
            const BTab& B = hb[(size_t)q * N + s];
            const size_t o = (size_t)r * stride + q;
            face[o] = B.face;
            for (int t = 0; t < 3; ++t) loc[3 * o + t] = B.loc[t];
            for (int t = 0; t < 8; ++t) bi[8 * o + t] = B.bi[t];
            for (int t = 0; t < 9; ++t) ig[9 * o + t] = B.ig[t];
        }
    }
    return 0;
}

int sph_dt(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    if (int rc = expect(c, PH_READY, "sph_dt")) return rc;
    HIPC(hipMemsetAsync(c->dtmin, 0x7F, 8 * c->nseg, c->stream));
    run_dt(c, 1, true);
    c->phase = PH_DT;
    return 0;
}

int sph_momentum(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    if (int rc = expect(c, PH_DT, "sph_momentum")) return rc;
    run_momentum(c);
    c->phase = PH_MOM;
    return 0;
}
int sph_leapfrog_mom(sph_ctx* c) {  // fused into sph_momentum
    if (!c) return fail(1, "null argument");
    return expect(c, PH_MOM, "sph_leapfrog_mom");
}
int sph_velocity_smoothing(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    if (int rc = expect(c, PH_MOM, "sph_velocity_smoothing")) return rc;
    run_vsmooth(c);
    c->phase = PH_VS;
    return 0;
}
int sph_leapfrog_traj(sph_ctx* c) {  // fused into sph_velocity_smoothing
    if (!c) return fail(1, "null argument");
    return expect(c, PH_VS, "sph_leapfrog_traj");
}
int sph_sort(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    if (int rc = expect(c, PH_VS, "sph_sort")) return rc;
    run_sort(c);
    c->phase = PH_SORT;
    return 0;
}
int sph_neighbours(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    if (c->P.rk1) return expect(c, c->phase, "sph_neighbours");
    if (c->phase != PH_SORT && c->phase != PH_SORTED)
        return fail(120, "sph_neighbours called out of the time_step_loop order");
    run_neighbours(c);
    c->phase = c->phase == PH_SORTED ? PH_SORTED : PH_NB;
    return 0;
}
int sph_boundary_tab(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    if (c->P.rk1) return expect(c, c->phase, "sph_boundary_tab");
    if (c->phase != PH_NB && c->phase != PH_SORTED)
        return fail(120, "sph_boundary_tab called out of the time_step_loop order");
    run_btab(c);
    c->phase = c->phase == PH_SORTED ? PH_SORTED : PH_BT;
    return 0;
}
int sph_continuity(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    if (int rc = expect(c, PH_BT, "sph_continuity")) return rc;
    run_continuity(c);
    c->phase = PH_CONT;
    return 0;
}
int sph_leapfrog_cont(sph_ctx* c) {  // fused into sph_continuity
    if (!c) return fail(1, "null argument");
    return expect(c, PH_CONT, "sph_leapfrog_cont");
}
int sph_eos(sph_ctx* c) {  // fused into sph_continuity
    if (!c) return fail(1, "null argument");
    return expect(c, PH_CONT, "sph_eos");
}
int sph_pressure_smoothing(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    if (int rc = expect(c, PH_CONT, "sph_pressure_smoothing")) return rc;
    run_psmooth(c);
    c->phase = PH_READY;
    HIPC(hipStreamSynchronize(c->stream));
    collect_timing(c);
    return check_flags(c);
}

int sph_preloop(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    if (c->phase != PH_SORTED) return fail(120, "sph_preloop needs a freshly uploaded state");
    run_neighbours(c);
    run_btab(c);
    if (c->P.dbsph) {
        // time_step_loop.f90:82-135 for bsph: CalcVarLength with Gamma from
        // sigma, then again after OrdGrid1 (nothing has moved since the
        // upload's sort, so that OrdGrid1 keeps every particle in place), and
        // the wall elements' state
        run_dbsph_lists(c, 0);
        run_dbsph_lists(c, 1);
        run_dbsph_walls(c);
    }
    HIPC(hipMemsetAsync(c->dtmin, 0x7F, 8 * c->nseg, c->stream));
    run_dt(c, 0, true);
    // the first step's time_step_duration sees the same velocities again
    HIPC(hipMemsetAsync(c->dtmin, 0x7F, 8 * c->nseg, c->stream));
    SEGL(k_dtmin, nblk(c->P.cap), 256, c->vel[0], c->dtmin, c->iscal);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(c->stream));
    collect_timing(c);
    c->phase = PH_READY;
    return check_flags(c);
}

static void enqueue_steps(sph_ctx* c, int nsteps) {
    if (c->P.dbsph) {
        // time_step_loop.f90:175-489 for bsph with RK1: momentum and continuity
        // on the lists of the previous step's end, Euler (velocity, smoothing,
        // positions; OrdGrid1; Gamma, CalcVarLength, density), EoS, pressure
        // smoothing on the new lists, the wall elements' state
        for (int s = 0; s < nsteps; ++s) {
            run_dt(c, 1, false);
            run_momentum_dbsph(c);
            run_continuity(c);
            run_vsmooth(c);
            run_dbsph_kinematics(c);
            run_sort(c);
            run_neighbours(c);
            run_dbsph_lists(c, 2);
            run_psmooth_dbsph(c);
            run_dbsph_walls(c);
        }
        return;
    }
    if (c->P.rk1) {
        // time_step_loop.f90:175-489 with time_split = 0, RKscheme = 1: the
        // lists and boundary table of a stage are built by the previous step's
        // closing sort (nothing moves between that OrdGrid1 and the next
        // stage's CalcVarLength); momentum and continuity on the stage's
        // state, Euler's velocity, smoothing, position and density updates
        // with EoS, pressure smoothing, then OrdGrid1
        for (int s = 0; s < nsteps; ++s) {
            run_dt(c, 1, false);
            run_momentum(c);
            run_continuity(c);
            run_vsmooth(c);
            run_psmooth(c);
            run_sort(c);
            run_neighbours(c);
            run_btab(c);
        }
        return;
    }
    for (int s = 0; s < nsteps; ++s) {
        run_dt(c, 1, false);
        run_momentum(c);
        run_vsmooth(c);
        run_sort(c);
        run_neighbours(c);
        run_btab(c);
        run_continuity(c);
        run_psmooth(c);
    }
}

int sph_step_async(sph_ctx* c, int32_t nsteps) {
    if (!c) return fail(1, "null argument");
    if (c->phase != PH_READY) return fail(120, "sph_step needs the state at a time-step boundary");
    enqueue_steps(c, nsteps);
    HIPC(hipGetLastError());
    return 0;
}

int sph_step_graph(sph_ctx* c, int32_t nsteps) {
    if (!c) return fail(1, "null argument");
    if (c->phase != PH_READY) return fail(120, "sph_step needs the state at a time-step boundary");
    if (nsteps <= 0) return 0;
    if (c->timing) return sph_step_async(c, nsteps);  // per-kernel events need eager launches
    hipGraphExec_t exec = nullptr;
    for (auto& g : c->graphs)
        if (g.first == nsteps) exec = g.second;
    if (!exec) {
        // every launch of a step uses the same arguments (each buffer role
        // returns to its buffer after a step) and no memset/copy nodes
        HIPC(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue_steps(c, nsteps);
        hipGraph_t g = nullptr;
        HIPC(hipStreamEndCapture(c->stream, &g));
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return fail(100, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        c->graphs.emplace_back(nsteps, exec);
    }
    HIPC(hipGraphLaunch(exec, c->stream));
    return 0;
}

int sph_sync(sph_ctx* c) {
    if (!c) return fail(1, "null argument");
    HIPC(hipStreamSynchronize(c->stream));
    collect_timing(c);
    return check_flags(c);
}

int sph_step(sph_ctx* c, int32_t nsteps) {
    int rc = sph_step_async(c, nsteps);
    if (rc) return rc;
    return sph_sync(c);
}

int sph_get_launch(sph_ctx* c, int32_t out[3]) {
    if (!c || !out) return fail(1, "null argument");
    out[0] = c->wg;
    out[1] = c->chunk;
    out[2] = c->grid;
    return 0;
}

int sph_set_timing(sph_ctx* c, int32_t on) {
    if (!c) return fail(1, "null argument");
    c->timing = on != 0;
    c->timers.clear();
    return 0;
}

int sph_get_timing(sph_ctx* c, char* names, int32_t names_len, double* ms, int32_t* calls,
                   int32_t max_kernels) {
    if (!c) return fail(1, "null argument");
    std::string all;
    int k = 0;
    for (auto& t : c->timers) {
        if (k >= max_kernels) break;
        all += t.name + ";";
        ms[k] = t.ms;
        calls[k] = t.calls;
        ++k;
    }
    if (names && names_len > 0) {
        std::strncpy(names, all.c_str(), names_len - 1);
        names[names_len - 1] = 0;
    }
    return k;
}

}  // extern "C"
