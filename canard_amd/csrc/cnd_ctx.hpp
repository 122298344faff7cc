// The context of libcnd (device arrays, tables, sweep shapes) and the
// launch helpers shared by the translation units that launch sweeps.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/cnd.h"
#include "cnd_dev.hpp"
#include "cnd_scheme.hpp"
#include "cnd_sweepk.hpp"

namespace cnd {
constexpr int kLine = kTile;  // pencils per face/sweep workgroup (one wave)
constexpr int kPts = 256;     // points per pointwise workgroup
// Every block's fields start on a multiple of this many points (2 KiB), so
// that workgroups reading or writing whole tiles of points touch whole
// cache lines of every field; the padding points are never read.
constexpr int64_t kAlignPts = 256;
// stored points from which a context runs independent sweeps side by side
constexpr int64_t kOverlapMinPts = int64_t(4) << 20;
}  // namespace cnd

// Internal header: every translation unit of libcnd works in namespace cnd.
using namespace cnd;

struct OutputState;   // output buffers and writer thread (cnd_output.hip)

// ==== context ===================================================================
// JunctionProgram on the device (one gather and one scatter launch)
struct JuncDev {
    int n = 0;
    int64_t *tgt = nullptr, *src = nullptr;
    int32_t *first = nullptr, *tc = nullptr, *sc = nullptr, *write = nullptr;
    double *scale = nullptr, *tmp = nullptr;
};

struct cnd_ctx {
    int ninst = 0, nblk = 0, precision = CND_FP64;
    std::vector<cnd_params> prm;
    std::vector<BlockSpec> spec;
    std::vector<BlkDev> hblk;
    std::vector<Tables> htab;
    std::vector<int64_t> off;
    int64_t NT = 0, FT[3] = {0, 0, 0};
    long long cells = 0;
    int64_t npts = 0;   // stored points (NT, the field stride, also counts the alignment padding)
    hipStream_t st = nullptr;
    // side stream for sweeps that do not depend on each other (eta with
    // zeta gradients, xi with eta fluxes); ev_fork/ev_join order it with st
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool overlap = true;   // side-by-side sweeps (contexts of >= kOverlapMinPts points; CANARD_OVERLAP)
    BlkDev *dblk = nullptr;
    TabDev *dtab = nullptr;
    InstDev *dinst = nullptr;
    double *dlu = nullptr;
    std::vector<Tile> htiles[3];
    Tile *dtiles[3] = {nullptr, nullptr, nullptr};
    // segmented sweeps: tiles per direction, segments per pencil (rounded to
    // whole waves; 0: the direction uses the tiled sweeps), homogeneous tables
    std::vector<Tile> hstiles[kNumP][3];            // tiles of P pencils
    Tile *dstiles[kNumP][3] = {};
    double *dseg[kNumL] = {};                       // homogeneous solutions per L
    // sweep shape per kind (CND_KIND_*) and direction: index into kShapes
    // or kTiled, with the segments per pencil of the launch (seg_sp)
    int shape[CND_NKIND][3] = {};
    int seg_sp[CND_NKIND][3] = {};
    bool tuned = false, tune_on = true, tiled_only = false;
    bool fuse_update = true;   // last viscous flux sweep writes the RK update (CANARD_FUSE_UPDATE=0: separate)
    int fuse_dir = 0;          // the direction of that sweep, 0 (xi) or 2 (zeta; CANARD_FUSE_DIR=2)
    bool recording = false;    // cnd_record: time averages accumulated in qb
    double *qb = nullptr;
    int lds_max = 65536;   // dynamic LDS per workgroup the device allows
    double tune_ms[CND_NKIND][3] = {};
    std::vector<Tile> hpt;
    Tile *dpt = nullptr;
    // de: right-hand side as three 5-field buffers, one per flux direction
    // (summed in the update); fl: the 15 stored fluxes F[5 d + m] of viscous
    // stages, scratch for metrics, single sweeps and the autotuner otherwise
    double *qa = nullptr, *qo = nullptr, *de = nullptr, *grad = nullptr, *fl = nullptr, *met = nullptr;
    double *ys = nullptr, *xyz = nullptr, *sponge = nullptr, *sp_par = nullptr;
    int nmax[3] = {0, 0, 0};   // longest pencil per direction
    bool part_other[3] = {};   // some block has another block as interface partner along d
    int64_t ys_len = 0;        // doubles in the forward-value scratch
    double *drva[3] = {}, *cmf[3] = {}, *surf[3] = {};
    // mixed precision: FP32 gradients and stresses; FP32 LU / homogeneous
    // tables and the FP64 banded LHS for the refinement
    float *gradf = nullptr, *dlu32 = nullptr, *dseg32[kNumL] = {};
    double *dlhs = nullptr;
    JuncDev jn;
    std::vector<char> grid_set;
    bool metrics_ready = false, state_in_qo = false;
    OutputState *out = nullptr;   // created by the first output call
    // the call the stage-by-stage API expects next: 0 a new step (cnd_filter
    // or cnd_step), 1..4 cnd_rk_stage(nk), 5 cnd_end_step
    int next_call = 0;
    bool any_gcbc = false, any_wall = false, any_avg = false, any_sponge = false, any_visc = false;
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    // section timers
    bool prof = false, capturing = false;
    struct Mark { int sec; hipEvent_t a, b; int launches; };
    std::vector<Mark> marks;
    std::vector<hipEvent_t> pool;
    double sec_ms[CND_NSEC] = {};
    long long sec_n[CND_NSEC] = {};
    std::string err;
};

namespace cnd {

// Record message m as the context's (or, without one, the thread's) last
// error and return code.
int fail(cnd_ctx *c, int code, const std::string &m);

#define HIPX(c, call)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return fail((c), CND_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)


template <class T>
int dalloc(cnd_ctx *c, T **p, size_t n)
{
    HIPX(c, hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    return CND_OK;
}


inline int check_launch(cnd_ctx *c)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, CND_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return CND_OK;
}

// Launch helpers ----------------------------------------------------------------
inline dim3 tiles_grid(cnd_ctx *c, int d, int y) { return dim3(unsigned(c->htiles[d].size()), unsigned(y)); }
inline dim3 sweep_grid(cnd_ctx *c, int d, int y) { return dim3(unsigned(xcd_width(int(c->htiles[d].size()))), unsigned(y)); }

// Launch a sweep kernel template K<C> with the chunk length of direction d.
#define LAUNCH_SWEEP(K, NV, c, d, y, block, ...)                                                           \
    do {                                                                                                   \
        if ((d) == 0)                                                                                      \
            hipLaunchKernelGGL(K<kChunkXi>, sweep_grid(c, d, y), dim3(block),                              \
                               (sweep_lds_bytes<NV, kChunkXi>()), (c)->st, __VA_ARGS__);                    \
        else                                                                                               \
            hipLaunchKernelGGL(K<kChunkEta>, sweep_grid(c, d, y), dim3(block),                             \
                               (sweep_lds_bytes<NV, kChunkEta>()), (c)->st, __VA_ARGS__);                   \
    } while (0)
inline dim3 pts_grid(cnd_ctx *c) { return dim3(unsigned(c->hpt.size())); }


inline int ntiles(cnd_ctx *c, int d) { return int(c->htiles[d].size()); }
// Launch the segmented sweep K<G, MIX> (NV variables) for kind along d with
// the shape chosen for them; the arguments after the common ones follow.
#define SEG_CASE(i, K, MIX, NV, c, kind, d, xm, y, ...)                                                    \
    case i: {                                                                                              \
        using G = Sh##i;                                                                                   \
        const int sp_ = (c)->seg_sp[kind][d];                                                              \
        const int pi_ = p_index(G::P);                                                                     \
        const int nt_ = int((c)->hstiles[pi_][d].size());                                                  \
        hipLaunchKernelGGL((K<G, MIX>), dim3(unsigned(xcd_width((xm) * nt_)), unsigned(y)),                \
                           dim3(unsigned((NV) * G::P * sp_)),                                              \
                           (SegShape<NV, 0, G::L, G::P, G::C, G::XI>::lds_bytes(sp_, G::SPMAX, (c)->nmax[d],     \
                                                                                (c)->part_other[d])),              \
                           (c)->st, (c)->dblk,                                                             \
                           (c)->dtab, seg_tabs(c, G::L), (c)->dstiles[pi_][d], nt_, sp_, d, __VA_ARGS__);   \
    } break;
#define LAUNCH_SEG_XM(K, MIX, NV, c, kind, d, xm, y, ...)                                                       \
    do {                                                                                                   \
        switch ((c)->shape[kind][d]) {                                                                     \
            SEG_CASE(0, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(1, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(2, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(3, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(4, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(5, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(6, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(7, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(8, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(9, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                       \
            SEG_CASE(10, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                      \
            SEG_CASE(11, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                      \
            SEG_CASE(12, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                      \
            SEG_CASE(13, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                      \
            SEG_CASE(14, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                      \
            SEG_CASE(15, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                      \
            SEG_CASE(16, K, MIX, NV, c, kind, d, xm, y, __VA_ARGS__)                                      \
            default: break;                                                                                \
        }                                                                                                  \
    } while (0)
// grid x covers xm * tiles workgroups (kernels that split a tile's work
// over xm workgroups themselves); LAUNCH_SEG: one workgroup per tile and y
#define LAUNCH_SEG(K, MIX, NV, c, kind, d, y, ...) LAUNCH_SEG_XM(K, MIX, NV, c, kind, d, 1, y, __VA_ARGS__)

inline SegTabs seg_tabs(cnd_ctx *c, int L)
{
    return SegTabs{c->dlu, c->dseg[l_index(L)], c->dlhs, c->dlu32, c->dseg32[l_index(L)]};
}
inline bool seg_on(cnd_ctx *c, int kind, int d) { return c->shape[kind][d] != kTiled; }


// Variables per workgroup of a sweep kind along d (viscous flux sweeps run
// one workgroup column per component of the stored fluxes; the eta and zeta
// gradient sweeps, one per primitive the xi sweep stored).
inline int kind_nv(const cnd_ctx *c, int kind, int d)
{
    if (kind == CND_KIND_GRAD) return d > 0 ? 1 : 4;
    return kind == CND_KIND_FLUX && !c->any_visc ? 5 : 1;
}


// Viscous stages without characteristic faces or sponge fold the update
// into the flux sweep of direction fuse_dir, run after the other two, when
// that sweep is segmented (GCBC needs all three flux derivatives before the
// update; the sponge stays in k_update).
inline bool fuse_possible(const cnd_ctx *c)
{
    return c->fuse_update && c->any_visc && !c->any_gcbc && !c->any_sponge;
}


// Sweep launches, one translation unit per sweep kind (cnd_run_*.hip).
// Filter (kind CND_KIND_FILTER) and plain derivative (CND_KIND_FIELD) sweeps
// of nvar fields; mix: FP32 solves with FP64 refinement (segmented only).
int run_filter_field(cnd_ctx *c, int d, const double *in, double *out, int nvar, bool mix = false);
int run_fd_field(cnd_ctx *c, int d, int n45, const double *in, int64_t ivs, double *out, int64_t ovs, int nvar,
                 bool mix = false);
// Gradient sweep along d of the state cur (kind CND_KIND_GRAD).  The xi
// sweep (d = 0) also stores the primitives u, v, w, T in the flux scratch
// fl, and the eta and zeta sweeps (segmented) differentiate those: the
// three must run in order, d = 0 first.
int run_grad(cnd_ctx *c, int d, const double *cur);
// Flux sweep along d into the d-th right-hand-side buffer (kind
// CND_KIND_FLUX).
int run_flux(cnd_ctx *c, int d, const double *cur);
// Flux sweep of the stored fluxes along fuse_dir (xi, or zeta) fused with the
// RK update of stage nk, written to qa; the other two directions' buffers
// must be complete.
int run_flux_update(cnd_ctx *c, int nk, double *qa);

// Joins an output write in flight and frees the output buffers (cnd_destroy).
void output_release(cnd_ctx *c);

}  // namespace cnd
