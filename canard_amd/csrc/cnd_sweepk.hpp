// Sources, sinks, sweep-shape table and sweep kernel templates of libcnd.
// Each sweep kind's launches (and so its template instantiations) live in a
// translation unit of their own (cnd_run_*.hip), so they compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "cnd_dev.hpp"
#include "cnd_kernels.hpp"
#include "cnd_seg.hpp"
#include "cnd_sweep.hpp"

namespace cnd {

// ---- sources (values a sweep differentiates or filters) ---------------------
template <int NV>
struct SrcField {   // NV plain fields, vs apart
    const double *a;
    int64_t vs;
    struct Raw { double v[NV]; };
    static constexpr bool kDirect = true;   // variable v reads field v only
    static constexpr bool kEmit = false;    // writes no copy of its values
    __device__ double field(int64_t g, int v) const { return a[v * vs + g]; }
    __device__ void load(int64_t g, int, Raw &r) const
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) r.v[k] = a[k * vs + g];
    }
    __device__ void finish(const Raw &r, int, double (&v)[NV]) const
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = r.v[k];
    }
};

// u, v, w, T of a state, exactly as the reference's primitives step
// (src/mo_canard_driver.f90:267-274).  With po set, xi sweeps also write the
// values they stage to po (four fields NT apart), for the eta and zeta
// sweeps to differentiate as plain fields.
struct SrcPrim {
    const double *q;
    int64_t NT;
    double *po = nullptr;
    struct Raw { double q[5]; };
    static constexpr bool kDirect = false;
    static constexpr bool kEmit = true;
    __device__ double field(int64_t, int) const { return 0.0; }
    __device__ void load(int64_t g, int, Raw &r) const
    {
#pragma unroll
        for (int m = 0; m < 5; ++m) r.q[m] = q[m * NT + g];
    }
    __device__ void finish(const Raw &r, int, double (&o)[4]) const
    {
        const double ri = 1.0 / r.q[0];
        const double q2 = r.q[1], q3 = r.q[2], q4 = r.q[3];
        const double u = q2 * ri, v = q3 * ri, w = q4 * ri;
        const double p = kGm1 * (r.q[4] - 0.5 * ((q2 * u + q3 * v) + q4 * w));
        o[0] = u; o[1] = v; o[2] = w;
        o[3] = (kGam * p) * ri;
    }
};

// The five flux components of direction d at one point
// (src/mo_physics.f90:319-382): the inviscid part from the state q1..q5 and
// the metric row m of d, then, with stresses t (txx, tyy, tzz, txy, tyz,
// tzx and the three heat-flux rows), the viscous part.
template <bool VIS>
__device__ __forceinline__ void dir_flux(double q1, double q2, double q3, double q4, double q5, double u, double v,
                                         double w, double p, const double (&m)[3], const double *umf,
                                         const double *t, double (&o)[5])
{
    const double m0 = m[0], m1 = m[1], m2 = m[2];
    const double s = (m0 * (u + umf[0]) + m1 * (v + umf[1])) + m2 * (w + umf[2]);
    o[0] = q1 * s;
    o[1] = q2 * s + m0 * p;
    o[2] = q3 * s + m1 * p;
    o[3] = q4 * s + m2 * p;
    o[4] = (q5 + p) * s - p * ((umf[0] * m0 + umf[1] * m1) + umf[2] * m2);
    if constexpr (VIS) {
        // stress rows x (txx, txy, tzx), y (txy, tyy, tyz), z (tzx, tyz, tzz)
        const double txx = t[0], tyy = t[1], tzz = t[2], txy = t[3], tyz = t[4], tzx = t[5];
        o[1] = o[1] - m0 * txx; o[1] = o[1] - m1 * txy; o[1] = o[1] - m2 * tzx;
        o[2] = o[2] - m0 * txy; o[2] = o[2] - m1 * tyy; o[2] = o[2] - m2 * tyz;
        o[3] = o[3] - m0 * tzx; o[3] = o[3] - m1 * tyz; o[3] = o[3] - m2 * tzz;
        o[4] = o[4] - m0 * t[6]; o[4] = o[4] - m1 * t[7]; o[4] = o[4] - m2 * t[8];
    }
}

// Inviscid fluxes of direction d evaluated in the sweep from q and the
// metric row of d (viscous runs differentiate the stored fluxes instead).
struct SrcFlux {
    const double *q;
    const double *met;
    const BlkDev *blks;
    const InstDev *inst;
    int64_t NT;
    int d;
    struct Raw { double q[5], m[3]; };
    static constexpr bool kDirect = false;
    static constexpr bool kEmit = false;
    __device__ double field(int64_t, int) const { return 0.0; }
    __device__ void load(int64_t g, int, Raw &r) const
    {
#pragma unroll
        for (int k = 0; k < 5; ++k) r.q[k] = q[k * NT + g];
#pragma unroll
        for (int k = 0; k < 3; ++k) r.m[k] = met[(3 * d + k) * NT + g];
    }
    __device__ void finish(const Raw &r, int b, double (&o)[5]) const
    {
        const double q1 = r.q[0], q2 = r.q[1], q3 = r.q[2], q4 = r.q[3], q5 = r.q[4];
        const double ri = 1.0 / q1;
        const double u = q2 * ri, v = q3 * ri, w = q4 * ri;
        const double p = kGm1 * (q5 - 0.5 * ((q2 * u + q3 * v) + q4 * w));
        dir_flux<false>(q1, q2, q3, q4, q5, u, v, w, p, r.m, inst[blks[b].inst].umf, nullptr, o);
    }
};

// ---- sinks ----------------------------------------------------------------
// store / put write results; fold(f, x) is the value to write for source f
// and result x, written with put_raw by sweeps that still hold the source in
// registers (the filter adds its correction to the source).
template <int NV, class TO = double>
struct SinkOut {   // out fields vs apart, stored as TO
    TO *o;
    int64_t vs;
    __device__ void store(int64_t g, int, const double (&x)[NV])
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) o[k * vs + g] = TO(x[k]);
    }
    __device__ void put(int64_t g, int, int k, double x) { o[k * vs + g] = TO(x); }
    static constexpr bool kReadBack = false;
    __device__ double src(int64_t, int) const { return 0.0; }
    __device__ double fold(double, double x) const { return x; }
    __device__ void put_raw(int64_t g, int b, int k, double x) { put(g, b, k, x); }
    __device__ void ends(int, int, int, double) {}
};

template <int NV>
struct SinkFilt {  // out = f + correction
    double *o;
    const double *f;
    int64_t vs;
    __device__ void store(int64_t g, int, const double (&x)[NV])
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) o[k * vs + g] = f[k * vs + g] + x[k];
    }
    __device__ void put(int64_t g, int, int k, double x) { o[k * vs + g] = f[k * vs + g] + x; }
    static constexpr bool kReadBack = true;   // results are added to the sources
    __device__ double src(int64_t g, int k) const { return f[k * vs + g]; }
    __device__ double fold(double fv, double x) const { return fv + x; }
    __device__ void put_raw(int64_t g, int, int k, double x) { o[k * vs + g] = x; }
    __device__ void ends(int, int, int, double) {}
};

struct SinkFlux {  // the five components of dF/dxi_d; pencil ends to drva
    double *de, *drva;
    int64_t NT, FT, fi0;
    __device__ void store(int64_t g, int, const double (&x)[5])
    {
#pragma unroll
        for (int m = 0; m < 5; ++m) de[m * NT + g] = x[m];
    }
    __device__ void put(int64_t g, int, int m, double x) { de[m * NT + g] = x; }
    static constexpr bool kReadBack = false;
    __device__ double src(int64_t, int) const { return 0.0; }
    __device__ double fold(double, double x) const { return x; }
    __device__ void put_raw(int64_t g, int b, int m, double x) { put(g, b, m, x); }
    __device__ void ends(int side, int lane, int m, double x) { drva[(side * 5 + m) * FT + fi0 + lane] = x; }
};

struct SinkFluxV {  // component v of dF/dxi_d; pencil ends to drva
    double *o, *drva;
    int v;
    int64_t FT, fi0;
    __device__ void store(int64_t g, int, const double (&x)[1]) { o[g] = x[0]; }
    __device__ void put(int64_t g, int, int, double x) { o[g] = x; }
    static constexpr bool kReadBack = false;
    __device__ double src(int64_t, int) const { return 0.0; }
    __device__ double fold(double, double x) const { return x; }
    __device__ void put_raw(int64_t g, int b, int k, double x) { put(g, b, k, x); }
    __device__ void ends(int side, int lane, int, double x) { drva[(side * 5 + v) * FT + fi0 + lane] = x; }
};

// Component v of the last flux derivative folded straight into the RK update
// (k_update's arithmetic without the sponge, same operation order, so results
// are bit-identical): qa = qo - dt/(5-nk) * yaco * ((de_xi + de_eta) + de_zeta),
// where the sweep's own result x is de_zeta (XFIRST false: da, db the xi and
// eta buffers) or de_xi (XFIRST true: da, db the eta and zeta buffers).
// Every operand is variable v at the sink's own point.  Pre holds one point's
// operands, so sweeps can load them for several points before storing any.
template <bool XFIRST>
struct SinkUpdateV {
    double *qa, *drva;
    const double *qo, *da, *db, *met;
    double dtk;   // dt / (5 - nk) of the tile's instance
    int v;
    int64_t NT, FT, fi0;
    struct Pre { double a, b, q, r; };
    __device__ void pre(int64_t g, Pre &o) const
    {
        o.a = da[g];
        o.b = db[g];
        o.q = qo[g];
        o.r = dtk * met[9 * NT + g];
    }
    __device__ void put_pre(int64_t g, double x, const Pre &o)
    {
        const double s = XFIRST ? (x + o.a) + o.b : (o.a + o.b) + x;
        qa[g] = o.q - o.r * s;
    }
    __device__ void put(int64_t g, int, int, double x)
    {
        Pre o;
        pre(g, o);
        put_pre(g, x, o);
    }
    __device__ void store(int64_t g, int b, const double (&x)[1]) { put(g, b, 0, x[0]); }
    static constexpr bool kReadBack = false;
    __device__ double src(int64_t, int) const { return 0.0; }
    __device__ double fold(double, double x) const { return x; }
    __device__ void put_raw(int64_t g, int b, int k, double x) { put(g, b, k, x); }
    __device__ void ends(int side, int lane, int, double x) { drva[(side * 5 + v) * FT + fi0 + lane] = x; }
};

// ---- sweep kernels -----------------------------------------------------------
// grid: (xcd_width(tiles of direction d), variable groups); NV waves per tile
// of 64 pencils; C = chunk_of(d); dynamic LDS = sweep_lds_bytes<NV, C>().
#define SWEEP_TILE_PROLOGUE(ntiles)        \
    const int tix = xcd_tile(ntiles);      \
    if (tix >= (ntiles)) return;           \
    extern __shared__ double lds[];

template <int C>
__global__ __launch_bounds__(kTile) void k_fd_field(const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs,
                                                   const double *__restrict__ lu, const Tile *__restrict__ tiles,
                                                   int ntiles, int d, int n45, int nmax, double *ys,
                                                   const double *in, int64_t in_vstride, double *out,
                                                   int64_t out_vstride)
{
    SWEEP_TILE_PROLOGUE(ntiles)
    const int v = blockIdx.y;
    SrcField<1> src{in + v * in_vstride, 0};
    SinkOut<1> sink{out + v * out_vstride, 0};
    sweep_tile<1, 0, C>(blks, tabs, lu, tiles[tix], int64_t(v) * ntiles + tix, d, n45, nmax, ys, src, sink, lds);
}

template <int C>
__global__ __launch_bounds__(kTile) void k_filter_field(const BlkDev *__restrict__ blks,
                                                       const TabDev *__restrict__ tabs,
                                                       const double *__restrict__ lu,
                                                       const Tile *__restrict__ tiles, int ntiles, int d, int nmax,
                                                       double *ys, const double *in, double *out, int64_t vstride)
{
    SWEEP_TILE_PROLOGUE(ntiles)
    const int v = blockIdx.y;
    SrcField<1> src{in + v * vstride, 0};
    SinkFilt<1> sink{out + v * vstride, in + v * vstride, 0};
    sweep_tile<1, 1, C>(blks, tabs, lu, tiles[tix], int64_t(v) * ntiles + tix, d, 0, nmax, ys, src, sink, lds);
}

// Velocity and temperature gradients along d (the 12 derivative sweeps of
// calc_viscous_shear_stress, src/mo_physics.f90:238-280).
template <int C>
__global__ __launch_bounds__(4 * kTile) void k_fd_grad(const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs,
                                                  const double *__restrict__ lu, const Tile *__restrict__ tiles,
                                                  int ntiles, int d, int nmax, double *ys, const double *q,
                                                  double *grad, int64_t NT)
{
    SWEEP_TILE_PROLOGUE(ntiles)
    SrcPrim src{q, NT};
    SinkOut<4> sink{grad + 4 * d * NT, NT};
    sweep_tile<4, 0, C>(blks, tabs, lu, tiles[tix], tix, d, 0, nmax, ys, src, sink, lds);
}

// Flux derivatives along d (src/mo_physics.f90:385-399) into the d-th
// right-hand-side buffer de: inviscid fluxes evaluated in the sweep ...
template <int C>
__global__ __launch_bounds__(5 * kTile) void k_fd_flux(const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs,
                                                  const InstDev *__restrict__ inst, const double *__restrict__ lu,
                                                  const Tile *__restrict__ tiles, int ntiles, int d, int nmax,
                                                  double *ys, const double *q, const double *met, double *de,
                                                  double *drva, int64_t NT, int64_t FT)
{
    SWEEP_TILE_PROLOGUE(ntiles)
    const Tile t = tiles[tix];
    SrcFlux src{q, met, blks, inst, NT, d};
    SinkFlux sink{de, drva, NT, FT, blks[t.b].foff[d] + t.q0};
    sweep_tile<5, 0, C>(blks, tabs, lu, t, tix, d, 0, nmax, ys, src, sink, lds);
}

// ... or stored fluxes F (component v = blockIdx.y), the viscous case.
template <int C>
__global__ __launch_bounds__(kTile) void k_fd_flux_field(const BlkDev *__restrict__ blks,
                                                        const TabDev *__restrict__ tabs,
                                                        const double *__restrict__ lu, const Tile *__restrict__ tiles,
                                                        int ntiles, int d, int nmax, double *ys, const double *F,
                                                        double *de, double *drva, int64_t NT, int64_t FT)
{
    SWEEP_TILE_PROLOGUE(ntiles)
    const int v = blockIdx.y;
    const Tile t = tiles[tix];
    SrcField<1> src{F + v * NT, 0};
    SinkFluxV sink{de + v * NT, drva, v, FT, blks[t.b].foff[d] + t.q0};
    sweep_tile<1, 0, C>(blks, tabs, lu, t, int64_t(v) * ntiles + tix, d, 0, nmax, ys, src, sink, lds);
}

// ---- segmented sweeps (cnd_seg.hpp) ---------------------------------------
// Compiled shapes: L points per segment (lanes per pencil = segments =
// ceil(n / L)), P pencils per workgroup, C-point staging runs, at most SPMAX
// segments per pencil; XI shapes stage runs along the pencil (xi sweeps),
// the others across P neighbouring pencils (eta, zeta sweeps).  The
// autotuner picks one per sweep kind and direction.
template <int L_, int P_, int C_, int SPMAX_, bool XI_>
struct SegG { static constexpr int L = L_, P = P_, C = C_, SPMAX = SPMAX_; static constexpr bool XI = XI_; };
struct ShapeInfo { int L, P, C, SPMAX; bool XI; };
constexpr int kNumShapes = 17;
constexpr ShapeInfo kShapes[kNumShapes] = {
    {17, 4, 16, 16, true},    // 0: xi default (pencils up to 272 points)
    {33, 4, 16, 8, true},     // 1
    {33, 16, 8, 8, false},    // 2: eta/zeta default (up to 264 points)
    {33, 16, 4, 8, false},    // 3
    {17, 8, 8, 16, false},    // 4
    {33, 4, 16, 16, true},    // 5: long xi pencils (up to 528 points)
    {33, 8, 8, 16, false},    // 6: long eta/zeta pencils (up to 528 points)
    {17, 8, 16, 16, true},    // 7: xi, wider tiles (single-field sweeps)
    {17, 16, 16, 16, true},   // 8
    {17, 16, 8, 16, false},   // 9: eta/zeta, short segments, wide tiles
    {33, 32, 8, 8, false},    // 10: eta/zeta, runs of 32 and 64 points
    {33, 64, 8, 8, false},    // 11
    {17, 32, 8, 16, false},   // 12
    {17, 64, 8, 16, false},   // 13
    {17, 32, 16, 16, true},   // 14: xi, 32-pencil tiles (single-field sweeps)
    {17, 20, 16, 16, true},   // 15: xi, tiles sized for three workgroups per CU
    {17, 24, 16, 16, true}};  // 16
using Sh0 = SegG<17, 4, 16, 16, true>;
using Sh1 = SegG<33, 4, 16, 8, true>;
using Sh2 = SegG<33, 16, 8, 8, false>;
using Sh3 = SegG<33, 16, 4, 8, false>;
using Sh4 = SegG<17, 8, 8, 16, false>;
using Sh5 = SegG<33, 4, 16, 16, true>;
using Sh6 = SegG<33, 8, 8, 16, false>;
using Sh7 = SegG<17, 8, 16, 16, true>;
using Sh8 = SegG<17, 16, 16, 16, true>;
using Sh9 = SegG<17, 16, 8, 16, false>;
using Sh10 = SegG<33, 32, 8, 8, false>;
using Sh11 = SegG<33, 64, 8, 8, false>;
using Sh12 = SegG<17, 32, 8, 16, false>;
using Sh13 = SegG<17, 64, 8, 16, false>;
using Sh14 = SegG<17, 32, 16, 16, true>;
using Sh15 = SegG<17, 20, 16, 16, true>;
using Sh16 = SegG<17, 24, 16, 16, true>;
// workgroup size bound of a kernel (shapes whose NV * P * SPMAX exceeds the
// 1024-thread limit are rejected at launch time by shape_sp)
__host__ __device__ constexpr int wg_bound(int x) { return x > 1024 ? 1024 : x; }
// waves per SIMD a segmented sweep is compiled for: shapes of 17-point
// segments are held to 128 VGPRs (4 waves, so two 512-thread workgroups
// share a CU), which their LDS also allows -- eta/zeta ones in FP64 only,
// where it costs no spills; the others are left to the compiler
template <class G, bool MIX>
__host__ __device__ constexpr int seg_waves() { return G::L <= 17 && (G::XI || !MIX) ? 4 : 1; }
constexpr int kTiled = -1;   // shape index of the tiled sweep
// pencils-per-workgroup values (tile lists) and segment lengths (tables)
constexpr int kNumP = 7, kPs[kNumP] = {4, 8, 16, 32, 64, 20, 24};
constexpr int kNumL = 2, kLs[kNumL] = {17, 33};
__host__ __device__ constexpr int p_index(int P)
{
    return P == 4 ? 0 : (P == 8 ? 1 : (P == 16 ? 2 : (P == 32 ? 3 : (P == 64 ? 4 : (P == 20 ? 5 : 6)))));
}
__host__ __device__ constexpr int l_index(int L) { return L == 17 ? 0 : 1; }

#define SEG_PROLOGUE(ntiles)               \
    const int tix = xcd_tile(ntiles);      \
    if (tix >= (ntiles)) return;           \
    extern __shared__ double lds[];

template <class G, bool MIX>
__global__ __launch_bounds__(wg_bound(G::P * G::SPMAX), (seg_waves<G, MIX>())) void k_seg_fd_field(
    const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs, SegTabs tb, const Tile *__restrict__ tiles,
    int ntiles, int SP, int d, int n45, const double *in, int64_t in_vstride, double *out, int64_t out_vstride)
{
    SEG_PROLOGUE(ntiles)
    const int v = blockIdx.y;
    SrcField<1> src{in + v * in_vstride, 0};
    SinkOut<1> sink{out + v * out_vstride, 0};
    seg_tile<1, 0, G::L, G::P, G::C, G::XI, MIX, G::SPMAX>(blks, tabs, tb, tiles[tix], d, n45, SP, src, sink, lds);
}

template <class G, bool MIX>
__global__ __launch_bounds__(wg_bound(G::P * G::SPMAX), (seg_waves<G, MIX>())) void k_seg_filter_field(
    const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs, SegTabs tb, const Tile *__restrict__ tiles,
    int ntiles, int SP, int d, const double *in, double *out, int64_t vstride)
{
    SEG_PROLOGUE(ntiles)
    const int v = blockIdx.y;
    SrcField<1> src{in + v * vstride, 0};
    SinkFilt<1> sink{out + v * vstride, in + v * vstride, 0};
    seg_tile<1, 1, G::L, G::P, G::C, G::XI, MIX, G::SPMAX>(blks, tabs, tb, tiles[tix], d, 0, SP, src, sink, lds);
}

// Velocity and temperature gradients along d; in mixed mode they are stored
// in FP32.  prim (xi sweeps, or null): the primitives u, v, w, T are also
// written there.
template <class G, bool MIX>
__global__ __launch_bounds__(wg_bound(4 * G::P * G::SPMAX), (seg_waves<G, MIX>())) void k_seg_grad(
    const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs, SegTabs tb, const Tile *__restrict__ tiles,
    int ntiles, int SP, int d, const double *q, typename std::conditional<MIX, float, double>::type *grad,
    int64_t NT, double *prim)
{
    SEG_PROLOGUE(ntiles)
    using TG = typename std::conditional<MIX, float, double>::type;
    SrcPrim src{q, NT, prim};
    SinkOut<4, TG> sink{grad + 4 * d * NT, NT};
    seg_tile<4, 0, G::L, G::P, G::C, G::XI, MIX, G::SPMAX>(blks, tabs, tb, tiles[tix], d, 0, SP, src, sink, lds);
}

// Gradients along d (eta, zeta) of the primitive v = blockIdx.y that the xi
// gradient sweep stored (prim, four fields NT apart), into grad; in mixed
// mode FP32.
template <class G, bool MIX>
__global__ __launch_bounds__(wg_bound(G::P * G::SPMAX), (seg_waves<G, MIX>())) void k_seg_grad_prim(
    const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs, SegTabs tb, const Tile *__restrict__ tiles,
    int ntiles, int SP, int d, const double *prim, typename std::conditional<MIX, float, double>::type *grad,
    int64_t NT)
{
    SEG_PROLOGUE(ntiles)
    using TG = typename std::conditional<MIX, float, double>::type;
    const int v = blockIdx.y;
    SrcField<1> src{prim + v * NT, 0};
    SinkOut<1, TG> sink{grad + (4 * d + v) * NT, 0};
    seg_tile<1, 0, G::L, G::P, G::C, G::XI, MIX, G::SPMAX>(blks, tabs, tb, tiles[tix], d, 0, SP, src, sink, lds);
}

template <class G, bool MIX>
__global__ __launch_bounds__(wg_bound(5 * G::P * G::SPMAX), (seg_waves<G, MIX>())) void k_seg_flux(
    const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs, SegTabs tb, const Tile *__restrict__ tiles,
    int ntiles, int SP, int d, const InstDev *__restrict__ inst, const double *q, const double *met, double *de,
    double *drva, int64_t NT, int64_t FT)
{
    SEG_PROLOGUE(ntiles)
    const Tile t = tiles[tix];
    SrcFlux src{q, met, blks, inst, NT, d};
    SinkFlux sink{de, drva, NT, FT, blks[t.b].foff[d] + t.q0};
    seg_tile<5, 0, G::L, G::P, G::C, G::XI, MIX, G::SPMAX>(blks, tabs, tb, t, d, 0, SP, src, sink, lds);
}

template <class G, bool MIX>
__global__ __launch_bounds__(wg_bound(G::P * G::SPMAX), (seg_waves<G, MIX>())) void k_seg_flux_field(
    const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs, SegTabs tb, const Tile *__restrict__ tiles,
    int ntiles, int SP, int d, const double *F, double *de, double *drva, int64_t NT, int64_t FT)
{
    SEG_PROLOGUE(ntiles)
    const int v = blockIdx.y;
    const Tile t = tiles[tix];
    SrcField<1> src{F + v * NT, 0};
    SinkFluxV sink{de + v * NT, drva, v, FT, blks[t.b].foff[d] + t.q0};
    seg_tile<1, 0, G::L, G::P, G::C, G::XI, MIX, G::SPMAX>(blks, tabs, tb, t, d, 0, SP, src, sink, lds);
}

// Last flux sweep of a stage fused with the RK update: sweep of the stored
// flux component v along d (xi or zeta), written as the updated state qa.
// Grid x: five workgroups (components) per tile.
template <class G, bool MIX>
__global__ __launch_bounds__(wg_bound(G::P * G::SPMAX), (seg_waves<G, MIX>())) void k_seg_flux_update(
    const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs, SegTabs tb, const Tile *__restrict__ tiles,
    int ntiles, int SP, int d, const InstDev *__restrict__ inst, const double *F, const double *de,
    const double *__restrict__ qo, const double *__restrict__ met, double *qa, double *drva, int64_t NT, int64_t FT,
    int nk)
{
    // the five component workgroups of a tile follow each other on one XCD,
    // so the Jacobian they all read is fetched once into that XCD's L2
    const int w = xcd_tile(5 * ntiles);
    if (w >= 5 * ntiles) return;
    extern __shared__ double lds[];
    const int tix = w / 5, v = w % 5;
    const Tile t = tiles[tix];
    SrcField<1> src{F + v * NT, 0};
    const double dtk = inst[blks[t.b].inst].dt / double(4 - nk + 1);
    // the two other directions' right-hand sides
    const double *da = de + ((G::XI ? 5 : 0) + v) * NT, *db = de + ((G::XI ? 10 : 5) + v) * NT;
    SinkUpdateV<G::XI> sink{qa + v * NT, drva, qo + v * NT, da, db, met, dtk, v, NT, FT, blks[t.b].foff[d] + t.q0};
    seg_tile<1, 0, G::L, G::P, G::C, G::XI, MIX, G::SPMAX>(blks, tabs, tb, t, d, 0, SP, src, sink, lds);
}

}  // namespace cnd
