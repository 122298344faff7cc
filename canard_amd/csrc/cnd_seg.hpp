// Segmented pencil sweeps: the pentadiagonal solves of the compact
// derivative (src/mo_numerics.f90:709-886) and compact filter
// (src/mo_numerics.f90:890-987) with every pencil split into S segments of
// at most L points, one lane per (pencil, segment, variable), so a lane keeps
// its whole segment in registers and the sweep reads each source point once
// and writes each result once (no forward-value scratch in HBM).
//
// The LU factors of the line (xl, xu of init_penta/penta) are the ones the
// sequential sweep uses.  Each lane runs each recurrence twice over its own
// rows: first with zero incoming values, which gives the two values the
// segment would pass on; a short scan over the segments of the pencil
// through LDS (a 2x2 transfer per segment, built from the homogeneous
// solutions phi1, phi2 / psi1, psi2 at the segment's ends) turns them into
// the true incoming values; the second run starts from those and is the
// sequential recurrence itself, so the result equals it up to rounding.
// Segments are held right-aligned in L registers; the registers in front of
// a shorter segment are pass-through rows (LU coefficients that swap the
// recurrence state exactly), so every lane runs the same branch-free code.
//
// Workgroup: NV * P * SP threads.  Thread t: variable v = t / (P * SP),
// pencil p = t % P, segment s = (t / P) % SP.  Sources reach the lanes in
// one of three ways: xi tiles (P contiguous pencils) are staged whole into
// LDS with consecutive threads on consecutive points; eta/zeta sweeps of
// plain fields load each lane's rows directly (the lanes of a wave hold P
// neighbouring pencils, so every load reads runs of P consecutive points);
// eta/zeta sweeps of derived sources (primitives, inviscid fluxes) evaluate
// them cooperatively chunk by chunk (C rows of every segment) into LDS and
// each lane picks its variable.  The lanes holding a pencil end keep the 13
// source values next to it in LDS (xi: they stay in the staged tile); the
// closure rows are formed from them after the stream (and, for a block that
// is its own periodic partner, the partner sums too, so no extra loads).
#pragma once

#include <type_traits>

#include "cnd_dev.hpp"
#include "cnd_kernels.hpp"

namespace cnd {

// Points of a pencil end used by the closure rows (12) plus the face point.
constexpr int kSegEnd = 12;
constexpr int kSegEndRows = kSegEnd + 1;

// Fewest rows of a segment (seg_count) and of a whole pencil (13).
constexpr int kSegMinRows = 12;

// LU values a thread stages per round trip.
constexpr int kLuBatch = 12;

// Points whose sources a thread reads back together at a xi store.
constexpr int kOutBatch = 8;

// Segments of an n-point pencil for lanes of at most L points: 0 when no
// partition fits (a segment holding a pencil end must contain the 12 points
// its closure rows use).  Segments are as even as possible.
__host__ __device__ inline int seg_count(int n, int L)
{
    const int S = (n + L - 1) / L;
    if (S > 1 && n / S < kSegMinRows) return 0;
    return S;
}

__host__ __device__ inline void seg_bounds(int n, int, int S, int s, int &a, int &len)
{
    const int q = n / S, r = n % S;
    a = s * q + (s < r ? s : r);
    len = q + (s < r ? 1 : 0);
}


template <int NV, int SCH, int L, int P, int C, bool XI>
struct SegShape {
    static constexpr int H = SCH == 0 ? 2 : 3;          // stencil half width
    static constexpr int NB = SCH == 0 ? 2 : 3;         // closure rows per end
    static constexpr int KR = L + 2 * H;                // rows streamed per lane
    static constexpr int NCH = (KR + C - 1) / C;        // staging chunks
    // xi tiles in LDS: pencil after pencil, each n | 1 points long (an odd
    // stride puts the lanes of neighbouring pencils in different banks),
    // variable after variable; for odd n a point's place is its index in
    // the tile
    __host__ __device__ static constexpr int pencil_stride(int n) { return n | 1; }
    // LDS: a work area (chunk buffer [C][NV][SP][P], then partner points
    // [2][13][NV][P], then the exchange buffer [4][NV][SP][P]) and the
    // pencil-end values [2][13][NV][P].
    __host__ __device__ static constexpr int work_doubles(int SP)
    {
        const int chunk = C * NV * SP * P, part = 2 * kSegEndRows * NV * P, ex = 4 * NV * SP * P;
        const int m = chunk > part ? chunk : part;
        return m > ex ? m : ex;
    }
    // then the segment transfer values [SPM][8] and the LU rows of the
    // launch's SP segments as their lanes hold them [SP][L][5]
    __host__ __device__ static constexpr int ends_doubles() { return 2 * kSegEndRows * NV * P; }
    // (xi: the whole tile's sources [NV][P][nmax | 1] replace the work area; once
    // the closure rows are formed, the LU rows and then the exchange buffer
    // [4][NV][SP][P] overlay the tile; the partner points [2][13][NV][P]
    // come last and only when has_part, some block having another block as
    // its partner along the sweep direction.)
    __host__ __device__ static constexpr size_t lds_bytes(int SP, int spm, int nmax, bool has_part)
    {
        const size_t tile = size_t(NV) * P * pencil_stride(nmax), ex4 = size_t(4) * NV * SP * P,
                     lu = size_t(5) * L * SP;
        const size_t over = lu + ex4;
        if (XI)
            return sizeof(double) * ((tile > over ? tile : over) + size_t(8) * spm +
                                     (has_part ? size_t(ends_doubles()) : size_t(0)));
        return sizeof(double) * (size_t(work_doubles(SP)) + lu + size_t(ends_doubles()) + size_t(8) * spm);
    }
};

// Chunk slot -> (o, s, p): xi runs along the pencil, eta/zeta across pencils.
template <int P, int C, bool XI>
__device__ __forceinline__ void seg_slot(int slot, int SP, int &o, int &s, int &p)
{
    if (XI) { o = slot % C; p = (slot / C) % P; s = slot / (C * P); }
    else { p = slot % P; s = (slot / P) % SP; o = slot / (P * SP); }
}

// xi sweeps (XI shapes run along xi only): the P pencils of a tile are
// consecutive rows of the block, so point e of the tile (pencil e / n, row
// e % n) sits at g0 + e, g0 the tile's first point; the LDS tile keeps the
// same order (pencils n | 1 points apart), so staging and stores address
// both by e.

// One lane's view of its segment: rows a .. a + len - 1 of the line, held
// right-aligned in Y[kb .. L - 1] (kb = L - len); Y[0 .. kb - 1] are
// pass-through registers.
struct SegLane {
    int a, len, kb, s, S, n, L, ok, exi, exs, v, p, P, SP;
};

// LU row of a pass-through register: the forward step y = (Y - l0 ya) - l1 yb
// with Y = 0 gives y = ya, the backward step x = (l2 Y - l3 xa) - l4 xb
// gives x = xb; either way the recurrence state (two values) is swapped
// exactly, and two such rows leave it unchanged.
__host__ __device__ constexpr double seg_pass_row(int c) { return c == 0 || c == 4 ? -1.0 : 0.0; }

// Solve the line for the lane's segment held in Y (right-hand side in,
// solution out; Y[0 .. kb - 1] must be zero) in arithmetic T.  LUs: the
// LU rows of every segment as the lanes hold them, [SP][L][5] in LDS
// (pass-through rows in front); SM: per-segment transfer values in LDS
// (phi1, phi2 at the segment's last two rows, psi1, psi2 at its first two);
// ex: LDS room for 2 x NV x SP x P doubles.  All threads must call it.
//
// Each recurrence runs twice over the lane's registers: first with zero
// incoming values, keeping only the two values that leave the segment;
// after the scan over the segments has turned those into the true incoming
// values, again from them, which is the sequential recurrence itself.  All
// passes run over all L registers with no branches: the segment's last row
// is always register L - 1, and the kb pass-through registers carry the
// state across the front of the segment (swapped kb times, which the start
// or end values undo).
template <class T, int L, int SPM>
__device__ __forceinline__ void seg_solve(T (&Y)[L], const SegLane &ln, const double *LUs, const double *SM,
                                          double *ex)
{
    const int s = ln.s;
    const bool odd = (ln.kb & 1) != 0;
    const double *lu = LUs + 5 * L * s;
    // forward, pass 1: zero incoming values, outgoing values only
    {
        T ya = T(0), yb = T(0);
#pragma unroll
        for (int k = 0; k < L; ++k) {
            const T y = (Y[k] - T(lu[5 * k]) * ya) - T(lu[5 * k + 1]) * yb;
            ya = yb;
            yb = y;
        }
        __syncthreads();
        ex[ln.exi] = double(ya);
        ex[ln.exs + ln.exi] = double(yb);
        __syncthreads();
    }
    // incoming values y(a-2), y(a-1): scan over the segments before this one
    T t0 = T(0), t1 = T(0);
    if (ln.ok) {
        for (int sp = 0; sp < s; ++sp) {
            const double *m = SM + 8 * sp;
            const int xi = ln.exi + (sp - s) * ln.P;
            const T n0 = (T(m[0]) * t0 + T(m[1]) * t1) + T(ex[xi]);
            const T n1 = (T(m[2]) * t0 + T(m[3]) * t1) + T(ex[ln.exs + xi]);
            t0 = n0;
            t1 = n1;
        }
    }
    // forward, pass 2: from the true incoming values
    {
        T ya = odd ? t1 : t0, yb = odd ? t0 : t1;
#pragma unroll
        for (int k = 0; k < L; ++k) {
            const T y = (Y[k] - T(lu[5 * k]) * ya) - T(lu[5 * k + 1]) * yb;
            Y[k] = y;
            ya = yb;
            yb = y;
        }
    }
    // backward, pass 1: zero incoming values, outgoing values only
    {
        T xa = T(0), xb = T(0);
#pragma unroll
        for (int k = L - 1; k >= 0; --k) {
            const T x = (T(lu[5 * k + 2]) * Y[k] - T(lu[5 * k + 3]) * xa) - T(lu[5 * k + 4]) * xb;
            xb = xa;
            xa = x;
        }
        __syncthreads();
        ex[ln.exi] = double(odd ? xb : xa);
        ex[ln.exs + ln.exi] = double(odd ? xa : xb);
        __syncthreads();
    }
    // incoming values x(b), x(b+1): scan over the segments after this one
    T h0 = T(0), h1 = T(0);
    if (ln.ok) {
        for (int sp = ln.S - 1; sp > s; --sp) {
            const double *m = SM + 8 * sp + 4;
            const int xi = ln.exi + (sp - s) * ln.P;
            const T n0 = (T(m[0]) * h0 + T(m[1]) * h1) + T(ex[xi]);
            const T n1 = (T(m[2]) * h0 + T(m[3]) * h1) + T(ex[ln.exs + xi]);
            h0 = n0;
            h1 = n1;
        }
    }
    // backward, pass 2
    {
        T xa = h0, xb = h1;
#pragma unroll
        for (int k = L - 1; k >= 0; --k) {
            const T x = (T(lu[5 * k + 2]) * Y[k] - T(lu[5 * k + 3]) * xa) - T(lu[5 * k + 4]) * xb;
            Y[k] = x;
            xb = xa;
            xa = x;
        }
    }
    // the solution is complete here: keep the compiler from sinking the
    // recurrence into the (per-lane conditional) stores that consume it
#pragma unroll
    for (int k = 0; k < L; ++k) asm volatile("" : "+v"(Y[k]));
}

// Coefficient tables of the segmented sweeps, indexed like the LU rows:
// LU factors and homogeneous solutions in FP64 and FP32, and the banded LHS
// (for the FP64 residual of the mixed-precision solve).
struct SegTabs {
    const double *lu, *seg, *lhs;
    const float *lu32, *seg32;
};

// Sinks that read operands of their own at each point (type Pre, filled by
// pre(g, Pre&), written with put_pre(g, x, Pre)): eta/zeta sweeps load them
// for kOutBatch points before storing any.
template <class S, class = void> struct sink_pre : std::false_type {};
template <class S> struct sink_pre<S, std::void_t<typename S::Pre>> : std::true_type {};

// One tile of P pencils of block tile.b along d.
//  Src: load(g, b, Raw&), finish(Raw, b, double (&)[NV])
//  Sink: store(g, b, const double (&)[NV]), put(g, b, v, x), ends(side, pencil, v, x);
//        fold(f, x): the value to write given the source f and the result x,
//        put_raw: write such a value (eta/zeta sweeps of plain fields, which
//        still hold the source in registers, use these instead of put);
//        optionally Pre, pre and put_pre (sink_pre)
// MIX: solve in FP32 and refine once with the FP64 residual r = b - A x
// (A the banded LHS), x += A^-1 r solved in FP32.
template <int NV, int SCH, int L, int P, int C, bool XI, bool MIX, int SPM, class Src, class Sink>
__device__ __forceinline__ void seg_tile(const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs,
                                         const SegTabs &tb, Tile tile, int d, int n45, int SP, const Src &src,
                                         Sink &sink, double *lds)
{
    using Sh = SegShape<NV, SCH, L, P, C, XI>;
    constexpr int H = Sh::H;
    const int nthr = NV * P * SP;
    const int t = threadIdx.x;
    // (one variable: t < P SP, so v = 0 and s = t / P without dividing by
    // the run-time SP)
    const int v = NV == 1 ? 0 : t / (P * SP), p = t % P, s = NV == 1 ? t / P : (t / P) % SP;
    const int b = tile.b;
    const BlkDev &B = blks[b];
    const TabDev &T = tabs[B.inst];
    const int npen = npencils(B, d);
    const int n = B.n[d], N = n - 1;
    const int S = seg_count(n, L);
    const double *LU = tb.lu + 5 * int64_t(B.lu[SCH][d]);
    const double *ST = tb.seg + 4 * int64_t(B.lu[SCH][d]);
    // LDS.  eta/zeta: work area (chunk buffer, then partner points, then the
    // exchange buffer), the pencil-end values, SM, LU rows.  xi: the whole
    // tile's sources [NV][P][n | 1] (once the closure rows are formed: the LU rows
    // and the exchange buffer; at the end: the results), SM, and the
    // partner points where some block has another block as partner along d.
    const int nlu = 5 * L * SP;
    double *buf = lds;
    double *ends = nullptr, *part = nullptr, *LUs = nullptr, *ex = nullptr;
    double *SM = nullptr;   // [SPM][8]
    if constexpr (XI) {
        const int fsz = NV * P * Sh::pencil_stride(n), exs4 = 4 * NV * SP * P;
        LUs = lds;
        ex = lds + nlu;
        SM = lds + (fsz > nlu + exs4 ? fsz : nlu + exs4);
        part = SM + 8 * SPM;
        ends = part;
    } else {
        ends = lds + Sh::work_doubles(SP);
        part = buf;
        ex = buf;
        SM = ends + Sh::ends_doubles();
        LUs = SM + 8 * SPM;
    }
    // xi: staged source of the lane's variable and pencil at row; the place
    // of tile point e (pencil e / n, row e % n) of variable 0; variable m
    // follows FS doubles further on
    const int NS = Sh::pencil_stride(n), FS = P * NS;
    auto F = [&](int row) { return buf[(v * P + p) * NS + row]; };
    auto place = [&](int e) { return (n & 1) ? e : e + e / n; };
    const bool pen_ok = tile.q0 + p < npen;
    const bool lane_ok = pen_ok && s < S;
    int a = 0, len = 0;
    if (s < S) seg_bounds(n, L, S, s, a, len);
    const int kb = L - len;                 // row a - kb + k is held in Y[k], kb <= k < L
    auto have = [&](int k) { return k >= kb; };
    auto pt = [&](int pp, int i) -> int64_t {
        int64_t pb, st;
        pencil_of(B, d, tile.q0 + pp, pb, st);
        return pb + int64_t(i) * st;
    };
    auto eidx = [&](int side, int i, int m, int pp) { return ((side * kSegEndRows + i) * NV + m) * P + pp; };

    // ---- 0. every segment's LU rows as its lanes hold them ([SP][L][5]:
    //         pass-through rows, then the segment's rows; a thread's loads are
    //         issued together, kLuBatch per round trip) and the segments'
    //         transfer values into LDS
    const int sq = n / (S > 0 ? S : 1), sr = n % (S > 0 ? S : 1);
    auto lu_value = [&](int i) {   // entry i of the [SP][L][5] table
        const int sg = i / (5 * L), k = (i / 5) % L, c = i % 5;
        const int sl = sg < S ? sq + (sg < sr ? 1 : 0) : 0;
        const int sa = sg * sq + (sg < sr ? sg : sr);
        const bool real = k >= L - sl;
        const double x = LU[real ? 5 * (sa - (L - sl) + k) + c : 0];
        return real ? x : seg_pass_row(c);
    };
    // xi: the first KPRE values per thread are loaded once the tile is
    // staged (their latency overlaps the right-hand sides and closure rows)
    // and written over the tile afterwards
    constexpr int KPRE0 = (5 * L + NV * P - 1) / (NV * P);
    constexpr int KPRE = KPRE0 < kLuBatch ? KPRE0 : kLuBatch;
    double lv[KPRE];
    auto lu_fetch = [&]() {
#pragma unroll
        for (int u = 0; u < KPRE; ++u) lv[u] = lu_value(t + u * nthr < nlu ? t + u * nthr : nlu - 1);
    };
    auto lu_store = [&]() {
#pragma unroll
        for (int u = 0; u < KPRE; ++u)
            if (t + u * nthr < nlu) LUs[t + u * nthr] = lv[u];
        for (int i0 = t + KPRE * nthr; i0 < nlu; i0 += kLuBatch * nthr) {
            double w2[kLuBatch];
#pragma unroll
            for (int u = 0; u < kLuBatch; ++u) w2[u] = lu_value(i0 + u * nthr < nlu ? i0 + u * nthr : nlu - 1);
#pragma unroll
            for (int u = 0; u < kLuBatch; ++u)
                if (i0 + u * nthr < nlu) LUs[i0 + u * nthr] = w2[u];
        }
    };
    if constexpr (!XI) {
        lu_fetch();
        lu_store();
    }
    for (int i = t; i < 8 * S; i += nthr) {
        const int sp = i / 8, e = i % 8;
        int sa, sl;
        seg_bounds(n, L, S, sp, sa, sl);
        // phi at the last two rows (e 0..3), psi at the first two (e 4..7)
        const int r = e < 4 ? sa + sl - 2 + (e >> 1) : sa + ((e - 4) >> 1);
        const int c = e < 4 ? (e & 1) : 2 + (e & 1);
        SM[i] = (r >= 0 && r < n) ? ST[4 * r + c] : 0.0;
    }

    // ---- 1. stream the segment (with H halo points each side): interior
    //         right-hand sides into registers, end values into LDS
    double Y[L];
#pragma unroll
    for (int k = 0; k < L; ++k) Y[k] = 0.0;
    double W[Sh::KR];   // eta/zeta plain fields: the lane's source rows
    double w[2 * H + 1];
#pragma unroll
    for (int m = 0; m <= 2 * H; ++m) w[m] = 0.0;
    const int nslots = C * SP * P;
    if constexpr (XI) {
        // stage all points of the tile at once, consecutive threads on
        // consecutive points: a thread has P n / (NV P SP) <= L / NV points
        // (segments hold at most L), so BX slots take them in one round
        constexpr int BX = (L + NV - 1) / NV;
        const int tot = P * n;
        const int64_t g0 = pt(0, 0);
        const int elim = (npen - tile.q0 < P ? npen - tile.q0 : P) * n;   // points of the tile inside the block
        __syncthreads();
        for (int e0 = t; e0 < tot; e0 += nthr * BX) {
            typename Src::Raw raw[BX];
            int dst[BX];
#pragma unroll
            for (int kk = 0; kk < BX; ++kk) {
                const int e = e0 + kk * nthr;
                const bool ok = e < elim;
                src.load(ok ? g0 + e : g0, b, raw[kk]);
                dst[kk] = ok ? place(e) : -1;
            }
#pragma unroll
            for (int kk = 0; kk < BX; ++kk) {
                if (dst[kk] < 0) continue;
                double f[NV];
                src.finish(raw[kk], b, f);
#pragma unroll
                for (int m = 0; m < NV; ++m) buf[dst[kk] + m * FS] = f[m];
            }
        }
        __syncthreads();
        lu_fetch();
        if constexpr (Src::kEmit) {
            // the staged values of the tile, written out point after point
            if (src.po) {
                for (int e = t; e < elim; e += nthr) {
                    const int li = place(e);
#pragma unroll
                    for (int m = 0; m < NV; ++m) src.po[m * src.NT + g0 + e] = buf[li + m * FS];
                }
            }
        }
#pragma unroll
        for (int kr = 0; kr < Sh::KR; ++kr) {
            const int row = a - kb - H + kr;
            // rows outside the line only reach the rows the closure rows
            // replace (and lanes that store nothing): a clamped row serves
            const double fv = F(row < 0 ? 0 : (row > N ? N : row));
#pragma unroll
            for (int m = 0; m < 2 * H; ++m) w[m] = w[m + 1];
            w[2 * H] = fv;
            const int k = kr - 2 * H;
            if (k < 0 || k >= L) continue;
            // interior rows; the closure rows are replaced below and the
            // pass-through registers cleared
            double rr;
            if constexpr (SCH == 0) {
                rr = kAa * (w[3] - w[1]) + kAb * (w[4] - w[0]);
            } else {
                const double resk = w[3] + w[3];
                rr = (T.fa * ((w[2] + w[4]) - resk) + T.fb * ((w[1] + w[5]) - resk)) +
                     T.fc * ((w[0] + w[6]) - resk);
            }
            Y[k] = rr;
        }
    } else if constexpr (Src::kDirect) {
        // eta/zeta, plain fields: each lane loads its own rows, all at once.
        // The lanes of a wave hold P neighbouring pencils, so one load
        // instruction reads runs of P consecutive points.  Rows outside the
        // lane's window are clamped into it (loaded but never used).
        const int ppc = tile.q0 + p < npen ? p : npen - 1 - tile.q0;
        int64_t pb, st;
        pencil_of(B, d, tile.q0 + ppc, pb, st);
        const int r0 = a - kb - H;                 // row held in W[0]
        const int rlo = (a - H) > 0 ? a - H : 0;
        const int rhi = (a + len + H - 1) < N ? a + len + H - 1 : N;
#pragma unroll
        for (int kr = 0; kr < Sh::KR; ++kr) {
            int row = r0 + kr;
            row = row < rlo ? rlo : (row > rhi ? rhi : row);
            W[kr] = src.field(pb + int64_t(row) * st, v);
        }
        // the 13 points next to each pencil end, for the closure rows
        if (lane_ok && (s == 0 || s == S - 1)) {
#pragma unroll
            for (int kr = 0; kr < Sh::KR; ++kr) {
                const int row = r0 + kr;
                if (row < rlo || row > rhi) continue;
                if (s == 0 && row <= kSegEnd) ends[eidx(0, row, v, p)] = W[kr];
                if (s == S - 1 && N - row <= kSegEnd) ends[eidx(1, N - row, v, p)] = W[kr];
            }
        }
        // row a - kb + k: stencil W[k .. k + 2H] (interior rows; the closure
        // rows are replaced below and the pass-through registers cleared)
#pragma unroll
        for (int k = 0; k < L; ++k) {
            double rr;
            if constexpr (SCH == 0) {
                rr = kAa * (W[k + 3] - W[k + 1]) + kAb * (W[k + 4] - W[k]);
            } else {
                const double resk = W[k + 3] + W[k + 3];
                rr = (T.fa * ((W[k + 2] + W[k + 4]) - resk) + T.fb * ((W[k + 1] + W[k + 5]) - resk)) +
                     T.fc * ((W[k] + W[k + 6]) - resk);
            }
            Y[k] = rr;
        }
    } else {
    constexpr int CPT = (C + NV - 1) / NV;   // staging slots per thread and chunk (= C / NV rounded up)
    // slots whose loads are issued together: as many as keep the loaded
    // values within about 24 doubles of registers
    constexpr int RAWD = int(sizeof(typename Src::Raw) / sizeof(double));
    constexpr int BAT = (24 / RAWD) < 1 ? 1 : ((24 / RAWD) < CPT ? 24 / RAWD : CPT);
    // Slot k of a thread stages, chunk after chunk, rows r0 + c C of one
    // segment's window [lo, hi] of pencil pb (stride st): these are the same
    // for every chunk and computed once.  The loads of slot k of chunk c
    // (addresses clamped into the block, so they are valid for every slot)
    // and the LDS place of its sources (-1: nothing to stage).
    constexpr int CPTB = (CPT + BAT - 1) / BAT * BAT;   // slots the batches below visit
    int64_t spb[CPTB];
    int sr0[CPTB], slo[CPTB], shi[CPTB], sdst[CPTB];
    int64_t sst = 0;
#pragma unroll
    for (int k = 0; k < CPTB; ++k) {
        const int slot = k < CPT ? t + k * nthr : nslots;
        int o, ss, pp;
        seg_slot<P, C, XI>(slot < nslots ? slot : 0, SP, o, ss, pp);
        const int ssc = ss < S ? ss : S - 1;
        int sa, sl;
        seg_bounds(n, L, S, ssc, sa, sl);
        const bool valid = slot < nslots && ss < S && tile.q0 + pp < npen;
        sr0[k] = sa - (L - sl) - H + o;
        slo[k] = sa - H > 0 ? sa - H : 0;
        shi[k] = sa + sl + H - 1 < N ? sa + sl + H - 1 : N;
        sdst[k] = valid ? ((o * NV) * SP + ss) * P + pp : -1;
        pencil_of(B, d, tile.q0 + (tile.q0 + pp < npen ? pp : npen - 1 - tile.q0), spb[k], sst);
    }
    auto issue = [&](int c, int k, typename Src::Raw &raw, int &dst) {
        const int row = sr0[k] + c * C;
        const int rowc = row < 0 ? 0 : (row > N ? N : row);
        src.load(spb[k] + int64_t(rowc) * sst, b, raw);
        dst = (sdst[k] >= 0 && row >= slo[k] && row <= shi[k]) ? sdst[k] : -1;
    };
    auto deposit = [&](const typename Src::Raw &raw, int dst) {
        if (dst < 0) return;
        double f[NV];
        src.finish(raw, b, f);
#pragma unroll
        for (int m = 0; m < NV; ++m) buf[dst + m * SP * P] = f[m];
    };
    // one batch per chunk: the next chunk's loads are issued before this
    // chunk's rows are consumed, so they are in flight meanwhile
    typename Src::Raw rawp[BAT];
    int dstp[BAT];
    if constexpr (BAT == CPT) {
#pragma unroll
        for (int kk = 0; kk < CPT; ++kk) issue(0, kk, rawp[kk], dstp[kk]);
    }
#pragma unroll
    for (int c = 0; c < Sh::NCH; ++c) {
        __syncthreads();
        if constexpr (BAT == CPT) {
#pragma unroll
            for (int kk = 0; kk < CPT; ++kk) deposit(rawp[kk], dstp[kk]);
            if (c + 1 < Sh::NCH) {
#pragma unroll
                for (int kk = 0; kk < CPT; ++kk) issue(c + 1, kk, rawp[kk], dstp[kk]);
            }
        } else {
#pragma unroll
            for (int k0 = 0; k0 < CPT; k0 += BAT) {
                typename Src::Raw raw[BAT];
                int dst[BAT];
#pragma unroll
                for (int kk = 0; kk < BAT; ++kk) issue(c, k0 + kk, raw[kk], dst[kk]);
#pragma unroll
                for (int kk = 0; kk < BAT; ++kk) deposit(raw[kk], dst[kk]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < C; ++o) {
            const int kr = c * C + o;            // streamed point a - kb - H + kr
            if (kr >= Sh::KR) break;
            const double fv = buf[((o * NV + v) * SP + s) * P + p];
            const int row = a - kb - H + kr;
            if (lane_ok && row >= 0 && row <= N && row >= a - H && row < a + len + H) {
                if (s == 0 && row <= kSegEnd) ends[eidx(0, row, v, p)] = fv;
                if (s == S - 1 && N - row <= kSegEnd) ends[eidx(1, N - row, v, p)] = fv;
            }
#pragma unroll
            for (int m = 0; m < 2 * H; ++m) w[m] = w[m + 1];
            w[2 * H] = fv;
            const int k = kr - 2 * H;            // row a - kb + k has its stencil complete
            if (k < 0 || k >= L) continue;
            double rr;
            if constexpr (SCH == 0) {
                rr = kAa * (w[3] - w[1]) + kAb * (w[4] - w[0]);
            } else {
                const double resk = w[3] + w[3];
                rr = (T.fa * ((w[2] + w[4]) - resk) + T.fb * ((w[1] + w[5]) - resk)) +
                     T.fc * ((w[0] + w[6]) - resk);
            }
            Y[k] = rr;
        }
    }

    }

    // the pass-through registers start at zero (a segment has at least
    // kSegMinRows rows, so only the first L - kSegMinRows can be such)
#pragma unroll
    for (int k = 0; k < L - kSegMinRows; ++k)
        if (k < kb) Y[k] = 0.0;

    // ---- 2. partner points of interfaces with another block (one round) ------
    bool self[2];
    for (int side = 0; side < 2; ++side) self[side] = B.qd[d][side] != 0 && B.part[d][side] == b;
    const bool stage_part = (B.qd[d][0] != 0 && !self[0]) || (B.qd[d][1] != 0 && !self[1]);
    __syncthreads();
    if (stage_part) {
        for (int slot = t; slot < 2 * kSegEndRows * P; slot += nthr) {
            const int pp = slot % P, o = (slot / P) % kSegEndRows, side = slot / (P * kSegEndRows);
            if (tile.q0 + pp >= npen || B.qd[d][side] == 0 || self[side]) continue;
            const int Pb = B.part[d][side];
            const BlkDev &BP = blks[Pb];
            const int ps = 1 - side, code = BP.bc[d][ps];
            const int iik = code == 40 ? 0 : 1;
            const int NP = BP.n[d] - 1;
            const int idx = o == kSegEnd ? (ps == 0 ? 0 : NP) : (ps == 0 ? o + iik : NP - (o + iik));
            int64_t pb, pst;
            pencil_of(BP, d, tile.q0 + pp, pb, pst);
            typename Src::Raw raw;
            src.load(pb + int64_t(idx) * pst, Pb, raw);
            double f[NV];
            src.finish(raw, Pb, f);
#pragma unroll
            for (int m = 0; m < NV; ++m) part[eidx(side, o, m, pp)] = f[m];
        }
        __syncthreads();
    }

    // ---- 3. closure rows at both ends (same arithmetic as the sequential sweep)
    double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        if (!(lane_ok && s == (side == 0 ? 0 : S - 1))) continue;
        // distance i from this end
        auto f = [&](int i) { return XI ? F(side == 0 ? i : N - i) : ends[eidx(side, i, v, p)]; };
        const bool qd = B.qd[d][side] != 0;
        double rows[3] = {0.0, 0.0, 0.0};
        if (qd) {
            const int Pb = B.part[d][side];
            const int ps = 1 - side, code = blks[Pb].bc[d][ps];
            const int iik = code == 40 ? 0 : 1;
            const double ra0k = code == 45 ? double(n45) : 0.0;
            // partner point o (o = 12: the partner's face point)
            auto g = [&](int o) {
                if (self[side]) {
                    const int i = o == kSegEnd ? 0 : o + iik;   // distance from the partner's end ps
                    return XI ? F(ps == 0 ? i : N - i) : ends[eidx(ps, i, v, p)];
                }
                return part[eidx(side, o, v, p)];
            };
            double r0 = 0.0, r1 = 0.0, sap0 = 0.0;
            for (int i = 0; i < kSegEnd; ++i) {
                const double fp = g(i);
                if (i == 0) sap0 = fp;
                r0 = r0 + T.qout[SCH][0][i] * fp;
                r1 = r1 + T.qout[SCH][1][i] * fp;
            }
            const double resk = ra0k != 0.0 ? ra0k * g(kSegEnd) : 0.0;
            r0 = r0 - resk * T.qsum[SCH][0];
            r1 = r1 - resk * T.qsum[SCH][1];
            double r2 = sap0 - resk;
            if (n45 == 1 && B.bc[d][side] == 45) {
                const double fo = f(0);
                r0 = r0 + fo * T.qsum[SCH][0];
                r1 = r1 + fo * T.qsum[SCH][1];
                r2 = r2 + double(SCH) * fo;
            }
            double a0 = 0.0, a1 = 0.0;
            for (int i = 0; i < kSegEnd; ++i) {
                a0 = a0 + T.qin[SCH][0][i] * f(i);
                a1 = a1 + T.qin[SCH][1][i] * f(i);
            }
            if (SCH == 0) {
                rows[0] = side == 0 ? a0 + r0 : -a0 - r0;
                rows[1] = side == 0 ? a1 + r1 : -a1 - r1;
            } else {
                rows[0] = a0 + r0;
                rows[1] = a1 + r1;
                const double ra2k = f(2) + f(2);
                if (side == 0)
                    rows[2] = (T.fa * ((f(1) + f(3)) - ra2k) + T.fb * ((f(0) + f(4)) - ra2k)) +
                              T.fc * ((r2 + f(5)) - ra2k);
                else
                    rows[2] = (T.fa * ((f(3) + f(1)) - ra2k) + T.fb * ((f(4) + f(0)) - ra2k)) +
                              T.fc * ((f(5) + r2) - ra2k);
            }
        } else if (SCH == 0) {
            double sum = 0.0;
            if (side == 0) {
                sum = sum + kA01 * (f(1) - f(0)); sum = sum + kA02 * (f(2) - f(0));
                sum = sum + kA03 * (f(3) - f(0)); sum = sum + kA04 * (f(4) - f(0));
                rows[0] = sum;
                sum = 0.0;
                sum = sum + kA10 * (f(0) - f(1)); sum = sum + kA12 * (f(2) - f(1));
                sum = sum + kA13 * (f(3) - f(1)); sum = sum + kA14 * (f(4) - f(1));
                rows[1] = sum;
            } else {
                sum = sum + kA01 * (f(0) - f(1)); sum = sum + kA02 * (f(0) - f(2));
                sum = sum + kA03 * (f(0) - f(3)); sum = sum + kA04 * (f(0) - f(4));
                rows[0] = sum;
                sum = 0.0;
                sum = sum + kA10 * (f(1) - f(0)); sum = sum + kA12 * (f(1) - f(2));
                sum = sum + kA13 * (f(1) - f(3)); sum = sum + kA14 * (f(1) - f(4));
                rows[1] = sum;
            }
        } else {
            const int o0[5] = {1, 2, 3, 4, 5}, o1[5] = {0, 2, 3, 4, 5}, o2[5] = {0, 1, 3, 4, 5};
            double sum = 0.0;
            for (int cc = 0; cc < 5; ++cc) sum = sum + T.fbc[0][cc] * (f(o0[cc]) - f(0));
            rows[0] = sum;
            sum = 0.0;
            for (int cc = 0; cc < 5; ++cc) sum = sum + T.fbc[1][cc] * (f(o1[cc]) - f(1));
            rows[1] = sum;
            sum = 0.0;
            for (int cc = 0; cc < 5; ++cc) sum = sum + T.fbc[2][cc] * (f(o2[cc]) - f(2));
            rows[2] = sum;
        }
        for (int r = 0; r < 3; ++r) (side == 0 ? lo : hi)[r] = rows[r];
    }
    // closure rows: rows 0..NB-1 open segment 0; rows N-r close the last one
    if (lane_ok && s == 0) {       // row r is Y[kb + r]
#pragma unroll
        for (int r = 0; r < Sh::NB; ++r)
#pragma unroll
            for (int k = 0; k <= L - kSegMinRows + r; ++k)
                if (k == kb + r) Y[k] = lo[r];
    }
    if (lane_ok && s == S - 1) {   // row N is Y[L - 1]
#pragma unroll
        for (int r = 0; r < Sh::NB; ++r) Y[L - 1 - r] = hi[r];
    }

    if constexpr (XI) {   // the tile is no longer read: LU rows over it
        __syncthreads();
        lu_store();
        __syncthreads();
    }

    // ---- 4. solve: FP64, or FP32 with one FP64 residual refinement ----------
    const SegLane ln{a, len, kb, s, S, n, L, lane_ok, (v * SP + s) * P + p, NV * SP * P, v, p, P, SP};
    if constexpr (!MIX) {
        seg_solve<double, L, SPM>(Y, ln, LUs, SM, ex);
    } else {
        float X[L];
#pragma unroll
        for (int k = 0; k < L; ++k) X[k] = float(Y[k]);
        seg_solve<float, L, SPM>(X, ln, LUs, SM, ex);
        // edge values of the neighbouring segments: x(a-2), x(a-1), x(b), x(b+1)
        double l2 = 0.0, l1 = 0.0, r1 = 0.0, r2 = 0.0;
        {
            double f0 = 0.0, f1 = 0.0, e0 = 0.0, e1 = 0.0;
#pragma unroll
            for (int k = 0; k < L; ++k) {
                if (k == kb) f0 = X[k];
                if (k == kb + 1) f1 = X[k];
                if (k == L - 2) e0 = X[k];
                if (k == L - 1) e1 = X[k];
            }
            __syncthreads();
            ex[ln.exi] = f0;
            ex[ln.exs + ln.exi] = f1;
            ex[2 * ln.exs + ln.exi] = e0;
            ex[3 * ln.exs + ln.exi] = e1;
            __syncthreads();
            if (lane_ok && s > 0) {
                const int xi = ln.exi - P;
                l2 = ex[2 * ln.exs + xi];
                l1 = ex[3 * ln.exs + xi];
            }
            if (lane_ok && s < S - 1) {
                const int xi = ln.exi + P;
                r1 = ex[xi];
                r2 = ex[ln.exs + xi];
            }
        }
        // r = b - A x in FP64 with the banded rows of the line
        const double *AR = tb.lhs + 5 * int64_t(B.lu[SCH][d]);
#pragma unroll
        for (int k = 0; k < L; ++k) {
            if (have(k)) {
                const int e = L;
                const double *ar = AR + 5 * (a - kb + k);
                const double xm2 = k - 2 >= kb ? double(X[k >= 2 ? k - 2 : 0]) : (k - 1 >= kb ? l1 : l2);
                const double xm1 = k - 1 >= kb ? double(X[k >= 1 ? k - 1 : 0]) : l1;
                const double xp1 = k + 1 < e ? double(X[k + 1 < L ? k + 1 : L - 1]) : r1;
                const double xp2 = k + 2 < e ? double(X[k + 2 < L ? k + 2 : L - 1]) : (k + 1 < e ? r1 : r2);
                const double ax = (((ar[0] * xm2 + ar[1] * xm1) + double(X[k])) + ar[3] * xp1) + ar[4] * xp2;
                Y[k] = Y[k] - ax;
            }
        }
        float D[L];
#pragma unroll
        for (int k = 0; k < L; ++k) D[k] = float(Y[k]);
        seg_solve<float, L, SPM>(D, ln, LUs, SM, ex);
#pragma unroll
        for (int k = 0; k < L; ++k) Y[k] = double(X[k]) + double(D[k]);
    }
    if (lane_ok && s == 0) {
        double x0 = 0.0;
#pragma unroll
        for (int k = 0; k <= L - kSegMinRows; ++k)
            if (k == kb) x0 = Y[k];
        sink.ends(0, p, v, x0);
    }
    if (lane_ok && s == S - 1) sink.ends(1, p, v, Y[L - 1]);

    // ---- 6. results ------------------------------------------------------------
    if constexpr (!XI) {
        // eta/zeta: lanes of a wave hold P neighbouring pencils, stores coalesce
        if (lane_ok) {
            int64_t pb, st;
            pencil_of(B, d, tile.q0 + p, pb, st);
            pb += int64_t(a - kb) * st;
            if constexpr (sink_pre<Sink>::value) {
                static_assert(!Sink::kReadBack, "prefetching sinks write results as they are");
#pragma unroll
                for (int k0 = 0; k0 < L; k0 += kOutBatch) {
                    typename Sink::Pre pr[kOutBatch];
#pragma unroll
                    for (int kk = 0; kk < kOutBatch; ++kk)
                        if (k0 + kk < L && have(k0 + kk)) sink.pre(pb + int64_t(k0 + kk) * st, pr[kk]);
#pragma unroll
                    for (int kk = 0; kk < kOutBatch; ++kk)
                        if (k0 + kk < L && have(k0 + kk)) sink.put_pre(pb + int64_t(k0 + kk) * st, Y[k0 + kk], pr[kk]);
                }
            } else {
#pragma unroll
                for (int k = 0; k < L; ++k)
                    if (have(k)) {
                        if constexpr (Src::kDirect)
                            sink.put_raw(pb + int64_t(k) * st, b, v, sink.fold(W[k + H], Y[k]));
                        else
                            sink.put(pb + int64_t(k) * st, b, v, Y[k]);
                    }
            }
        }
    } else {
        // xi: results into the tile buffer, then stored along the pencils
        // (the filter's sink adds them to the sources it reads back)
        __syncthreads();
        if (lane_ok) {
#pragma unroll
            for (int k = 0; k < L; ++k)
                if (have(k)) {
                    buf[(v * P + p) * NS + (a - kb + k)] = Y[k];
                }
        }
        __syncthreads();
        const int tot = P * n;
        const int64_t g0 = pt(0, 0);
        const int elim = (npen - tile.q0 < P ? npen - tile.q0 : P) * n;   // points of the tile inside the block
        if constexpr (Sink::kReadBack) {
            // the sink adds the results to its sources: read those back for
            // kOutBatch points at a time, all loads in flight together
            for (int e0 = t; e0 < tot; e0 += kOutBatch * nthr) {
                double fs[kOutBatch][NV];
                int li[kOutBatch];   // LDS place of the point's results, -1: none
#pragma unroll
                for (int kk = 0; kk < kOutBatch; ++kk) {
                    const int e = e0 + kk * nthr;
                    const bool ok = e < elim;
                    li[kk] = ok ? place(e) : -1;
#pragma unroll
                    for (int m = 0; m < NV; ++m) fs[kk][m] = sink.src(ok ? g0 + e : g0, m);
                }
#pragma unroll
                for (int kk = 0; kk < kOutBatch; ++kk) {
                    if (li[kk] < 0) continue;
                    const int64_t g = g0 + (e0 + kk * nthr);
#pragma unroll
                    for (int m = 0; m < NV; ++m) sink.put_raw(g, b, m, sink.fold(fs[kk][m], buf[li[kk] + m * FS]));
                }
            }
        } else if constexpr (sink_pre<Sink>::value) {
            // the sink's own operands of kOutBatch points in flight together
            static_assert(NV == 1, "prefetching sinks take one variable");
            for (int e0 = t; e0 < tot; e0 += kOutBatch * nthr) {
                typename Sink::Pre pr[kOutBatch];
                int li[kOutBatch];
#pragma unroll
                for (int kk = 0; kk < kOutBatch; ++kk) {
                    const int e = e0 + kk * nthr;
                    const bool ok = e < elim;
                    li[kk] = ok ? place(e) : -1;
                    sink.pre(ok ? g0 + e : g0, pr[kk]);
                }
#pragma unroll
                for (int kk = 0; kk < kOutBatch; ++kk)
                    if (li[kk] >= 0) sink.put_pre(g0 + (e0 + kk * nthr), buf[li[kk]], pr[kk]);
            }
        } else {
            for (int e = t; e < elim; e += nthr) {
                const int li = place(e);
                double x[NV];
#pragma unroll
                for (int m = 0; m < NV; ++m) x[m] = buf[li + m * FS];
                sink.store(g0 + e, b, x);
            }
        }
    }

}

}  // namespace cnd
