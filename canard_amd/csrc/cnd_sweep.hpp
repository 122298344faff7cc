// Tiled pencil sweeps: a workgroup of NV waves solves 64 pencils of one
// block along direction d; wave v carries variable v of all 64 pencils, one
// lane per pencil.
//
// The compact derivative (src/mo_numerics.f90:709-886) and compact filter
// (src/mo_numerics.f90:890-987) share one structure: boundary rows from the
// pencil ends (and, at interfaces, the partner block's partial sums),
// interior rows from a sliding window, forward substitution, back
// substitution.  Here:
//   1. the workgroup evaluates the source (a field, the primitive variables
//      or the flux components) at the 12 points next to each pencil end and
//      at the partner pencil's 12 points, stages them in LDS, and each wave
//      forms its variable's boundary rows in the reference's order;
//   2. the pencil is streamed in chunks of C points (chunk_of): the workgroup
//      evaluates the source for chunk x 64 points with coalesced loads (xi:
//      runs of consecutive points of one pencil; eta/zeta: one point of 64
//      neighbouring pencils) into LDS, and each wave advances its forward
//      recurrence, parking forward values in a coalesced scratch array;
//   3. the backward pass walks the chunks in reverse, stages the final values
//      in LDS, and the workgroup stores them with coalesced writes.
#pragma once

#include "cnd_dev.hpp"
#include "cnd_kernels.hpp"

namespace cnd {

constexpr int kTile = 64;    // pencils per tile (one wave per variable)
constexpr int kEnd = 12;     // points staged per pencil end

// Points per chunk along the pencil.  xi pencils are contiguous, so a chunk
// reads chunk x 8 bytes of each pencil: 16 points fill a 128-byte cache
// line; eta/zeta chunks read whole 64-pencil rows and stay at 8 points.
#ifndef CND_CHUNK_XI
#define CND_CHUNK_XI 16
#endif
#ifndef CND_CHUNK_ETA
#define CND_CHUNK_ETA 8
#endif
constexpr int kChunkXi = CND_CHUNK_XI, kChunkEta = CND_CHUNK_ETA;
__host__ __device__ constexpr int chunk_of(int d) { return d == 0 ? kChunkXi : kChunkEta; }

// LDS block [o][v][p], one double of padding per o-row.
template <int NV>
struct LdsRows {
    static constexpr int kPitch = NV * kTile + 1;
    double *s;
    __device__ double &at(int o, int v, int p) const { return s[o * kPitch + v * kTile + p]; }
};

template <int NV, int C>
constexpr size_t sweep_lds_bytes() { return sizeof(double) * (C > kEnd ? C : kEnd) * (NV * kTile + 1); }

// Workgroups are handed to the 8 XCDs (each with its own L2) round robin.
// Workgroup w takes tile (w % 8) * per + w / 8, so consecutive tiles, which
// share the cache lines at their edges, run on one XCD at about the same
// time.  The grid is 8 * per wide; workgroups past the last tile return.
__host__ __device__ inline int xcd_width(int ntiles) { return 8 * ((ntiles + 7) / 8); }
__device__ inline int xcd_tile(int ntiles)
{
    const int per = (ntiles + 7) / 8;
    return int(blockIdx.x % 8) * per + int(blockIdx.x / 8);
}

// Staging slot k of thread t: which pencil p of the tile and which offset o
// of an nrows-point window.  xi pencils are contiguous, so consecutive
// threads take consecutive points of one pencil; eta/zeta pencils are
// interleaved, so consecutive threads take neighbouring pencils.
__device__ __forceinline__ void stage_slot(int d, int nrows, int slot, int &p, int &o)
{
    if (d == 0) { p = slot / nrows; o = slot % nrows; }
    else { p = slot % kTile; o = slot / kTile; }
}

// Scratch index of the forward value of row j, variable v, lane p of tile.
__device__ __forceinline__ int64_t yidx(int64_t tile, int NV, int nmax, int v, int j, int p)
{
    return ((tile * NV + v) * int64_t(nmax) + j) * kTile + p;
}

// Evaluate the source for all 64 pencils at nrows points each and stage
// them.  where(p, o, g, blk) gives the point of (pencil p, offset o).
// (Issuing all of a thread's loads before using any was measured slower:
// the extra live registers cost more occupancy than the loads gain.)
template <int NV, class Src, class Where>
__device__ __forceinline__ void stage(int d, int nrows, const Src &src, const Where &where, LdsRows<NV> lds)
{
    const int nthr = NV * kTile;
    for (int slot = threadIdx.x; slot < nrows * kTile; slot += nthr) {
        int p, o;
        stage_slot(d, nrows, slot, p, o);
        int64_t g;
        int blk;
        if (where(p, o, g, blk)) {
            typename Src::Raw raw;
            src.load(g, blk, raw);
            double v[NV];
            src.finish(raw, blk, v);
#pragma unroll
            for (int k = 0; k < NV; ++k) lds.at(o, k, p) = v[k];
        }
    }
}

// One tile of a sweep.  S = 0 derivative, 1 filter.
//  Src::eval(g, b, double (&v)[NV])            source values at point g
//  Sink::store(g, b, const double (&x)[NV])     final values (cooperative)
//  Sink::ends(side, lane, v, x)                 pencil end value of variable v
template <int NV, int S, int C, class Src, class Sink>
__device__ __forceinline__ void sweep_tile(const BlkDev *__restrict__ blks, const TabDev *__restrict__ tabs,
                                           const double *__restrict__ lu, Tile tile, int64_t tile_id, int d,
                                           int n45, int nmax, double *__restrict__ ys, const Src &src, Sink &sink,
                                           double *lds_raw)
{
    constexpr int H = S == 0 ? 2 : 3;   // stencil half width
    const int lane = threadIdx.x % kTile, wv = threadIdx.x / kTile;   // wave = variable
    const int b = tile.b;
    const BlkDev &B = blks[b];
    const TabDev &T = tabs[B.inst];
    const int npen = npencils(B, d);
    const int n = B.n[d], N = n - 1;
    const double *L = lu + 5 * int64_t(B.lu[S][d]);
    LdsRows<NV> lds{lds_raw};
    const int q = tile.q0 + lane;
    const bool active = q < npen;
    int64_t st;
    {
        int64_t b0;
        pencil_of(B, d, tile.q0, b0, st);
    }
    auto own = [&](int p, int i) -> int64_t {
        int64_t pb, s_;
        pencil_of(B, d, tile.q0 + p, pb, s_);
        return pb + int64_t(i) * st;
    };

    // ---- 1. boundary rows: lo[r] = row r, hi[r] = row N-r ----------------------
    double lo[3], hi[3];
    for (int side = 0; side < 2; ++side) {
        const int sg = side == 0 ? 1 : -1, i0 = side == 0 ? 0 : N;
        const bool qd = B.qd[d][side] != 0;
        double r0 = 0.0, r1 = 0.0, r2 = 0.0;
        if (qd) {
            // partner pencil's 12 points next to its face, plus its face point
            const int P = B.part[d][side];
            const BlkDev &BP = blks[P];
            const int ps = 1 - side, code = BP.bc[d][ps];
            const int iik = code == 40 ? 0 : 1;
            const double ra0k = code == 45 ? double(n45) : 0.0;
            const int NP = BP.n[d] - 1;
            auto partner_pt = [&](int p, int o, int64_t &g, int &blk) {
                if (tile.q0 + p >= npen) return false;
                int64_t pb, pst;
                pencil_of(BP, d, tile.q0 + p, pb, pst);
                const int idx = o == kEnd ? (ps == 0 ? 0 : NP) : (ps == 0 ? o + iik : NP - (o + iik));
                g = pb + int64_t(idx) * pst;
                blk = P;
                return true;
            };
            __syncthreads();
            stage<NV>(d, kEnd, src, [&](int p, int o, int64_t &g, int &blk) { return partner_pt(p, o, g, blk); }, lds);
            __syncthreads();
            double sap0 = 0.0;
            if (active) {
                for (int i = 0; i < kEnd; ++i) {
                    const double v = lds.at(i, wv, lane);
                    if (i == 0) sap0 = v;
                    r0 = r0 + T.qout[S][0][i] * v;
                    r1 = r1 + T.qout[S][1][i] * v;
                }
            }
            // the partner's face value (resk) when the periodic jump is asked for
            double resk = 0.0;
            if (ra0k != 0.0) {
                __syncthreads();
                stage<NV>(d, 1, src, [&](int p, int, int64_t &g, int &blk) { return partner_pt(p, kEnd, g, blk); }, lds);
                __syncthreads();
                if (active) resk = ra0k * lds.at(0, wv, lane);
            }
            r0 = r0 - resk * T.qsum[S][0];
            r1 = r1 - resk * T.qsum[S][1];
            r2 = sap0 - resk;
        }
        // own points next to this end
        const int nown = qd ? kEnd : (S == 0 ? 5 : 6);
        __syncthreads();
        stage<NV>(d, nown, src, [&](int p, int o, int64_t &g, int &blk) {
            if (tile.q0 + p >= npen) return false;
            g = own(p, i0 + sg * o);
            blk = b;
            return true;
        }, lds);
        __syncthreads();
        double rows[3] = {0.0, 0.0, 0.0};
        if (active) {
            auto f = [&](int i) { return lds.at(i, wv, lane); };   // value at distance i from the end
            if (qd) {
                if (n45 == 1 && B.bc[d][side] == 45) {
                    const double fo = f(0);
                    r0 = r0 + fo * T.qsum[S][0];
                    r1 = r1 + fo * T.qsum[S][1];
                    r2 = r2 + double(S) * fo;
                }
                double a0 = 0.0, a1 = 0.0;
                for (int i = 0; i < kEnd; ++i) {
                    a0 = a0 + T.qin[S][0][i] * f(i);
                    a1 = a1 + T.qin[S][1][i] * f(i);
                }
                if (S == 0) {
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
            } else if (S == 0) {
                double s = 0.0;
                if (side == 0) {
                    s = s + kA01 * (f(1) - f(0)); s = s + kA02 * (f(2) - f(0));
                    s = s + kA03 * (f(3) - f(0)); s = s + kA04 * (f(4) - f(0));
                    rows[0] = s;
                    s = 0.0;
                    s = s + kA10 * (f(0) - f(1)); s = s + kA12 * (f(2) - f(1));
                    s = s + kA13 * (f(3) - f(1)); s = s + kA14 * (f(4) - f(1));
                    rows[1] = s;
                } else {
                    s = s + kA01 * (f(0) - f(1)); s = s + kA02 * (f(0) - f(2));
                    s = s + kA03 * (f(0) - f(3)); s = s + kA04 * (f(0) - f(4));
                    rows[0] = s;
                    s = 0.0;
                    s = s + kA10 * (f(1) - f(0)); s = s + kA12 * (f(1) - f(2));
                    s = s + kA13 * (f(1) - f(3)); s = s + kA14 * (f(1) - f(4));
                    rows[1] = s;
                }
            } else {
                const int o0[5] = {1, 2, 3, 4, 5}, o1[5] = {0, 2, 3, 4, 5}, o2[5] = {0, 1, 3, 4, 5};
                double s = 0.0;
                for (int c = 0; c < 5; ++c) s = s + T.fbc[0][c] * (f(o0[c]) - f(0));
                rows[0] = s;
                s = 0.0;
                for (int c = 0; c < 5; ++c) s = s + T.fbc[1][c] * (f(o1[c]) - f(1));
                rows[1] = s;
                s = 0.0;
                for (int c = 0; c < 5; ++c) s = s + T.fbc[2][c] * (f(o2[c]) - f(2));
                rows[2] = s;
            }
        }
        for (int r = 0; r < 3; ++r) (side == 0 ? lo : hi)[r] = rows[r];
    }

    // ---- 2. forward substitution, chunk by chunk -----------------------------
    const int nch = (n + C - 1) / C;
    double w[2 * H + 1];     // window f(i-2H..i)
    double ya = 0.0, yb = 0.0;
    double *yrow = ys + yidx(tile_id, NV, nmax, wv, 0, lane);
    if (active) {
        ya = lo[0];
        yrow[0] = ya;
        yb = lo[1] - L[5 * 1 + 1] * ya;
        yrow[kTile] = yb;
        if (S == 1) {
            const double y = (lo[2] - L[5 * 2] * ya) - L[5 * 2 + 1] * yb;
            ya = yb; yb = y;
            yrow[2 * kTile] = y;
        }
    }
    const int jfirst = S == 0 ? 2 : 3, jlast = S == 0 ? N - 2 : N - 3;
    for (int c = 0; c < nch; ++c) {
        __syncthreads();
        stage<NV>(d, C, src, [&](int p, int o, int64_t &g, int &blk) {
            const int i = c * C + o;
            if (tile.q0 + p >= npen || i >= n) return false;
            g = own(p, i);
            blk = b;
            return true;
        }, lds);
        __syncthreads();
        if (!active) continue;
#pragma unroll
        for (int o = 0; o < C; ++o) {
            const int i = c * C + o;
            if (i >= n) break;
#pragma unroll
            for (int m = 0; m < 2 * H; ++m) w[m] = w[m + 1];
            w[2 * H] = lds.at(o, wv, lane);
            const int j = i - H;   // row whose stencil is now complete
            if (j < jfirst || j > jlast) continue;
            double rr;
            if constexpr (S == 0) {
                rr = kAa * (w[3] - w[1]) + kAb * (w[4] - w[0]);
            } else {
                const double resk = w[3] + w[3];
                rr = (T.fa * ((w[2] + w[4]) - resk) + T.fb * ((w[1] + w[5]) - resk)) + T.fc * ((w[0] + w[6]) - resk);
            }
            const double y = (rr - L[5 * j] * ya) - L[5 * j + 1] * yb;
            ya = yb; yb = y;
            yrow[int64_t(j) * kTile] = y;
        }
    }
    // end rows N-2 (filter), N-1, N; then x(N), x(N-1)
    double xa = 0.0, xb = 0.0;   // x(j+1), x(j+2) in the backward pass
    if (active) {
        if (S == 1) {
            const double y = (hi[2] - L[5 * (N - 2)] * ya) - L[5 * (N - 2) + 1] * yb;
            ya = yb; yb = y;
            yrow[int64_t(N - 2) * kTile] = y;
        }
        const double y1 = (hi[1] - L[5 * (N - 1)] * ya) - L[5 * (N - 1) + 1] * yb;
        const double y0 = (hi[0] - L[5 * N] * yb) - L[5 * N + 1] * y1;
        xb = L[5 * N + 2] * y0;
        xa = L[5 * (N - 1) + 2] * y1 - L[5 * (N - 1) + 3] * xb;
        sink.ends(1, lane, wv, xb);
    }
    // ---- 3. back substitution, chunks in reverse -------------------------------
    for (int c = nch - 1; c >= 0; --c) {
        if (active) {
#pragma unroll
            for (int o = C - 1; o >= 0; --o) {
                const int j = c * C + o;
                if (j > N) continue;
                double x;
                if (j == N) {
                    x = xb;
                } else if (j == N - 1) {
                    x = xa;
                } else {
                    const double y = yrow[int64_t(j) * kTile];
                    x = (L[5 * j + 2] * y - L[5 * j + 3] * xa) - L[5 * j + 4] * xb;
                    xb = xa; xa = x;
                    if (j == 0) sink.ends(0, lane, wv, x);
                }
                lds.at(o, wv, lane) = x;
            }
        }
        __syncthreads();
        for (int slot = threadIdx.x; slot < C * kTile; slot += NV * kTile) {
            int p, o;
            stage_slot(d, C, slot, p, o);
            const int j = c * C + o;
            if (tile.q0 + p < npen && j < n) {
                double x[NV];
#pragma unroll
                for (int k = 0; k < NV; ++k) x[k] = lds.at(o, k, p);
                sink.store(own(p, j), b, x);
            }
        }
        __syncthreads();
    }
}

}  // namespace cnd
