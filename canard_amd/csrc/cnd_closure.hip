// Learned closure: a three-layer 3x3x3 convolutional network (5 -> 32 -> 32
// -> 5 channels, ReLU between layers) that maps the coarse state after a
// time step to a correction of it, trained against box-filtered fine-grid
// runs.  It has no counterpart in the reference; the compact filter stays in
// every step, so a context without a network is the plain scheme.
//
// Every convolution is an implicit GEMM on the gfx950 BF16 matrix cores: one
// wave computes a 32 (voxels) x 32 (channels) tile with
// v_mfma_f32_32x32x16_bf16 (FP32 accumulate), stepping K over the 27 taps x
// input channels, its workgroup's input halo and weights staged in LDS;
// activations are BF16 channel-last [voxel][channel].
//   forward        out[v][o]  = sum_{t,c} a[nb(v,t)][c] W[o][t][c] (+ b[o])
//   data gradient  g[v][c]    = sum_{t,o} go[nb(v,t)][o] W[o][26-t][c]
//   weight grad.   gW[o][t][c] = sum_v go[v][o] a[nb(v,t)][c]   (partial sums
//                  per workgroup of voxel tiles, then summed in a fixed order)
// Parameters are FP32 (Adam); BF16 copies in both layouts are rebuilt after
// every update.  The grid is periodic, n^3 voxels (the context's stored
// duplicate planes are filled from their originals).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cnd.h"
#include "cnd_internal.hpp"

using namespace cnd;

namespace {

using bf16 = unsigned short;   // bfloat16 bit pattern
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int kC0 = 5, kC0P = 8, kC1 = 32, kC2 = 32, kC3 = 5, kCP = 32, kTaps = 27;
// flat FP32 parameters: W1 [32][27][5], b1 [32], W2 [32][27][32], b2 [32],
// W3 [5][27][32], b3 [5]
constexpr int kOW1 = 0, kOB1 = kOW1 + kC1 * kTaps * kC0, kOW2 = kOB1 + kC1, kOB2 = kOW2 + kC2 * kTaps * kC1,
              kOW3 = kOB2 + kC2, kOB3 = kOW3 + kC3 * kTaps * kC2, kNParam = kOB3 + kC3;
constexpr int kM1 = kTaps * kC0P, kM2 = kTaps * kCP;   // weight-gradient row lengths
constexpr int kWgSlices = 8;                           // slices of the first weight-gradient reduction level

__host__ __device__ inline uint32_t f2u(float f)
{
    union { float f; uint32_t u; } x;
    x.f = f;
    return x.u;
}
__host__ __device__ inline float u2f(uint32_t u)
{
    union { float f; uint32_t u; } x;
    x.u = u;
    return x.f;
}
// round to nearest even
__host__ __device__ inline bf16 to_bf16(float f)
{
    uint32_t u = f2u(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return bf16((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return bf16(u >> 16);
}
__host__ __device__ inline float from_bf16(bf16 h) { return u2f(uint32_t(h) << 16); }

__device__ inline int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }
// neighbour of voxel (x, y, z) at tap t = (dx+1) + 3 (dy+1) + 9 (dz+1)
__device__ inline int nbr(int x, int y, int z, int t, int n)
{
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    return wrap(x + dx, n) + n * (wrap(y + dy, n) + n * wrap(z + dz, n));
}

// Convolution as an implicit GEMM on v_mfma_f32_32x32x16_bf16 (gfx950: 16
// K values per instruction, 8 per lane).  A workgroup of 8 waves owns a
// tile of X (x) by 2R (y) by 4 (z) voxels, X = 32, 16 or 8 dividing n and
// R = 32 / X; wave (wy, wz) computes the 32 voxels of its R rows of X at
// plane wz against all 32 output channels.  The tile's input halo (one voxel
// each side, periodic) and the layer's weights are staged in LDS first, so
// the 27 taps read LDS only; voxel and weight rows are padded by 8 BF16
// values, which spreads a wave's 16-byte reads over distinct banks.
//   K order: for CIP = 32, tap t and channel block cb of 16 (lanes kh = 0, 1
//   take its halves); for CIP = 8, one instruction covers the 8 channels of
//   taps 2 tt (lanes kh = 0) and 2 tt + 1 (kh = 1), the 28th tap being zero.
// EPI 0: + bias, ReLU, BF16 [V][32]; EPI 1: (+ bias) FP32 [V][ostride] for
// o < cout; EPI 2: times (mask[v][o] > 0), BF16 [V][32].
constexpr int kConvWaves = 8;

__host__ __device__ constexpr int conv_vs(int cip) { return cip + 8; }           // LDS BF16 per halo voxel
__host__ __device__ constexpr int conv_ws(int cip) { return kTaps * cip + 8; }   // LDS BF16 per weight row
// x extent of a wave's voxel tile for an n^3 grid (n a multiple of 8)
__host__ __device__ inline int conv_x(int n) { return n % 32 == 0 ? 32 : (n % 16 == 0 ? 16 : 8); }
__host__ inline size_t conv_lds_bytes(int cip, int n)
{
    const int X = conv_x(n), R = 32 / X;
    return sizeof(bf16) * (size_t(32) * conv_ws(cip) + size_t(X + 2) * (2 * R + 2) * 6 * conv_vs(cip));
}

typedef short s8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f16v mfma16(const s8v &a, const s8v &b, const f16v &c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8v, a), __builtin_bit_cast(bf8v, b), c, 0, 0,
                                                  0);
}

template <int CIP, int EPI>
__global__ __launch_bounds__(64 * kConvWaves) void k_conv(const bf16 *__restrict__ in, const bf16 *__restrict__ w,
                                                         const float *__restrict__ bias,
                                                         const bf16 *__restrict__ mask, void *__restrict__ out, int n,
                                                         int cout, int ostride)
{
    constexpr int VS = conv_vs(CIP), WS = conv_ws(CIP), KT = kTaps * CIP;
    extern __shared__ unsigned short conv_lds[];
    bf16 *sw = conv_lds;             // weights [32][WS]
    bf16 *sa = conv_lds + 32 * WS;   // halo [HZ][HY][HX][VS]
    const int X = conv_x(n), R = 32 / X;
    const int TY = 2 * R, HX = X + 2, HY = TY + 2, HZ = 6;
    int wg = blockIdx.x;
    const int ntx = n / X, nty = n / TY;
    const int x0 = (wg % ntx) * X;
    wg /= ntx;
    const int y0 = (wg % nty) * TY, z0 = (wg / nty) * 4;
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int e = tid; e < 32 * (KT / 8); e += nthr) {
        const int o = e / (KT / 8), k8 = e - o * (KT / 8);
        *reinterpret_cast<s8v *>(sw + o * WS + 8 * k8) = *reinterpret_cast<const s8v *>(w + o * KT + 8 * k8);
    }
    const int HV = HX * HY * HZ;
    for (int e = tid; e < HV * (CIP / 8); e += nthr) {
        const int hv = e / (CIP / 8), c8 = e - hv * (CIP / 8);
        const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
        const int64_t gv = wrap(x0 + hx - 1, n) + int64_t(n) * (wrap(y0 + hy - 1, n) + int64_t(n) * wrap(z0 + hz - 1, n));
        *reinterpret_cast<s8v *>(sa + hv * VS + 8 * c8) = *reinterpret_cast<const s8v *>(in + gv * CIP + 8 * c8);
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const int wy = wave & 1, wz = wave >> 1;
    const int i = lane & 31, kh = lane >> 5;
    const int lx = i % X, ly = wy * R + i / X;   // voxel i of the wave in the tile (plane wz)
    f16v acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    const bf16 *wrow = sw + i * WS;   // B column: output channel i
    auto halo = [&](int t) { return ((wz + t / 9) * HY + (ly + (t / 3) % 3)) * HX + (lx + t % 3); };
    if constexpr (CIP == 8) {
#pragma unroll 2
        for (int tt = 0; tt < (kTaps + 1) / 2; ++tt) {
            const int t = 2 * tt + kh;
            s8v a, b;
#pragma unroll
            for (int e = 0; e < 8; ++e) { a[e] = 0; b[e] = 0; }
            if (t < kTaps) {
                a = *reinterpret_cast<const s8v *>(sa + halo(t) * VS);
                b = *reinterpret_cast<const s8v *>(wrow + t * CIP);
            }
            acc = mfma16(a, b, acc);
        }
    } else {
        for (int t = 0; t < kTaps; ++t) {
            const bf16 *arow = sa + halo(t) * VS + 8 * kh;
            const bf16 *brow = wrow + t * CIP + 8 * kh;
#pragma unroll
            for (int cb = 0; cb < CIP / 16; ++cb)
                acc = mfma16(*reinterpret_cast<const s8v *>(arow + 16 * cb), *reinterpret_cast<const s8v *>(brow + 16 * cb),
                             acc);
        }
    }
    const int o = i;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int vl = 8 * (r >> 2) + 4 * kh + (r & 3);   // row of the result: voxel vl of the wave
        const int gx = x0 + vl % X, gy = y0 + wy * R + vl / X, gz = z0 + wz;
        const int64_t v = gx + int64_t(n) * (gy + int64_t(n) * gz);
        float val = acc[r];
        if constexpr (EPI == 0) {
            val += bias[o];
            static_cast<bf16 *>(out)[v * kCP + o] = to_bf16(val > 0.f ? val : 0.f);
        } else if constexpr (EPI == 1) {
            if (o < cout) {
                if (bias) val += bias[o];
                static_cast<float *>(out)[v * ostride + o] = val;
            }
        } else {
            const float m = from_bf16(mask[v * kCP + o]);
            static_cast<bf16 *>(out)[v * kCP + o] = to_bf16(m > 0.f ? val : 0.f);
        }
    }
}

// Weight and bias gradients of one layer as a GEMM over voxels on
// v_mfma_f32_32x32x16_bf16: dW[o][m] = sum_v go[v][o] a[nb(v, m / CIP)][m % CIP]
// (M = 27 CIP columns, in 32-column tiles) and db[o] = sum_v go[v][o].
// Workgroups of 8 waves walk the 256-voxel tiles of k_conv (tile =
// blockIdx.x, blockIdx.x + gridDim.x, ...).  Per tile they stage the
// activation halo (VS BF16 per voxel: the two lane halves, 8 voxels apart,
// then read different LDS banks) and the output gradient transposed,
// goT[o][vl] with vl = lx + X (ly + TY lz), in LDS.  Wave w accumulates the
// column tiles w, w + 8, ... over all voxels of its workgroup, 16 per
// instruction: the A operand (lane: o = lane % 32, 8 consecutive voxels
// along x) is one 16-byte LDS read shared by all of the wave's columns, the
// B operand the 8 neighbour values of the lane's (tap, channel).  Wave 7
// also sums the A operands for the bias gradient.  The workgroup's sums go
// to part[block][o][m] and bpart[block][o]; k_grad_part and k_grad_reduce
// add the blocks in a fixed order (no atomics: the gradient is the same on
// every run).
constexpr int kWgGS = 256 + 8;   // goT row stride (voxels)
constexpr int kWgBlocks = 256;   // workgroups of a weight-gradient launch (at most)
__host__ __device__ constexpr int wg_vs(int cip) { return cip + 4; }
__host__ inline size_t wgrad_lds_bytes(int cip, int n)
{
    const int X = conv_x(n), R = 32 / X;
    return sizeof(bf16) * (size_t(32) * kWgGS + size_t(X + 2) * (2 * R + 2) * 6 * wg_vs(cip));
}
__host__ inline int wgrad_blocks(int n)
{
    const int tiles = n * n * n / 256;
    return tiles < kWgBlocks ? tiles : kWgBlocks;
}
typedef short s4v __attribute__((ext_vector_type(4)));

template <int CIP>
__global__ __launch_bounds__(512) void k_wgrad(const bf16 *__restrict__ go, const bf16 *__restrict__ act,
                                               float *__restrict__ part, float *__restrict__ bpart, int n)
{
    constexpr int VS = wg_vs(CIP), M = kTaps * CIP, NT = (M + 31) / 32, TPW = (NT + 7) / 8;
    extern __shared__ unsigned short wg_lds[];
    bf16 *goT = wg_lds;               // [32][kWgGS]
    bf16 *sa = wg_lds + 32 * kWgGS;   // halo [HZ][HY][HX][VS]
    const int X = conv_x(n), R = 32 / X;
    const int TY = 2 * R, HX = X + 2, HY = TY + 2, HZ = 6, HV = HX * HY * HZ;
    const int ntx = n / X, nty = n / TY;
    const int ntiles = ntx * nty * (n / 4);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int wave = tid >> 6, lane = tid & 63, i = lane & 31, kh = lane >> 5;
    // the wave's column tiles j = wave + 8 u: the lane's column m = 32 j + i,
    // its tap's halo offset and its channel
    int toff[TPW], coff[TPW];
    bool jok[TPW], mok[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const int j = wave + 8 * u, m = 32 * j + i;
        jok[u] = j < NT;
        mok[u] = jok[u] && m < M;
        const int t = mok[u] ? m / CIP : 0;
        toff[u] = ((t / 9) * HY + (t / 3) % 3) * HX + t % 3;
        coff[u] = mok[u] ? m % CIP : 0;
    }
    f16v acc[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[u][r] = 0.f;
    float bsum = 0.f;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int w = tile;
        const int x0 = (w % ntx) * X;
        w /= ntx;
        const int y0 = (w % nty) * TY, z0 = (w / nty) * 4;
        __syncthreads();   // the previous tile's operands are read
        for (int e = tid; e < HV * (CIP / 4); e += nthr) {
            const int hv = e / (CIP / 4), c4 = e - hv * (CIP / 4);
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int64_t gv = wrap(x0 + hx - 1, n) + int64_t(n) * (wrap(y0 + hy - 1, n) + int64_t(n) * wrap(z0 + hz - 1, n));
            *reinterpret_cast<s4v *>(sa + hv * VS + 4 * c4) = *reinterpret_cast<const s4v *>(act + gv * CIP + 4 * c4);
        }
        for (int e = tid; e < 256 * (kCP / 8); e += nthr) {
            const int vl = e / (kCP / 8), c8 = e - vl * (kCP / 8);
            const int lx = vl % X, ly = (vl / X) % TY, lz = vl / (X * TY);
            const int64_t v = (x0 + lx) + int64_t(n) * ((y0 + ly) + int64_t(n) * (z0 + lz));
            const s8v g = *reinterpret_cast<const s8v *>(go + v * kCP + 8 * c8);
#pragma unroll
            for (int k = 0; k < 8; ++k) goT[(8 * c8 + k) * kWgGS + vl] = bf16(g[k]);
        }
        __syncthreads();
        for (int q = 0; q < 16; ++q) {
            const int vl0 = 16 * q + 8 * kh;   // the lane's 8 voxels: vl0 .. vl0 + 7, along x
            const s8v a = *reinterpret_cast<const s8v *>(goT + i * kWgGS + vl0);
            if (wave == 7) {
#pragma unroll
                for (int e = 0; e < 8; ++e) bsum += from_bf16(bf16(a[e]));
            }
            const int lx = vl0 % X, ly = (vl0 / X) % TY, lz = vl0 / (X * TY);
            const int hb = (lz * HY + ly) * HX + lx;
#pragma unroll
            for (int u = 0; u < TPW; ++u) {
                if (!jok[u]) continue;
                const bf16 *bp = sa + (hb + toff[u]) * VS + coff[u];
                s8v bv;
#pragma unroll
                for (int e = 0; e < 8; ++e) bv[e] = mok[u] ? short(bp[e * VS]) : short(0);
                acc[u] = mfma16(a, bv, acc[u]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        if (!mok[u]) continue;
        const int m = 32 * (wave + 8 * u) + i;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int o = 8 * (r >> 2) + 4 * kh + (r & 3);
            part[(int64_t(blockIdx.x) * kCP + o) * M + m] = acc[u][r];
        }
    }
    if (wave == 7) {
        bsum += __shfl_xor(bsum, 32);
        if (kh == 0) bpart[blockIdx.x * kCP + i] = bsum;
    }
}

// Bias gradient: gb[o] = sum of part[c][o] over the nch blocks c; one
// workgroup, eight slices of blocks summed in order.
__global__ __launch_bounds__(256) void k_bias_reduce(const float *__restrict__ part, float *__restrict__ gb, int nch)
{
    __shared__ float sl[8][kCP];
    const int o = threadIdx.x & (kCP - 1), k = threadIdx.x / kCP;
    float s = 0.f;
    for (int c = k; c < nch; c += 8) s += part[c * kCP + o];
    sl[k][o] = s;
    __syncthreads();
    if (threadIdx.x < kCP) {
        float t = 0.f;
        for (int j = 0; j < 8; ++j) t += sl[j][o];
        gb[o] = t;
    }
}

// First level of the block reductions: out[s][k] = sum of part[c][k] over
// blocks c = s per .. (s + 1) per - 1 (below nch), k < len; grid (len / 256,
// slices).
__global__ __launch_bounds__(256) void k_grad_part(const float *__restrict__ part, float *__restrict__ out, int len,
                                                   int nch, int per)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= len) return;
    const int c0 = blockIdx.y * per, c1 = min(nch, c0 + per);
    float s = 0.f;
    for (int c = c0; c < c1; ++c) s += part[int64_t(c) * len + k];
    out[int64_t(blockIdx.y) * len + k] = s;
}

// g[k] = sum over nch chunks of part[chunk][k], k < len (second level of the
// gradient reductions).
__global__ __launch_bounds__(256) void k_grad_reduce(const float *__restrict__ part, float *__restrict__ g, int len,
                                                     int nch)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= len) return;
    float s = 0.f;
    for (int ch = 0; ch < nch; ++ch) s += part[int64_t(ch) * len + k];
    g[k] = s;
}

// Mean-square loss over V x 5 values: g = 2 (y - t) / (5 V) as BF16 [V][32]
// (channels 5.. zero), and each workgroup's sum of (y - t)^2 in
// lpart[block] (k_loss_add adds them to the running loss in order).
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ y, const float *__restrict__ tg,
                                             bf16 *__restrict__ g, double *__restrict__ lpart, int V)
{
    __shared__ double ws[4];
    const int v = blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (v < V) {
        const float sc = 2.0f / (5.0f * float(V));
        s8v g0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            float gv = 0.f;
            if (c < kC3) {
                const float d = y[int64_t(v) * kC3 + c] - tg[int64_t(v) * kC3 + c];
                s += double(d) * double(d);
                gv = sc * d;
            }
            g0[c] = short(to_bf16(gv));
        }
        s8v z;
#pragma unroll
        for (int c = 0; c < 8; ++c) z[c] = 0;
        s8v *gp = reinterpret_cast<s8v *>(g + int64_t(v) * kCP);
        gp[0] = g0;
#pragma unroll
        for (int k = 1; k < kCP / 8; ++k) gp[k] = z;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) lpart[blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

// loss += the sum of lpart[0 .. nb - 1] (one workgroup, slices summed in order).
__global__ __launch_bounds__(256) void k_loss_add(const double *__restrict__ lpart, double *__restrict__ loss, int nb)
{
    __shared__ double sl[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) s += lpart[b];
    sl[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) sl[threadIdx.x] += sl[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss += sl[0];
}

// BF16 [V][32] from an FP32 [V][5] gradient (test hook).
__global__ void k_pad_grad(const float *__restrict__ gy, bf16 *__restrict__ g, int V)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    for (int c = 0; c < kCP; ++c) g[int64_t(v) * kCP + c] = to_bf16(c < kC3 ? gy[int64_t(v) * kC3 + c] : 0.f);
}

// BF16 [V][8] input from FP32 [V][5] (test hook).
__global__ void k_pack_host_in(const float *__restrict__ x, bf16 *__restrict__ a, int V)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    for (int c = 0; c < kC0P; ++c) a[int64_t(v) * kC0P + c] = to_bf16(c < kC0 ? x[int64_t(v) * kC0 + c] : 0.f);
}

struct Norm {
    float in_mean[5], in_scale[5], out_scale[5];
};

// Normalised network input from the state of a block with n+1 stored points
// per direction (unique points 0..n-1).
__global__ void k_pack_state(const double *__restrict__ q, int64_t NT, int64_t off, int n, Norm nm,
                             bf16 *__restrict__ a)
{
    const int V = n * n * n;
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    const int x = v % n, y = (v / n) % n, z = v / (n * n);
    const int64_t g = off + x + int64_t(n + 1) * (y + int64_t(n + 1) * z);
    for (int c = 0; c < kC0P; ++c) {
        float f = 0.f;
        if (c < kC0) f = float((q[c * NT + g] - double(nm.in_mean[c])) / double(nm.in_scale[c]));
        a[int64_t(v) * kC0P + c] = to_bf16(f);
    }
}

// q += correction (denormalised) at every stored point of the block.
__global__ void k_add_correction(double *__restrict__ q, int64_t NT, int64_t off, int n, Norm nm,
                                 const float *__restrict__ y)
{
    const int np = n + 1;
    const int64_t P = int64_t(np) * np * np;
    const int64_t l = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (l >= P) return;
    const int x = int(l % np), yy = int((l / np) % np), z = int(l / (int64_t(np) * np));
    const int64_t v = (x % n) + int64_t(n) * ((yy % n) + int64_t(n) * (z % n));
    for (int c = 0; c < kC3; ++c) q[c * NT + off + l] += double(y[v * kC3 + c]) * double(nm.out_scale[c]);
}

// Top-hat filter of width r fine cells (trapezoidal weights 1/2, 1, .., 1,
// 1/2 over r+1 points, divided by r per direction) of a periodic fine block
// (nf+1 points per direction) sampled at the coarse points, into a coarse
// block with n+1 points per direction; nf = r n.
__global__ void k_box(const double *__restrict__ qf, int64_t NTf, int64_t offf, int nf, double *__restrict__ qc,
                      int64_t NTc, int64_t offc, int n, int r)
{
    const int np = n + 1, nfp = nf + 1;
    const int64_t P = int64_t(np) * np * np;
    const int64_t l = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (l >= P) return;
    const int x = int(l % np), y = int((l / np) % np), z = int(l / (int64_t(np) * np));
    const int h = r / 2;
    double s[5] = {0, 0, 0, 0, 0};
    for (int c = -h; c <= h; ++c) {
        const double wc = (c == -h || c == h) ? 0.5 : 1.0;
        const int fz = ((z * r + c) % nf + nf) % nf;
        for (int b = -h; b <= h; ++b) {
            const double wb = wc * ((b == -h || b == h) ? 0.5 : 1.0);
            const int fy = ((y * r + b) % nf + nf) % nf;
            for (int a = -h; a <= h; ++a) {
                const double w = wb * ((a == -h || a == h) ? 0.5 : 1.0);
                const int fx = ((x * r + a) % nf + nf) % nf;
                const int64_t g = offf + fx + int64_t(nfp) * (fy + int64_t(nfp) * fz);
                for (int m = 0; m < 5; ++m) s[m] += w * qf[m * NTf + g];
            }
        }
    }
    const double inv = 1.0 / (double(r) * r * r);
    for (int m = 0; m < 5; ++m) qc[m * NTc + offc + l] = s[m] * inv;
}

// Training pair: input = packed coarse state, target = (filtered fine - coarse) / out_scale.
__global__ void k_target(const double *__restrict__ qc, int64_t NTc, int64_t offc, const double *__restrict__ qt,
                         int64_t NTt, int64_t offt, int n, Norm nm, float *__restrict__ tg)
{
    const int V = n * n * n;
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    const int x = v % n, y = (v / n) % n, z = v / (n * n);
    const int64_t l = x + int64_t(n + 1) * (y + int64_t(n + 1) * z);
    for (int c = 0; c < kC3; ++c)
        tg[int64_t(v) * kC3 + c] =
            float((qt[c * NTt + offt + l] - qc[c * NTc + offc + l]) / double(nm.out_scale[c]));
}

// BF16 weight layouts from the FP32 parameters.
//  wf1 [32][27][8], wf2 [32][27][32], wf3 [32][27][32] (o < 5 used): forward
//  wd2 [32][27][32] = W2[o][26-t][c] transposed (c, t, o): data gradient
//  through layer 2; wd3 (o < 5) through layer 3; wd1 (c < 5) to the input
//  bb1..3 [32] FP32 biases (zero padded)
struct Wts {
    bf16 *wf1, *wf2, *wf3, *wd1, *wd2, *wd3;
    float *bb1, *bb2, *bb3;
};
__global__ void k_prep(const float *__restrict__ p, Wts W)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    constexpr int S8 = kCP * kTaps * kC0P, S32 = kCP * kTaps * kCP;
    if (e < S8) {   // wf1
        const int o = e / (kTaps * kC0P), t = (e / kC0P) % kTaps, c = e % kC0P;
        W.wf1[e] = to_bf16(c < kC0 ? p[kOW1 + (o * kTaps + t) * kC0 + c] : 0.f);
    }
    if (e < S32) {
        const int a = e / (kTaps * kCP), t = (e / kCP) % kTaps, b = e % kCP;   // [a][t][b]
        W.wf2[e] = to_bf16(p[kOW2 + (a * kTaps + t) * kC1 + b]);
        W.wf3[e] = to_bf16(a < kC3 ? p[kOW3 + (a * kTaps + t) * kC2 + b] : 0.f);
        // data gradient layouts: row a = input channel c, column b = output channel o
        const int tf = kTaps - 1 - t;
        W.wd2[e] = to_bf16(p[kOW2 + (b * kTaps + tf) * kC1 + a]);
        W.wd3[e] = to_bf16(b < kC3 ? p[kOW3 + (b * kTaps + tf) * kC2 + a] : 0.f);
        W.wd1[e] = to_bf16(a < kC0 ? p[kOW1 + (b * kTaps + tf) * kC0 + a] : 0.f);
    }
    if (e < kCP) {
        W.bb1[e] = p[kOB1 + e];
        W.bb2[e] = p[kOB2 + e];
        W.bb3[e] = e < kC3 ? p[kOB3 + e] : 0.f;
    }
}

// Flat parameter gradient from the accumulation buffers.
__device__ inline float param_grad(int k, const float *gw1, const float *gw2, const float *gw3, const float *gb)
{
    if (k < kOB1) { const int o = k / (kTaps * kC0), r = k % (kTaps * kC0); return gw1[o * kM1 + (r / kC0) * kC0P + r % kC0]; }
    if (k < kOW2) return gb[k - kOB1];
    if (k < kOB2) { const int o = (k - kOW2) / kM2, r = (k - kOW2) % kM2; return gw2[o * kM2 + r]; }
    if (k < kOW3) return gb[kCP + k - kOB2];
    if (k < kOB3) { const int o = (k - kOW3) / kM2, r = (k - kOW3) % kM2; return gw3[o * kM2 + r]; }
    return gb[2 * kCP + k - kOB3];
}

__global__ void k_gather_grad(const float *gw1, const float *gw2, const float *gw3, const float *gb, float *g)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < kNParam) g[k] = param_grad(k, gw1, gw2, gw3, gb);
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float c1, float c2)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= kNParam) return;
    constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gk = g[k];
    const float a = b1 * m1[k] + (1.f - b1) * gk, b = b2 * m2[k] + (1.f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = b;
    p[k] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void k_fill(float *a, float v, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) a[i] = v;
}

}  // namespace

// Device scratch of cnd_net_apply_noalloc: the network input, the two
// hidden activations and the output of one apply on an n^3 grid with n up
// to kScratchN (V = n^3 voxels: BF16 [V][8], [V][32], [V][32], FP32 [V][5],
// each part aligned to 256 bytes), 41 MiB.  Every GPU the library runs on
// holds its own copy until the process ends; host code reaches it only
// through the address hipGetSymbolAddress gives for the current device.
constexpr int kScratchN = 64;
constexpr size_t kScratchV = size_t(kScratchN) * kScratchN * kScratchN;
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kScratchBytes = align256(kScratchV * 8 * 2) + 2 * align256(kScratchV * 32 * 2) +
                                 align256(kScratchV * 5 * 4);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct cnd_net {
    int n = 0, V = 0, cap = 0, npairs = 0;
    hipStream_t st = nullptr;
    // recorded after every use of the network's activation buffers, on
    // whichever stream ran it; the next use waits for it first, so calls on
    // different streams (one per context) never overlap on those buffers
    hipEvent_t done = nullptr;
    Norm nm{};
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;   // parameters, gradient, Adam moments
    float *gacc = nullptr;                                            // gw1 | gw2 | gw3 | gb (3 x 32)
    // weight and bias gradients per workgroup, and per slice of workgroups
    float *gpart = nullptr, *gpart2 = nullptr, *bpart = nullptr;
    Wts W{};
    bf16 *x0 = nullptr, *a1 = nullptr, *a2 = nullptr, *g3 = nullptr, *g2 = nullptr, *g1 = nullptr;
    float *y = nullptr, *gx = nullptr, *hx = nullptr;
    double *loss = nullptr, *lpart = nullptr;   // running loss, per-workgroup sums of one pair
    bf16 *din = nullptr;    // pairs: inputs [cap][V][8]
    float *dtg = nullptr;   // pairs: targets [cap][V][5]
    long long adam_t = 0;
    std::string err;
};

namespace {

thread_local std::string g_net_err;

int nfail(cnd_net *net, int code, const std::string &m)
{
    if (net) net->err = m; else g_net_err = m;
    return code;
}

#define NHIP(net, call)                                                                             \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return nfail((net), CND_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

inline unsigned blocks_for(int64_t n, int per) { return unsigned((n + per - 1) / per); }

float *gw1(cnd_net *t) { return t->gacc; }
float *gw2(cnd_net *t) { return t->gacc + kCP * kM1; }
float *gw3(cnd_net *t) { return t->gacc + kCP * kM1 + kCP * kM2; }
float *gbias(cnd_net *t) { return t->gacc + kCP * kM1 + 2 * kCP * kM2; }
constexpr int64_t kGacc = int64_t(kCP) * kM1 + 2 * int64_t(kCP) * kM2 + 3 * kCP;

int net_begin(cnd_net *t, hipStream_t st);
int net_end(cnd_net *t, hipStream_t st);

int prep(cnd_net *t, hipStream_t st)
{
    hipLaunchKernelGGL(k_prep, dim3(blocks_for(kCP * kTaps * kCP, 256)), dim3(256), 0, st, t->p, t->W);
    NHIP(t, hipGetLastError());
    return CND_OK;
}

// forward of the input in x0 into y (normalised correction, FP32 [V][5]),
// with the hidden activations in a1, a2 (BF16 [V][32])
int forward_into(cnd_net *t, const bf16 *x0, bf16 *a1, bf16 *a2, float *y, hipStream_t st)
{
    const dim3 gv(unsigned(t->V / (32 * kConvWaves))), b(64 * kConvWaves);
    const size_t l8 = conv_lds_bytes(kC0P, t->n), l32 = conv_lds_bytes(kCP, t->n);
    hipLaunchKernelGGL((k_conv<kC0P, 0>), gv, b, l8, st, x0, t->W.wf1, t->W.bb1, nullptr, a1, t->n, kCP, kCP);
    hipLaunchKernelGGL((k_conv<kCP, 0>), gv, b, l32, st, a1, t->W.wf2, t->W.bb2, nullptr, a2, t->n, kCP, kCP);
    hipLaunchKernelGGL((k_conv<kCP, 1>), gv, b, l32, st, a2, t->W.wf3, t->W.bb3, nullptr, y, t->n, kC3, kC3);
    NHIP(t, hipGetLastError());
    return CND_OK;
}

// ... with the network's own buffers
int forward(cnd_net *t, const bf16 *x0, hipStream_t st) { return forward_into(t, x0, t->a1, t->a2, t->y, st); }

// q += correction(q) on the coarse block C with the buffers x0, a1, a2, y
// (cnd_net_apply and its variant without allocations)
int apply_into(cnd_net *t, const CtxView &C, bf16 *x0, bf16 *a1, bf16 *a2, float *y)
{
    int r;
    if ((r = net_begin(t, C.st))) return r;
    const int64_t V = t->V;
    hipLaunchKernelGGL(k_pack_state, dim3(blocks_for(V, 256)), dim3(256), 0, C.st, C.q, C.NT, C.off, t->n, t->nm, x0);
    if ((r = forward_into(t, x0, a1, a2, y, C.st))) return r;
    const int64_t P = int64_t(t->n + 1) * (t->n + 1) * (t->n + 1);
    hipLaunchKernelGGL(k_add_correction, dim3(blocks_for(P, 256)), dim3(256), 0, C.st, C.q, C.NT, C.off, t->n, t->nm,
                       y);
    return net_end(t, C.st);
}

// backward from g3 (dL/dy, BF16 [V][32]) with the activations of the last
// forward: parameter gradient into t->g; with gx, dL/dx0 into gx (FP32 [V][5]).
int backward(cnd_net *t, const bf16 *x0, float *gx, hipStream_t st)
{
    const int V = t->V;
    const dim3 b(256), gv(unsigned(V / (32 * kConvWaves))), bc(64 * kConvWaves);
    const size_t l32 = conv_lds_bytes(kCP, t->n);
    const int G = wgrad_blocks(t->n);
    // weight and bias gradients of a layer from its output gradient go and
    // input activations act (CIP channels): G block partials, then
    // kWgSlices slice sums, then the total
    auto wgrad = [&](auto cip, const bf16 *go, const bf16 *act, float *gw, float *gb) {
        constexpr int CIP = decltype(cip)::value, M = kTaps * CIP;
        hipLaunchKernelGGL(k_wgrad<CIP>, dim3(unsigned(G)), dim3(512), wgrad_lds_bytes(CIP, t->n), st, go, act,
                           t->gpart, t->bpart, t->n);
        const int per = (G + kWgSlices - 1) / kWgSlices;
        hipLaunchKernelGGL(k_grad_part, dim3(blocks_for(int64_t(kCP) * M, 256), kWgSlices), b, 0, st, t->gpart,
                           t->gpart2, kCP * M, G, per);
        hipLaunchKernelGGL(k_grad_reduce, dim3(blocks_for(int64_t(kCP) * M, 256)), b, 0, st, t->gpart2, gw, kCP * M,
                           kWgSlices);
        hipLaunchKernelGGL(k_bias_reduce, dim3(1), b, 0, st, t->bpart, gb, G);
    };
    using C32 = std::integral_constant<int, kCP>;
    using C8 = std::integral_constant<int, kC0P>;
    // layer 3
    wgrad(C32{}, t->g3, t->a2, gw3(t), gbias(t) + 2 * kCP);
    hipLaunchKernelGGL((k_conv<kCP, 2>), gv, bc, l32, st, t->g3, t->W.wd3, nullptr, t->a2, t->g2, t->n, kCP, kCP);
    // layer 2
    wgrad(C32{}, t->g2, t->a1, gw2(t), gbias(t) + kCP);
    hipLaunchKernelGGL((k_conv<kCP, 2>), gv, bc, l32, st, t->g2, t->W.wd2, nullptr, t->a1, t->g1, t->n, kCP, kCP);
    // layer 1
    wgrad(C8{}, t->g1, x0, gw1(t), gbias(t));
    if (gx)
        hipLaunchKernelGGL((k_conv<kCP, 1>), gv, bc, l32, st, t->g1, t->W.wd1, nullptr, nullptr, gx, t->n, kC0, kC0);
    hipLaunchKernelGGL(k_gather_grad, dim3(blocks_for(kNParam, 256)), b, 0, st, gw1(t), gw2(t), gw3(t), gbias(t),
                       t->g);
    NHIP(t, hipGetLastError());
    return CND_OK;
}

template <class T>
int nalloc(cnd_net *t, T **p, int64_t n)
{
    NHIP(t, hipMalloc((void **)p, size_t(std::max<int64_t>(n, 1)) * sizeof(T)));
    return CND_OK;
}

// Order work on st after every earlier use of the network's buffers, and
// mark the end of this use.
int net_begin(cnd_net *t, hipStream_t st)
{
    NHIP(t, hipStreamWaitEvent(st, t->done, 0));
    return CND_OK;
}
int net_end(cnd_net *t, hipStream_t st)
{
    NHIP(t, hipGetLastError());
    NHIP(t, hipEventRecord(t->done, st));
    return CND_OK;
}

// The address of scratch on the current device, looked up once per device
// (the first call on it) and kept; nullptr and the error if that fails.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

void *scratch_address(cnd_net *t)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        nfail(t, CND_EDEVICE, "cnd_net_apply_noalloc: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            nfail(t, CND_EDEVICE, std::string("cnd_net_apply_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[dev] = ptr;
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" {

int cnd_net_param_count(void) { return kNParam; }

const char *cnd_net_error(const cnd_net *net) { return net ? net->err.c_str() : g_net_err.c_str(); }

int cnd_net_create(int n, int capacity, unsigned long long seed, cnd_net **out)
{
    if (!out) return nfail(nullptr, CND_EINVAL, "cnd_net_create: null output");
    *out = nullptr;
    if (n < 8 || n % 8 != 0 || n > 512) return nfail(nullptr, CND_EINVAL, "cnd_net_create: n must be a multiple of 8 in 8..512");
    if (capacity < 0) return nfail(nullptr, CND_EINVAL, "cnd_net_create: negative capacity");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return nfail(nullptr, CND_EDEVICE, "no HIP device");
    {
        int dev = 0, lds = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            return nfail(nullptr, CND_EDEVICE, "cnd_net_create: device attributes");
        const size_t need = std::max(conv_lds_bytes(kCP, n), wgrad_lds_bytes(kCP, n));
        if (need > size_t(lds))
            return nfail(nullptr, CND_EDEVICE, "cnd_net_create: the convolution tiles need " + std::to_string(need) +
                                                   " bytes of LDS, the device has " + std::to_string(lds));
    }
    cnd_net *t = new cnd_net;
    t->n = n;
    t->V = n * n * n;
    t->cap = capacity;
    const int64_t V = t->V;
    int r = CND_OK;
    auto ok = [&](int rc) { if (rc) r = rc; return rc == CND_OK; };
    const int64_t S8 = int64_t(kCP) * kTaps * kC0P, S32 = int64_t(kCP) * kTaps * kCP;
    bool good = hipStreamCreateWithFlags(&t->st, hipStreamNonBlocking) == hipSuccess &&
                hipEventCreateWithFlags(&t->done, hipEventDisableTiming) == hipSuccess;
    if (!good) r = nfail(t, CND_EDEVICE, "stream or event creation failed");
    good = good && ok(nalloc(t, &t->p, kNParam)) && ok(nalloc(t, &t->g, kNParam)) && ok(nalloc(t, &t->m1, kNParam)) &&
           ok(nalloc(t, &t->m2, kNParam)) && ok(nalloc(t, &t->gacc, kGacc)) && ok(nalloc(t, &t->W.wf1, S8)) &&
           ok(nalloc(t, &t->gpart, int64_t(wgrad_blocks(n)) * kCP * kM2)) &&
           ok(nalloc(t, &t->gpart2, int64_t(kWgSlices) * kCP * kM2)) &&
           ok(nalloc(t, &t->bpart, int64_t(wgrad_blocks(n)) * kCP)) &&
           ok(nalloc(t, &t->W.wf2, S32)) && ok(nalloc(t, &t->W.wf3, S32)) && ok(nalloc(t, &t->W.wd1, S32)) &&
           ok(nalloc(t, &t->W.wd2, S32)) && ok(nalloc(t, &t->W.wd3, S32)) && ok(nalloc(t, &t->W.bb1, kCP)) &&
           ok(nalloc(t, &t->W.bb2, kCP)) && ok(nalloc(t, &t->W.bb3, kCP)) && ok(nalloc(t, &t->x0, V * kC0P)) &&
           ok(nalloc(t, &t->a1, V * kCP)) && ok(nalloc(t, &t->a2, V * kCP)) && ok(nalloc(t, &t->g3, V * kCP)) &&
           ok(nalloc(t, &t->g2, V * kCP)) && ok(nalloc(t, &t->g1, V * kCP)) && ok(nalloc(t, &t->y, V * kC3)) &&
           ok(nalloc(t, &t->gx, V * kC0)) && ok(nalloc(t, &t->hx, V * kC0)) && ok(nalloc(t, &t->loss, 1)) &&
           ok(nalloc(t, &t->lpart, (V + 255) / 256)) &&
           ok(nalloc(t, &t->din, int64_t(capacity) * V * kC0P)) && ok(nalloc(t, &t->dtg, int64_t(capacity) * V * kC3));
    if (!good) {
        std::string m = t->err;
        cnd_net_destroy(t);
        return nfail(nullptr, r ? r : CND_ENOMEM, m.empty() ? "allocation failed" : m);
    }
    // He-uniform initial weights, the last layer scaled down so the initial
    // correction is small; zero biases
    std::vector<float> h(kNParam, 0.f);
    std::mt19937_64 rng(seed);
    auto fillu = [&](int off, int cnt, int fan_in, float gain) {
        const float a = gain * std::sqrt(6.0f / float(fan_in));
        std::uniform_real_distribution<float> u(-a, a);
        for (int k = 0; k < cnt; ++k) h[off + k] = u(rng);
    };
    fillu(kOW1, kOB1 - kOW1, kTaps * kC0, 1.0f);
    fillu(kOW2, kOB2 - kOW2, kTaps * kC1, 1.0f);
    fillu(kOW3, kOB3 - kOW3, kTaps * kC2, 0.1f);
    for (int c = 0; c < 5; ++c) { t->nm.in_mean[c] = 0.f; t->nm.in_scale[c] = 1.f; t->nm.out_scale[c] = 1.f; }
    if (hipMemcpy(t->p, h.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(t->m1, 0, sizeof(float) * kNParam) != hipSuccess ||
        hipMemset(t->m2, 0, sizeof(float) * kNParam) != hipSuccess || prep(t, t->st) != CND_OK ||
        hipEventRecord(t->done, t->st) != hipSuccess || hipStreamSynchronize(t->st) != hipSuccess) {
        cnd_net_destroy(t);
        return nfail(nullptr, CND_EDEVICE, "initial upload failed");
    }
    *out = t;
    return CND_OK;
}

void cnd_net_destroy(cnd_net *t)
{
    if (!t) return;
    if (t->done) (void)hipEventSynchronize(t->done);
    if (t->st) (void)hipStreamSynchronize(t->st);
    void *ptrs[] = {t->p, t->g, t->m1, t->m2, t->gacc, t->gpart, t->gpart2, t->bpart, t->W.wf1, t->W.wf2, t->W.wf3, t->W.wd1, t->W.wd2, t->W.wd3,
                    t->W.bb1, t->W.bb2, t->W.bb3, t->x0, t->a1, t->a2, t->g3, t->g2, t->g1, t->y, t->gx, t->hx,
                    t->loss, t->lpart, t->din, t->dtg};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    if (t->done) (void)hipEventDestroy(t->done);
    if (t->st) (void)hipStreamDestroy(t->st);
    delete t;
}

int cnd_net_get_params(cnd_net *t, float *w)
{
    if (!t || !w) return nfail(t, CND_EINVAL, "cnd_net_get_params: bad argument");
    NHIP(t, hipStreamSynchronize(t->st));
    NHIP(t, hipMemcpy(w, t->p, sizeof(float) * kNParam, hipMemcpyDeviceToHost));
    return CND_OK;
}

int cnd_net_set_params(cnd_net *t, const float *w)
{
    if (!t || !w) return nfail(t, CND_EINVAL, "cnd_net_set_params: bad argument");
    NHIP(t, hipEventSynchronize(t->done));
    NHIP(t, hipStreamSynchronize(t->st));
    NHIP(t, hipMemcpy(t->p, w, sizeof(float) * kNParam, hipMemcpyHostToDevice));
    int r;
    if ((r = prep(t, t->st)) || (r = net_end(t, t->st))) return r;
    NHIP(t, hipStreamSynchronize(t->st));
    return CND_OK;
}

int cnd_net_set_norm(cnd_net *t, const double *in_mean, const double *in_scale, const double *out_scale)
{
    if (!t || !in_mean || !in_scale || !out_scale) return nfail(t, CND_EINVAL, "cnd_net_set_norm: bad argument");
    for (int c = 0; c < 5; ++c) {
        if (!(in_scale[c] > 0.0) || !(out_scale[c] > 0.0)) return nfail(t, CND_EINVAL, "cnd_net_set_norm: scales must be > 0");
        t->nm.in_mean[c] = float(in_mean[c]);
        t->nm.in_scale[c] = float(in_scale[c]);
        t->nm.out_scale[c] = float(out_scale[c]);
    }
    return CND_OK;
}

static int coarse_view(cnd_net *t, cnd_ctx *c, int b, CtxView &v, const char *what)
{
    if (ctx_block_view(c, b, v) != CND_OK) return nfail(t, CND_EINVAL, std::string(what) + ": bad context or block");
    if (v.n[0] != t->n + 1 || v.n[1] != t->n + 1 || v.n[2] != t->n + 1)
        return nfail(t, CND_EINVAL, std::string(what) + ": the block must have n+1 points per direction");
    return CND_OK;
}

int cnd_box_filter(cnd_ctx *fine, int bf, cnd_ctx *coarse, int bc)
{
    CtxView F, Cv;
    if (ctx_block_view(fine, bf, F) != CND_OK || ctx_block_view(coarse, bc, Cv) != CND_OK) return CND_EINVAL;
    const int nf = F.n[0] - 1, n = Cv.n[0] - 1;
    if (F.n[1] != F.n[0] || F.n[2] != F.n[0] || Cv.n[1] != Cv.n[0] || Cv.n[2] != Cv.n[0] || n < 1 || nf % n != 0 ||
        (nf / n) % 2 != 0)
        return CND_EINVAL;
    if (hipStreamSynchronize(F.st) != hipSuccess) return CND_EDEVICE;
    const int64_t P = int64_t(n + 1) * (n + 1) * (n + 1);
    hipLaunchKernelGGL(k_box, dim3(blocks_for(P, 256)), dim3(256), 0, Cv.st, F.q, F.NT, F.off, nf, Cv.q, Cv.NT, Cv.off,
                       n, nf / n);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(Cv.st) != hipSuccess) return CND_EDEVICE;
    return CND_OK;
}

int cnd_net_add_pair(cnd_net *t, cnd_ctx *coarse, int bc, cnd_ctx *target, int bt)
{
    if (!t) return nfail(t, CND_EINVAL, "cnd_net_add_pair: null network");
    if (t->npairs >= t->cap) return nfail(t, CND_EINVAL, "cnd_net_add_pair: the pair store is full");
    CtxView C, T;
    int r;
    if ((r = coarse_view(t, coarse, bc, C, "cnd_net_add_pair"))) return r;
    if ((r = coarse_view(t, target, bt, T, "cnd_net_add_pair (target)"))) return r;
    NHIP(t, hipStreamSynchronize(T.st));
    if ((r = net_begin(t, C.st))) return r;
    const int64_t V = t->V;
    hipLaunchKernelGGL(k_pack_state, dim3(blocks_for(V, 256)), dim3(256), 0, C.st, C.q, C.NT, C.off, t->n, t->nm,
                       t->din + int64_t(t->npairs) * V * kC0P);
    hipLaunchKernelGGL(k_target, dim3(blocks_for(V, 256)), dim3(256), 0, C.st, C.q, C.NT, C.off, T.q, T.NT, T.off,
                       t->n, t->nm, t->dtg + int64_t(t->npairs) * V * kC3);
    if ((r = net_end(t, C.st))) return r;
    NHIP(t, hipStreamSynchronize(C.st));
    return t->npairs++;
}

int cnd_net_add_pair_host(cnd_net *t, const float *x, const float *tg)
{
    if (!t || !x || !tg) return nfail(t, CND_EINVAL, "cnd_net_add_pair_host: bad argument");
    if (t->npairs >= t->cap) return nfail(t, CND_EINVAL, "cnd_net_add_pair_host: the pair store is full");
    const int64_t V = t->V;
    hipStream_t st = t->st;
    int r;
    if ((r = net_begin(t, st))) return r;
    NHIP(t, hipMemcpyAsync(t->hx, x, sizeof(float) * V * kC0, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pack_host_in, dim3(blocks_for(V, 256)), dim3(256), 0, st, t->hx,
                       t->din + int64_t(t->npairs) * V * kC0P, int(V));
    NHIP(t, hipMemcpyAsync(t->dtg + int64_t(t->npairs) * V * kC3, tg, sizeof(float) * V * kC3, hipMemcpyHostToDevice, st));
    if ((r = net_end(t, st))) return r;
    NHIP(t, hipStreamSynchronize(st));
    return t->npairs++;
}

int cnd_net_pairs(const cnd_net *t) { return t ? t->npairs : CND_EINVAL; }

int cnd_net_train(cnd_net *t, int iters, double lr, unsigned long long seed, double *mean_loss)
{
    if (!t || iters < 0 || !(lr > 0.0)) return nfail(t, CND_EINVAL, "cnd_net_train: bad argument");
    if (t->npairs < 1) return nfail(t, CND_ESTATE, "cnd_net_train: no training pairs");
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> pick(0, t->npairs - 1);
    const int64_t V = t->V;
    hipStream_t st = t->st;
    int r;
    if ((r = net_begin(t, st))) return r;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, reinterpret_cast<float *>(t->loss), 0.f, 2);
    for (int it = 0; it < iters; ++it) {
        const int k = pick(rng);
        const bf16 *x0 = t->din + int64_t(k) * V * kC0P;
        if ((r = forward(t, x0, st))) return r;
        hipLaunchKernelGGL(k_mse, dim3(blocks_for(V, 256)), dim3(256), 0, st, t->y, t->dtg + int64_t(k) * V * kC3,
                           t->g3, t->lpart, int(V));
        hipLaunchKernelGGL(k_loss_add, dim3(1), dim3(256), 0, st, t->lpart, t->loss, int(blocks_for(V, 256)));
        if ((r = backward(t, x0, nullptr, st))) return r;
        ++t->adam_t;
        const float c1 = 1.f - std::pow(0.9f, float(t->adam_t)), c2 = 1.f - std::pow(0.999f, float(t->adam_t));
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(kNParam, 256)), dim3(256), 0, st, t->p, t->g, t->m1, t->m2,
                           float(lr), c1, c2);
        if ((r = prep(t, st))) return r;
    }
    double s = 0.0;
    NHIP(t, hipMemcpyAsync(&s, t->loss, sizeof(double), hipMemcpyDeviceToHost, st));
    if ((r = net_end(t, st))) return r;
    NHIP(t, hipStreamSynchronize(st));
    if (mean_loss) *mean_loss = iters > 0 ? s / (double(iters) * 5.0 * double(V)) : 0.0;
    return CND_OK;
}

int cnd_net_eval(cnd_net *t, int first, int count, double *mean_loss)
{
    if (!t || first < 0 || count < 1 || first + count > t->npairs || !mean_loss)
        return nfail(t, CND_EINVAL, "cnd_net_eval: bad range");
    const int64_t V = t->V;
    hipStream_t st = t->st;
    int r;
    if ((r = net_begin(t, st))) return r;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, st, reinterpret_cast<float *>(t->loss), 0.f, 2);
    for (int k = first; k < first + count; ++k) {
        if ((r = forward(t, t->din + int64_t(k) * V * kC0P, st))) return r;
        hipLaunchKernelGGL(k_mse, dim3(blocks_for(V, 256)), dim3(256), 0, st, t->y, t->dtg + int64_t(k) * V * kC3,
                           t->g3, t->lpart, int(V));
        hipLaunchKernelGGL(k_loss_add, dim3(1), dim3(256), 0, st, t->lpart, t->loss, int(blocks_for(V, 256)));
    }
    double s = 0.0;
    NHIP(t, hipMemcpyAsync(&s, t->loss, sizeof(double), hipMemcpyDeviceToHost, st));
    if ((r = net_end(t, st))) return r;
    NHIP(t, hipStreamSynchronize(st));
    *mean_loss = s / (double(count) * 5.0 * double(V));
    return CND_OK;
}

int cnd_net_apply(cnd_net *t, cnd_ctx *ctx, int b)
{
    if (!t) return nfail(t, CND_EINVAL, "cnd_net_apply: null network");
    CtxView C;
    int r;
    if ((r = coarse_view(t, ctx, b, C, "cnd_net_apply"))) return r;
    return apply_into(t, C, t->x0, t->a1, t->a2, t->y);
}

int cnd_net_apply_noalloc(cnd_net *t, cnd_ctx *ctx, int b)
{
    if (!t) return nfail(t, CND_EINVAL, "cnd_net_apply_noalloc: null network");
    if (t->n > kScratchN)
        return nfail(t, CND_EINVAL, "cnd_net_apply_noalloc: grid of " + std::to_string(t->n) +
                                        "^3 cells exceeds the scratch limit of " + std::to_string(kScratchN) + "^3");
    CtxView C;
    int r;
    if ((r = coarse_view(t, ctx, b, C, "cnd_net_apply_noalloc"))) return r;
    unsigned char *base = static_cast<unsigned char *>(scratch_address(t));
    if (!base) return CND_EDEVICE;
    // every part is written in full before it is read, so nothing is cleared
    const size_t V = size_t(t->V);
    bf16 *x0 = reinterpret_cast<bf16 *>(base);
    bf16 *a1 = reinterpret_cast<bf16 *>(base + align256(V * 8 * 2));
    bf16 *a2 = reinterpret_cast<bf16 *>(base + align256(V * 8 * 2) + align256(V * 32 * 2));
    float *y = reinterpret_cast<float *>(base + align256(V * 8 * 2) + 2 * align256(V * 32 * 2));
    return apply_into(t, C, x0, a1, a2, y);
}

int cnd_net_forward_host(cnd_net *t, const float *x, float *y)
{
    if (!t || !x || !y) return nfail(t, CND_EINVAL, "cnd_net_forward_host: bad argument");
    const int64_t V = t->V;
    hipStream_t st = t->st;
    int r;
    if ((r = net_begin(t, st))) return r;
    NHIP(t, hipMemcpyAsync(t->hx, x, sizeof(float) * V * kC0, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pack_host_in, dim3(blocks_for(V, 256)), dim3(256), 0, st, t->hx, t->x0, int(V));
    if ((r = forward(t, t->x0, st))) return r;
    NHIP(t, hipMemcpyAsync(y, t->y, sizeof(float) * V * kC3, hipMemcpyDeviceToHost, st));
    if ((r = net_end(t, st))) return r;
    NHIP(t, hipStreamSynchronize(st));
    return CND_OK;
}

int cnd_net_grad_host(cnd_net *t, const float *x, const float *gy, float *gx, float *gp)
{
    if (!t || !x || !gy || !gx || !gp) return nfail(t, CND_EINVAL, "cnd_net_grad_host: bad argument");
    const int64_t V = t->V;
    hipStream_t st = t->st;
    int r;
    if ((r = net_begin(t, st))) return r;
    NHIP(t, hipMemcpyAsync(t->hx, x, sizeof(float) * V * kC0, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pack_host_in, dim3(blocks_for(V, 256)), dim3(256), 0, st, t->hx, t->x0, int(V));
    if ((r = forward(t, t->x0, st))) return r;
    NHIP(t, hipMemcpyAsync(t->gx, gy, sizeof(float) * V * kC3, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pad_grad, dim3(blocks_for(V, 256)), dim3(256), 0, st, t->gx, t->g3, int(V));
    if ((r = backward(t, t->x0, t->hx, st))) return r;
    NHIP(t, hipMemcpyAsync(gx, t->hx, sizeof(float) * V * kC0, hipMemcpyDeviceToHost, st));
    NHIP(t, hipMemcpyAsync(gp, t->g, sizeof(float) * kNParam, hipMemcpyDeviceToHost, st));
    if ((r = net_end(t, st))) return r;
    NHIP(t, hipStreamSynchronize(st));
    return CND_OK;
}

}  // extern "C"
