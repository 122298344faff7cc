// libcnd: MI355X (gfx950) implementation of Canard's compressible ILES time
// step behind the C ABI of include/cnd.h.
//
// Data layout in HBM: every field of every block is concatenated into one
// structure-of-arrays vector of NT points (block b starts at off[b], xi
// fastest inside a block), so one launch covers all blocks of all ensemble
// instances.  Pencil sweeps run one lane per (pencil, variable); lanes of a
// wave hold consecutive pencils, which makes eta and zeta sweeps coalesced.
// Interface ("halo") partial sums are read directly from the partner block's
// pencil, so the MPI exchange of the reference becomes plain device loads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cnd.h"
#include "cnd_dev.hpp"
#include "cnd_internal.hpp"
#include "cnd_ctx.hpp"


using namespace cnd;

namespace {

// ---- pointwise kernels ------------------------------------------------------
#define POINT_PROLOGUE                                                   \
    const Tile t = tiles[blockIdx.x];                                    \
    const BlkDev &B = blks[t.b];                                         \
    const int64_t npt = int64_t(B.n[0]) * B.n[1] * B.n[2];              \
    const int64_t lp = int64_t(t.q0) + threadIdx.x;                     \
    if (lp >= npt) return;                                               \
    const int64_t g = B.off + lp;

// Viscous stresses and heat flux (src/mo_physics.f90:245-299) and, from
// them, the fluxes of all three directions (src/mo_physics.f90:319-382),
// stored as F[5 d + m] for the flux sweeps.  TG: storage type of the
// gradients (float in mixed mode); stresses and fluxes are FP64.
template <class TG>
__global__ __launch_bounds__(kPts) void k_visc_flux(const BlkDev *__restrict__ blks, const InstDev *__restrict__ inst,
                                                   const Tile *__restrict__ tiles, const double *__restrict__ q,
                                                   const TG *__restrict__ grad, const double *__restrict__ met,
                                                   double *__restrict__ F, int64_t NT)
{
    POINT_PROLOGUE
    const InstDev &I = inst[B.inst];
    const double q1 = q[g], q5 = q[4 * NT + g];
    const double ri = 1.0 / q1;
    const double q2 = q[NT + g], q3 = q[2 * NT + g], q4 = q[3 * NT + g];
    const double u = q2 * ri, v = q3 * ri, w = q4 * ri;
    const double p = kGm1 * (q5 - 0.5 * ((q2 * u + q3 * v) + q4 * w));
    const double T = (kGam * p) * ri;
    const double mu = (I.srefp1dre * pow(T, 1.5)) / (T + I.srefoo);
    double mt[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) mt[c] = met[c * NT + g];
    auto G = [&](int d, int var) { return double(grad[(4 * d + var) * NT + g]); };
    auto CT = [&](int c, int var) { return (mt[c] * G(0, var) + mt[3 + c] * G(1, var)) + mt[6 + c] * G(2, var); };
    const double vtxx = CT(0, 0), vhzz = CT(1, 0), vtzx = CT(2, 0);
    const double vtxy = CT(0, 1), vtyy = CT(1, 1), vhxx = CT(2, 1);
    const double vhyy = CT(0, 2), vtyz = CT(1, 2), vtzz = CT(2, 2);
    const double s1 = CT(0, 3), s2 = CT(1, 3), s3 = CT(2, 3);
    const double r1 = mu * met[9 * NT + g];
    const double r2 = kGm1PrI * r1;
    const double d5 = kTwoThirds * ((vtxx + vtyy) + vtzz);
    double ts[9];
    const double txx = r1 * ((vtxx + vtxx) - d5), tyy = r1 * ((vtyy + vtyy) - d5), tzz = r1 * ((vtzz + vtzz) - d5);
    const double txy = r1 * (vtxy + vhzz), tyz = r1 * (vtyz + vhxx), tzx = r1 * (vtzx + vhyy);
    ts[0] = txx; ts[1] = tyy; ts[2] = tzz; ts[3] = txy; ts[4] = tyz; ts[5] = tzx;
    ts[6] = ((r2 * s1 + u * txx) + v * txy) + w * tzx;
    ts[7] = ((r2 * s2 + u * txy) + v * tyy) + w * tyz;
    ts[8] = ((r2 * s3 + u * tzx) + v * tyz) + w * tzz;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double m[3] = {mt[3 * d], mt[3 * d + 1], mt[3 * d + 2]};
        double o[5];
        dir_flux<true>(q1, q2, q3, q4, q5, u, v, w, p, m, I.umf, ts, o);
#pragma unroll
        for (int c = 0; c < 5; ++c) F[(5 * d + c) * NT + g] = o[c];
    }
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// Spectral-radius maxima of calc_time_step (src/mo_physics.f90:185-207);
// runs only on steps where mod(n,10) == 1.  Each workgroup (one point tile)
// folds its two maxima into its instance's (I.red, cleared at stage 1) with
// an atomic max of their bit patterns, which order like the values for
// non-negative doubles: the result does not depend on the order of the
// workgroups.  A workgroup whose maximum does not exceed what it reads
// there already skips the atomic.
__global__ __launch_bounds__(kPts) void k_tstep(const BlkDev *__restrict__ blks, InstDev *__restrict__ inst,
                                               const Tile *__restrict__ tiles, const double *__restrict__ q,
                                               const double *__restrict__ met, int64_t NT)
{
    __shared__ double red[2][kPts / 64];
    const Tile t = tiles[blockIdx.x];
    const BlkDev &B = blks[t.b];
    InstDev &I = inst[B.inst];
    if (I.n % 10 != 1) return;
    const int64_t npt = int64_t(B.n[0]) * B.n[1] * B.n[2];
    const int64_t lp = int64_t(t.q0) + threadIdx.x;
    double v1 = 0.0, v2 = 0.0;
    if (lp < npt) {
        const int64_t g = B.off + lp;
        const double ri = 1.0 / q[g];
        const double q2 = q[NT + g], q3 = q[2 * NT + g], q4 = q[3 * NT + g];
        const double u = q2 * ri, v = q3 * ri, w = q4 * ri;
        const double p = kGm1 * (q[4 * NT + g] - 0.5 * ((q2 * u + q3 * v) + q4 * w));
        const double T = (kGam * p) * ri;
        const double mu = (I.srefp1dre * pow(T, 1.5)) / (T + I.srefoo);
        double mt[9];
        for (int c = 0; c < 9; ++c) mt[c] = met[c * NT + g];
        double r1 = mt[0] * mt[0] + mt[1] * mt[1];
        for (int c = 2; c < 9; ++c) r1 = r1 + mt[c] * mt[c];
        const double u0 = u + I.umf[0], u1 = v + I.umf[1], u2 = w + I.umf[2];
        double r2 = fabs((mt[0] * u0 + mt[1] * u1) + mt[2] * u2) + fabs((mt[3] * u0 + mt[4] * u1) + mt[5] * u2);
        r2 = r2 + fabs((mt[6] * u0 + mt[7] * u1) + mt[8] * u2);
        const double ssi = fabs(met[9 * NT + g]);
        v1 = (sqrt(T * r1) + r2) * ssi;
        v2 = (((ri * mu) * r1) * ssi) * ssi;
    }
    v1 = wave_max(v1);
    v2 = wave_max(v2);
    if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x / 64] = v1; red[1][threadIdx.x / 64] = v2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red[0][0], c = red[1][0];
        for (int k = 1; k < kPts / 64; ++k) { a = fmax(a, red[0][k]); c = fmax(c, red[1][k]); }
        const unsigned long long ab = (unsigned long long)__double_as_longlong(a),
                                 cb = (unsigned long long)__double_as_longlong(c);
        volatile unsigned long long *seen = I.red;
        if (a >= 0.0 && ab > seen[0]) atomicMax(&I.red[0], ab);
        if (c >= 0.0 && cb > seen[1]) atomicMax(&I.red[1], cb);
    }
}

// moving-frame velocity (src/mo_physics.f90:144-169)
__device__ void movef(InstDev &I, double dtko, double dtk)
{
    if (I.nsmf == 0) {
        const double ra0 = 3.141592653589793 / I.timf;
        const double ra1 = ra0 * fmin(I.timo, I.timf), ra2 = ra0 * fmin(I.timo + dtko, I.timf);
        double fctr = 1.0 - cos(ra1), dfdt = ra0 * sin(ra2);
        double pr = 0.5 * (fctr + dtk * dfdt);
        for (int c = 0; c < 3; ++c) I.umf[c] = pr * I.uoo[c];
        fctr = sin(ra1); dfdt = ra0 * cos(ra2);
        pr = (0.5 * ra0) * (fctr + dtk * dfdt);
        for (int c = 0; c < 3; ++c) I.dudtmf[c] = pr * I.uoo[c];
    } else {
        for (int c = 0; c < 3; ++c) { I.umf[c] = I.uoo[c]; I.dudtmf[c] = 0.0; }
    }
}

// Start of stage nk: frame velocity for fluxes (mo_canard_driver.f90:261-263)
// and, in stage 1, clearing of the time-step maxima.
__global__ void k_stage_begin(InstDev *inst, int ninst, int nk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    InstDev &I = inst[i];
    const int a = nk - 2 > 0 ? (nk - 2 < 1 ? nk - 2 : 1) : 0, c = nk - 1 < 1 ? nk - 1 : 1;
    movef(I, (I.dt * double(a)) / double(4 - nk + 3), (I.dt * double(c)) / double(4 - nk + 2));
    if (nk == 1) { I.red[0] = 0ull; I.red[1] = 0ull; }
}

// dt schedule (mo_canard_driver.f90:278-285).
__global__ void k_sched(InstDev *inst, const TabDev *tabs, int ninst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    InstDev &I = inst[i];
    if (I.n % 10 == 1) {
        I.ndt = I.n;
        I.dts = I.dte;
        const double ra0 = I.cfl / __longlong_as_double((long long)I.red[0]);
        double ra1 = ra0;
        if (I.viscous) ra1 = 0.5 / __longlong_as_double((long long)I.red[1]);
        I.dte = fmin(ra0, ra1);
    }
    const int k = I.n - I.ndt;
    double r;
    if (k >= 0 && k < 64) r = tabs[i].ramp[k];
    else { const double s = sin((0.05 * 3.141592653589793) * double(k)); r = s * s; }
    I.dt = I.dts + (I.dte - I.dts) * r;
    // output time (mo_canard_driver.f90:286-291)
    I.nout = 0;
    if (I.ndata > 0) {
        const double res = (double(I.ndati + 1) * I.tmax) / double(I.ndata);
        if ((I.timo - res) * ((I.timo + I.dt) - res) <= 0.0) {
            I.nout = 1;
            I.ndati = I.ndati + 1;
        }
    }
}

// frame velocity for the update half of the stage (mo_canard_driver.f90:325-327)
__device__ void stage_mid(InstDev &I, int nk)
{
    const int c = nk - 1 < 1 ? nk - 1 : 1;
    movef(I, (I.dt * double(c)) / double(4 - nk + 2), I.dt / double(4 - nk + 1));
}

__global__ void k_stage_mid(InstDev *inst, int ninst, int nk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    stage_mid(inst[i], nk);
}

// End of step; mid4: first the frame velocity of stage 4's update half,
// where no kernel of stage 4 reads it.
__global__ void k_end_step(InstDev *inst, int ninst, int mid4)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    if (mid4) stage_mid(inst[i], 4);
    inst[i].n = inst[i].n + 1;
    inst[i].timo = inst[i].timo + inst[i].dt;
}

// Recording of intermediate results (mo_canard_driver.f90:373-378): the
// running trapezoid sum qb += dt/2 (qo + qa) over the step just taken and,
// per instance, the recorded time dtsum += dt.
__global__ __launch_bounds__(kPts) void k_record(const BlkDev *__restrict__ blks, const InstDev *__restrict__ inst,
                                                const Tile *__restrict__ tiles, const double *__restrict__ qo,
                                                const double *__restrict__ qa, double *__restrict__ qb, int64_t NT)
{
    POINT_PROLOGUE
    const double fctr = 0.5 * inst[B.inst].dt;
#pragma unroll
    for (int m = 0; m < 5; ++m) qb[m * NT + g] = qb[m * NT + g] + fctr * (qo[m * NT + g] + qa[m * NT + g]);
}

__global__ void k_record_clock(InstDev *inst, int ninst, int reset)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ninst) return;
    inst[i].dtsum = reset ? 0.0 : inst[i].dtsum + inst[i].dt;
}

// Sponge source and RK update (mo_sponge.f90:75-88,
// mo_canard_driver.f90:329-334): qa = qo - dt/(5-nk) * yaco * de.
__global__ __launch_bounds__(kPts) void k_update(const BlkDev *__restrict__ blks, const InstDev *__restrict__ inst,
                                                const Tile *__restrict__ tiles, const double *cur,
                                                const double *__restrict__ qo, double *__restrict__ de,
                                                const double *__restrict__ met, const double *__restrict__ sponge,
                                                double *qa, int64_t NT, int nk)
{
    POINT_PROLOGUE
    const InstDev &I = inst[B.inst];
    const double dtk = I.dt / double(4 - nk + 1);
    const double r = dtk * met[9 * NT + g];
    double d[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) d[m] = (de[m * NT + g] + de[(5 + m) * NT + g]) + de[(10 + m) * NT + g];
    if (I.sponge) {
        const double a = sponge[g], bb = sponge[NT + g];
        if (a != 0.0) {
            d[0] = d[0] + a * (cur[g] - 1.0);
            for (int m = 1; m < 4; ++m) d[m] = d[m] + bb * (cur[m * NT + g] - 0.0);
            d[4] = d[4] + a * (cur[4 * NT + g] - kHamHamm1);
        }
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) qa[m * NT + g] = qo[m * NT + g] - r * d[m];
}

// ---- characteristic boundaries (mo_gcbc.f90:133-340, 504-611) ---------------
struct Elem { double ao, aoi, ve[3], hv2, vn, vs; };

__device__ Elem eleme(const double cm[3], double q1, const double q24[3], double pp, const double umf[3])
{
    Elem e;
    const double rhoi = 1.0 / q1;
    e.ao = sqrt((kGam * rhoi) * pp);
    e.aoi = 1.0 / e.ao;
    for (int c = 0; c < 3; ++c) e.ve[c] = rhoi * q24[c];
    e.hv2 = 0.5 * ((e.ve[0] * e.ve[0] + e.ve[1] * e.ve[1]) + e.ve[2] * e.ve[2]);
    e.vn = (cm[0] * e.ve[0] + cm[1] * e.ve[1]) + cm[2] * e.ve[2];
    e.vs = (cm[0] * umf[0] + cm[1] * umf[1]) + cm[2] * umf[2];
    return e;
}

__device__ void xtq2r(const Elem &e, const double cm[3], double xt[5][5])
{
    const double ho = (kGm1 * e.aoi) * e.aoi, bo = 1.0 - ho * e.hv2, co = e.aoi * e.vn;
    const double *ve = e.ve;
    double dm[3], rv[3];
    for (int c = 0; c < 3; ++c) { dm[c] = e.aoi * cm[c]; rv[c] = ho * ve[c]; }
    xt[0][0] = (bo * cm[0] + dm[1] * ve[2]) - dm[2] * ve[1];
    xt[0][1] = cm[0] * rv[0]; xt[0][2] = cm[0] * rv[1] + dm[2]; xt[0][3] = cm[0] * rv[2] - dm[1]; xt[0][4] = -ho * cm[0];
    xt[1][0] = (bo * cm[1] + dm[2] * ve[0]) - dm[0] * ve[2];
    xt[1][1] = cm[1] * rv[0] - dm[2]; xt[1][2] = cm[1] * rv[1]; xt[1][3] = cm[1] * rv[2] + dm[0]; xt[1][4] = -ho * cm[1];
    xt[2][0] = (bo * cm[2] + dm[0] * ve[1]) - dm[1] * ve[0];
    xt[2][1] = cm[2] * rv[0] + dm[1]; xt[2][2] = cm[2] * rv[1] - dm[0]; xt[2][3] = cm[2] * rv[2]; xt[2][4] = -ho * cm[2];
    xt[3][0] = (1.0 - bo) - co; xt[3][1] = dm[0] - rv[0]; xt[3][2] = dm[1] - rv[1]; xt[3][3] = dm[2] - rv[2]; xt[3][4] = ho;
    xt[4][0] = (1.0 - bo) + co; xt[4][1] = -dm[0] - rv[0]; xt[4][2] = -dm[1] - rv[1]; xt[4][3] = -dm[2] - rv[2]; xt[4][4] = ho;
}

__device__ void xtr2q(const Elem &e, const double cm[3], double xt[5][5])
{
    const double bo = e.hv2 + (kHamm1 * e.ao) * e.ao, co = e.ao * e.vn;
    const double *ve = e.ve;
    double dm[3];
    for (int c = 0; c < 3; ++c) dm[c] = e.ao * cm[c];
    xt[0][0] = cm[0]; xt[0][1] = cm[1]; xt[0][2] = cm[2]; xt[0][3] = 0.5; xt[0][4] = 0.5;
    xt[1][0] = cm[0] * ve[0]; xt[1][1] = cm[1] * ve[0] - dm[2]; xt[1][2] = cm[2] * ve[0] + dm[1];
    xt[1][3] = 0.5 * (ve[0] + dm[0]); xt[1][4] = xt[1][3] - dm[0];
    xt[2][0] = cm[0] * ve[1] + dm[2]; xt[2][1] = cm[1] * ve[1]; xt[2][2] = cm[2] * ve[1] - dm[0];
    xt[2][3] = 0.5 * (ve[1] + dm[1]); xt[2][4] = xt[2][3] - dm[1];
    xt[3][0] = cm[0] * ve[2] - dm[1]; xt[3][1] = cm[1] * ve[2] + dm[0]; xt[3][2] = cm[2] * ve[2];
    xt[3][3] = 0.5 * (ve[2] + dm[2]); xt[3][4] = xt[3][3] - dm[2];
    xt[4][0] = (e.hv2 * cm[0] + dm[2] * ve[1]) - dm[1] * ve[2];
    xt[4][1] = (e.hv2 * cm[1] + dm[0] * ve[2]) - dm[2] * ve[0];
    xt[4][2] = (e.hv2 * cm[2] + dm[1] * ve[0]) - dm[0] * ve[1];
    xt[4][3] = 0.5 * (bo + co); xt[4][4] = xt[4][3] - co;
}

__device__ __forceinline__ bool is_gcbc(int c) { return c == 10 || c == 20 || c == 25 || c == 30; }
__device__ __forceinline__ bool is_avg(int c) { return c == 30 || c == 35 || c == 45; }

__device__ __forceinline__ void face_point(const BlkDev &B, int d, int q, int side, int64_t &g, int64_t &st)
{
    int64_t base;
    pencil_of(B, d, q, base, st);
    g = base + int64_t(side == 0 ? 0 : B.n[d] - 1) * st;
}

__device__ __forceinline__ double pressure(const double *q, int64_t NT, int64_t g)
{
    const double ri = 1.0 / q[g];
    const double q2 = q[NT + g], q3 = q[2 * NT + g], q4 = q[3 * NT + g];
    return kGm1 * (q[4 * NT + g] - 0.5 * ((q2 * (q2 * ri) + q3 * (q3 * ri)) + q4 * (q4 * ri)));
}

// gcbc_setup: characteristic transform of the face derivative (BC 10) or of
// the full right-hand side (walls, curvilinear interfaces).  grid (tiles, side).
__global__ __launch_bounds__(kLine) void k_gcbc_setup(const BlkDev *__restrict__ blks,
                                                     const InstDev *__restrict__ inst,
                                                     const Tile *__restrict__ tiles, int d, const double *q,
                                                     const double *de, const double *met, const double *cmf,
                                                     double *drva, int64_t NT, int64_t FT)
{
    const Tile t = tiles[blockIdx.x];
    const BlkDev &B = blks[t.b];
    const int side = blockIdx.y, qq = t.q0 + threadIdx.x;
    const int code = B.bc[d][side];
    if (qq >= npencils(B, d) || !is_gcbc(code)) return;
    int64_t g, st;
    face_point(B, d, qq, side, g, st);
    const int64_t fi = B.foff[d] + qq;
    const double ra0 = double(((20 - code) * (25 - code) * (30 - code)) / 3000), ra1 = 1.0 - ra0;
    double cm[3], q24[3], cha[5], yc[5], xt[5][5];
    for (int c = 0; c < 3; ++c) { cm[c] = cmf[(side * 3 + c) * FT + fi]; q24[c] = q[(c + 1) * NT + g]; }
    const Elem e = eleme(cm, q[g], q24, pressure(q, NT, g), inst[B.inst].umf);
    xtq2r(e, cm, xt);
    for (int c = 0; c < 5; ++c)
        cha[c] = ra0 * drva[(side * 5 + c) * FT + fi] + ra1 * ((de[c * NT + g] + de[(5 + c) * NT + g]) + de[(10 + c) * NT + g]);
    const double y = met[9 * NT + g];
    for (int c = 0; c < 5; ++c) yc[c] = y * cha[c];
    for (int r = 0; r < 5; ++r) {
        double s = 0.0;
        for (int c = 0; c < 5; ++c) s = s + xt[r][c] * yc[c];
        drva[(side * 5 + r) * FT + fi] = s;
    }
}

// gcbc_update of face (d, side); faces are applied in the reference's order.
__global__ __launch_bounds__(kLine) void k_gcbc_update(const BlkDev *__restrict__ blks,
                                                      const InstDev *__restrict__ inst,
                                                      const TabDev *__restrict__ tabs,
                                                      const Tile *__restrict__ tiles, int d, int side,
                                                      const double *q, double *de, const double *met,
                                                      const double *cmf, const double *drva, int64_t NT, int64_t FT)
{
    const Tile t = tiles[blockIdx.x];
    const BlkDev &B = blks[t.b];
    const int qq = t.q0 + threadIdx.x;
    const int code = B.bc[d][side];
    if (qq >= npencils(B, d) || !is_gcbc(code)) return;
    const InstDev &I = inst[B.inst];
    int64_t g, st;
    face_point(B, d, qq, side, g, st);
    const int64_t fi = B.foff[d] + qq;
    const double ra0 = side == 0 ? 1.0 : -1.0;
    double cm[3], q24[3], cha[5], dha[5], dif[5], xt[5][5];
    for (int c = 0; c < 3; ++c) { cm[c] = cmf[(side * 3 + c) * FT + fi]; q24[c] = q[(c + 1) * NT + g]; }
    const Elem e = eleme(cm, q[g], q24, pressure(q, NT, g), I.umf);
    for (int c = 0; c < 5; ++c) cha[c] = drva[(side * 5 + c) * FT + fi];
    if (code == 10) {
        if (ra0 * ((e.vn + e.vs) + e.ao) > 0.0) cha[3] = 0.0;
        if (ra0 * ((e.vn + e.vs) - e.ao) > 0.0) cha[4] = 0.0;
    } else if (code == 20 || code == 25) {
        const double dtwi = 1.0 / (4.0 * I.dt + kSml);
        double sdu = 0.0;
        for (int c = 0; c < 3; ++c) sdu = sdu + cm[c] * I.dudtmf[c];
        cha[3 + side] = cha[4 - side] + (((2.0 * ra0) * e.aoi) * q[g]) * (sdu + dtwi * (e.vn + e.vs));
    } else {
        const BlkDev &P = blks[B.part[d][side]];
        const int64_t pi = P.foff[d] + qq;
        for (int c = 0; c < 5; ++c) dha[c] = drva[((1 - side) * 5 + c) * FT + pi];
        if (ra0 * (e.vn + e.vs) > 0.0) { cha[0] = dha[0]; cha[1] = dha[1]; cha[2] = dha[2]; }
        if (ra0 * ((e.vn + e.vs) + e.ao) > 0.0) cha[3] = dha[3];
        if (ra0 * ((e.vn + e.vs) - e.ao) > 0.0) cha[4] = dha[4];
    }
    xtr2q(e, cm, xt);
    for (int c = 0; c < 5; ++c) dif[c] = cha[c] - drva[(side * 5 + c) * FT + fi];
    for (int r = 0; r < 5; ++r) {
        double s = 0.0;
        for (int c = 0; c < 5; ++c) s = s + xt[r][c] * dif[c];
        dha[r] = s;
    }
    // spreading weights (gcbc_init, mo_gcbc.f90:82-129): 1/yaco of the face
    // point, scaled by sbci for inward points, divided by the number of
    // characteristic faces whose 5-point band covers the point.
    const double res = 1.0 / met[9 * NT + g];
    const int sgn = side == 0 ? 1 : -1;
    int64_t base; int64_t st2;
    pencil_of(B, d, qq, base, st2);
    const int i0 = side == 0 ? 0 : B.n[d] - 1;
    for (int ii = 0; ii <= 4; ++ii) {
        const int ip = i0 + sgn * ii;
        const int64_t gl = base + int64_t(ip) * st2;
        const int64_t lloc = gl - B.off;
        int idx[3];
        idx[0] = int(lloc % B.n[0]);
        idx[1] = int((lloc / B.n[0]) % B.n[1]);
        idx[2] = int(lloc / (int64_t(B.n[0]) * B.n[1]));
        int cnt = 0;
        for (int dd = 0; dd < 3; ++dd)
            for (int s = 0; s < 2; ++s)
                if (is_gcbc(B.bc[dd][s])) {
                    const int dist = s == 0 ? idx[dd] : B.n[dd] - 1 - idx[dd];
                    if (dist <= 4) ++cnt;
                }
        const double wv = ii == 0 ? res : res * tabs[B.inst].gcbc_w[ii - 1];
        const double sbcc = wv / double(cnt);
        // the right-hand side is the sum of the three direction buffers: the
        // corrected sum goes to the first, the others are cleared
        for (int c = 0; c < 5; ++c) {
            const double tot = (de[c * NT + gl] + de[(5 + c) * NT + gl]) + de[(10 + c) * NT + gl];
            de[c * NT + gl] = tot + sbcc * dha[c];
            de[(5 + c) * NT + gl] = 0.0;
            de[(10 + c) * NT + gl] = 0.0;
        }
    }
}

// wall_condition_update (mo_gcbc.f90:424-459) on face (d, side)
__global__ __launch_bounds__(kLine) void k_wall(const BlkDev *__restrict__ blks, const InstDev *__restrict__ inst,
                                               const Tile *__restrict__ tiles, int d, int side, double *q, int64_t NT)
{
    const Tile t = tiles[blockIdx.x];
    const BlkDev &B = blks[t.b];
    const int qq = t.q0 + threadIdx.x;
    const int code = B.bc[d][side];
    if (qq >= npencils(B, d) || (code != 20 && code != 25)) return;
    int64_t g, st;
    face_point(B, d, qq, side, g, st);
    if (code == 20) {
        double s = 0.0;
        for (int c = 1; c <= 3; ++c) s = s + q[c * NT + g] * q[c * NT + g];
        q[4 * NT + g] = 0.0 * q[4 * NT + g] + 1.0 * (kHamHamm1 * pow(q[g], kGam) + (0.5 * s) / q[g]);
    } else {
        const double *umf = inst[B.inst].umf;
        const double fctr = 1.0 * q[g];
        for (int c = 1; c <= 3; ++c) q[c * NT + g] = 0.0 * q[c * NT + g] - fctr * umf[c - 1];
        double s = 0.0;
        for (int c = 1; c <= 3; ++c) s = s + q[c * NT + g] * q[c * NT + g];
        q[4 * NT + g] = 0.0 * q[4 * NT + g] + 1.0 * (kHamHamm1 * q[g] + (0.5 * s) / q[g]);
    }
}

// ---- averaging of duplicated points ------------------------------------------
// Junction averaging (average_point, then average_line: mo_domdcomp.f90:439-559)
// as one gather and one scatter over a JunctionProgram: an edge entry reads a
// point that corner averaging changes through that corner's entry, which is
// recomputed in the same order of additions.
__device__ __forceinline__ double junc_corner(const JuncDev &J, const double *q, int j)
{
    double acc = q[J.tgt[j]];
    for (int k = J.first[j]; k < J.first[j + 1]; ++k) acc = acc + q[J.src[k]];
    return J.scale[j] * acc;
}

// value of source k of an entry (variable plane q): the point itself, or a
// corner's average
__device__ __forceinline__ double junc_src(const JuncDev &J, const double *q, int k)
{
    const int c = J.sc[k];
    return c < 0 ? q[J.src[k]] : junc_corner(J, q, c);
}

// One thread per (entry, variable = blockIdx.y).  The sources are loaded a
// chunk at a time before they are added, in order, so their loads overlap.
__global__ void k_junc_gather(const JuncDev J, const double *__restrict__ q, int64_t NT)
{
    constexpr int kChunk = 4;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, m = blockIdx.y;
    if (i >= J.n || !J.write[i]) return;
    const double *qm = q + m * NT;
    const int k0 = J.first[i], k1 = J.first[i + 1];
    double acc = J.tc[i] < 0 ? qm[J.tgt[i]] : junc_corner(J, qm, J.tc[i]);
    for (int kb = k0; kb < k1; kb += kChunk) {
        double v[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) v[j] = kb + j < k1 ? junc_src(J, qm, kb + j) : 0.0;
#pragma unroll
        for (int j = 0; j < kChunk; ++j)
            if (kb + j < k1) acc = acc + v[j];
    }
    J.tmp[m * int64_t(J.n) + i] = J.scale[i] * acc;
}

__global__ void k_junc_scatter(const JuncDev J, double *q, int64_t NT)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, m = blockIdx.y;
    if (i >= J.n || !J.write[i]) return;
    q[m * NT + J.tgt[i]] = J.tmp[m * int64_t(J.n) + i];
}

// The face tiles of all three directions in one launch: blockIdx.x runs over
// the tiles of direction 0, then 1, then 2.
struct FaceSet {
    const Tile *tiles[3];
    int n0, n01;
    double *sv[3];
    int64_t FT[3];
    __device__ Tile tile(int bx, int &d) const
    {
        d = bx < n0 ? 0 : (bx < n01 ? 1 : 2);
        return tiles[d][bx - (d == 0 ? 0 : (d == 1 ? n0 : n01))];
    }
};

// pencil of direction d through point (i, j, k) (the inverse of pencil_of)
__device__ __forceinline__ int pencil_at(const BlkDev &B, int d, const int idx[3])
{
    return d == 0 ? idx[1] + B.n[1] * idx[2] : (d == 1 ? idx[0] + B.n[0] * idx[2] : idx[0] + B.n[0] * idx[1]);
}

__device__ __forceinline__ bool on_avg_face(const BlkDev &B, int d, int s, const int idx[3])
{
    return is_avg(B.bc[d][s]) && idx[d] == (s == 0 ? 0 : B.n[d] - 1);
}

// average_surface (mo_gcbc.f90:342-422): snapshot of every averaged face, then
// running means face by face in the reference's (direction, side) order.
__global__ __launch_bounds__(kLine) void k_surf_snap(const BlkDev *__restrict__ blks, const FaceSet fs,
                                                    const double *q, int64_t NT)
{
    int d;
    const Tile t = fs.tile(blockIdx.x, d);
    const BlkDev &B = blks[t.b];
    const int side = blockIdx.y, qq = t.q0 + threadIdx.x;
    if (qq >= npencils(B, d) || !is_avg(B.bc[d][side])) return;
    int64_t g, st;
    face_point(B, d, qq, side, g, st);
    const int64_t fi = B.foff[d] + qq;
    for (int m = 0; m < 5; ++m) fs.sv[d][(side * 5 + m) * fs.FT[d] + fi] = q[m * NT + g];
}

// One thread per (face, point); a point on several averaged faces of its
// block (an edge or corner) is written once, by the thread of its last face,
// which applies the running means of all of them in order.
__global__ __launch_bounds__(kLine) void k_surf_avg(const BlkDev *__restrict__ blks, const FaceSet fs, double *q,
                                                   int64_t NT)
{
    int d;
    const Tile t = fs.tile(blockIdx.x, d);
    const BlkDev &B = blks[t.b];
    const int side = blockIdx.y, qq = t.q0 + threadIdx.x;
    if (qq >= npencils(B, d) || !is_avg(B.bc[d][side])) return;
    int64_t g, st;
    face_point(B, d, qq, side, g, st);
    const int64_t lloc = g - B.off;
    int idx[3];
    idx[0] = int(lloc % B.n[0]);
    idx[1] = int((lloc / B.n[0]) % B.n[1]);
    idx[2] = int(lloc / (int64_t(B.n[0]) * B.n[1]));
    for (int dd = d + 1; dd < 3; ++dd)
        if (on_avg_face(B, dd, 0, idx) || on_avg_face(B, dd, 1, idx)) return;
    // the point's value: its own face's snapshot (nothing writes q between the
    // two launches), read coalesced where q's face points are a pencil apart
    const double *own = fs.sv[d] + side * 5 * fs.FT[d] + B.foff[d] + qq;
    double v[5];
    for (int m = 0; m < 5; ++m) v[m] = own[m * fs.FT[d]];
    double cnt = 1.0;
    for (int f = 0; f <= 2 * d + side; ++f) {
        const int dd = f / 2, s = f % 2;
        if (!on_avg_face(B, dd, s, idx)) continue;
        cnt = cnt + 1.0;
        const double r2 = 1.0 / cnt;
        const int64_t pi = blks[B.part[dd][s]].foff[dd] + pencil_at(B, dd, idx);
        const double *sv = fs.sv[dd] + (1 - s) * 5 * fs.FT[dd] + pi;
        for (int m = 0; m < 5; ++m) v[m] = r2 * ((cnt - 1.0) * v[m] + sv[m * fs.FT[dd]]) + 0.0 * v[m];
    }
    for (int m = 0; m < 5; ++m) q[m * NT + g] = v[m];
}

// ---- grid metrics (mo_grid.f90:107-186) ----------------------------------------
// dx[(3c+d)] = d(coord c)/d(xi_d)
__global__ __launch_bounds__(kPts) void k_cross(const BlkDev *__restrict__ blks, const Tile *__restrict__ tiles,
                                               const double *__restrict__ dx, double *__restrict__ met, int64_t NT)
{
    POINT_PROLOGUE
    auto D = [&](int c, int d) { return dx[(3 * c + d) * NT + g]; };
    met[0 * NT + g] = D(1, 1) * D(2, 2) - D(2, 1) * D(1, 2);
    met[1 * NT + g] = D(2, 1) * D(0, 2) - D(0, 1) * D(2, 2);
    met[2 * NT + g] = D(0, 1) * D(1, 2) - D(1, 1) * D(0, 2);
    met[3 * NT + g] = D(1, 2) * D(2, 0) - D(2, 2) * D(1, 0);
    met[4 * NT + g] = D(2, 2) * D(0, 0) - D(0, 2) * D(2, 0);
    met[5 * NT + g] = D(0, 2) * D(1, 0) - D(1, 2) * D(0, 0);
    met[6 * NT + g] = D(1, 0) * D(2, 1) - D(2, 0) * D(1, 1);
    met[7 * NT + g] = D(2, 0) * D(0, 1) - D(0, 0) * D(2, 1);
    met[8 * NT + g] = D(0, 0) * D(1, 1) - D(1, 0) * D(0, 1);
}

__global__ __launch_bounds__(kPts) void k_yaco(const BlkDev *__restrict__ blks, const Tile *__restrict__ tiles,
                                              const double *__restrict__ dx, double *__restrict__ met, int64_t NT)
{
    POINT_PROLOGUE
    auto D = [&](int c, int d) { return dx[(3 * c + d) * NT + g]; };
    auto M = [&](int k) { return met[k * NT + g]; };
    double s = D(0, 0) * M(0);
    s = s + D(0, 1) * M(3);
    s = s + D(0, 2) * M(6);
    s = s + D(1, 0) * M(1);
    s = s + D(1, 1) * M(4);
    s = s + D(1, 2) * M(7);
    s = s + D(2, 0) * M(2);
    s = s + D(2, 1) * M(5);
    s = s + D(2, 2) * M(8);
    met[9 * NT + g] = 3.0 / s;
}

__global__ __launch_bounds__(kLine) void k_normals(const BlkDev *__restrict__ blks, const Tile *__restrict__ tiles,
                                                  int d, const double *met, double *cmf, int64_t NT, int64_t FT)
{
    const Tile t = tiles[blockIdx.x];
    const BlkDev &B = blks[t.b];
    const int side = blockIdx.y, qq = t.q0 + threadIdx.x;
    if (qq >= npencils(B, d)) return;
    int64_t g, st;
    face_point(B, d, qq, side, g, st);
    const int64_t fi = B.foff[d] + qq;
    double rv[3];
    for (int c = 0; c < 3; ++c) rv[c] = met[9 * NT + g] * met[(3 * d + c) * NT + g];
    const double f = 1.0 / sqrt((rv[0] * rv[0] + rv[1] * rv[1]) + rv[2] * rv[2]);
    for (int c = 0; c < 3; ++c) cmf[(side * 3 + c) * FT + fi] = f * rv[c];
}

// Sponge coefficients (spongeup, mo_sponge.f90:34-71): sigma/yaco on the
// points whose sigma exceeds 1e-8, zero elsewhere.
__global__ __launch_bounds__(kPts) void k_sponge_init(const BlkDev *__restrict__ blks,
                                                     const Tile *__restrict__ tiles, const double *__restrict__ xyz,
                                                     const double *__restrict__ met, const double *__restrict__ sp_par,
                                                     double *__restrict__ sponge, int64_t NT)
{
    POINT_PROLOGUE
    const double *P = sp_par + 5 * B.inst;   // szco, szth0, szth1, doml0, doml1
    const double pi = 3.141592653589793;
    const double tmpa = pi / P[1], tmpb = pi / P[2];
    const double x = xyz[g], ra3 = 0.0 * xyz[NT + g];
    const double ra0 = tmpa * (x - ((ra3 - P[3]) + P[1]));
    const double ra1 = tmpb * (((ra3 + P[4]) - P[2]) - x);
    const double c0 = cos(fmax(fmin(ra0, pi), 0.0)), c1 = cos(fmax(fmin(ra1, pi), 0.0));
    const double s1 = (P[0] * 0.5) * ((2.0 + c0) + c1);
    const double s2 = (P[0] * 0.5) * (1.0 + c0);
    const bool on = s1 > kSml;
    sponge[g] = on ? s1 / met[9 * NT + g] : 0.0;
    sponge[NT + g] = on ? s2 / met[9 * NT + g] : 0.0;
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

}  // namespace

namespace {
thread_local std::string g_noctx_err;
}  // namespace

int cnd::fail(cnd_ctx *c, int code, const std::string &m)
{
    if (c) c->err = m; else g_noctx_err = m;
    return code;
}

namespace {

int upload_inst(cnd_ctx *c)
{
    std::vector<InstDev> h(c->ninst);
    for (int i = 0; i < c->ninst; ++i) {
        const cnd_params &p = c->prm[i];
        InstDev &I = h[i];
        std::memset(&I, 0, sizeof I);
        // mo_physics.f90:101-111
        const double am = std::sqrt((p.amach1 * p.amach1 + p.amach2 * p.amach2) + p.amach3 * p.amach3);
        I.reoo = p.reoo;
        if (am > kSml) I.reoo = I.reoo / am;
        I.srefoo = 111 / p.tempoo;
        I.srefp1dre = (I.srefoo + 1.0) / I.reoo;
        I.uoo[0] = p.amach1; I.uoo[1] = p.amach2; I.uoo[2] = p.amach3;
        I.timf = p.timf; I.cfl = p.cfl; I.nsmf = p.nsmf; I.viscous = p.viscous; I.sponge = p.sponge;
        I.ndata = 0;
        I.ndati = -1;
    }
    HIPX(c, hipMemcpy(c->dinst, h.data(), sizeof(InstDev) * c->ninst, hipMemcpyHostToDevice));
    return CND_OK;
}


hipEvent_t pool_event(cnd_ctx *c)
{
    if (!c->pool.empty()) { hipEvent_t e = c->pool.back(); c->pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Bracket a kernel group of section sec (no-op unless profiling eagerly).
struct Section {
    cnd_ctx *c;
    int sec, launches;
    hipEvent_t a = nullptr;
    Section(cnd_ctx *c_, int s, int n) : c(c_), sec(s), launches(n)
    {
        if (c->prof && !c->capturing && (a = pool_event(c))) (void)hipEventRecord(a, c->st);
    }
    ~Section()
    {
        if (!a) return;
        hipEvent_t b = pool_event(c);
        if (!b) { c->pool.push_back(a); return; }
        (void)hipEventRecord(b, c->st);
        c->marks.push_back({sec, a, b, launches});
    }
};


// Segments per pencil a launch of shape sh along d needs (the largest over
// blocks, rounded up to whole waves per variable where that fits), or 0 when
// some block's pencils do not fit the shape.
// LDS bytes of a segmented launch (cnd_seg.hpp SegShape::lds_bytes).
size_t seg_lds_bytes(const ShapeInfo &G, int nv, int sp, int nmax, bool has_part)
{
    const size_t chunk = size_t(G.C) * nv * sp * G.P, part = size_t(2) * kSegEndRows * nv * G.P,
                 ex = size_t(4) * nv * sp * G.P, lu = size_t(5) * G.L * sp;
    const size_t tile = size_t(nv) * G.P * size_t(nmax | 1);   // pencils n | 1 apart (SegShape::pencil_stride)
    if (G.XI) return sizeof(double) * (std::max(tile, lu + ex) + size_t(8) * G.SPMAX + (has_part ? part : 0));
    return sizeof(double) * (std::max(std::max(chunk, part), ex) + lu + part + size_t(8) * G.SPMAX);
}

int shape_sp(cnd_ctx *c, int sh, int d, int nv)
{
    const ShapeInfo &G = kShapes[sh];
    if (G.XI != (d == 0)) return 0;
    int sp = 1;
    for (const BlockSpec &B : c->spec) {
        const int S = seg_count(B.n[d], G.L);
        if (S == 0) return 0;
        sp = std::max(sp, S);
    }
    if (sp > G.SPMAX) return 0;
    const int per_wave = 64 / G.P;
    const int up = (sp + per_wave - 1) / per_wave * per_wave;
    sp = up <= G.SPMAX ? up : sp;
    if (nv * G.P * sp > 1024) return 0;
    if (seg_lds_bytes(G, nv, sp, c->nmax[d], c->part_other[d]) > size_t(c->lds_max)) return 0;
    return sp;
}

// Filter phase (mo_canard_driver.f90:233-252): eta, zeta, xi sweeps of all
// five variables, out of place so partner reads see the pre-sweep field; the
// filtered state (which is also qo) ends in qo.
int enqueue_filter(cnd_ctx *c)
{
    Section sec(c, CND_SEC_FILTER, 3);
    int r;
    const bool mix = c->precision == CND_MIXED;
    if ((r = run_filter_field(c, 1, c->qa, c->qo, 5, mix))) return r;
    if ((r = run_filter_field(c, 2, c->qo, c->qa, 5, mix))) return r;
    if ((r = run_filter_field(c, 0, c->qa, c->qo, 5, mix))) return r;
    return CND_OK;
}

// Sweeps that neither read what the other writes nor share scratch (the
// tiled sweeps' ys) run side by side: launch() goes to the side stream
// after everything queued so far on st; join_side makes st wait for it.
template <class F>
int on_side(cnd_ctx *c, F launch)
{
    HIPX(c, hipEventRecord(c->ev_fork, c->st));
    HIPX(c, hipStreamWaitEvent(c->st2, c->ev_fork, 0));
    std::swap(c->st, c->st2);
    const int r = launch();
    std::swap(c->st, c->st2);
    if (r) return r;
    HIPX(c, hipEventRecord(c->ev_join, c->st2));
    return CND_OK;
}

int join_side(cnd_ctx *c)
{
    HIPX(c, hipStreamWaitEvent(c->st, c->ev_join, 0));
    return CND_OK;
}

int enqueue_stage(cnd_ctx *c, int nk)
{
    const double *cur = nk == 1 ? c->qo : c->qa;
    const unsigned ib = unsigned((c->ninst + 63) / 64);
    hipLaunchKernelGGL(k_stage_begin, dim3(ib), dim3(64), 0, c->st, c->dinst, c->ninst, nk);
    int r;
    // the time step of stage 1 runs beside the xi gradient sweep of viscous
    // stages, which reads neither dt nor the maxima (its time then counts in
    // the gradient section)
    const bool tstep_side = nk == 1 && c->any_visc && c->overlap;
    if (nk == 1) {
        auto tstep = [&] {
            hipLaunchKernelGGL(k_tstep, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dinst, c->dpt, cur, c->met,
                               c->NT);
            hipLaunchKernelGGL(k_sched, dim3(ib), dim3(64), 0, c->st, c->dinst, c->dtab, c->ninst);
            return check_launch(c);
        };
        Section sec(c, CND_SEC_TSTEP, 1);
        if ((r = tstep_side ? on_side(c, tstep) : tstep())) return r;
    }
    if (c->any_visc) {
        // the eta sweep beside the zeta sweep when it is a segmented one
        Section sec(c, CND_SEC_GRAD, 3);
        if ((r = run_grad(c, 0, cur))) return r;
        if (tstep_side && (r = join_side(c))) return r;
        if (c->overlap && (c->precision == CND_MIXED || seg_on(c, CND_KIND_GRAD, 1))) {
            if ((r = on_side(c, [&] { return run_grad(c, 1, cur); })) || (r = run_grad(c, 2, cur)) ||
                (r = join_side(c)))
                return r;
        } else {
            for (int d = 1; d < 3; ++d)
                if ((r = run_grad(c, d, cur))) return r;
        }
    }
    if (c->any_visc) {
        Section sec(c, CND_SEC_STRESS, 1);
        if (c->precision == CND_MIXED)
            hipLaunchKernelGGL(k_visc_flux<float>, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dinst, c->dpt, cur,
                               c->gradf, c->met, c->fl, c->NT);
        else
            hipLaunchKernelGGL(k_visc_flux<double>, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dinst, c->dpt,
                               cur, c->grad, c->met, c->fl, c->NT);
    }
    const bool fused = fuse_possible(c) && (c->precision == CND_MIXED || seg_on(c, CND_KIND_FLUX, c->fuse_dir));
    {
    // the flux sweeps the update does not take in, the eta one beside the
    // others when it is a segmented one
    Section sec(c, CND_SEC_FLUX, fused ? 2 : 3);
    const int other = fused && c->fuse_dir == 0 ? 2 : 0;
    if (c->overlap && (c->precision == CND_MIXED || seg_on(c, CND_KIND_FLUX, 1))) {
        if ((r = on_side(c, [&] { return run_flux(c, 1, cur); })) || (r = run_flux(c, other, cur)) ||
            (!fused && (r = run_flux(c, 2, cur))) || (r = join_side(c)))
            return r;
    } else {
        for (int d = 0; d < 3; ++d)
            if ((!fused || d != c->fuse_dir) && (r = run_flux(c, d, cur))) return r;
    }
    }
    if (c->any_gcbc) {
        Section sec(c, CND_SEC_GCBC, 9);
        for (int d = 0; d < 3; ++d)
            hipLaunchKernelGGL(k_gcbc_setup, tiles_grid(c, d, 2), dim3(kLine), 0, c->st, c->dblk, c->dinst,
                               c->dtiles[d], d, cur, c->de, c->met, c->cmf[d], c->drva[d], c->NT, c->FT[d]);
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s)
                hipLaunchKernelGGL(k_gcbc_update, tiles_grid(c, d, 1), dim3(kLine), 0, c->st, c->dblk, c->dinst,
                                   c->dtab, c->dtiles[d], d, s, cur, c->de, c->met, c->cmf[d], c->drva[d], c->NT,
                                   c->FT[d]);
    }
    if (fused) {
        Section sec(c, CND_SEC_UPDATE, 1);
        if ((r = run_flux_update(c, nk, c->qa))) return r;
    } else {
        Section sec(c, CND_SEC_UPDATE, 1);
        hipLaunchKernelGGL(k_update, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dinst, c->dpt, cur, c->qo, c->de,
                           c->met, c->sponge, c->qa, c->NT, nk);
    }
    // The update half's frame velocity is read only by the wall condition
    // (and, after the step, by the output); without walls the next stage's
    // frame velocity replaces it unread, and stage 4's is set at the end of
    // the step.
    if (c->any_wall) hipLaunchKernelGGL(k_stage_mid, dim3(ib), dim3(64), 0, c->st, c->dinst, c->ninst, nk);
    Section avg(c, CND_SEC_AVG, 1);
    if (c->any_wall)
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s)
                hipLaunchKernelGGL(k_wall, tiles_grid(c, d, 1), dim3(kLine), 0, c->st, c->dblk, c->dinst,
                                   c->dtiles[d], d, s, c->qa, c->NT);
    if (c->jn.n > 0) {
        const unsigned nb = unsigned((c->jn.n + 127) / 128);
        hipLaunchKernelGGL(k_junc_gather, dim3(nb, 5), dim3(128), 0, c->st, c->jn, c->qa, c->NT);
        hipLaunchKernelGGL(k_junc_scatter, dim3(nb, 5), dim3(128), 0, c->st, c->jn, c->qa, c->NT);
    }
    if (c->any_avg) {
        FaceSet fs;
        for (int d = 0; d < 3; ++d) { fs.tiles[d] = c->dtiles[d]; fs.sv[d] = c->surf[d]; fs.FT[d] = c->FT[d]; }
        fs.n0 = ntiles(c, 0);
        fs.n01 = ntiles(c, 0) + ntiles(c, 1);
        const dim3 grid(unsigned(fs.n01 + ntiles(c, 2)), 2);
        hipLaunchKernelGGL(k_surf_snap, grid, dim3(kLine), 0, c->st, c->dblk, fs, c->qa, c->NT);
        hipLaunchKernelGGL(k_surf_avg, grid, dim3(kLine), 0, c->st, c->dblk, fs, c->qa, c->NT);
    }
    return check_launch(c);
}

int enqueue_end(cnd_ctx *c)
{
    const unsigned ib = unsigned((c->ninst + 63) / 64);
    hipLaunchKernelGGL(k_end_step, dim3(ib), dim3(64), 0, c->st, c->dinst, c->ninst, c->any_wall ? 0 : 1);
    if (c->recording) {
        hipLaunchKernelGGL(k_record, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dinst, c->dpt, c->qo, c->qa,
                           c->qb, c->NT);
        hipLaunchKernelGGL(k_record_clock, dim3(ib), dim3(64), 0, c->st, c->dinst, c->ninst, 0);
    }
    return check_launch(c);
}

int enqueue_step(cnd_ctx *c)
{
    int r;
    if ((r = enqueue_filter(c))) return r;
    for (int nk = 1; nk <= 4; ++nk)
        if ((r = enqueue_stage(c, nk))) return r;
    return enqueue_end(c);
}

}  // namespace

// ==== C ABI =========================================================================
int cnd::ctx_block_view(cnd_ctx *c, int b, CtxView &v)
{
    if (!c || b < 0 || b >= c->nblk) return CND_EINVAL;
    v.q = c->state_in_qo ? c->qo : c->qa;
    v.NT = c->NT;
    v.off = c->off[b];
    for (int d = 0; d < 3; ++d) v.n[d] = c->spec[b].n[d];
    v.st = c->st;
    return CND_OK;
}

extern "C" {

const char *cnd_last_error(const cnd_ctx *c) { return c ? c->err.c_str() : g_noctx_err.c_str(); }

int cnd_create(const cnd_params *params, int ninst, const cnd_block *blocks, int nblocks, int precision,
               cnd_ctx **out)
{
    if (!out) return fail(nullptr, CND_EINVAL, "null output pointer");
    *out = nullptr;
    if (!params || ninst < 1 || !blocks || nblocks < 1) return fail(nullptr, CND_EINVAL, "empty case");
    if (precision != CND_FP64 && precision != CND_MIXED) return fail(nullptr, CND_EINVAL, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, CND_EDEVICE, "no HIP device");
    cnd_ctx *c = new cnd_ctx();
    c->ninst = ninst; c->nblk = nblocks; c->precision = precision;
    c->prm.assign(params, params + ninst);
    for (int i = 0; i < ninst; ++i) {
        const cnd_params &p = params[i];
        if (!(p.fltrbc > 0.0) || p.fltrbc == 1.0 || !(p.tempoo > 0.0) || !(p.reoo > 0.0) || !(p.cfl > 0.0)) {
            delete c;
            return fail(nullptr, CND_EINVAL, "instance " + std::to_string(i) + ": bad parameters");
        }
        if (p.sponge && (!(p.szth0 > 0.0) || !(p.szth1 > 0.0))) {
            delete c;
            return fail(nullptr, CND_EINVAL, "sponge needs szth0 > 0 and szth1 > 0");
        }
        if (p.nsmf == 0 && !(p.timf > 0.0)) { delete c; return fail(nullptr, CND_EINVAL, "nsmf = 0 needs timf > 0"); }
    }
    for (int b = 0; b < nblocks; ++b) {
        BlockSpec s;
        s.n[0] = blocks[b].lxi + 1; s.n[1] = blocks[b].let + 1; s.n[2] = blocks[b].lze + 1;
        std::memcpy(s.bc, blocks[b].nbc, sizeof s.bc);
        std::memcpy(s.partner, blocks[b].mcd, sizeof s.partner);
        s.inst = blocks[b].inst;
        if (s.inst < 0 || s.inst >= ninst) { delete c; return fail(nullptr, CND_EINVAL, "block instance out of range"); }
        c->spec.push_back(s);
    }
    std::string terr;
    if (!check_topology(c->spec, terr)) { delete c; return fail(nullptr, CND_EINVAL, terr); }
    c->htab.resize(ninst);
    for (int i = 0; i < ninst; ++i) build_tables(params[i].fltk, params[i].fltrbc, c->htab[i]);

    // layout
    std::vector<double> lu, segtab[kNumL], lhs;
    c->hblk.resize(nblocks);
    int64_t off = 0, foff[3] = {0, 0, 0};
    for (int b = 0; b < nblocks; ++b) {
        const BlockSpec &s = c->spec[b];
        BlkDev &B = c->hblk[b];
        std::memset(&B, 0, sizeof B);
        for (int d = 0; d < 3; ++d) {
            B.n[d] = s.n[d];
            for (int e = 0; e < 2; ++e) {
                B.bc[d][e] = s.bc[d][e];
                B.part[d][e] = s.partner[d][e];
                const int cd = s.bc[d][e];
                B.qd[d][e] = (cd == kInterStraight || cd == kInterSub || cd == kPeriodic) ? 1 : 0;
                if (B.qd[d][e] && B.part[d][e] != b) c->part_other[d] = true;
                if (cd == kNonReflective || cd == kWallInviscid || cd == kWallViscous || cd == kInterCurv) {
                    c->any_gcbc = true; B.gcbc_any = 1;
                }
                if (cd == kWallInviscid || cd == kWallViscous) c->any_wall = true;
                if (cd == kInterCurv || cd == kInterStraight || cd == kPeriodic) c->any_avg = true;
            }
        }
        B.inst = s.inst;
        B.off = off;
        c->off.push_back(off);
        const int64_t npt = int64_t(s.n[0]) * s.n[1] * s.n[2];
        c->npts += npt;
        off += (npt + kAlignPts - 1) / kAlignPts * kAlignPts;   // next block's fields start aligned
        c->cells += (long long)(s.n[0] - 1) * (s.n[1] - 1) * (s.n[2] - 1);
        for (int d = 0; d < 3; ++d) {
            B.foff[d] = foff[d];
            const int np = npencils(B, d);
            for (int q0 = 0; q0 < np; q0 += kLine) c->htiles[d].push_back(Tile{b, q0});
            foff[d] += np;
            for (int pi = 0; pi < kNumP; ++pi)
                for (int q0 = 0; q0 < np; q0 += kPs[pi]) c->hstiles[pi][d].push_back(Tile{b, q0});
            for (int sch = 0; sch < 2; ++sch) {
                B.lu[sch][d] = int(lu.size() / 5);
                const size_t at = lu.size();
                lu.resize(at + 5 * size_t(s.n[d]));
                lu_factors(c->htab[s.inst], sch, s.n[d], B.qd[d][0] != 0, B.qd[d][1] != 0, lu.data() + at);
                lhs.resize(lu.size());
                lhs_rows(c->htab[s.inst], sch, s.n[d], B.qd[d][0] != 0, B.qd[d][1] != 0, lhs.data() + at);
                const int nd = s.n[d];
                for (int li = 0; li < kNumL; ++li) {
                    const int SL = kLs[li], S = seg_count(nd, SL);
                    segtab[li].resize(4 * (at / 5 + size_t(nd)), 0.0);
                    if (S > 0)
                        seg_tables(lu.data() + at, S,
                                   [&](int k, int &a0, int &l0) { seg_bounds(nd, SL, S, k, a0, l0); },
                                   segtab[li].data() + 4 * (at / 5));
                }
            }
        }
        for (int64_t p0 = 0; p0 < npt; p0 += kPts) c->hpt.push_back(Tile{b, int(p0)});
        if (npt > int64_t(2000000000)) { delete c; return fail(nullptr, CND_EINVAL, "block too large"); }
    }
    c->NT = off;
    for (int d = 0; d < 3; ++d) c->FT[d] = foff[d];
    for (int b = 0; b < nblocks; ++b)
        for (int d = 0; d < 3; ++d) c->nmax[d] = std::max(c->nmax[d], c->spec[b].n[d]);
    // sweep shapes: the default segmented shape of each direction where every
    // block's pencils fit it (else another compiled shape, else the tiled
    // sweep); CANARD_SWEEP=tiled keeps the tiled sweeps everywhere and
    // CANARD_TUNE=0 keeps these defaults (no autotuning)
    {
        const char *sw = std::getenv("CANARD_SWEEP");
        c->tiled_only = sw && std::string(sw) == "tiled";
        const char *tu = std::getenv("CANARD_TUNE");
        c->tune_on = !c->tiled_only && !(tu && std::string(tu) == "0");
        const char *fu = std::getenv("CANARD_FUSE_UPDATE");
        c->fuse_update = !(fu && std::string(fu) == "0");
        // Side-by-side sweeps pay in a context whose sweeps fill the GPU; in
        // many small contexts stepped together (each its own stream pair on
        // a few hardware queues) the cross-stream waits cost far more than
        // they save.  CANARD_OVERLAP=0 / 1 forces either way.
        const char *ov = std::getenv("CANARD_OVERLAP");
        c->overlap = ov ? std::string(ov) != "0" : c->NT >= kOverlapMinPts;
        {
            int dev = 0, lds = 0;
            if (hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
                c->lds_max = lds;
        }
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < CND_NKIND; ++k) {
                const int nv = kind_nv(c, k, d), dflt = d == 0 ? 0 : 2;
                int pick = kTiled;
                if (!c->tiled_only) {
                    if (shape_sp(c, dflt, d, nv) > 0) pick = dflt;
                    for (int sh = 0; pick == kTiled && sh < kNumShapes; ++sh)
                        if (shape_sp(c, sh, d, nv) > 0) pick = sh;
                }
                c->shape[k][d] = pick;
                c->seg_sp[k][d] = pick == kTiled ? 0 : shape_sp(c, pick, d, nv);
                if (precision == CND_MIXED && pick == kTiled) {
                    delete c;
                    return fail(nullptr, CND_EINVAL,
                                "mixed precision needs the segmented sweeps in every direction (pencils of 13..528 "
                                "points, and CANARD_SWEEP not set to tiled)");
                }
            }
        // the update goes with the xi flux sweep, whose tile is one contiguous
        // run of points (its operands stream like a pointwise kernel's), unless
        // only the zeta one is segmented or CANARD_FUSE_DIR=2
        const char *fd = std::getenv("CANARD_FUSE_DIR");
        c->fuse_dir = (fd && std::string(fd) == "2") || !seg_on(c, CND_KIND_FLUX, 0) ? 2 : 0;
    }
    // forward values of up to 9 variables per pencil (metric filtering)
    for (int d = 0; d < 3; ++d)
        c->ys_len = std::max(c->ys_len, int64_t(9) * int64_t(c->htiles[d].size()) * kTile * c->nmax[d]);
    for (const cnd_params &p : c->prm) { c->any_sponge |= p.sponge != 0; c->any_visc |= p.viscous != 0; }
    for (const cnd_params &p : c->prm)
        if ((p.viscous != 0) != c->any_visc) { delete c; return fail(nullptr, CND_EINVAL, "mixed viscous/Euler instances"); }

    int r = CND_OK;
    auto ok = [&](int rc) { if (rc) r = rc; return rc == CND_OK; };
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, CND_EDEVICE, "stream creation failed");
    }
    // the side stream only where sweeps run side by side: every stream takes
    // a share of the few hardware queues that many contexts' streams share
    if (c->overlap && (hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) != hipSuccess ||
                       hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
                       hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)) {
        cnd_destroy(c);
        return fail(nullptr, CND_EDEVICE, "side stream creation failed");
    }
    const int64_t NT = c->NT;
    bool good = ok(dalloc(c, &c->dblk, nblocks)) && ok(dalloc(c, &c->dtab, ninst)) && ok(dalloc(c, &c->dinst, ninst)) &&
                ok(dalloc(c, &c->dlu, lu.size())) && ok(dalloc(c, &c->qa, 5 * NT)) && ok(dalloc(c, &c->qo, 5 * NT)) &&
                ok(dalloc(c, &c->de, 15 * NT)) && ok(dalloc(c, &c->grad, 12 * NT)) && ok(dalloc(c, &c->fl, 15 * NT)) &&
                ok(dalloc(c, &c->met, 10 * NT)) && ok(dalloc(c, &c->ys, c->ys_len)) && ok(dalloc(c, &c->xyz, 3 * NT)) &&
                ok(dalloc(c, &c->sp_par, 5 * ninst));
    for (int li = 0; good && li < kNumL; ++li) good = ok(dalloc(c, &c->dseg[li], segtab[li].size()));
    if (good && c->any_sponge) good = ok(dalloc(c, &c->sponge, 2 * NT));
    if (good && precision == CND_MIXED)
        good = ok(dalloc(c, &c->gradf, 12 * NT)) &&
               ok(dalloc(c, &c->dlu32, lu.size())) && ok(dalloc(c, &c->dlhs, lhs.size()));
    for (int li = 0; good && precision == CND_MIXED && li < kNumL; ++li)
        good = ok(dalloc(c, &c->dseg32[li], segtab[li].size()));
    for (int d = 0; good && d < 3; ++d)
        good = ok(dalloc(c, &c->dtiles[d], c->htiles[d].size())) && ok(dalloc(c, &c->drva[d], 10 * c->FT[d])) &&
               ok(dalloc(c, &c->cmf[d], 6 * c->FT[d])) && ok(dalloc(c, &c->surf[d], 10 * c->FT[d]));
    for (int d = 0; good && d < 3; ++d)
        for (int pi = 0; good && pi < kNumP; ++pi) good = ok(dalloc(c, &c->dstiles[pi][d], c->hstiles[pi][d].size()));
    if (good)
        good = ok(dalloc(c, &c->dpt, c->hpt.size()));
    if (!good) { std::string m = c->err; cnd_destroy(c); return fail(nullptr, r ? r : CND_ENOMEM, m); }
    std::vector<TabDev> ht(ninst);
    for (int i = 0; i < ninst; ++i) {
        const Tables &T = c->htab[i];
        TabDev &D = ht[i];
        D.fa = T.fa; D.fb = T.fb; D.fc = T.fc;
        std::memcpy(D.fbc, T.fbc, sizeof D.fbc);
        std::memcpy(D.qin, T.qin, sizeof D.qin);
        std::memcpy(D.qout, T.qout, sizeof D.qout);
        std::memcpy(D.qsum, T.qsum, sizeof D.qsum);
        std::memcpy(D.gcbc_w, T.gcbc_w, sizeof D.gcbc_w);
        std::memcpy(D.ramp, T.ramp, sizeof D.ramp);
    }
    std::vector<double> sp(5 * ninst);
    for (int i = 0; i < ninst; ++i) {
        sp[5 * i] = params[i].szco; sp[5 * i + 1] = params[i].szth0; sp[5 * i + 2] = params[i].szth1;
        sp[5 * i + 3] = params[i].doml0; sp[5 * i + 4] = params[i].doml1;
    }
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (r) return;
        hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) r = fail(c, CND_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
    };
    up(c->dblk, c->hblk.data(), sizeof(BlkDev) * nblocks);
    up(c->dtab, ht.data(), sizeof(TabDev) * ninst);
    up(c->dlu, lu.data(), sizeof(double) * lu.size());
    up(c->sp_par, sp.data(), sizeof(double) * sp.size());
    for (int d = 0; d < 3; ++d) up(c->dtiles[d], c->htiles[d].data(), sizeof(Tile) * c->htiles[d].size());
    for (int pi = 0; pi < kNumP; ++pi)
        for (int d = 0; d < 3; ++d)
            up(c->dstiles[pi][d], c->hstiles[pi][d].data(), sizeof(Tile) * c->hstiles[pi][d].size());
    for (int li = 0; li < kNumL; ++li) up(c->dseg[li], segtab[li].data(), sizeof(double) * segtab[li].size());
    if (precision == CND_MIXED) {
        const std::vector<float> lu32(lu.begin(), lu.end());
        up(c->dlu32, lu32.data(), sizeof(float) * lu32.size());
        for (int li = 0; li < kNumL; ++li) {
            const std::vector<float> seg32(segtab[li].begin(), segtab[li].end());
            up(c->dseg32[li], seg32.data(), sizeof(float) * seg32.size());
        }
        up(c->dlhs, lhs.data(), sizeof(double) * lhs.size());
    }
    up(c->dpt, c->hpt.data(), sizeof(Tile) * c->hpt.size());
    if (!r) r = upload_inst(c);
    if (!r) {
        hipError_t e = hipMemset(c->de, 0, sizeof(double) * 15 * NT);
        if (e == hipSuccess) e = hipMemset(c->fl, 0, sizeof(double) * 15 * NT);
        if (e != hipSuccess) r = fail(c, CND_EDEVICE, "memset");
    }
    // junction averaging (search_point / search_line, average_point / average_line)
    if (!r) {
        JunctionPlan pc, pe;
        junction_plans(c->spec, c->off, pc, pe);
        const JunctionProgram p = junction_program(pc, pe);
        JuncDev &j = c->jn;
        j.n = int(p.target.size());
        if (j.n > 0 &&
            !((r = dalloc(c, &j.tgt, p.target.size())) || (r = dalloc(c, &j.src, p.src.size())) ||
              (r = dalloc(c, &j.first, p.first.size())) || (r = dalloc(c, &j.tc, p.tc.size())) ||
              (r = dalloc(c, &j.sc, p.sc.size())) || (r = dalloc(c, &j.write, p.write.size())) ||
              (r = dalloc(c, &j.scale, p.scale.size())) || (r = dalloc(c, &j.tmp, 5 * size_t(j.n))))) {
            up(j.tgt, p.target.data(), sizeof(int64_t) * p.target.size());
            up(j.src, p.src.data(), sizeof(int64_t) * p.src.size());
            up(j.first, p.first.data(), sizeof(int32_t) * p.first.size());
            up(j.tc, p.tc.data(), sizeof(int32_t) * p.tc.size());
            up(j.sc, p.sc.data(), sizeof(int32_t) * p.sc.size());
            up(j.write, p.write.data(), sizeof(int32_t) * p.write.size());
            up(j.scale, p.scale.data(), sizeof(double) * p.scale.size());
        }
    }
    if (r) { std::string m = c->err; cnd_destroy(c); return fail(nullptr, r, m); }
    c->grid_set.assign(nblocks, 0);
    *out = c;
    return CND_OK;
}

void cnd_destroy(cnd_ctx *c)
{
    if (!c) return;
    if (c->st) hipStreamSynchronize(c->st);
    if (c->st2) hipStreamSynchronize(c->st2);
    output_release(c);
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    for (auto &m : c->marks) { hipEventDestroy(m.a); hipEventDestroy(m.b); }
    for (hipEvent_t e : c->pool) hipEventDestroy(e);
    void *ptrs[] = {c->dblk, c->dtab, c->dinst, c->dlu, c->qa, c->qo, c->de, c->grad, c->fl, c->met, c->ys,
                    c->xyz, c->sponge, c->sp_par, c->dpt, c->gradf, c->dlu32, c->dlhs, c->jn.tgt, c->jn.src,
                    c->jn.first, c->jn.tc, c->jn.sc, c->jn.write, c->jn.scale, c->jn.tmp, c->qb};
    for (void *p : ptrs) if (p) hipFree(p);
    for (int li = 0; li < kNumL; ++li) {
        if (c->dseg[li]) hipFree(c->dseg[li]);
        if (c->dseg32[li]) hipFree(c->dseg32[li]);
    }
    for (int d = 0; d < 3; ++d) {
        if (c->dtiles[d]) hipFree(c->dtiles[d]);
        for (int pi = 0; pi < kNumP; ++pi)
            if (c->dstiles[pi][d]) hipFree(c->dstiles[pi][d]);
        if (c->drva[d]) hipFree(c->drva[d]);
        if (c->cmf[d]) hipFree(c->cmf[d]);
        if (c->surf[d]) hipFree(c->surf[d]);
    }
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    if (c->st2) hipStreamDestroy(c->st2);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
}

int cnd_set_grid(cnd_ctx *c, int b, const double *xyz)
{
    if (!c || !xyz || b < 0 || b >= c->nblk) return fail(c, CND_EINVAL, "cnd_set_grid: bad argument");
    const int64_t n = int64_t(c->spec[b].n[0]) * c->spec[b].n[1] * c->spec[b].n[2];
    for (int k = 0; k < 3; ++k)
        HIPX(c, hipMemcpyAsync(c->xyz + k * c->NT + c->off[b], xyz + k * n, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPX(c, hipStreamSynchronize(c->st));
    c->grid_set[b] = 1;
    c->metrics_ready = false;
    return CND_OK;
}

int cnd_init_metrics(cnd_ctx *c)
{
    if (!c) return fail(c, CND_EINVAL, "null context");
    for (int b = 0; b < c->nblk; ++b)
        if (!c->grid_set[b]) return fail(c, CND_ESTATE, "cnd_init_metrics: grid of block " + std::to_string(b) + " not set");
    int r;
    // coordinate derivatives with the periodic jump (mpigo n45go + deriv):
    // dx[(3c+d)] = d(coord c)/d(xi_d) in the scratch fl
    for (int d = 0; d < 3; ++d)
        if ((r = run_fd_field(c, d, 1, c->xyz, c->NT, c->fl + d * c->NT, 3 * c->NT, 3))) return r;
    hipLaunchKernelGGL(k_cross, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dpt, c->fl, c->met, c->NT);
    // filter the nine metric arrays in eta, zeta, xi (mo_grid.f90:117-156)
    if ((r = run_filter_field(c, 1, c->met, c->grad, 9))) return r;
    if ((r = run_filter_field(c, 2, c->grad, c->met, 9))) return r;
    if ((r = run_filter_field(c, 0, c->met, c->grad, 9))) return r;
    hipLaunchKernelGGL(k_copy, dim3(unsigned((9 * c->NT + 255) / 256)), dim3(256), 0, c->st, c->grad, c->met, 9 * c->NT);
    hipLaunchKernelGGL(k_yaco, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dpt, c->fl, c->met, c->NT);
    for (int d = 0; d < 3; ++d)
        hipLaunchKernelGGL(k_normals, tiles_grid(c, d, 2), dim3(kLine), 0, c->st, c->dblk, c->dtiles[d], d, c->met,
                           c->cmf[d], c->NT, c->FT[d]);
    if (c->any_sponge)
        hipLaunchKernelGGL(k_sponge_init, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dpt, c->xyz, c->met,
                           c->sp_par, c->sponge, c->NT);
    if ((r = check_launch(c))) return r;
    HIPX(c, hipStreamSynchronize(c->st));
    c->metrics_ready = true;
    if (c->gexec) { hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    return CND_OK;
}

int cnd_get_metrics(cnd_ctx *c, int b, double *xim, double *etm, double *zem, double *yaco)
{
    if (!c || b < 0 || b >= c->nblk) return fail(c, CND_EINVAL, "cnd_get_metrics: bad argument");
    if (!c->metrics_ready) return fail(c, CND_ESTATE, "metrics not initialised");
    const int64_t n = int64_t(c->spec[b].n[0]) * c->spec[b].n[1] * c->spec[b].n[2];
    double *dst[4] = {xim, etm, zem, yaco};
    for (int w = 0; w < 4; ++w) {
        if (!dst[w]) continue;
        const int ncomp = w < 3 ? 3 : 1;
        for (int k = 0; k < ncomp; ++k)
            HIPX(c, hipMemcpyAsync(dst[w] + k * n, c->met + (3 * w + k) * c->NT + c->off[b], sizeof(double) * n,
                                   hipMemcpyDeviceToHost, c->st));
    }
    HIPX(c, hipStreamSynchronize(c->st));
    return CND_OK;
}

int cnd_set_state(cnd_ctx *c, int b, const double *q)
{
    if (!c || !q || b < 0 || b >= c->nblk) return fail(c, CND_EINVAL, "cnd_set_state: bad argument");
    const int64_t n = int64_t(c->spec[b].n[0]) * c->spec[b].n[1] * c->spec[b].n[2];
    double *dst = c->state_in_qo ? c->qo : c->qa;
    for (int m = 0; m < 5; ++m)
        HIPX(c, hipMemcpyAsync(dst + m * c->NT + c->off[b], q + m * n, sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    HIPX(c, hipStreamSynchronize(c->st));
    return CND_OK;
}

int cnd_get_state(cnd_ctx *c, int b, double *q)
{
    if (!c || !q || b < 0 || b >= c->nblk) return fail(c, CND_EINVAL, "cnd_get_state: bad argument");
    const int64_t n = int64_t(c->spec[b].n[0]) * c->spec[b].n[1] * c->spec[b].n[2];
    const double *srcp = c->state_in_qo ? c->qo : c->qa;
    for (int m = 0; m < 5; ++m)
        HIPX(c, hipMemcpyAsync(q + m * n, srcp + m * c->NT + c->off[b], sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPX(c, hipStreamSynchronize(c->st));
    return CND_OK;
}

int cnd_set_clock(cnd_ctx *c, int inst, int n, int ndt, double dt, double dts, double dte, double timo)
{
    if (!c || inst < 0 || inst >= c->ninst) return fail(c, CND_EINVAL, "cnd_set_clock: bad instance");
    HIPX(c, hipStreamSynchronize(c->st));
    InstDev I;
    HIPX(c, hipMemcpy(&I, c->dinst + inst, sizeof I, hipMemcpyDeviceToHost));
    I.n = n; I.ndt = ndt; I.dt = dt; I.dts = dts; I.dte = dte; I.timo = timo;
    HIPX(c, hipMemcpy(c->dinst + inst, &I, sizeof I, hipMemcpyHostToDevice));
    return CND_OK;
}

int cnd_get_clock(cnd_ctx *c, int inst, int *n, int *ndt, double *dt, double *dts, double *dte, double *timo)
{
    if (!c || inst < 0 || inst >= c->ninst) return fail(c, CND_EINVAL, "cnd_get_clock: bad instance");
    HIPX(c, hipStreamSynchronize(c->st));
    InstDev I;
    HIPX(c, hipMemcpy(&I, c->dinst + inst, sizeof I, hipMemcpyDeviceToHost));
    if (n) *n = I.n;
    if (ndt) *ndt = I.ndt;
    if (dt) *dt = I.dt;
    if (dts) *dts = I.dts;
    if (dte) *dte = I.dte;
    if (timo) *timo = I.timo;
    return CND_OK;
}

static int ready(cnd_ctx *c)
{
    if (!c) return fail(c, CND_EINVAL, "null context");
    if (!c->metrics_ready) return fail(c, CND_ESTATE, "grid metrics not initialised (cnd_init_metrics)");
    return CND_OK;
}

// Autotuner: for each sweep kind of the step and each direction, times every
// compiled segmented shape that fits the blocks (and, in FP64, the tiled
// sweep) on the context's own arrays and keeps the fastest.  All candidates
// compute the same values up to rounding.  Runs once, before the first step.
static int tune(cnd_ctx *c)
{
    if (c->tuned) return CND_OK;
    c->tuned = true;
    if (!c->tune_on) return CND_OK;
    const bool mix = c->precision == CND_MIXED;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPX(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return fail(c, CND_EDEVICE, "tune: event"); }
    int r = CND_OK;
    // the flux sweep that carries the update is timed as the step runs it:
    // fused with the update (into the gradient scratch) where fusion applies
    const bool fz = fuse_possible(c);
    auto run = [&](int kind, int d) {
        if (kind == CND_KIND_FILTER) return run_filter_field(c, d, c->qa, c->fl, 5, mix);
        if (kind == CND_KIND_GRAD) return run_grad(c, d, c->qa);
        if (d == c->fuse_dir && fz) return run_flux_update(c, 1, c->grad);
        return run_flux(c, d, c->qa);
    };
    for (int kind : {int(CND_KIND_FILTER), int(CND_KIND_GRAD), int(CND_KIND_FLUX)}) {
        if (kind == CND_KIND_GRAD && !c->any_visc) continue;
        for (int d = 0; d < 3 && !r; ++d) {
            std::vector<int> cands;
            if (!mix && !(kind == CND_KIND_FLUX && d == c->fuse_dir && fz)) cands.push_back(kTiled);   // no fused tiled form
            for (int sh = 0; sh < kNumShapes; ++sh)
                if (shape_sp(c, sh, d, kind_nv(c, kind, d)) > 0) cands.push_back(sh);
            if (cands.size() < 2) continue;
            int best = c->shape[kind][d];
            double best_ms = 1e300;
            for (int cand : cands) {
                c->shape[kind][d] = cand;
                c->seg_sp[kind][d] = cand == kTiled ? 0 : shape_sp(c, cand, d, kind_nv(c, kind, d));
                constexpr int reps = 3;
                if ((r = run(kind, d))) break;
                (void)hipEventRecord(e0, c->st);
                for (int i = 0; i < reps && !r; ++i) r = run(kind, d);
                (void)hipEventRecord(e1, c->st);
                if (r) break;
                float ms = 0.f;
                if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    r = fail(c, CND_EDEVICE, "tune: timing");
                    break;
                }
                if (ms / reps < best_ms) { best_ms = ms / reps; best = cand; }
            }
            c->shape[kind][d] = best;
            c->seg_sp[kind][d] = best == kTiled ? 0 : shape_sp(c, best, d, kind_nv(c, kind, d));
            c->tune_ms[kind][d] = best_ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return r;
}

int cnd_tune(cnd_ctx *c)
{
    int r;
    if ((r = ready(c))) return r;
    if (c->next_call != 0) return fail(c, CND_ESTATE, "cnd_tune: a step is half done");
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    c->tuned = false;
    return tune(c);
}

int cnd_shape_info(int shape, int *seg_len, int *pencils, int *run, int *xi)
{
    if (shape < 0 || shape >= kNumShapes) return CND_EINVAL;
    const ShapeInfo &G = kShapes[shape];
    if (seg_len) *seg_len = G.L;
    if (pencils) *pencils = G.P;
    if (run) *run = G.C;
    if (xi) *xi = G.XI ? 1 : 0;
    return CND_OK;
}

int cnd_set_shape(cnd_ctx *c, int kind, int dir, int shape)
{
    if (!c || kind < 0 || kind >= CND_NKIND || dir < 0 || dir > 2 || shape < -1 || shape >= kNumShapes)
        return fail(c, CND_EINVAL, "cnd_set_shape: bad argument");
    if (c->next_call != 0) return fail(c, CND_ESTATE, "cnd_set_shape: a step is half done");
    if (shape == kTiled && c->precision == CND_MIXED)
        return fail(c, CND_EINVAL, "cnd_set_shape: mixed precision runs on the segmented sweeps only");
    const int sp = shape == kTiled ? 0 : shape_sp(c, shape, dir, kind_nv(c, kind, dir));
    if (shape != kTiled && sp == 0)
        return fail(c, CND_EINVAL, "cnd_set_shape: the shape does not fit the pencils of this direction");
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    c->shape[kind][dir] = shape;
    c->seg_sp[kind][dir] = sp;
    c->tuned = true;   // an explicit choice is kept
    c->tune_ms[kind][dir] = 0.0;
    return CND_OK;
}

int cnd_tuning(const cnd_ctx *c, int kind, int dir, int *shape, int *seg_len, int *pencils, int *segments,
               double *ms)
{
    if (!c || kind < 0 || kind >= CND_NKIND || dir < 0 || dir > 2) return CND_EINVAL;
    const int sh = c->shape[kind][dir];
    if (shape) *shape = sh;
    if (seg_len) *seg_len = sh == kTiled ? 0 : kShapes[sh].L;
    if (pencils) *pencils = sh == kTiled ? kTile : kShapes[sh].P;
    if (segments) *segments = sh == kTiled ? 1 : c->seg_sp[kind][dir];
    if (ms) *ms = c->tune_ms[kind][dir];
    return CND_OK;
}

// What the stage-by-stage API expects next, for refusal messages.
static std::string expected_call(const cnd_ctx *c)
{
    if (c->next_call == 0) return "cnd_filter or cnd_step";
    if (c->next_call == 5) return "cnd_end_step";
    return "cnd_rk_stage(" + std::to_string(c->next_call) + ")";
}

int cnd_filter(cnd_ctx *c)
{
    int r;
    if ((r = ready(c))) return r;
    if (c->next_call == 1) return fail(c, CND_ESTATE, "cnd_filter: the filter phase already ran for this step");
    if (c->next_call != 0)
        return fail(c, CND_ESTATE, "cnd_filter: a step is half done (expected " + expected_call(c) + ")");
    if ((r = tune(c))) return r;
    if ((r = enqueue_filter(c))) return r;
    c->state_in_qo = true;
    c->next_call = 1;
    return CND_OK;
}

int cnd_rk_stage(cnd_ctx *c, int nk)
{
    int r;
    if ((r = ready(c))) return r;
    if (nk < 1 || nk > 4) return fail(c, CND_EINVAL, "cnd_rk_stage: nk must be 1..4");
    if (nk == 1 && c->next_call == 0) return fail(c, CND_ESTATE, "cnd_rk_stage: stage 1 must follow cnd_filter");
    if (nk != c->next_call)
        return fail(c, CND_ESTATE, "cnd_rk_stage: stage " + std::to_string(nk) + " out of order (expected " +
                                       expected_call(c) + ")");
    if ((r = tune(c))) return r;
    if ((r = enqueue_stage(c, nk))) return r;
    c->state_in_qo = false;
    c->next_call = nk + 1;
    return CND_OK;
}

int cnd_end_step(cnd_ctx *c)
{
    int r;
    if ((r = ready(c))) return r;
    if (c->next_call != 5)
        return fail(c, CND_ESTATE, "cnd_end_step: the four stages of the step have not run (expected " +
                                       expected_call(c) + ")");
    if ((r = enqueue_end(c))) return r;
    c->next_call = 0;
    return CND_OK;
}

int cnd_record_reset(cnd_ctx *c)
{
    if (!c) return fail(c, CND_EINVAL, "null context");
    if (!c->qb) return CND_OK;
    HIPX(c, hipMemsetAsync(c->qb, 0, sizeof(double) * 5 * c->NT, c->st));
    hipLaunchKernelGGL(k_record_clock, dim3(unsigned((c->ninst + 63) / 64)), dim3(64), 0, c->st, c->dinst, c->ninst,
                       1);
    return check_launch(c);
}

int cnd_record(cnd_ctx *c, int on)
{
    if (!c) return fail(c, CND_EINVAL, "null context");
    if (c->next_call != 0) return fail(c, CND_ESTATE, "cnd_record: a step is half done");
    if (on && !c->qb) {
        int r;
        if ((r = dalloc(c, &c->qb, 5 * c->NT))) return r;
        if ((r = cnd_record_reset(c))) return r;
    }
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }   // the step's launches change
    c->recording = on != 0;
    return CND_OK;
}

int cnd_use_graph(cnd_ctx *c, int enable)
{
    if (!c) return fail(c, CND_EINVAL, "null context");
    c->use_graph = enable != 0;
    return CND_OK;
}

int cnd_step(cnd_ctx *c, int nsteps)
{
    int r;
    if ((r = ready(c))) return r;
    if (c->next_call != 0)
        return fail(c, CND_ESTATE, "cnd_step: a step is half done (expected " + expected_call(c) + ")");
    if (nsteps < 0) return fail(c, CND_EINVAL, "cnd_step: negative step count");
    if ((r = tune(c))) return r;
    if (!c->use_graph) {
        for (int s = 0; s < nsteps; ++s)
            if ((r = enqueue_step(c))) return r;
        return CND_OK;
    }
    if (!c->gexec) {
        hipGraph_t g;
        HIPX(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        r = enqueue_step(c);
        c->capturing = false;
        hipError_t e = hipStreamEndCapture(c->st, &g);
        if (r) { if (e == hipSuccess) hipGraphDestroy(g); return r; }
        if (e != hipSuccess) return fail(c, CND_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) { c->gexec = nullptr; return fail(c, CND_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e)); }
    }
    for (int s = 0; s < nsteps; ++s) HIPX(c, hipGraphLaunch(c->gexec, c->st));
    return CND_OK;
}

int cnd_synchronize(cnd_ctx *c)
{
    if (!c) return fail(c, CND_EINVAL, "null context");
    HIPX(c, hipStreamSynchronize(c->st));
    return CND_OK;
}

static int sweep_io(cnd_ctx *c, const double *const *in, double *dev)
{
    for (int b = 0; b < c->nblk; ++b) {
        if (!in[b]) return fail(c, CND_EINVAL, "null field pointer");
        const int64_t n = int64_t(c->spec[b].n[0]) * c->spec[b].n[1] * c->spec[b].n[2];
        HIPX(c, hipMemcpyAsync(dev + c->off[b], in[b], sizeof(double) * n, hipMemcpyHostToDevice, c->st));
    }
    return CND_OK;
}

static int sweep_out(cnd_ctx *c, const double *dev, double *const *out)
{
    for (int b = 0; b < c->nblk; ++b) {
        if (!out[b]) return fail(c, CND_EINVAL, "null field pointer");
        const int64_t n = int64_t(c->spec[b].n[0]) * c->spec[b].n[1] * c->spec[b].n[2];
        HIPX(c, hipMemcpyAsync(out[b], dev + c->off[b], sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    }
    HIPX(c, hipStreamSynchronize(c->st));
    return CND_OK;
}

int cnd_deriv_sweep(cnd_ctx *c, int dir, const double *const *in, double *const *out)
{
    if (!c || !in || !out || dir < 0 || dir > 2) return fail(c, CND_EINVAL, "cnd_deriv_sweep: bad argument");
    int r;
    double *a = c->fl, *o = c->fl + c->NT;   // scratch fields
    if ((r = sweep_io(c, in, a))) return r;
    if ((r = run_fd_field(c, dir, 0, a, 0, o, 0, 1, c->precision == CND_MIXED))) return r;
    return sweep_out(c, o, out);
}

int cnd_filter_sweep(cnd_ctx *c, int dir, double *const *f)
{
    if (!c || !f || dir < 0 || dir > 2) return fail(c, CND_EINVAL, "cnd_filter_sweep: bad argument");
    int r;
    double *a = c->fl, *o = c->fl + c->NT;   // scratch fields
    if ((r = sweep_io(c, (const double *const *)f, a))) return r;
    if ((r = run_filter_field(c, dir, a, o, 1, c->precision == CND_MIXED))) return r;
    return sweep_out(c, o, f);
}

static int collect_marks(cnd_ctx *c)
{
    HIPX(c, hipStreamSynchronize(c->st));
    for (auto &m : c->marks) {
        float ms = 0.f;
        HIPX(c, hipEventElapsedTime(&ms, m.a, m.b));
        c->sec_ms[m.sec] += ms;
        c->sec_n[m.sec] += m.launches;
        c->pool.push_back(m.a);
        c->pool.push_back(m.b);
    }
    c->marks.clear();
    return CND_OK;
}

int cnd_profile(cnd_ctx *c, int enable)
{
    if (!c) return fail(c, CND_EINVAL, "null context");
    int r;
    if ((r = collect_marks(c))) return r;
    if (enable) {
        for (int s = 0; s < CND_NSEC; ++s) { c->sec_ms[s] = 0.0; c->sec_n[s] = 0; }
    }
    c->prof = enable != 0;
    return CND_OK;
}

int cnd_profile_read(cnd_ctx *c, int section, double *ms, long long *launches)
{
    if (!c || section < 0 || section >= CND_NSEC) return fail(c, CND_EINVAL, "cnd_profile_read: bad section");
    int r;
    if ((r = collect_marks(c))) return r;
    if (ms) *ms = c->sec_ms[section];
    if (launches) *launches = c->sec_n[section];
    return CND_OK;
}

void *cnd_stream(cnd_ctx *c) { return c ? (void *)c->st : nullptr; }
long long cnd_points(const cnd_ctx *c) { return c ? (long long)c->npts : 0; }
long long cnd_cells(const cnd_ctx *c) { return c ? c->cells : 0; }

}  // extern "C"
