// Device-side data structures and small helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cnd {

// Gas constants (src/mo_parameters.f90:52-58), folded at compile time with
// round-to-nearest exactly as the reference's parameter expressions.
constexpr double kGam = 1.4;
constexpr double kGm1 = kGam - 1.0;
constexpr double kHamm1 = 1.0 / kGm1;
constexpr double kHamHamm1 = (1.0 / kGam) * kHamm1;
constexpr double kGm1PrI = 1.0 / (kGm1 * 0.71);
constexpr double kTwoThirds = 2.0 / 3.0;
constexpr double kSml = 1.0e-8;
// Interior derivative stencil (src/mo_parameters.f90:63-64)
constexpr double kAa = 20.0 / 27.0, kAb = 25.0 / 216.0;

// One block as the kernels see it.
struct BlkDev {
    int n[3];           // points per direction
    int bc[3][2];       // boundary code per face
    int part[3][2];     // partner block per face (-1: none)
    int qd[3][2];       // 1: quasi-disjoint closure rows at this face
    int inst;           // ensemble instance
    int gcbc_any;       // any characteristic face in this block
    int64_t off;        // first global point index
    int64_t foff[3];    // first pencil of this block in per-direction face buffers
    int lu[2][3];       // first LU row (scheme, direction) in the LU table
};

// Per-instance scheme constants.
struct TabDev {
    double fa, fb, fc;
    double fbc[3][5];
    double qin[2][2][12], qout[2][2][12], qsum[2][2];
    double gcbc_w[4];
    double ramp[64];
};

// Per-instance physics parameters and time-loop clock (all on device so a
// whole step can be replayed as a graph without host round trips).
struct InstDev {
    double reoo, srefoo, srefp1dre, uoo[3], timf, cfl;
    int nsmf, viscous, sponge, pad;
    int n, ndt;
    double dt, dts, dte, timo;
    double umf[3], dudtmf[3];
    unsigned long long red[2];   // running maxima (bit patterns of >= 0 doubles)
    double dtsum;                // recorded time since the last reset (cnd_record)
    // output schedule (nml_canard ndata, tmax): output ndati is due on the
    // step whose interval holds ndati * tmax / ndata; nout = 1 on that step
    double tmax;
    int ndata, ndati, nout, pad2;
};

struct Tile { int b, q0; };

__host__ __device__ inline int npencils(const BlkDev &B, int d)
{
    return d == 0 ? B.n[1] * B.n[2] : (d == 1 ? B.n[0] * B.n[2] : B.n[0] * B.n[1]);
}

// Pencil q of direction d: global index of its first point and the stride
// between consecutive points.  q depends only on the transverse indices, so
// the same q addresses the matching pencil of a conforming partner block.
__device__ inline void pencil_of(const BlkDev &B, int d, int q, int64_t &base, int64_t &stride)
{
    const int nx = B.n[0], ny = B.n[1];
    if (d == 0) { base = B.off + int64_t(q) * nx; stride = 1; }
    else if (d == 1) { const int xi = q % nx, ze = q / nx; base = B.off + xi + int64_t(ze) * nx * ny; stride = nx; }
    else { base = B.off + q; stride = int64_t(nx) * ny; }
}

}  // namespace cnd
