// Grid generation on the device: the reference's generators gridtoy and
// gridtoy_multi (src/mo_gridgen.f90:510-654) and the sinusoidally sheared
// Cartesian blocks of the multi-block benchmark case, written straight into
// the context's coordinate arrays, so large and ensemble cases never build
// their grids on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "cnd_ctx.hpp"

namespace {

// One thread per point of block b: x, y, z = X(i, j, k) of the generator.
//  kind CND_GRID_TOY / CND_GRID_TOY_MULTI: x = -doml0 + (i + is - ra2) ra0,
//      y = -domh + (j + js - ra3) ra1, z = span ((lze0 - k) / lze0 - 1/2)
//  kind CND_GRID_SHEARED: X = x0 + i h, Y = j h, Z = k h,
//      x = X + amp lx sin(2 pi Y / ly)
struct GridGen {
    int kind, lze0, is, js;
    double ra0, ra1, ra2, ra3, doml0, domh, span;   // toy generators
    double x0, h, amp, lx, ly;                      // sheared
};

__global__ __launch_bounds__(kPts) void k_make_grid(const BlkDev *__restrict__ blks, int b, GridGen G,
                                                   double *__restrict__ xyz, int64_t NT)
{
    const BlkDev &B = blks[b];
    const int64_t npt = int64_t(B.n[0]) * B.n[1] * B.n[2];
    const int64_t l = int64_t(blockIdx.x) * kPts + threadIdx.x;
    if (l >= npt) return;
    const int i = int(l % B.n[0]), j = int((l / B.n[0]) % B.n[1]), k = int(l / (int64_t(B.n[0]) * B.n[1]));
    double x, y, z;
    if (G.kind == CND_GRID_SHEARED) {
        const double X = G.x0 + double(i) * G.h, Y = double(j) * G.h;
        z = double(k) * G.h;
        y = Y;
        x = X + (G.amp * G.lx) * sin(((2.0 * 3.141592653589793) * Y) / G.ly);
    } else {
        x = -G.doml0 + (double(i + G.is) - G.ra2) * G.ra0;
        y = -G.domh + (double(j + G.js) - G.ra3) * G.ra1;
        z = G.span * (double(G.lze0 - k) / double(G.lze0) - 0.5);
    }
    const int64_t g = B.off + l;
    xyz[g] = x;
    xyz[NT + g] = y;
    xyz[2 * NT + g] = z;
}

}  // namespace

extern "C" {

int cnd_make_grid(cnd_ctx *c, int b, int kind, const double *par, int npar)
{
    if (!c || b < 0 || b >= c->nblk || (npar > 0 && !par)) return fail(c, CND_EINVAL, "cnd_make_grid: bad argument");
    GridGen G{};
    G.kind = kind;
    auto need = [&](int n) { return npar == n; };
    if (kind == CND_GRID_TOY) {
        // par: lxi0, let0, lze0, doml0, doml1, domh, span (mo_gridgen.f90:510-567)
        if (!need(7) || !(par[0] >= 1) || !(par[1] >= 1) || !(par[2] >= 1))
            return fail(c, CND_EINVAL, "cnd_make_grid: gridtoy takes lxi0, let0, lze0, doml0, doml1, domh, span");
        G.ra0 = (par[4] + par[3]) / par[0];
        G.ra1 = 2.0 * par[5] / par[1];
        G.lze0 = int(par[2]);
        G.doml0 = par[3]; G.domh = par[5]; G.span = par[6];
    } else if (kind == CND_GRID_TOY_MULTI) {
        // par: lxi0, lxi1, lxi2, let0, lze0, doml0, doml1, domh, span, mb (mo_gridgen.f90:569-654)
        if (!need(10) || !(par[9] >= 0 && par[9] <= 5) || !(par[3] >= 1) || !(par[4] >= 1))
            return fail(c, CND_EINVAL,
                        "cnd_make_grid: gridtoy_multi takes lxi0, lxi1, lxi2, let0, lze0, doml0, doml1, domh, span, mb");
        const int lxi0 = int(par[0]), lxi1 = int(par[1]), lxi2 = int(par[2]), let0 = int(par[3]), mb = int(par[9]);
        G.ra0 = (par[6] + par[5]) / double(lxi0 + lxi1 + lxi2);
        G.ra1 = par[7] / (2.0 * double(let0)) * 2.0;
        const int col = mb % 3;
        G.is = col == 0 ? 0 : (col == 1 ? lxi0 + 1 : lxi0 + lxi1 + 2);
        G.ra2 = double(col);
        G.js = mb < 3 ? 0 : let0 + 1;
        G.ra3 = mb < 3 ? 0.0 : 1.0;
        G.lze0 = int(par[4]);
        G.doml0 = par[5]; G.domh = par[7]; G.span = par[8];
    } else if (kind == CND_GRID_SHEARED) {
        // par: x0, h, amp, lx, ly
        if (!need(5) || !(par[1] > 0.0) || !(par[4] > 0.0))
            return fail(c, CND_EINVAL, "cnd_make_grid: the sheared grid takes x0, h, amp, lx, ly");
        G.x0 = par[0]; G.h = par[1]; G.amp = par[2]; G.lx = par[3]; G.ly = par[4];
    } else {
        return fail(c, CND_EINVAL, "cnd_make_grid: unknown generator");
    }
    const int64_t n = int64_t(c->spec[b].n[0]) * c->spec[b].n[1] * c->spec[b].n[2];
    hipLaunchKernelGGL(k_make_grid, dim3(unsigned((n + kPts - 1) / kPts)), dim3(kPts), 0, c->st, c->dblk, b, G,
                       c->xyz, c->NT);
    int r;
    if ((r = check_launch(c))) return r;
    HIPX(c, hipStreamSynchronize(c->st));
    c->grid_set[b] = 1;
    c->metrics_ready = false;
    return CND_OK;
}

int cnd_get_grid(cnd_ctx *c, int b, double *xyz)
{
    if (!c || !xyz || b < 0 || b >= c->nblk) return fail(c, CND_EINVAL, "cnd_get_grid: bad argument");
    if (!c->grid_set[b]) return fail(c, CND_ESTATE, "cnd_get_grid: grid of block " + std::to_string(b) + " not set");
    const int64_t n = int64_t(c->spec[b].n[0]) * c->spec[b].n[1] * c->spec[b].n[2];
    for (int k = 0; k < 3; ++k)
        HIPX(c, hipMemcpyAsync(xyz + k * n, c->xyz + k * c->NT + c->off[b], sizeof(double) * n, hipMemcpyDeviceToHost,
                               c->st));
    HIPX(c, hipStreamSynchronize(c->st));
    return CND_OK;
}

}  // extern "C"
