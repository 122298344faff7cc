// Velocity and temperature gradient sweeps of the viscous stages.
#include "cnd_ctx.hpp"

namespace cnd {

namespace {

// u, v, w, T of the state q at every stored point into prim (four fields NT
// apart), the arithmetic of SrcPrim: the primitives of the eta and zeta
// sweeps when the xi sweep is the tiled one, which stores none.
__global__ __launch_bounds__(kPts) void k_prims(const BlkDev *__restrict__ blks, const Tile *__restrict__ tiles,
                                               const double *__restrict__ q, double *__restrict__ prim, int64_t NT)
{
    const Tile t = tiles[blockIdx.x];
    const BlkDev &B = blks[t.b];
    const int64_t npt = int64_t(B.n[0]) * B.n[1] * B.n[2];
    const int64_t lp = int64_t(t.q0) + threadIdx.x;
    if (lp >= npt) return;
    const int64_t g = B.off + lp;
    SrcPrim src{q, NT};
    SrcPrim::Raw raw;
    src.load(g, t.b, raw);
    double f[4];
    src.finish(raw, t.b, f);
#pragma unroll
    for (int m = 0; m < 4; ++m) prim[m * NT + g] = f[m];
}

}  // namespace

int run_grad(cnd_ctx *c, int d, const double *cur)
{
    const int K = CND_KIND_GRAD;
    const bool mix = c->precision == CND_MIXED;   // FP32 solves and gradients (segmented only)
    double *prim = c->fl;   // the fluxes of the previous stage are consumed by now
    if (d == 0) {
        if (mix) {
            LAUNCH_SEG(k_seg_grad, true, 4, c, K, 0, 1, cur, c->gradf, c->NT, prim);
        } else if (seg_on(c, K, 0)) {
            LAUNCH_SEG(k_seg_grad, false, 4, c, K, 0, 1, cur, c->grad, c->NT, prim);
        } else {
            LAUNCH_SWEEP(k_fd_grad, 4, c, 0, 1, 4 * kTile, c->dblk, c->dtab, c->dlu, c->dtiles[0], ntiles(c, 0), 0,
                         c->nmax[0], c->ys, cur, c->grad, c->NT);
            if (seg_on(c, K, 1) || seg_on(c, K, 2))
                hipLaunchKernelGGL(k_prims, pts_grid(c), dim3(kPts), 0, c->st, c->dblk, c->dpt, cur, prim, c->NT);
        }
    } else if (mix) {
        LAUNCH_SEG(k_seg_grad_prim, true, 1, c, K, d, 4, prim, c->gradf, c->NT);
    } else if (seg_on(c, K, d)) {
        LAUNCH_SEG(k_seg_grad_prim, false, 1, c, K, d, 4, prim, c->grad, c->NT);
    } else {
        LAUNCH_SWEEP(k_fd_grad, 4, c, d, 1, 4 * kTile, c->dblk, c->dtab, c->dlu, c->dtiles[d], ntiles(c, d), d,
                     c->nmax[d], c->ys, cur, c->grad, c->NT);
    }
    return check_launch(c);
}

}  // namespace cnd
