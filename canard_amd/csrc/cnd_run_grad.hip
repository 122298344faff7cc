// Velocity and temperature gradient sweeps of the viscous stages.
#include "cnd_ctx.hpp"

namespace cnd {

int run_grad(cnd_ctx *c, int d, const double *cur)
{
    const int K = CND_KIND_GRAD;
    if (c->precision == CND_MIXED)
        LAUNCH_SEG(k_seg_grad, true, 4, c, K, d, 1, cur, c->gradf, c->NT);
    else if (seg_on(c, K, d))
        LAUNCH_SEG(k_seg_grad, false, 4, c, K, d, 1, cur, c->grad, c->NT);
    else
        LAUNCH_SWEEP(k_fd_grad, 4, c, d, 1, 4 * kTile, c->dblk, c->dtab, c->dlu, c->dtiles[d], ntiles(c, d), d,
                     c->nmax[d], c->ys, cur, c->grad, c->NT);
    return check_launch(c);
}

}  // namespace cnd
