// The last flux sweep of a viscous stage (direction fuse_dir) fused with the
// RK update.
#include "cnd_ctx.hpp"

namespace cnd {

int run_flux_update(cnd_ctx *c, int nk, double *qa)
{
    const int K = CND_KIND_FLUX, d = c->fuse_dir;
    const double *F = c->fl + 5 * d * c->NT;
    if (c->precision == CND_MIXED)
        LAUNCH_SEG_XM(k_seg_flux_update, true, 1, c, K, d, 5, 1, c->dinst, F, c->de, c->qo, c->met, qa,
                      c->drva[d], c->NT, c->FT[d], nk);
    else
        LAUNCH_SEG_XM(k_seg_flux_update, false, 1, c, K, d, 5, 1, c->dinst, F, c->de, c->qo, c->met, qa,
                      c->drva[d], c->NT, c->FT[d], nk);
    return check_launch(c);
}

}  // namespace cnd
