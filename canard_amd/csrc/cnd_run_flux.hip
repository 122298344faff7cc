// Flux derivative sweeps into the right-hand-side buffers.
#include "cnd_ctx.hpp"

namespace cnd {

int run_flux(cnd_ctx *c, int d, const double *cur)
{
    const int K = CND_KIND_FLUX;
    const bool mix = c->precision == CND_MIXED;
    double *de = c->de + 5 * d * c->NT;
    if (c->any_visc) {
        const double *F = c->fl + 5 * d * c->NT;
        if (mix)
            LAUNCH_SEG(k_seg_flux_field, true, 1, c, K, d, 5, F, de, c->drva[d], c->NT, c->FT[d]);
        else if (seg_on(c, K, d))
            LAUNCH_SEG(k_seg_flux_field, false, 1, c, K, d, 5, F, de, c->drva[d], c->NT, c->FT[d]);
        else
            LAUNCH_SWEEP(k_fd_flux_field, 1, c, d, 5, kTile, c->dblk, c->dtab, c->dlu, c->dtiles[d], ntiles(c, d), d,
                         c->nmax[d], c->ys, F, de, c->drva[d], c->NT, c->FT[d]);
    } else {
        if (mix)
            LAUNCH_SEG(k_seg_flux, true, 5, c, K, d, 1, c->dinst, cur, c->met, de, c->drva[d], c->NT, c->FT[d]);
        else if (seg_on(c, K, d))
            LAUNCH_SEG(k_seg_flux, false, 5, c, K, d, 1, c->dinst, cur, c->met, de, c->drva[d], c->NT, c->FT[d]);
        else
            LAUNCH_SWEEP(k_fd_flux, 5, c, d, 1, 5 * kTile, c->dblk, c->dtab, c->dinst, c->dlu, c->dtiles[d],
                         ntiles(c, d), d, c->nmax[d], c->ys, cur, c->met, de, c->drva[d], c->NT, c->FT[d]);
    }
    return check_launch(c);
}

}  // namespace cnd
