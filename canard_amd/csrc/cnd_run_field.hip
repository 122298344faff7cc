// Filter and plain derivative sweeps of stored fields (filter phase,
// metrics, single sweeps).
#include "cnd_ctx.hpp"

namespace cnd {

int run_filter_field(cnd_ctx *c, int d, const double *in, double *out, int nvar, bool mix)
{
    const int K = CND_KIND_FILTER;
    if (mix)
        LAUNCH_SEG(k_seg_filter_field, true, 1, c, K, d, nvar, in, out, c->NT);
    else if (seg_on(c, K, d))
        LAUNCH_SEG(k_seg_filter_field, false, 1, c, K, d, nvar, in, out, c->NT);
    else
        LAUNCH_SWEEP(k_filter_field, 1, c, d, nvar, kTile, c->dblk, c->dtab, c->dlu, c->dtiles[d], ntiles(c, d), d,
                     c->nmax[d], c->ys, in, out, c->NT);
    return check_launch(c);
}

int run_fd_field(cnd_ctx *c, int d, int n45, const double *in, int64_t ivs, double *out, int64_t ovs, int nvar,
                 bool mix)
{
    const int K = CND_KIND_FIELD;
    if (mix)
        LAUNCH_SEG(k_seg_fd_field, true, 1, c, K, d, nvar, n45, in, ivs, out, ovs);
    else if (seg_on(c, K, d))
        LAUNCH_SEG(k_seg_fd_field, false, 1, c, K, d, nvar, n45, in, ivs, out, ovs);
    else
        LAUNCH_SWEEP(k_fd_field, 1, c, d, nvar, kTile, c->dblk, c->dtab, c->dlu, c->dtiles[d], ntiles(c, d), d, n45,
                     c->nmax[d], c->ys, in, ivs, out, ovs);
    return check_launch(c);
}

}  // namespace cnd
