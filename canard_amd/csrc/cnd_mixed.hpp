// Mixed-precision sweeps: FP32 line solves with one FP64 residual
// refinement pass.  Implemented in cnd_mixed.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "cnd_dev.hpp"

namespace cnd {

int64_t mixed_scratch_floats(int64_t NT);
// filter phase sweep of all five state variables along d (in -> out)
int mixed_filter(hipStream_t st, const BlkDev *blks, const TabDev *tabs, const double *lu, Tile *const *dtiles,
                 const std::vector<Tile> *htiles, const double *in, double *out, float *scr, int64_t NT, int d);
int mixed_filter_one(hipStream_t st, const BlkDev *blks, const TabDev *tabs, const double *lu, const Tile *tiles,
                     unsigned ntiles, int d, const double *in, double *out, float *scr, int64_t NT);
int mixed_fd_field(hipStream_t st, const BlkDev *blks, const TabDev *tabs, const double *lu, const Tile *tiles,
                   unsigned ntiles, int d, const double *in, double *out, float *scr, int64_t NT);
int mixed_grad(hipStream_t st, const BlkDev *blks, const TabDev *tabs, const double *lu, const Tile *tiles,
               unsigned ntiles, int d, const double *q, double *grad, float *scr, int64_t NT);
int mixed_flux(hipStream_t st, const BlkDev *blks, const TabDev *tabs, const InstDev *inst, const double *lu,
               const Tile *tiles, unsigned ntiles, int d, const double *q, const double *tau, const double *met,
               double *de, double *ybuf, double *drva, float *scr, int64_t NT, int64_t FT, int vis);

}  // namespace cnd
