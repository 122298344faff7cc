// Mixed-precision sweeps (FP32 line solves + one FP64 refinement pass).
#include "cnd_mixed.hpp"

#include "../../include/cnd.h"

namespace cnd {

int64_t mixed_scratch_floats(int64_t NT) { return NT; }
int mixed_filter(hipStream_t, const BlkDev *, const TabDev *, const double *, Tile *const *,
                 const std::vector<Tile> *, const double *, double *, float *, int64_t, int) { return CND_EINVAL; }
int mixed_filter_one(hipStream_t, const BlkDev *, const TabDev *, const double *, const Tile *, unsigned, int,
                     const double *, double *, float *, int64_t) { return CND_EINVAL; }
int mixed_fd_field(hipStream_t, const BlkDev *, const TabDev *, const double *, const Tile *, unsigned, int,
                   const double *, double *, float *, int64_t) { return CND_EINVAL; }
int mixed_grad(hipStream_t, const BlkDev *, const TabDev *, const double *, const Tile *, unsigned, int,
               const double *, double *, float *, int64_t) { return CND_EINVAL; }
int mixed_flux(hipStream_t, const BlkDev *, const TabDev *, const InstDev *, const double *, const Tile *, unsigned,
               int, const double *, const double *, const double *, double *, double *, double *, float *, int64_t,
               int64_t, int) { return CND_EINVAL; }

}  // namespace cnd
