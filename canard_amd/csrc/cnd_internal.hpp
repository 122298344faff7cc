// Views of a context that the other translation units of libcnd use.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cnd.h"

namespace cnd {

// Block b of a context: its state vector (5 fields NT apart, the block from
// off), sizes and the context's stream.
struct CtxView {
    double *q;
    int64_t NT, off;
    int n[3];
    hipStream_t st;
};
int ctx_block_view(cnd_ctx *c, int b, CtxView &v);

}  // namespace cnd
