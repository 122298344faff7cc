// Constants of the one-sided derivative rows shared by the sweep kernels.
#pragma once

#include "cnd_dev.hpp"

namespace cnd {

// Optimised one-sided derivative rows (src/mo_parameters.f90:66-79); the
// published 16-digit values are kept exactly.
constexpr double kA10 = -0.3506944444280639, kA12 = 0.8249999999974599, kA13 = 0.7444444444773477,
                 kA14 = 0.014583333334139971;
constexpr double kA01 = -6.666666667794732, kA02 = 9.000000000052879, kA03 = 1.333333333178555,
                 kA04 = -0.0833333333455545;

}  // namespace cnd
