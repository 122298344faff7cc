// Host-side set-up of the compact schemes and of the block topology.
//
// Coefficient tables are computed once per context on the host in FP64 and
// uploaded; every value follows the reference's recipe (cited per function
// in cnd_scheme.cpp) so that device sweeps reproduce its arithmetic.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cnd {

// Boundary codes, src/mo_parameters.f90:43-49.
enum Bc { kNonReflective = 10, kWallInviscid = 20, kWallViscous = 25, kInterCurv = 30,
          kInterStraight = 35, kInterSub = 40, kPeriodic = 45 };

// Number of one-sided points in a quasi-disjoint interface stencil
// (lmd = lmf = 11 in src/mo_numerics.f90:21, i.e. 12 points).
constexpr int kQd = 12;
// Points spread by the characteristic boundary correction (mbci = 4 inward
// plus the face point, src/mo_gcbc.f90:22).
constexpr int kGcbcDepth = 4;

struct Tables {
    // compact filter (nml_numerics fltk, fltrbc)
    double alf = 0, bef = 0, fa = 0, fb = 0, fc = 0;   // interior filter
    double flhs[3][5] = {};   // one-sided filter LHS rows 0..2 (offsets -2..2)
    double fbc[3][5] = {};    // one-sided filter RHS weights rows 0..2
    // quasi-disjoint interface closures, [scheme 0=FD 1=filter][row][i]
    double qin[2][2][kQd] = {};   // weights on the block's own points
    double qout[2][2][kQd] = {};  // weights on the partner's points
    double qsum[2][2] = {};       // sum of qout (periodic coordinate jump)
    // characteristic boundary spreading weights for 4 inward points
    double gcbc_w[kGcbcDepth] = {};
    // dt ramp sin^2(0.05*pi*k), k = 0..63 (mo_canard_driver.f90:285)
    double ramp[64] = {};
};

void build_tables(double fltk_over_pi, double fltrbc, Tables &t);

// LU factors of one pentadiagonal line of n points. scheme 0 = derivative,
// 1 = filter; qd_lo/qd_hi select the quasi-disjoint rows at each end.
// Output rows: (xl1, xl2, xu1, xu2, xu3) per point.
void lu_factors(const Tables &t, int scheme, int n, bool qd_lo, bool qd_hi, double *rows5);

// Dense inverse by Gauss-Jordan elimination with row pivoting (n x n,
// column-major).
void gauss_jordan_inverse(const double *a, double *inv, int n);

struct BlockSpec {
    int n[3];          // points per direction
    int bc[3][2];
    int partner[3][2];
    int inst;
};

// Checks that interfaces are reciprocal and conforming; fills err.
bool check_topology(const std::vector<BlockSpec> &blocks, std::string &err);

// Junction averaging plan (corners first, then edges): for each target
// point, the points whose values are summed onto it, in summation order.
struct JunctionPlan {
    std::vector<int64_t> target;      // global point index
    std::vector<int32_t> first;       // into src, size target+1
    std::vector<int64_t> src;         // global point index
    std::vector<double> scale;        // 1/(count+1)
};
// offsets[b] = global index of block b's first point.
void junction_plans(const std::vector<BlockSpec> &blocks, const std::vector<int64_t> &offsets,
                    JunctionPlan &corners, JunctionPlan &edges);

}  // namespace cnd
