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
// The banded LHS itself, 5 entries per row at offsets -2..2 (diagonal 1).
void lhs_rows(const Tables &t, int scheme, int n, bool qd_lo, bool qd_hi, double *rows5);

// Homogeneous solutions of the LU recurrences of one line (rows5 as from
// lu_factors) restarted at every segment of an S-segment partition: per row
// (phi1, phi2, psi1, psi2).  phi solves the forward recurrence of the
// segment with incoming (y[a-2], y[a-1]) = (1, 0) / (0, 1); psi solves the
// backward recurrence with incoming (x[b], x[b+1]) = (1, 0) / (0, 1).
// bounds(s, a, len) gives segment s.
template <class Bounds>
void seg_tables(const double *rows5, int S, Bounds bounds, double *out4)
{
    for (int s = 0; s < S; ++s) {
        int a, len;
        bounds(s, a, len);
        double p1a = 1.0, p1b = 0.0, p2a = 0.0, p2b = 1.0;   // (j-2, j-1)
        for (int j = a; j < a + len; ++j) {
            const double *l = rows5 + 5 * j;
            const double y1 = (0.0 - l[0] * p1a) - l[1] * p1b;
            const double y2 = (0.0 - l[0] * p2a) - l[1] * p2b;
            out4[4 * j] = y1;
            out4[4 * j + 1] = y2;
            p1a = p1b; p1b = y1;
            p2a = p2b; p2b = y2;
        }
        double q1a = 1.0, q1b = 0.0, q2a = 0.0, q2b = 1.0;   // (j+1, j+2)
        for (int j = a + len - 1; j >= a; --j) {
            const double *l = rows5 + 5 * j;
            const double x1 = (l[2] * 0.0 - l[3] * q1a) - l[4] * q1b;
            const double x2 = (l[2] * 0.0 - l[3] * q2a) - l[4] * q2b;
            out4[4 * j + 2] = x1;
            out4[4 * j + 3] = x2;
            q1b = q1a; q1a = x1;
            q2b = q2a; q2a = x2;
        }
    }
}

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

// Corner and edge averaging as one gather and one scatter.  Entries
// [0, ncorner) are the corner plan's, the rest the edge plan's.  An edge
// entry reads a corner target through that corner's entry (tc for its
// target, sc per source; -1: the point itself), i.e. the value corner
// averaging leaves there.  Of several edge entries for one point only the
// last (the one the reference writes last) is kept; a corner entry whose
// point an edge entry also writes has write = 0.
struct JunctionProgram {
    int ncorner = 0;
    std::vector<int64_t> target, src;
    std::vector<int32_t> first, tc, sc, write;
    std::vector<double> scale;
};
JunctionProgram junction_program(const JunctionPlan &corners, const JunctionPlan &edges);

}  // namespace cnd
