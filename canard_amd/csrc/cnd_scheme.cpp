// Host-side coefficient tables, LU factors and topology plans.
// Reference recipes are cited per function (paths relative to the reference
// repository root).  All arithmetic is FP64 with separately rounded products
// (the library is built with -ffp-contract=off) so that the tables equal the
// reference's bit for bit where its evaluation order is known.
#include "cnd_scheme.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

namespace cnd {

namespace {
constexpr double kPi = 3.141592653589793;
// Lele's pentadiagonal 8th-order scheme and the optimised one-sided rows,
// src/mo_parameters.f90:61-79 (values must be identical for parity).
constexpr double kAl = 4.0 / 9.0, kBe = 1.0 / 36.0;
constexpr double kAl10 = 0.09166666665902128, kAl12 = 1.3500000000438646, kBe13 = 0.2666666666794667;
constexpr double kA10 = -0.3506944444280639, kA12 = 0.8249999999974599, kA13 = 0.7444444444773477,
                 kA14 = 0.014583333334139971;
constexpr double kAl01 = 8.000000000449464, kBe02 = 6.000000000542;
constexpr double kA01 = -6.666666667794732, kA02 = 9.000000000052879, kA03 = 1.333333333178555,
                 kA04 = -0.0833333333455545;

// Interior filter of cut-off kc and parameter r (src/mo_numerics.f90:306-326).
struct Filter5 { double alpha, beta, a, b, c; };
Filter5 interior_filter(double kc, double r)
{
    const double c1 = std::cos(kc), c2 = std::cos(2 * kc), c3 = std::cos(3 * kc);
    const double den = ((30.0 + (5.0 * (7.0 - 16.0 * r)) * c1) + (2.0 * (1.0 + 8.0 * r)) * c2) - 3.0 * c3;
    const double inv = 1.0 / den;
    Filter5 f;
    f.alpha = inv * (((20.0 * (2.0 * r - 1.0) - 30.0 * c1) + (12.0 * (2.0 * r - 1.0)) * c2) - 2.0 * c3);
    f.beta = (0.5 * inv) * (((2.0 * (13.0 - 8.0 * r) + (33.0 - 48.0 * r) * c1) + 6.0 * c2) - c3);
    const double h = std::cos(0.5 * kc), h2 = h * h;
    f.a = ((60.0 * (1.0 - r)) * (h2 * h2)) * inv;
    f.b = (-2.0 * f.a) / 5.0;
    f.c = f.a / 15.0;
    return f;
}

struct Mat {  // small dense column-major matrix, 1-based accessors
    int n;
    std::vector<double> v;
    explicit Mat(int n_) : n(n_), v(size_t(n_) * n_, 0.0) {}
    double &operator()(int i, int j) { return v[size_t(i - 1) + size_t(n) * (j - 1)]; }
    double operator()(int i, int j) const { return v[size_t(i - 1) + size_t(n) * (j - 1)]; }
};

Mat product(const Mat &a, const Mat &b)
{
    Mat c(a.n);
    for (int j = 1; j <= a.n; ++j)
        for (int i = 1; i <= a.n; ++i) {
            double s = 0.0;
            for (int k = 1; k <= a.n; ++k) s = s + a(i, k) * b(k, j);
            c(i, j) = s;
        }
    return c;
}
}  // namespace

// Gauss-Jordan with partial pivoting and in-place column bookkeeping, the
// elimination order of src/mo_utils.f90:15-50.
void gauss_jordan_inverse(const double *a, double *inv, int n)
{
    std::vector<double> w(a, a + size_t(n) * n), col(n);
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    auto W = [&](int r, int c) -> double & { return w[size_t(r) + size_t(n) * c]; };
    for (int p = 0; p < n; ++p) {
        int piv = p;
        for (int r = p + 1; r < n; ++r)
            if (std::fabs(W(r, p)) > std::fabs(W(piv, p))) piv = r;
        if (piv != p) {
            std::swap(perm[piv], perm[p]);
            for (int c = 0; c < n; ++c) std::swap(W(piv, c), W(p, c));
        }
        const double rp = 1.0 / W(p, p);
        for (int r = 0; r < n; ++r) col[r] = W(r, p);
        for (int c = 0; c < n; ++c) {
            const double f = rp * W(p, c);
            for (int r = 0; r < n; ++r) W(r, c) = W(r, c) - f * col[r];
            W(p, c) = f;
        }
        for (int r = 0; r < n; ++r) W(r, p) = -rp * col[r];
        W(p, p) = rp;
    }
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r) inv[size_t(r) + size_t(n) * perm[c]] = W(r, c);
}

void build_tables(double fltk_over_pi, double fltrbc, Tables &t)
{
    const double kc = kPi * fltk_over_pi;   // src/mo_numerics.f90:107
    // One-sided boundary filter family (src/mo_numerics.f90:253-302).
    {
        const Filter5 z = interior_filter(kc, 0.5);
        const double lg = std::log(fltrbc);
        const double nrm = 1.0 / ((1.0 + z.alpha * fltrbc) + z.beta * (fltrbc * fltrbc));
        const double r0[5] = {0, 0, 1, z.alpha * nrm, z.beta * nrm};
        const double r1[5] = {0, z.alpha + z.beta * fltrbc, 1, z.alpha, z.beta};
        const double r2[5] = {z.beta, z.alpha, 1, z.alpha, z.beta};
        std::memcpy(t.flhs[0], r0, sizeof r0);
        std::memcpy(t.flhs[1], r1, sizeof r1);
        std::memcpy(t.flhs[2], r2, sizeof r2);
        double g = ((fltrbc - 1) * ((z.a + z.c) + (fltrbc + 1) * (z.b + fltrbc * z.c))) / lg;
        t.fbc[0][0] = (z.a - (5 * g) / 3) * nrm;
        t.fbc[0][1] = (z.b + (10 * g) / 21) * nrm;
        t.fbc[0][2] = (z.c - (5 * g) / 42) * nrm;
        t.fbc[0][3] = ((5 * g) / 252) * nrm;
        t.fbc[0][4] = (-g / 630) * nrm;
        g = ((fltrbc - 1) * (z.b + z.c * (fltrbc + 1))) / lg;
        t.fbc[1][0] = ((z.a + z.b) + z.c) + (1627 * g) / 1260;
        t.fbc[1][1] = z.a + (10 * g) / 21;
        t.fbc[1][2] = z.b - (5 * g) / 42;
        t.fbc[1][3] = z.c + (5 * g) / 252;
        t.fbc[1][4] = -g / 630;
        g = (z.c * (fltrbc - 1)) / lg;
        t.fbc[2][0] = (z.b + z.c) + (1627 * g) / 1260;
        t.fbc[2][1] = z.a - (5 * g) / 3;
        t.fbc[2][2] = z.a - (5 * g) / 42;
        t.fbc[2][3] = z.b + (5 * g) / 252;
        t.fbc[2][4] = z.c - g / 630;
    }
    const Filter5 f = interior_filter(kc, 0.5);
    t.alf = f.alpha; t.bef = f.beta; t.fa = f.a; t.fb = f.b; t.fc = f.c;

    // Quasi-disjoint closures (src/mo_numerics.f90:330-425): two 12-point
    // half-lines meet at an interface; the global compact system is solved
    // exactly, and the truncated boundary rows applied to that solution give
    // the RHS weights on the own side (qin) and on the partner side (qout).
    for (int s = 0; s < 2; ++s) {
        const int n = 2 * kQd, lo = 1, hi = n;
        Mat A(n), B(n);
        if (s == 0) {
            A(lo, lo) = 1; A(lo, lo + 1) = kAl01; A(lo, lo + 2) = kBe02;
            const double b0[5] = {-(((kA01 + kA02) + kA03) + kA04), kA01, kA02, kA03, kA04};
            for (int c = 0; c < 5; ++c) B(lo, lo + c) = b0[c];
            A(lo + 1, lo) = kAl10; A(lo + 1, lo + 1) = 1; A(lo + 1, lo + 2) = kAl12; A(lo + 1, lo + 3) = kBe13;
            const double b1[5] = {kA10, -(((kA10 + kA12) + kA13) + kA14), kA12, kA13, kA14};
            for (int c = 0; c < 5; ++c) B(lo + 1, lo + c) = b1[c];
            for (int i = lo + 2; i <= hi - 2; ++i) {
                const double ar[5] = {kBe, kAl, 1, kAl, kBe}, br[5] = {-25.0 / 216.0, -20.0 / 27.0, 0,
                                                                       20.0 / 27.0, 25.0 / 216.0};
                for (int c = 0; c < 5; ++c) { A(i, i - 2 + c) = ar[c]; B(i, i - 2 + c) = br[c]; }
            }
            for (int c = 0; c < 4; ++c) A(hi - 1, hi - c) = A(lo + 1, lo + c);
            for (int c = 0; c < 5; ++c) B(hi - 1, hi - c) = -B(lo + 1, lo + c);
            for (int c = 0; c < 3; ++c) A(hi, hi - c) = A(lo, lo + c);
            for (int c = 0; c < 5; ++c) B(hi, hi - c) = -B(lo, lo + c);
        } else {
            const int cols[3][5] = {{1, 2, 3, 4, 5}, {0, 2, 3, 4, 5}, {0, 1, 3, 4, 5}};
            for (int r = 0; r < 3; ++r) {
                // rows 0..2 hold offsets -r..2 of the one-sided filter LHS
                for (int c = 0; c <= r + 2; ++c) A(lo + r, lo + c) = t.flhs[r][2 - r + c];
                double sum = 0;
                for (int c = 0; c < 5; ++c) { B(lo + r, lo + cols[r][c]) = t.fbc[r][c]; sum = sum + t.fbc[r][c]; }
                B(lo + r, lo + r) = -sum;
            }
            for (int i = lo + 3; i <= hi - 3; ++i) {
                const double ar[5] = {t.bef, t.alf, 1, t.alf, t.bef};
                const double br[7] = {t.fc, t.fb, t.fa, -2 * ((t.fa + t.fb) + t.fc), t.fa, t.fb, t.fc};
                for (int c = 0; c < 5; ++c) A(i, i - 2 + c) = ar[c];
                for (int c = 0; c < 7; ++c) B(i, i - 3 + c) = br[c];
            }
            for (int r = 0; r < 3; ++r) {
                for (int c = 0; c <= r + 2; ++c) A(hi - r, hi - c) = A(lo + r, lo + c);
                for (int c = 0; c < 6; ++c) B(hi - r, hi - c) = B(lo + r, lo + c);
            }
        }
        Mat Ainv(n);
        gauss_jordan_inverse(A.v.data(), Ainv.v.data(), n);
        Mat R = A;   // truncate the LHS couplings across the interface
        int m = hi / 2 - 1;
        R(m, m + 2) = 0; R(m + 1, m + 2) = 0; R(m + 1, m + 3) = 0;
        m = hi / 2 + 2;
        R(m, m - 2) = 0; R(m - 1, m - 2) = 0; R(m - 1, m - 3) = 0;
        const Mat W = product(R, product(Ainv, B));
        for (int row = 0; row < 2; ++row) {
            const int r = hi / 2 + 1 + row;
            double sum = 0;
            for (int i = 0; i < kQd; ++i) {
                t.qout[s][row][i] = W(r, kQd - i);
                t.qin[s][row][i] = W(r, kQd + 1 + i);
            }
            for (int i = 0; i < kQd; ++i) sum = sum + t.qout[s][row][i];
            t.qsum[s][row] = sum;
        }
    }

    // Characteristic-boundary spreading weights (src/mo_gcbc.f90:56-76).
    {
        constexpr int m = kGcbcDepth;
        double c[m * m] = {}, ci[m * m];
        auto C = [&](int i, int j) -> double & { return c[(i - 1) + m * (j - 1)]; };
        C(1, 1) = kAl01; C(1, 2) = kBe02;
        C(2, 1) = 1; C(2, 2) = kAl12; C(2, 3) = kBe13;
        C(3, 1) = kAl; C(3, 2) = 1; C(3, 3) = kAl; C(3, 4) = kBe;
        for (int i = 4; i <= m; ++i) {
            C(i, i - 3) = kBe; C(i, i - 2) = kAl; C(i, i - 1) = 1; C(i, i) = kAl;
            if (i < m) C(i, i + 1) = kBe;
        }
        const double rhs[m] = {1, kAl10, kBe, 0};
        gauss_jordan_inverse(c, ci, m);
        double acc = 0;
        for (int i = 0; i < m; ++i) {
            double s = 0;
            for (int k = 0; k < m; ++k) s = s + ci[i + m * k] * rhs[k];
            acc = acc + 1.0;
            t.gcbc_w[i] = (0.5 * -s) * (1.0 + std::cos(acc * (kPi / (m + 1))));
        }
    }
    for (int k = 0; k < 64; ++k) {
        const double s = std::sin((0.05 * kPi) * double(k));
        t.ramp[k] = s * s;
    }
}

// LU factorisation of the pentadiagonal LHS (src/mo_numerics.f90:179-249):
// rows (xl1, xl2) are the lower factors, (xu1, xu2, xu3) the reciprocal
// pivot and the two scaled upper entries.
// LHS rows 0..2 next to a line end (offsets -2..2 from the row), for the
// derivative (scheme 0: one-sided rows of mo_parameters.f90:66-79, or the
// truncated interior rows at an interface) or the filter (scheme 1).
static void end_rows(const Tables &t, int scheme, bool qd, double (*r)[5])
{
    if (scheme == 0) {
        const double one_sided[3][5] = {{0, 0, 1, kAl01, kBe02}, {0, kAl10, 1, kAl12, kBe13},
                                        {kBe, kAl, 1, kAl, kBe}};
        const double qd_rows[3][5] = {{0, 0, 1, kAl, kBe}, {0, kAl, 1, kAl, kBe}, {kBe, kAl, 1, kAl, kBe}};
        std::memcpy(r, qd ? qd_rows : one_sided, sizeof(double) * 15);
    } else {
        const double qd_rows[3][5] = {{0, 0, 1, t.alf, t.bef}, {0, t.alf, 1, t.alf, t.bef},
                                      {t.bef, t.alf, 1, t.alf, t.bef}};
        if (qd) std::memcpy(r, qd_rows, sizeof(double) * 15);
        else std::memcpy(r, t.flhs, sizeof(double) * 15);
    }
}

void lhs_rows(const Tables &t, int scheme, int n, bool qd_lo, bool qd_hi, double *o)
{
    double lo[3][5], hi[3][5];
    end_rows(t, scheme, qd_lo, lo);
    end_rows(t, scheme, qd_hi, hi);
    const double al = scheme == 0 ? kAl : t.alf, be = scheme == 0 ? kBe : t.bef;
    for (int i = 0; i < n; ++i) {
        double *r = o + 5 * i;
        r[0] = be; r[1] = al; r[2] = 1.0; r[3] = al; r[4] = be;
    }
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 5; ++c) {
            o[5 * i + c] = lo[i][c];
            o[5 * (n - 1 - i) + c] = hi[i][4 - c];   // the far end mirrors the near-end rows
        }
}

void lu_factors(const Tables &t, int scheme, int n, bool qd_lo, bool qd_hi, double *o)
{
    // LHS rows 0..2 for the end family (offsets -2..2)
    double lo[3][5], hi[3][5], al, be;
    end_rows(t, scheme, qd_lo, lo);
    end_rows(t, scheme, qd_hi, hi);
    if (scheme == 0) { al = kAl; be = kBe; } else { al = t.alf; be = t.bef; }
    auto L1 = [&](int i) -> double & { return o[5 * i + 0]; };
    auto L2 = [&](int i) -> double & { return o[5 * i + 1]; };
    auto U1 = [&](int i) -> double & { return o[5 * i + 2]; };
    auto U2 = [&](int i) -> double & { return o[5 * i + 3]; };
    auto U3 = [&](int i) -> double & { return o[5 * i + 4]; };
    for (int i = 0; i < n; ++i) for (int c = 0; c < 5; ++c) o[5 * i + c] = 1.0;
    // generic elimination step for row i with left entries (b2, b1) and
    // right entries (a1, a2)
    auto step = [&](int i, double b2, double b1, double a1, double a2, bool has_a2) {
        L1(i) = b2 * U1(i - 2);
        L2(i) = (b1 - U2(i - 2) * L1(i)) * U1(i - 1);
        U1(i) = 1.0 / ((1.0 - U3(i - 2) * L1(i)) - U2(i - 1) * L2(i));
        U2(i) = a1 - U3(i - 1) * L2(i);
        if (has_a2) U3(i) = a2;
    };
    U1(0) = 1.0; U2(0) = lo[0][3]; U3(0) = lo[0][4];
    L2(1) = lo[1][1] * U1(0);
    U1(1) = 1.0 / (1.0 - U2(0) * L2(1));
    U2(1) = lo[1][3] - U3(0) * L2(1);
    U3(1) = lo[1][4];
    step(2, lo[2][0], lo[2][1], lo[2][3], lo[2][4], true);
    for (int i = 3; i <= n - 4; ++i) step(i, be, al, al, be, true);
    // the far end mirrors the near-end rows
    step(n - 3, hi[2][4], hi[2][3], hi[2][1], hi[2][0], true);
    step(n - 2, hi[1][4], hi[1][3], hi[1][1], 0.0, false);
    {
        const int i = n - 1;
        L1(i) = hi[0][4] * U1(i - 2);
        L2(i) = (hi[0][3] - U2(i - 2) * L1(i)) * U1(i - 1);
        U1(i) = 1.0 / ((1.0 - U3(i - 2) * L1(i)) - U2(i - 1) * L2(i));
    }
    for (int i = 0; i < n; ++i) { U2(i) = U2(i) * U1(i); U3(i) = U3(i) * U1(i); }
}

bool check_topology(const std::vector<BlockSpec> &blocks, std::string &err)
{
    const int nb = int(blocks.size());
    for (int b = 0; b < nb; ++b) {
        const BlockSpec &B = blocks[b];
        for (int d = 0; d < 3; ++d) {
            if (B.n[d] < kQd + 1) { err = "block " + std::to_string(b) + ": every direction needs >= 13 points"; return false; }
            if (B.n[d] > 2112) { err = "block " + std::to_string(b) + ": at most 2112 points per direction"; return false; }
            for (int s = 0; s < 2; ++s) {
                const int c = B.bc[d][s];
                const bool iface = c == kInterCurv || c == kInterStraight || c == kInterSub || c == kPeriodic;
                if (!(iface || c == kNonReflective || c == kWallInviscid || c == kWallViscous)) {
                    err = "block " + std::to_string(b) + ": unknown boundary code " + std::to_string(c);
                    return false;
                }
                if (!iface) continue;
                const int p = B.partner[d][s];
                if (p < 0 || p >= nb) { err = "block " + std::to_string(b) + ": interface face without partner"; return false; }
                const BlockSpec &P = blocks[p];
                if (P.partner[d][1 - s] != b || P.bc[d][1 - s] != c) {
                    err = "block " + std::to_string(b) + ": partner face is not reciprocal";
                    return false;
                }
                if (P.inst != B.inst) { err = "interface between two instances"; return false; }
                for (int e = 0; e < 3; ++e)
                    if (e != d && P.n[e] != B.n[e]) { err = "block " + std::to_string(b) + ": non-conforming interface"; return false; }
            }
        }
    }
    return true;
}

// ---- junction plans (src/mo_domdcomp.f90:269-559) ------------------------
// The reference finds junction partners with a one-pass merge of per-face
// partner lists (search_point / search_line) and exchanges corner and edge
// values with isend/irecv under a shared tag; with one rank per block, a
// receive is served by the sender's matching send in posting order.  The
// plan below replays both steps so the device kernel only gathers.
namespace {
constexpr int kListCap = 101;   // ljpl+1 entries per list

struct Lists {
    std::vector<int> count;               // entries beyond the first
    std::vector<std::vector<int>> ids;    // kListCap entries each
    std::vector<int> tag;
};

void squeeze(std::vector<int> &v, int &count, int last)
{
    for (int a = 0; a < last; ++a)
        for (int b = a + 1; b <= last; ++b)
            if (v[b] == v[a]) v[b] = -1;
    count = last;
    for (int a = 0; a <= last; ++a)
        if (v[a] == -1) {
            --count;
            if (a + 1 < kListCap) { v[a] = v[a + 1]; v[a + 1] = -1; }
            else v[a] = -1;
        }
}

Lists build_lists(int nitems, const std::vector<BlockSpec> &blocks, int per_block,
                  const int table[4][3][2])
{
    Lists L;
    L.count.assign(nitems, 0);
    L.ids.assign(nitems, std::vector<int>(kListCap, -1));
    L.tag.assign(nitems, 0);
    for (int i = 0; i < nitems; ++i) L.ids[i][0] = i;
    for (int b = 0; b < int(blocks.size()); ++b)
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s) {
                const int p = blocks[b].partner[d][s];
                if (p < 0) continue;
                for (int c = 0; c < 4; ++c) {
                    const int mine = per_block * b + table[c][d][s];
                    const int theirs = per_block * p + table[c][d][1 - s];
                    ++L.count[mine];
                    L.ids[mine][L.count[mine]] = L.ids[theirs][0];
                }
            }
    for (int i = 0; i < nitems; ++i)
        for (int j = 0; j < nitems; ++j) {
            if (i == j || L.count[i] == 0 || L.count[j] == 0) continue;
            bool shared = false;
            for (int a = 0; a <= L.count[i] && !shared; ++a)
                for (int b = 0; b <= L.count[j]; ++b)
                    if (L.ids[j][b] == L.ids[i][a]) { shared = true; break; }
            if (!shared) continue;
            const int first = L.count[i] + 1, last = first + L.count[j];
            for (int c = 0; c <= last - first; ++c) L.ids[i][first + c] = L.ids[j][c];
            squeeze(L.ids[i], L.count[i], last);
        }
    for (int i = 0; i < nitems; ++i)
        L.tag[i] = *std::min_element(L.ids[i].begin(), L.ids[i].begin() + L.count[i] + 1);
    return L;
}

struct Post { int peer, tag, item; };
}  // namespace

void junction_plans(const std::vector<BlockSpec> &blocks, const std::vector<int64_t> &offsets,
                    JunctionPlan &corners, JunctionPlan &edges)
{
    const int nb = int(blocks.size());
    // corner c = ic + 2 jc + 4 kc; face (d, s) holds these four corners
    int ctab[4][3][2], etab[4][3][2];
    for (int s = 0; s < 2; ++s) {
        const int a[4] = {0, 2, 4, 6}, b[4] = {0, 4, 1, 5}, c[4] = {0, 1, 2, 3};
        const int ea[4] = {0, 1, 8, 10}, eb[4] = {4, 5, 0, 2}, ec[4] = {8, 9, 4, 6}, w[4] = {2, 2, 1, 1};
        for (int q = 0; q < 4; ++q) {
            ctab[q][0][s] = a[q] + s; ctab[q][1][s] = b[q] + 2 * s; ctab[q][2][s] = c[q] + 4 * s;
            etab[q][0][s] = ea[q] + s * w[q]; etab[q][1][s] = eb[q] + s * w[q]; etab[q][2][s] = ec[q] + s * w[q];
        }
    }
    auto point_of = [&](int b, int kind, int item, int k) -> int64_t {
        const BlockSpec &B = blocks[b];
        int ijk[3];
        if (kind == 0) {
            ijk[0] = (item % 2) * (B.n[0] - 1);
            ijk[1] = ((item % 4) / 2) * (B.n[1] - 1);
            ijk[2] = (item / 4) * (B.n[2] - 1);
        } else {
            // edge e: runs along direction dl; (ip, jp) select its position
            const int nn = item / 4, ip = (item % 4) / 2, jp = item % 2;
            const int da = nn, db = (nn + 1) % 3, dl = (nn + 2) % 3;
            ijk[da] = ip * (B.n[da] - 1);
            ijk[db] = jp * (B.n[db] - 1);
            ijk[dl] = k;
        }
        return offsets[b] + ijk[0] + int64_t(B.n[0]) * (ijk[1] + int64_t(B.n[1]) * ijk[2]);
    };
    auto edge_len = [&](int b, int item) { return blocks[b].n[(item / 4 + 2) % 3]; };

    for (int kind = 0; kind < 2; ++kind) {
        const int per = kind == 0 ? 8 : 12, need = kind == 0 ? 3 : 2;
        const Lists L = build_lists(per * nb, blocks, per, kind == 0 ? ctab : etab);
        JunctionPlan &plan = kind == 0 ? corners : edges;
        plan = JunctionPlan();
        plan.first.push_back(0);
        std::vector<std::vector<Post>> sends(nb), recvs(nb);
        for (int b = 0; b < nb; ++b)
            for (int it = 0; it < per; ++it) {
                const int id = per * b + it;
                if (L.count[id] < need) continue;
                for (int j = 1; j <= L.count[id]; ++j) {
                    const int peer = L.ids[id][j] / per;   // one rank per block
                    sends[b].push_back({peer, L.tag[id], it});
                    recvs[b].push_back({peer, L.tag[id], it});
                }
            }
        for (int b = 0; b < nb; ++b) {
            size_t q = 0;
            for (int it = 0; it < per; ++it) {
                const int id = per * b + it;
                if (L.count[id] < need) continue;
                const int len = kind == 0 ? 1 : edge_len(b, it);
                std::vector<std::pair<int, int>> from;   // (block, item) per partner
                for (int j = 1; j <= L.count[id]; ++j, ++q) {
                    const Post &r = recvs[b][q];
                    int order = 0;
                    for (size_t e = 0; e < q; ++e)
                        if (recvs[b][e].peer == r.peer && recvs[b][e].tag == r.tag) ++order;
                    int hit = -1;
                    for (const Post &s : sends[r.peer])
                        if (s.peer == b && s.tag == r.tag && order-- == 0) { hit = s.item; break; }
                    if (hit < 0 || (kind == 1 && edge_len(r.peer, hit) != len)) hit = -1;
                    from.push_back({r.peer, hit});
                }
                for (int k = 0; k < len; ++k) {
                    plan.target.push_back(point_of(b, kind, it, k));
                    for (auto &f : from)
                        plan.src.push_back(f.second < 0 ? point_of(b, kind, it, k) : point_of(f.first, kind, f.second, k));
                    plan.first.push_back(int32_t(plan.src.size()));
                    plan.scale.push_back(1.0 / double(L.count[id] + 1));
                }
            }
        }
    }
}

// average_line packs every edge before it averages any (mo_domdcomp.f90:505-529)
// and then writes edge by edge (:533-557), so a block corner on several
// averaged edges ends with the last edge's value; average_point runs before
// it (mo_canard_driver.f90:347-351).
JunctionProgram junction_program(const JunctionPlan &corners, const JunctionPlan &edges)
{
    JunctionProgram p;
    const int nc = int(corners.target.size()), ne = int(edges.target.size());
    std::map<int64_t, int> corner_of, last_edge;
    for (int i = 0; i < nc; ++i) corner_of[corners.target[i]] = i;
    for (int i = 0; i < ne; ++i) last_edge[edges.target[i]] = i;
    auto cidx = [&](int64_t g) { auto it = corner_of.find(g); return it == corner_of.end() ? -1 : it->second; };
    p.ncorner = nc;
    p.first.push_back(0);
    for (int i = 0; i < nc; ++i) {
        p.target.push_back(corners.target[i]);
        for (int k = corners.first[i]; k < corners.first[i + 1]; ++k) {
            p.src.push_back(corners.src[k]);
            p.sc.push_back(-1);
        }
        p.first.push_back(int32_t(p.src.size()));
        p.scale.push_back(corners.scale[i]);
        p.tc.push_back(-1);
        p.write.push_back(last_edge.count(corners.target[i]) ? 0 : 1);
    }
    for (int i = 0; i < ne; ++i) {
        if (last_edge[edges.target[i]] != i) continue;
        p.target.push_back(edges.target[i]);
        for (int k = edges.first[i]; k < edges.first[i + 1]; ++k) {
            p.src.push_back(edges.src[k]);
            p.sc.push_back(cidx(edges.src[k]));
        }
        p.first.push_back(int32_t(p.src.size()));
        p.scale.push_back(edges.scale[i]);
        p.tc.push_back(cidx(edges.target[i]));
        p.write.push_back(1);
    }
    return p;
}

}  // namespace cnd
