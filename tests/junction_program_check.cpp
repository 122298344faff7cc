// Junction averaging as the library runs it (cnd::junction_program: one
// gather and one scatter) against the corner plan and then the edge plan
// applied one after the other, each entry written in plan order so that a
// point on several edges keeps the last edge's value (the order in which
// average_line writes, mo_domdcomp.f90:533-557).  Built and run by
// test_junction_program.py; prints one line per topology and exits nonzero
// on any bitwise difference.
#include <cstdio>
#include <cstring>
#include <random>

#include "cnd_scheme.hpp"

using namespace cnd;

namespace {
void apply_plan(const JunctionPlan &p, std::vector<double> &q, int64_t NT)
{
    const int n = int(p.target.size());
    std::vector<double> tmp(5 * size_t(n));
    for (int i = 0; i < n; ++i)
        for (int m = 0; m < 5; ++m) {
            double acc = q[m * NT + p.target[i]];
            for (int k = p.first[i]; k < p.first[i + 1]; ++k) acc = acc + q[m * NT + p.src[k]];
            tmp[m * size_t(n) + i] = p.scale[i] * acc;
        }
    for (int i = 0; i < n; ++i)
        for (int m = 0; m < 5; ++m) q[m * NT + p.target[i]] = tmp[m * size_t(n) + i];
}

double corner_value(const JunctionProgram &J, const std::vector<double> &q, int j, int m, int64_t NT)
{
    double acc = q[m * NT + J.target[j]];
    for (int k = J.first[j]; k < J.first[j + 1]; ++k) acc = acc + q[m * NT + J.src[k]];
    return J.scale[j] * acc;
}

// what k_junc_gather + k_junc_scatter do, thread by thread
void apply_program(const JunctionProgram &J, std::vector<double> &q, int64_t NT)
{
    const int n = int(J.target.size());
    std::vector<double> tmp(5 * size_t(n));
    for (int i = 0; i < n; ++i) {
        if (!J.write[i]) continue;
        for (int m = 0; m < 5; ++m) {
            double acc = J.tc[i] < 0 ? q[m * NT + J.target[i]] : corner_value(J, q, J.tc[i], m, NT);
            for (int k = J.first[i]; k < J.first[i + 1]; ++k)
                acc = acc + (J.sc[k] < 0 ? q[m * NT + J.src[k]] : corner_value(J, q, J.sc[k], m, NT));
            tmp[m * size_t(n) + i] = J.scale[i] * acc;
        }
    }
    for (int i = 0; i < n; ++i)
        if (J.write[i])
            for (int m = 0; m < 5; ++m) q[m * NT + J.target[i]] = tmp[m * size_t(n) + i];
}

int check(const char *name, const std::vector<BlockSpec> &blocks)
{
    std::vector<int64_t> off;
    int64_t NT = 0;
    for (const BlockSpec &b : blocks) { off.push_back(NT); NT += int64_t(b.n[0]) * b.n[1] * b.n[2]; }
    std::string err;
    if (!check_topology(blocks, err)) { std::printf("%s: topology refused: %s\n", name, err.c_str()); return 1; }
    JunctionPlan pc, pe;
    junction_plans(blocks, off, pc, pe);
    const JunctionProgram J = junction_program(pc, pe);
    std::mt19937_64 rng(11);
    std::uniform_real_distribution<double> u(0.5, 1.5);
    std::vector<double> q(5 * size_t(NT));
    for (double &x : q) x = u(rng);
    std::vector<double> a = q, b = q;
    apply_plan(pc, a, NT);
    apply_plan(pe, a, NT);
    apply_program(J, b, NT);
    long bad = 0, changed = 0;
    for (size_t i = 0; i < q.size(); ++i) {
        bad += std::memcmp(&a[i], &b[i], sizeof(double)) != 0;
        changed += a[i] != q[i];
    }
    std::printf("%s: corners %zu edges %zu kept %zu changed %ld mismatches %ld\n", name, pc.target.size(),
                pe.target.size(), J.target.size() - size_t(J.ncorner), changed, bad);
    return bad != 0 || changed == 0;
}

BlockSpec blk(int nx, int ny, int nz, int bc0, int bc1, int bc2)
{
    BlockSpec b{{nx, ny, nz}, {{bc0, bc0}, {bc1, bc1}, {bc2, bc2}}, {{-1, -1}, {-1, -1}, {-1, -1}}, 0};
    return b;
}
}  // namespace

int main()
{
    int fails = 0;
    {   // one block, periodic in every direction
        BlockSpec b = blk(13, 14, 15, 45, 45, 45);
        for (int d = 0; d < 3; ++d) b.partner[d][0] = b.partner[d][1] = 0;
        fails += check("periodic_block", {b});
    }
    for (int nzb : {1, 2}) {   // 2 x 2 x nzb blocks, periodic
        std::vector<BlockSpec> v;
        auto id = [&](int i, int j, int k) { return i + 2 * j + 4 * k; };
        for (int k = 0; k < nzb; ++k)
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < 2; ++i) {
                    BlockSpec b = blk(13, 14, 15, 35, 35, nzb == 1 ? 45 : 35);
                    b.partner[0][0] = b.partner[0][1] = id(1 - i, j, k);
                    b.partner[1][0] = b.partner[1][1] = id(i, 1 - j, k);
                    b.partner[2][0] = b.partner[2][1] = id(i, j, nzb == 1 ? k : 1 - k);
                    v.push_back(b);
                }
        fails += check(nzb == 1 ? "periodic_2x2x1" : "periodic_2x2x2", v);
    }
    {   // two blocks along xi, characteristic faces elsewhere, periodic in zeta
        BlockSpec a = blk(20, 14, 15, 10, 10, 45), b = blk(17, 14, 15, 10, 10, 45);
        a.bc[0][1] = 35; a.partner[0][1] = 1; a.partner[2][0] = a.partner[2][1] = 0;
        b.bc[0][0] = 35; b.partner[0][0] = 0; b.partner[2][0] = b.partner[2][1] = 1;
        fails += check("two_blocks_xi", {a, b});
    }
    std::printf("failures %d\n", fails);
    return fails;
}
