/*
 * cpu_ref.cpp -- the reference's CPU mat-vec path restated in C++/OpenMP, as
 * the CPU baseline of the measurements.  TEST INFRASTRUCTURE ONLY: loaded by
 * tests/ and the cpu_baseline leg of bench.py, never by the product.
 *
 * What it restates (the reference's CACHE_GPU_RESULTS CPU branch):
 *   precompute   BEEP::precalc_neighbour_interactions, src/beep/beep.cpp:768-956:
 *                one 24-byte LocalIntegrations record {u32 src, u32 targ, float A,
 *                B, C, D} per explicit pair (src/bem/local_integrations.h:17-25),
 *                the integrand in FP32 on FP32 points (bem_kernels.h:134-251 with
 *                float QuadPoints, quad_point.h:231), OpenMP over leaves; the FP64
 *                singular diagonal (oracle.c orc_singular)
 *   near product LocalIntegrations::evaluate_local_contributions,
 *                local_integrations.h:123-130 (scatter, FP64 sums)
 *   far field    FMM_BEM_Octree::solve, src/fmm/fmm_bem_hybrid.h:184-642: a Yukawa /
 *                Laplace plane-wave FMM in FP64, expansion order 9 with the 9-node
 *                plane-wave quadrature (the reference's <9,9,67>), on the tables of
 *                beep3_amd/csrc/pw_tables.cpp, three channels instead of the
 *                reference's twelve components
 *   solve        BEEP::gmres, src/beep/beep.cpp:1262-1483 (no restart, CGS, |y_j| test)
 *
 * The octree of the far field is uniform (leaves of about leaf_target
 * patches), as on the GPU, so its leaf pass runs over the neighbouring-leaf
 * pairs that are not explicit, and explicit pairs outside neighbouring leaves
 * store the full rule minus the 1-point interaction (which the FMM adds):
 * the operator is the oracle's (explicit pairs by the full rule, every other
 * pair by the 1-point interaction) up to the FMM and FP32 errors.
 */
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <unordered_map>
#include <vector>

#include "../beep3_amd/csrc/pw.hpp"
#include "oracle.h"

using beepdev::PwTableSet;

namespace {

constexpr double FOUR_PI_INV = 0.07957747154594767;

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// LocalIntegrations record (local_integrations.h:17-25)
struct Lint {
    uint32_t src, targ;
    float A, B, C, D;
};
static_assert(sizeof(Lint) == 24, "24-byte record");

// FP32 integrand over the qualocation points of the row patch and the
// quadrature points of the column patch, on FP32 absolute coordinates
// (bem_kernels.h:161-248; the same terms as oracle.c integrate()).
void integrate_f(float kappa, float inv_eps, const float* src, int nsrc, const float* targ, int ntarg,
                 float out[4]) {
    const float sigm = 1e-9f, g0 = float(FOUR_PI_INV);
    float Ao = 0, Bo = 0, Co = 0, Do = 0;
    for (int s = 0; s < nsrc; ++s) {
        const float* q = src + 7 * s;
        float A = 0, B = 0, C = 0, D = 0;
        for (int t = 0; t < ntarg; ++t) {
            const float* p = targ + 7 * t;
            const float wt = p[6];
            const float dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (std::fabs(r2) < sigm) continue;
            const float r = std::sqrt(r2);
            const float ir2 = 1.0f / r2, ir = 1.0f / r;
            const float ir3 = ir2 * ir, ir4 = ir2 * ir2, ir5 = ir4 * ir;
            const float gr1 = g0 * ir, gr3 = g0 * ir3;
            const float ur0 = std::exp(-kappa * r) * g0;
            const float ur1 = ur0 * ir, ur2 = ur0 * ir2, ur3 = ur0 * ir3;
            const float ur3ur2 = ur3 + kappa * ur2;
            const float gd[3] = {-dx * gr3, -dy * gr3, -dz * gr3};
            const float ud[3] = {-dx * ur3ur2, -dy * ur3ur2, -dz * ur3ur2};
            const float* n = p + 3;
            const float* n0 = q + 3;
            A += wt * (gr1 - ur1);
            B += wt * (n[0] * (inv_eps * gd[0] - ud[0]) + n[1] * (inv_eps * gd[1] - ud[1]) +
                       n[2] * (inv_eps * gd[2] - ud[2]));
            C += -wt * (n0[0] * (gd[0] - ud[0] * inv_eps) + n0[1] * (gd[1] - ud[1] * inv_eps) +
                        n0[2] * (gd[2] - ud[2] * inv_eps));
            if (kappa != 0) {
                const float gr5 = g0 * ir5, ur4 = ur0 * ir4, ur5 = ur0 * ir5;
                const float pur4ur3 = kappa * (3.0f * ur4 + kappa * ur3) + 3.0f * ur5;
                const float d[3] = {dx, dy, dz};
                for (int jj = 0; jj < 3; ++jj)
                    for (int kk = 0; kk < 3; ++kk) {
                        const float kr = jj == kk ? 1.0f : 0.0f;
                        const float djk = d[jj] * d[kk];
                        D -= wt * n0[kk] * n[jj] * ((3.0f * djk * gr5 - kr * gr3) - (djk * pur4ur3 - kr * ur3ur2));
                    }
            }
        }
        Ao += A * q[6];
        Bo += B * q[6];
        Co += C * q[6];
        Do += D * q[6];
    }
    out[0] = Ao;
    out[1] = Bo;
    out[2] = Co;
    out[3] = Do;
}

struct Box {
    uint32_t x, y, z;
};

uint64_t bkey(uint32_t x, uint32_t y, uint32_t z) { return (uint64_t(x) << 42) | (uint64_t(y) << 21) | z; }

struct CRef {
    const orc_geom* g = nullptr;
    int N = 0;
    double kappa = 0;
    std::vector<Lint> lints;
    std::vector<size_t> leaf_rec;  // per BEM leaf: first record (+ end)
    std::vector<double> diag;
    // uniform far-field octree
    int D = 0, p = 9, nlam = 9, nj = 0, jp = 0, nwp = 0;
    double edge = 0, x0[3] = {0, 0, 0};
    std::vector<std::vector<Box>> lev;
    std::vector<int> lvl_first;
    std::vector<double> box_cx, box_cy, box_cz, box_h;
    std::vector<int> parent, oct, child8;
    std::vector<int> leaf_ptr, leaf_pat;  // per leaf box: patches
    std::vector<int> xl_ptr, xl_src, xl_code;  // per box (levels >= 2) and direction
    std::vector<int> res_ptr, res_src;         // residual leaf pass per patch
    PwTableSet lap, yuk;
    std::vector<double> M, L, W, Wacc;  // [box][3][jp], level buffers [box][3][nwp]
    double setup_ms = 0, near_ms = 0, tables_ms = 0;
};

void build_tree(CRef& c, int leaf_target) {
    const orc_geom* g = c.g;
    const int N = c.N;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], g->centroid[3 * i + k]);
            hi[k] = std::max(hi[k], g->centroid[3 * i + k]);
        }
    double edge = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    edge = edge * (1.0 + 1e-9) + 1e-9;
    for (int k = 0; k < 3; ++k) c.x0[k] = (lo[k] + hi[k]) / 2 - edge / 2;
    c.edge = edge;
    auto cell = [&](int i, int l, int k) {
        const long cells = 1L << l;
        long v = long(std::floor((g->centroid[3 * i + k] - c.x0[k]) / edge * cells));
        return uint32_t(std::min(cells - 1, std::max(0L, v)));
    };
    int D = 0;
    for (; D < 12; ++D) {
        std::vector<uint64_t> keys(N);
        for (int i = 0; i < N; ++i) keys[i] = bkey(cell(i, D, 0), cell(i, D, 1), cell(i, D, 2));
        std::sort(keys.begin(), keys.end());
        const size_t occ = size_t(std::unique(keys.begin(), keys.end()) - keys.begin());
        if (double(N) / double(occ) <= leaf_target) break;
    }
    c.D = D;
    std::vector<std::unordered_map<uint64_t, int>> idx(D + 1);
    c.lev.assign(D + 1, {});
    std::vector<std::vector<int>> leaf_members;
    for (int i = 0; i < N; ++i) {
        const Box b{cell(i, D, 0), cell(i, D, 1), cell(i, D, 2)};
        auto it = idx[D].find(bkey(b.x, b.y, b.z));
        if (it == idx[D].end()) {
            it = idx[D].emplace(bkey(b.x, b.y, b.z), int(c.lev[D].size())).first;
            c.lev[D].push_back(b);
            leaf_members.emplace_back();
        }
        leaf_members[it->second].push_back(i);
    }
    for (int l = D - 1; l >= 0; --l)
        for (const Box& b : c.lev[l + 1]) {
            const Box pb{b.x >> 1, b.y >> 1, b.z >> 1};
            if (idx[l].emplace(bkey(pb.x, pb.y, pb.z), int(c.lev[l].size())).second) c.lev[l].push_back(pb);
        }
    c.lvl_first.assign(D + 2, 0);
    for (int l = 0; l <= D; ++l) c.lvl_first[l + 1] = c.lvl_first[l] + int(c.lev[l].size());
    const int nbox = c.lvl_first[D + 1];
    c.box_cx.resize(nbox);
    c.box_cy.resize(nbox);
    c.box_cz.resize(nbox);
    c.box_h.resize(D + 1);
    c.parent.assign(nbox, -1);
    c.oct.assign(nbox, 0);
    c.child8.assign(size_t(nbox) * 8, -1);
    for (int l = 0; l <= D; ++l) {
        const double h = edge / double(1L << l);
        c.box_h[l] = h;
        for (size_t q = 0; q < c.lev[l].size(); ++q) {
            const Box& b = c.lev[l][q];
            const int gi = c.lvl_first[l] + int(q);
            c.box_cx[gi] = c.x0[0] + (b.x + 0.5) * h;
            c.box_cy[gi] = c.x0[1] + (b.y + 0.5) * h;
            c.box_cz[gi] = c.x0[2] + (b.z + 0.5) * h;
            if (l > 0) {
                const int par = c.lvl_first[l - 1] + idx[l - 1].at(bkey(b.x >> 1, b.y >> 1, b.z >> 1));
                const int o = int(b.x & 1) | (int(b.y & 1) << 1) | (int(b.z & 1) << 2);
                c.parent[gi] = par;
                c.oct[gi] = o;
                c.child8[size_t(par) * 8 + o] = gi;
            }
        }
    }
    c.leaf_ptr.assign(c.lev[D].size() + 1, 0);
    for (size_t q = 0; q < c.lev[D].size(); ++q) {
        for (int i : leaf_members[q]) c.leaf_pat.push_back(i);
        c.leaf_ptr[q + 1] = int(c.leaf_pat.size());
    }
    // X2X lists (levels >= 2)
    double rot[6][3][3];
    beepdev::pw_direction_frames(rot);
    const int base = D >= 2 ? c.lvl_first[2] : nbox;
    c.xl_ptr.assign(size_t(nbox - base) * 6 + 1, 0);
    for (int l = 2; l <= D; ++l) {
        const long pcells = 1L << (l - 1);
        for (size_t q = 0; q < c.lev[l].size(); ++q) {
            const Box& b = c.lev[l][q];
            std::vector<std::pair<int, int>> byd[6];
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const long X = long(b.x >> 1) + dx, Y = long(b.y >> 1) + dy, Z = long(b.z >> 1) + dz;
                        if (X < 0 || Y < 0 || Z < 0 || X >= pcells || Y >= pcells || Z >= pcells) continue;
                        auto it = idx[l - 1].find(bkey(uint32_t(X), uint32_t(Y), uint32_t(Z)));
                        if (it == idx[l - 1].end()) continue;
                        const int pg = c.lvl_first[l - 1] + it->second;
                        for (int o = 0; o < 8; ++o) {
                            const int sg = c.child8[size_t(pg) * 8 + o];
                            if (sg < 0) continue;
                            const Box& sb = c.lev[l][sg - c.lvl_first[l]];
                            const int tx = int(long(b.x) - long(sb.x)), ty = int(long(b.y) - long(sb.y)),
                                      tz = int(long(b.z) - long(sb.z));
                            const int dir = beepdev::pw_direction(tx, ty, tz);
                            if (dir < 0) continue;
                            int tl[3];
                            for (int a = 0; a < 3; ++a)
                                tl[a] = int(std::lround(rot[dir][0][a] * tx + rot[dir][1][a] * ty + rot[dir][2][a] * tz));
                            byd[dir].push_back({sg, (tl[2] - 2) * 49 + (tl[1] + 3) * 7 + (tl[0] + 3)});
                        }
                    }
            const size_t row = size_t(c.lvl_first[l] + int(q) - base) * 6;
            for (int dir = 0; dir < 6; ++dir) {
                for (auto& e : byd[dir]) {
                    c.xl_src.push_back(e.first);
                    c.xl_code.push_back(e.second);
                }
                c.xl_ptr[row + dir + 1] = int(c.xl_src.size());
            }
        }
    }
    // leaf cell of every patch, and the residual leaf pass: pairs in
    // neighbouring leaf boxes that are not explicit
    std::vector<int> leafbox(N);
    for (size_t q = 0; q + 1 < c.leaf_ptr.size(); ++q)
        for (int k = c.leaf_ptr[q]; k < c.leaf_ptr[q + 1]; ++k) leafbox[c.leaf_pat[k]] = int(q);
    std::vector<int> bem_leaf(N);
    for (int l = 0; l < g->n_leaves; ++l)
        for (int k = g->leaf_ptr[l]; k < g->leaf_ptr[l + 1]; ++k) bem_leaf[g->leaf_idx[k]] = l;
    std::vector<std::vector<int>> res(N);
#pragma omp parallel
    {
        std::vector<char> mark(N, 0);
        int cur = -1;
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            const int bl = bem_leaf[i];
            if (bl != cur) {
                if (cur >= 0)
                    for (int k = g->nb_ptr[cur]; k < g->nb_ptr[cur + 1]; ++k) mark[g->nb_idx[k]] = 0;
                for (int k = g->nb_ptr[bl]; k < g->nb_ptr[bl + 1]; ++k) mark[g->nb_idx[k]] = 1;
                cur = bl;
            }
            const Box& b = c.lev[D][leafbox[i]];
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const long X = long(b.x) + dx, Y = long(b.y) + dy, Z = long(b.z) + dz;
                        const long cells = 1L << D;
                        if (X < 0 || Y < 0 || Z < 0 || X >= cells || Y >= cells || Z >= cells) continue;
                        auto it = idx[D].find(bkey(uint32_t(X), uint32_t(Y), uint32_t(Z)));
                        if (it == idx[D].end()) continue;
                        for (int k = c.leaf_ptr[it->second]; k < c.leaf_ptr[it->second + 1]; ++k) {
                            const int j = c.leaf_pat[k];
                            if (j != i && !mark[j]) res[i].push_back(j);
                        }
                    }
        }
    }
    c.res_ptr.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) c.res_ptr[i + 1] = c.res_ptr[i] + int(res[i].size());
    c.res_src.resize(c.res_ptr[N]);
    for (int i = 0; i < N; ++i) std::copy(res[i].begin(), res[i].end(), c.res_src.begin() + c.res_ptr[i]);
}

// near coefficients: explicit pairs, FP32; full rule when the pair's leaf
// cells are adjacent (the leaf pass skips it), else full minus 1-point
void precompute(CRef& c, int stride) {
    const orc_geom* g = c.g;
    const int N = c.N;
    std::vector<float> qual(size_t(g->qual_ptr[N]) * 7), quad(size_t(g->quad_ptr[N]) * 7);
    for (size_t k = 0; k < qual.size(); ++k) qual[k] = float(g->qual[k]);
    for (size_t k = 0; k < quad.size(); ++k) quad[k] = float(g->quad[k]);
    std::vector<uint32_t> cx(N), cy(N), cz(N);
    const double hD = c.box_h[c.D];
    for (int i = 0; i < N; ++i) {
        const long cells = 1L << c.D;
        auto cl = [&](int k) {
            long v = long(std::floor((g->centroid[3 * i + k] - c.x0[k]) / hD));
            return uint32_t(std::min(cells - 1, std::max(0L, v)));
        };
        cx[i] = cl(0);
        cy[i] = cl(1);
        cz[i] = cl(2);
    }
    c.leaf_rec.assign(g->n_leaves + 1, 0);
    for (int l = 0; l < g->n_leaves; ++l)
        c.leaf_rec[l + 1] = c.leaf_rec[l] + size_t(g->leaf_ptr[l + 1] - g->leaf_ptr[l]) * (g->nb_ptr[l + 1] - g->nb_ptr[l]);
    c.lints.resize(c.leaf_rec[g->n_leaves]);
    const float kf = float(c.kappa);
#pragma omp parallel for schedule(dynamic, 1)
    for (int l = 0; l < g->n_leaves; ++l) {
        size_t r = c.leaf_rec[l];
        const bool sampled = l % stride == 0;
        for (int a = g->leaf_ptr[l]; a < g->leaf_ptr[l + 1]; ++a) {
            const int i = g->leaf_idx[a];
            for (int k = g->nb_ptr[l]; k < g->nb_ptr[l + 1]; ++k) {
                const int j = g->nb_idx[k];
                Lint& li = c.lints[r++];
                li.src = uint32_t(i);
                li.targ = uint32_t(j);
                if (i == j || !sampled) {
                    li.A = li.B = li.C = li.D = 0.0f;
                    continue;
                }
                const float inv_eps = float(1.0 / g->eps[j]);
                float v[4];
                integrate_f(kf, inv_eps, &qual[size_t(g->qual_ptr[i]) * 7], g->qual_ptr[i + 1] - g->qual_ptr[i],
                            &quad[size_t(g->quad_ptr[j]) * 7], g->quad_ptr[j + 1] - g->quad_ptr[j], v);
                v[3] *= inv_eps;
                const bool adjacent = std::abs(int(cx[i]) - int(cx[j])) <= 1 && std::abs(int(cy[i]) - int(cy[j])) <= 1 &&
                                      std::abs(int(cz[i]) - int(cz[j])) <= 1;
                if (!adjacent) {
                    double o[4];
                    orc_pair_1pt(g, c.kappa, i, j, o);
                    for (int q = 0; q < 4; ++q) v[q] = float(double(v[q]) - o[q]);
                }
                li.A = v[0];
                li.B = v[1];
                li.C = v[2];
                li.D = v[3];
            }
        }
    }
    c.diag.resize(size_t(N) * 4);
    orc_singular(g, c.kappa, c.diag.data());
}

void matvec(CRef& c, const double* x, double* y) {
    const orc_geom* g = c.g;
    const int N = c.N;
    const double* f = x;
    const double* h = x + N;
    std::fill(y, y + 2 * size_t(N), 0.0);
    // near field: the records of a leaf write only its own rows
#pragma omp parallel for schedule(dynamic, 4)
    for (int l = 0; l < g->n_leaves; ++l)
        for (size_t r = c.leaf_rec[l]; r < c.leaf_rec[l + 1]; ++r) {
            const Lint& li = c.lints[r];
            y[li.src] += double(li.B) * f[li.targ] - double(li.A) * h[li.targ];
            y[N + li.src] += double(li.D) * f[li.targ] - double(li.C) * h[li.targ];
        }
#pragma omp parallel for
    for (int i = 0; i < N; ++i) {
        const double* d = &c.diag[size_t(i) * 4];
        y[i] += d[1] * f[i] - d[0] * h[i];
        y[N + i] += d[3] * f[i] - d[2] * h[i];
    }
    // far field: plane-wave FMM
    const int D = c.D, jp = c.jp, nj = c.nj, nwp = c.nwp, p = c.p;
    const int nbox = c.lvl_first[D + 1];
    std::fill(c.L.begin(), c.L.end(), 0.0);
    const double hD = c.box_h[D];
    const int nleaf = int(c.lev[D].size());
    const double bL = c.kappa * hD;
    // P2M
#pragma omp parallel
    {
        std::vector<double> psL(nj), psY(nj), gL(3 * nj), gY(3 * nj);
#pragma omp for schedule(dynamic, 8)
        for (int q = 0; q < nleaf; ++q) {
            const int b = c.lvl_first[D] + q;
            double* Mb = &c.M[size_t(b) * 3 * jp];
            std::fill(Mb, Mb + 3 * jp, 0.0);
            for (int k = c.leaf_ptr[q]; k < c.leaf_ptr[q + 1]; ++k) {
                const int j = c.leaf_pat[k];
                const double v[3] = {(g->centroid[3 * j] - c.box_cx[b]) / hD, (g->centroid[3 * j + 1] - c.box_cy[b]) / hD,
                                     (g->centroid[3 * j + 2] - c.box_cz[b]) / hD};
                beepdev::pw_basis(p, 0.0, v, psL.data(), reinterpret_cast<double(*)[3]>(gL.data()));
                beepdev::pw_basis(p, bL, v, psY.data(), reinterpret_cast<double(*)[3]>(gY.data()));
                const double a = g->area[j], ie = 1.0 / g->eps[j];
                const double sf = a * f[j] / hD, sh = a * h[j];
                const double* n = &g->normal[3 * j];
                for (int t = 0; t < nj; ++t) {
                    const double dL = n[0] * gL[3 * t] + n[1] * gL[3 * t + 1] + n[2] * gL[3 * t + 2];
                    const double dY = n[0] * gY[3 * t] + n[1] * gY[3 * t + 1] + n[2] * gY[3 * t + 2];
                    Mb[t] += -sh * psL[t] + ie * sf * dL;
                    const double c1 = sh * psY[t] - sf * dY;
                    Mb[jp + t] += c1;
                    Mb[2 * jp + t] += ie * c1;
                }
            }
        }
    }
    // M2M
    for (int l = D - 1; l >= 2; --l) {
#pragma omp parallel for schedule(dynamic, 8)
        for (int b = c.lvl_first[l]; b < c.lvl_first[l + 1]; ++b) {
            double* Mp = &c.M[size_t(b) * 3 * jp];
            std::fill(Mp, Mp + 3 * jp, 0.0);
            for (int o = 0; o < 8; ++o) {
                const int ch = c.child8[size_t(b) * 8 + o];
                if (ch < 0) continue;
                for (int e = 0; e < 3; ++e) {
                    const double* A = (e == 0 ? c.lap : c.yuk).lev[l].m2mT.data() + size_t(o) * jp * jp;
                    const double* Mc = &c.M[(size_t(ch) * 3 + e) * jp];
                    double* out = Mp + e * jp;
                    for (int jj = 0; jj < nj; ++jj) {
                        const double m = Mc[jj];
                        const double* row = A + size_t(jj) * jp;
                        for (int t = 0; t < nj; ++t) out[t] += row[t] * m;
                    }
                }
            }
        }
    }
    // downward pass
    const int base = D >= 2 ? c.lvl_first[2] : nbox;
    for (int l = 2; l <= D; ++l) {
        const int first = c.lvl_first[l], cnt = c.lvl_first[l + 1] - first;
        const double ih = 1.0 / c.box_h[l];
        if (l >= 3) {
#pragma omp parallel for schedule(dynamic, 8)
            for (int b = first; b < first + cnt; ++b) {
                const int par = c.parent[b], o = c.oct[b];
                for (int e = 0; e < 3; ++e) {
                    const double* A = (e == 0 ? c.lap : c.yuk).lev[l - 1].l2lT.data() + size_t(o) * jp * jp;
                    const double* Lp = &c.L[(size_t(par) * 3 + e) * jp];
                    double* out = &c.L[(size_t(b) * 3 + e) * jp];
                    for (int jj = 0; jj < nj; ++jj) {
                        const double lv = Lp[jj];
                        const double* row = A + size_t(jj) * jp;
                        for (int t = 0; t < nj; ++t) out[t] += lv * row[t];
                    }
                }
            }
        }
        // M2X
#pragma omp parallel for schedule(dynamic, 8)
        for (int q = 0; q < cnt; ++q) {
            for (int e = 0; e < 3; ++e) {
                const beepdev::PwLevelTables& t = (e == 0 ? c.lap : c.yuk).lev[l];
                const int tnwp = (e == 0 ? c.lap : c.yuk).nwp, ncol = 12 * t.q.nw;
                const double* Mb = &c.M[(size_t(first + q) * 3 + e) * jp];
                double* w = &c.W[(size_t(q) * 3 + e) * nwp];
                std::fill(w, w + ncol, 0.0);
                for (int jj = 0; jj < nj; ++jj) {
                    const double m = Mb[jj] * ih;
                    const double* row = t.m2x.data() + size_t(jj) * tnwp;
                    for (int k = 0; k < ncol; ++k) w[k] += m * row[k];
                }
            }
        }
        // X2X, then X2L
#pragma omp parallel
        {
            std::vector<double> acc(nwp);
#pragma omp for schedule(dynamic, 8)
            for (int q = 0; q < cnt; ++q) {
                const int b = first + q;
                for (int e = 0; e < 3; ++e) {
                    const beepdev::PwLevelTables& t = (e == 0 ? c.lap : c.yuk).lev[l];
                    const int nw = t.q.nw, tnwp = (e == 0 ? c.lap : c.yuk).nwp;
                    std::fill(acc.begin(), acc.end(), 0.0);
                    for (int dir = 0; dir < 6; ++dir) {
                        const size_t row = size_t(b - base) * 6 + dir;
                        for (int x2 = c.xl_ptr[row]; x2 < c.xl_ptr[row + 1]; ++x2) {
                            const double* w = &c.W[(size_t(c.xl_src[x2] - first) * 3 + e) * nwp + size_t(dir) * nw * 2];
                            const double* sh = t.shift.data() + size_t(c.xl_code[x2]) * nw * 2;
                            double* a = acc.data() + size_t(dir) * nw * 2;
                            for (int k = 0; k < nw; ++k) {
                                const double wr = w[2 * k], wi = w[2 * k + 1], sr = sh[2 * k], si = sh[2 * k + 1];
                                a[2 * k] += wr * sr - wi * si;
                                a[2 * k + 1] += wr * si + wi * sr;
                            }
                        }
                    }
                    double* out = &c.L[(size_t(b) * 3 + e) * jp];
                    for (int k = 0; k < 12 * nw; ++k) {
                        const double av = acc[k];
                        if (av == 0.0) continue;
                        const double* row = t.x2l.data() + size_t(k) * jp;
                        for (int jj = 0; jj < nj; ++jj) out[jj] += av * row[jj];
                    }
                    (void)tnwp;
                }
            }
        }
    }
    // L2P and the residual leaf pass
    if (D >= 2) {
#pragma omp parallel
        {
            std::vector<double> psL(nj), psY(nj), gL(3 * nj), gY(3 * nj);
#pragma omp for schedule(dynamic, 8)
            for (int q = 0; q < nleaf; ++q) {
                const int b = c.lvl_first[D] + q;
                const double* Lb = &c.L[size_t(b) * 3 * jp];
                for (int k = c.leaf_ptr[q]; k < c.leaf_ptr[q + 1]; ++k) {
                    const int i = c.leaf_pat[k];
                    const double v[3] = {(g->centroid[3 * i] - c.box_cx[b]) / hD, (g->centroid[3 * i + 1] - c.box_cy[b]) / hD,
                                         (g->centroid[3 * i + 2] - c.box_cz[b]) / hD};
                    beepdev::pw_basis(p, 0.0, v, psL.data(), reinterpret_cast<double(*)[3]>(gL.data()));
                    beepdev::pw_basis(p, bL, v, psY.data(), reinterpret_cast<double(*)[3]>(gY.data()));
                    double pot = 0, gr[3] = {0, 0, 0};
                    for (int t = 0; t < nj; ++t) {
                        pot += Lb[t] * psL[t] + Lb[jp + t] * psY[t];
                        for (int a = 0; a < 3; ++a) gr[a] += Lb[t] * gL[3 * t + a] + Lb[2 * jp + t] * gY[3 * t + a];
                    }
                    const double* n = &g->normal[3 * i];
                    y[i] += pot * FOUR_PI_INV;
                    y[N + i] += (n[0] * gr[0] + n[1] * gr[1] + n[2] * gr[2]) / hD * FOUR_PI_INV;
                }
            }
        }
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
        double yf = 0, yh = 0, o[4];
        for (int k = c.res_ptr[i]; k < c.res_ptr[i + 1]; ++k) {
            const int j = c.res_src[k];
            orc_pair_1pt(g, c.kappa, i, j, o);
            yf += o[1] * f[j] - o[0] * h[j];
            yh += o[3] * f[j] - o[2] * h[j];
        }
        y[i] += yf;
        y[N + i] += yh;
    }
}

double dot(const double* a, const double* b, size_t n) {
    double s = 0;
#pragma omp parallel for reduction(+ : s)
    for (long k = 0; k < long(n); ++k) s += a[k] * b[k];
    return s;
}

}  // namespace

extern "C" {

/* Build the CPU reference path for a scene: far-field octree, plane-wave
 * tables (order, nodes), residual leaf pass, FP32 LocalIntegrations and the
 * singular diagonal.  stride > 1 computes the coefficients of every
 * stride-th leaf only (a timing sample: the other records keep their
 * indices, with zero values, so a mat-vec still walks every record).
 * info[9]: setup ms (all), near precompute ms, records, residual pairs,
 * leaf level, boxes, table ms, threads, records computed. */
void* cref_create(const orc_geom* g, double kappa, int order, int nodes, int leaf_target, int stride,
                  double* info) {
    auto* c = new CRef();
    c->g = g;
    c->N = g->n;
    c->kappa = kappa;
    c->p = order;
    c->nlam = nodes;
    const double t0 = now_ms();
    build_tree(*c, leaf_target);
    const double t1 = now_ms();
    c->lap = beepdev::pw_build_tables(order, nodes, 0.0, c->edge, c->D);
    c->yuk = beepdev::pw_build_tables(order, nodes, kappa, c->edge, c->D);
    c->nj = (order + 1) * (order + 1);
    c->jp = c->lap.jp;
    c->nwp = std::max(c->lap.nwp, c->yuk.nwp);
    const int nbox = c->lvl_first[c->D + 1];
    int maxlev = 1;
    for (int l = 2; l <= c->D; ++l) maxlev = std::max(maxlev, int(c->lev[l].size()));
    c->M.assign(size_t(nbox) * 3 * c->jp, 0.0);
    c->L.assign(size_t(nbox) * 3 * c->jp, 0.0);
    c->W.assign(size_t(maxlev) * 3 * c->nwp, 0.0);
    const double t2 = now_ms();
    stride = std::max(1, stride);
    precompute(*c, stride);
    const double t3 = now_ms();
    c->tables_ms = t2 - t1;
    c->near_ms = t3 - t2;
    c->setup_ms = t3 - t0;
    if (info) {
        info[0] = c->setup_ms;
        info[1] = c->near_ms;
        info[2] = double(c->lints.size());
        info[3] = double(c->res_src.size());
        info[4] = c->D;
        info[5] = nbox;
        info[6] = c->tables_ms;
        info[7] = omp_get_max_threads();
        size_t computed = 0;
        for (int l = 0; l < g->n_leaves; l += stride) computed += c->leaf_rec[l + 1] - c->leaf_rec[l];
        info[8] = double(computed);
    }
    return c;
}

/* y = K x (x = [f; h]); returns the wall milliseconds. */
double cref_matvec(void* h, const double* x, double* y) {
    const double t0 = now_ms();
    matvec(*static_cast<CRef*>(h), x, y);
    return now_ms() - t0;
}

/* GMRES of the reference (beep.cpp:1262-1483): x holds the initial guess;
 * returns the mat-vec count (its + 1); *ms the wall time. */
int cref_gmres(void* h, const double* b, double* x, double eps, int m, double* ms) {
    CRef& c = *static_cast<CRef*>(h);
    const double t0 = now_ms();
    const size_t n = 2 * size_t(c.N);
    std::vector<double> V(n * (m + 1)), U(size_t(m) * (m + 1) / 2 + 1), r(n), yv(m + 1), cs(m + 1), sn(m + 1);
    auto near0 = [](double v) { return v == 0 ? 2.2250738585072014e-308 : v; };
    const double nb = std::sqrt(dot(b, b, n));
    matvec(c, x, r.data());
    for (size_t k = 0; k < n; ++k) r[k] -= b[k];
    const double beta = std::sqrt(dot(r.data(), r.data(), n));
    for (size_t k = 0; k < n; ++k) V[k] = r[k] / near0(beta);
    yv[0] = beta;
    int j = 0, uij = 0;
    do {
        const int u0j = uij;
        double* vj = &V[size_t(j) * n];
        double* vn = &V[size_t(j + 1) * n];
        matvec(c, vj, vn);
        for (int q = 0; q <= j; ++q) U[u0j + q] = dot(&V[size_t(q) * n], vn, n);
#pragma omp parallel for
        for (long k = 0; k < long(n); ++k) {
            double s = vn[k];
            for (int q = 0; q <= j; ++q) s -= U[u0j + q] * V[size_t(q) * n + k];
            vn[k] = s;
        }
        const double hn = std::sqrt(dot(vn, vn, n));
#pragma omp parallel for
        for (long k = 0; k < long(n); ++k) vn[k] /= near0(hn);
        for (int i = 0; i < j; ++i) {
            const double tmp = cs[i] * U[uij] - sn[i] * U[uij + 1];
            U[uij + 1] = sn[i] * U[uij] + cs[i] * U[uij + 1];
            U[uij] = tmp;
            ++uij;
        }
        const double rd = U[uij], dd = std::sqrt(rd * rd + hn * hn);
        cs[j] = rd / near0(dd);
        sn[j] = -hn / near0(dd);
        U[uij] = dd;
        ++uij;
        yv[j + 1] = sn[j] * yv[j];
        yv[j] = cs[j] * yv[j];
        ++j;
    } while (j < m && std::fabs(yv[j]) > eps * nb);
    for (int i = j - 1; i >= 0; --i) {
        double acc = yv[i];
        for (int q = i + 1; q < j; ++q) acc -= U[size_t(q) * (q + 1) / 2 + i] * yv[q];
        yv[i] = acc / U[size_t(i) * (i + 1) / 2 + i];
    }
#pragma omp parallel for
    for (long k = 0; k < long(n); ++k) {
        double s = x[k];
        for (int q = 0; q < j; ++q) s -= V[size_t(q) * n + k] * yv[q];
        x[k] = s;
    }
    if (ms) *ms = now_ms() - t0;
    return j + 1;
}

void cref_destroy(void* h) { delete static_cast<CRef*>(h); }

}  // extern "C"
