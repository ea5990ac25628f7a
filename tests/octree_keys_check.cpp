// CPU check of the device octree build (beep_hip_build_tree): descent keys
// (tree_key.hpp), sorted; the tree's shape from the sorted keys
// (Octree::build_sorted); the leaf and neighbourhood lists gathered from
// per-level stable sorts by key prefix, as octree_dev.hip does on the device.
// Every list must equal the host build's (Octree::build + leaves_csr), which
// tests/test_octree.py pins to the oracle.  Point sets: a bumpy sphere, a
// dense blob beside it, points on box boundaries (coordinates 0), and small
// neighbourhood targets (deep trees).  Prints one line per case; exit status
// = the number of cases that differ.
#include <cmath>
#include <cstdio>
#include <numeric>
#include <random>

#include "host.hpp"
#include "tree_key.hpp"

using namespace beep;

static int check(int npts, int nbsize, unsigned seed, bool blob) {
    std::vector<Vec3> pts;
    std::mt19937 g(seed);
    std::normal_distribution<double> nd(0, 1);
    std::uniform_real_distribution<double> ud(-1, 1);
    for (int i = 0; i < npts; ++i) {
        Vec3 v(nd(g), nd(g), nd(g));
        const double r = std::sqrt(v.dot(v));
        v = v * (25.0 * (1 + 0.1 * std::sin(3 * v.x / r)) / r);
        if (blob && i % 3 == 0) v = Vec3(40 + ud(g), ud(g), ud(g));
        if (i % 97 == 0) v = Vec3(0.0, 0.0, v.z);
        pts.push_back(v);
    }
    const Vec3 c(7.5, 0.0, 0.0);
    const double e = 140.0;
    Octree host;
    host.build(pts, c, e, nbsize);
    std::vector<int> lp, li, np_, ni;
    host.leaves_csr(lp, li, np_, ni);

    std::vector<uint64_t> key(npts);
    for (int i = 0; i < npts; ++i)
        key[i] = beeptree::descent_key(pts[i].x, pts[i].y, pts[i].z, c.x, c.y, c.z, e);
    std::vector<int> ord(npts);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::vector<uint64_t> sorted(npts);
    for (int i = 0; i < npts; ++i) sorted[i] = key[ord[i]];
    Octree dev;
    dev.build_sorted(sorted.data(), npts, c, e, nbsize);
    std::vector<int> lev, sptr, slo, shi;
    dev.leaf_segments(lev, sptr, slo, shi);
    std::vector<std::vector<int>> by_level(beeptree::KEY_LEVELS + 1);
    std::vector<int> lp2(1, 0), li2, np2(1, 0), ni2;
    for (size_t f = 0; f < lev.size(); ++f) {
        const int l = lev[f];
        std::vector<int>& S = by_level[l];
        if (S.empty()) {
            S.resize(npts);
            std::iota(S.begin(), S.end(), 0);
            std::stable_sort(S.begin(), S.end(), [&](int a, int b) {
                return beeptree::key_prefix(key[a], l) < beeptree::key_prefix(key[b], l);
            });
        }
        const int g0 = sptr[f];
        li2.insert(li2.end(), S.begin() + slo[g0], S.begin() + shi[g0]);
        lp2.push_back(int(li2.size()));
        for (int q = g0 + 1; q < sptr[f + 1]; ++q) ni2.insert(ni2.end(), S.begin() + slo[q], S.begin() + shi[q]);
        np2.push_back(int(ni2.size()));
    }
    const bool same = lp == lp2 && li == li2 && np_ == np2 && ni == ni2;
    std::printf("points %d nbsize %d blob %d: leaves %zu / %zu, neighbourhood entries %zu / %zu, depth %ld, %s\n",
                npts, nbsize, int(blob), lp.size() - 1, lp2.size() - 1, ni.size(), ni2.size(),
                host.bottom_level(), same ? "identical" : "DIFFERENT");
    return same ? 0 : 1;
}

int main() {
    int bad = 0;
    bad += check(20000, 1500, 1, false);
    bad += check(20000, 120, 2, false);
    bad += check(60000, 300, 3, true);
    bad += check(5000, 40, 5, true);
    bad += check(3000, 1500, 6, false);  // one box: no split at all
    return bad;
}
