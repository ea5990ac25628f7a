// Adaptive octree of the BEM-FMM scene, restating src/fmm/octree.h under the
// policy of BEEP::init_fmm_octree (src/beep/beep.cpp:1585-1641), kept here as
// the checker of the product's host octree (beep3_amd/csrc/host_octree.cpp):
// test infrastructure only.  Written for clarity, not speed:
//   relaxed insertion (split a childless node at > 4*nbsize, octree.h:1134-1142),
//   build_neighbourhoods (octree.h:780-805),
//   optimize_neighbourhoods (octree.h:867-986),
//   remove_empty_nodes (octree.h:1052-1082),
//   leaves and neighbourhood contents (get_neighbourhood_contents, octree.h:627-648).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <vector>

namespace orc_tree {

struct Vec3 {
    double x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
};

struct OctNode {
    int level;
    uint32_t ix, iy, iz;
    Vec3 centre;
    double edge;
    int parent = -1;
    int children[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    bool has_children = false, is_leaf = true, is_shadow = false, is_deleted = false;
    bool removed = false;
    long cached_nb = -1;
    std::vector<int> contents;
};

struct Octree {
    Vec3 centre;
    double edge = 0;
    int nbsize = 1500;
    int max_depth = 20;
    std::vector<OctNode> nodes;
    std::vector<std::map<uint64_t, int>> levels;
    std::vector<int> leaf_nodes;
    static uint64_t key(uint32_t x, uint32_t y, uint32_t z) {
        return (uint64_t(x) << 42) | (uint64_t(y) << 21) | uint64_t(z);
    }
    int find(int level, long x, long y, long z) const;
    int neighbour(int node, int nb) const;
    void build(const std::vector<Vec3>& pts, const Vec3& c, double e, int nb);
    void leaves_csr(std::vector<int>& leaf_ptr, std::vector<int>& leaf_idx,
                    std::vector<int>& nb_ptr, std::vector<int>& nb_idx) const;
    std::vector<Vec3> const* pts_ = nullptr;
    int new_node(int level, uint32_t x, uint32_t y, uint32_t z, const Vec3& c, double e, int parent);
    void subdivide(int n);
    void subdivide_neighbourhood(int n);
    void add(int item, int node);
    long count_nb(int n) const;
    long cached_nb(int n);
    double interacts_per_item(long total_items);
};


int Octree::new_node(int level, uint32_t x, uint32_t y, uint32_t z, const Vec3& c, double e,
                     int parent) {
    OctNode n;
    n.level = level;
    n.ix = x; n.iy = y; n.iz = z;
    n.centre = c;
    n.edge = e;
    n.parent = parent;
    nodes.push_back(std::move(n));
    int id = int(nodes.size()) - 1;
    while (int(levels.size()) <= level) levels.emplace_back();
    levels[level][key(x, y, z)] = id;
    return id;
}

int Octree::find(int level, long x, long y, long z) const {
    if (level < 0 || level >= int(levels.size())) return -1;
    const long cells = 1L << level;
    if (x < 0 || y < 0 || z < 0 || x >= cells || y >= cells || z >= cells) return -1;
    auto it = levels[level].find(key(uint32_t(x), uint32_t(y), uint32_t(z)));
    if (it == levels[level].end()) return -1;
    return nodes[it->second].removed ? -1 : it->second;
}

// neighbour id nb in [0, 27): dz = nb/9 - 1, dy = (nb%9)/3 - 1, dx = nb%3 - 1
// (OctreeIndexer::get_neighbour_idxer, src/common/octree_indexer.h:350-390)
int Octree::neighbour(int n, int nb) const {
    const OctNode& o = nodes[n];
    long dz = nb / 9 - 1, dy = (nb % 9) / 3 - 1, dx = nb % 3 - 1;
    return find(o.level, long(o.ix) + dx, long(o.iy) + dy, long(o.iz) + dz);
}

// child ids 1..8 of the reference map to (x, y, z) bits: id-1 = x + 2y + 4z
void Octree::subdivide(int n) {
    if (nodes[n].has_children) return;
    nodes[n].has_children = true;
    const int level = nodes[n].level + 1;
    for (int c = 0; c < 8; ++c) {
        const OctNode& p = nodes[n];
        const double h = p.edge / 2.0;
        Vec3 cc = p.centre;
        cc.x += (c & 1) ? h / 2.0 : -h / 2.0;
        cc.y += (c & 2) ? h / 2.0 : -h / 2.0;
        cc.z += (c & 4) ? h / 2.0 : -h / 2.0;
        int id = new_node(level, p.ix * 2 + (c & 1), p.iy * 2 + ((c >> 1) & 1),
                          p.iz * 2 + ((c >> 2) & 1), cc, h, n);
        nodes[n].children[c] = id;
        // the child receives the parent's items that fall in its octant
        // (Node child constructor, octree.h:120-131), in the parent's order
        const OctNode& pp = nodes[n];
        const Vec3& pc = pp.centre;
        for (int item : pp.contents) {
            const Vec3& q = (*pts_)[item];
            int cid = (q.x > pc.x ? 1 : 0) + (q.y > pc.y ? 2 : 0) + (q.z > pc.z ? 4 : 0);
            if (cid == c) nodes[id].contents.push_back(item);
        }
    }
}

void Octree::add(int item, int n) {
    for (;;) {
        nodes[n].contents.push_back(item);
        if (!nodes[n].has_children) {
            if (long(nodes[n].contents.size()) > 4L * nbsize && nodes[n].level < max_depth) {
                subdivide(n);
                nodes[n].is_leaf = false;
            }
            return;
        }
        const Vec3& q = (*pts_)[item];
        const Vec3& c = nodes[n].centre;
        int cid = (q.x > c.x ? 1 : 0) + (q.y > c.y ? 2 : 0) + (q.z > c.z ? 4 : 0);
        n = nodes[n].children[cid];
    }
}

void Octree::subdivide_neighbourhood(int n) {
    for (int nb = 0; nb < 27; ++nb) {
        int m = neighbour(n, nb);
        if (m < 0) continue;
        if (!nodes[m].has_children) {
            subdivide(m);
            for (int c = 0; c < 8; ++c) {
                OctNode& ch = nodes[nodes[m].children[c]];
                ch.is_leaf = false;
                ch.is_shadow = true;
            }
        } else {
            for (int c = 0; c < 8; ++c) {
                int ch = nodes[m].children[c];
                if (ch >= 0) nodes[ch].is_deleted = false;
            }
        }
    }
}

long Octree::count_nb(int n) const {
    long r = long(nodes[n].contents.size());
    for (int nb = 0; nb < 27; ++nb) {
        if (nb == 13) continue;
        int m = neighbour(n, nb);
        if (m >= 0) r += long(nodes[m].contents.size());
    }
    return r;
}

long Octree::cached_nb(int n) {
    if (nodes[n].cached_nb < 0) nodes[n].cached_nb = count_nb(n);
    return nodes[n].cached_nb;
}

double Octree::interacts_per_item(long total_items) {
    long total = 0;
    for (int level = 0; level < int(levels.size()) && level < max_depth; ++level)
        for (const auto& kv : levels[level]) {
            const OctNode& o = nodes[kv.second];
            if (o.is_leaf && !o.is_shadow && !o.is_deleted)
                total += cached_nb(kv.second) * long(o.contents.size());
        }
    return double(total) / double(total_items);
}

void Octree::build(const std::vector<Vec3>& pts, const Vec3& c, double e, int nb) {
    nodes.clear();
    levels.clear();
    leaf_nodes.clear();
    centre = c;
    edge = e;
    nbsize = nb;
    pts_ = &pts;
    new_node(0, 0, 0, 0, c, e, -1);
    for (int i = 0; i < int(pts.size()); ++i) add(i, 0);

    // build_neighbourhoods: every non-shadow node with children gets a
    // complete set of same-level neighbours with children
    for (int level = 0; level < int(levels.size()); ++level) {
        std::vector<int> with_children;
        for (const auto& kv : levels[level]) {
            const OctNode& o = nodes[kv.second];
            if (o.has_children && !o.is_shadow) with_children.push_back(kv.second);
        }
        for (int n : with_children) subdivide_neighbourhood(n);
    }

    // optimize_neighbourhoods
    long total_items = 0;
    for (int level = 0; level < int(levels.size()); ++level)
        for (const auto& kv : levels[level]) {
            const OctNode& o = nodes[kv.second];
            if (!o.is_shadow && o.is_leaf && !o.is_deleted) total_items += long(o.contents.size());
        }
    double avg = interacts_per_item(total_items);
    while (std::floor(avg) > nbsize) {
        std::vector<int> fat;
        for (int level = 0; level < int(levels.size()); ++level)
            for (const auto& kv : levels[level]) {
                const OctNode& o = nodes[kv.second];
                if (!o.is_shadow && o.is_leaf && !o.contents.empty() && cached_nb(kv.second) > nbsize)
                    fat.push_back(kv.second);
            }
        // fattest neighbourhoods first; ties keep level/key order
        std::stable_sort(fat.begin(), fat.end(),
                         [&](int a, int b) { return nodes[a].cached_nb > nodes[b].cached_nb; });
        const int estimate =
            1 + int(std::ceil(double(fat.size()) * (avg - nbsize) / avg));
        int done = 0;
        for (int n : fat) {
            nodes[n].is_leaf = false;
            if (nodes[n].has_children) {
                for (int c = 0; c < 8; ++c) {
                    int ch = nodes[n].children[c];
                    if (ch < 0) continue;
                    nodes[ch].is_shadow = false;
                    nodes[ch].is_leaf = true;
                }
            } else {
                subdivide(n);
            }
            subdivide_neighbourhood(n);
            if (++done >= estimate) break;
        }
        avg = interacts_per_item(total_items);
    }

    // remove_empty_nodes (bottom-up), keeping the root
    for (int level = int(levels.size()) - 1; level >= 0; --level)
        for (const auto& kv : levels[level]) {
            OctNode& o = nodes[kv.second];
            if (o.contents.empty() && kv.second != 0 && !o.removed) {
                o.removed = true;
                if (o.parent >= 0)
                    for (int c = 0; c < 8; ++c)
                        if (nodes[o.parent].children[c] == kv.second) nodes[o.parent].children[c] = -1;
            }
        }

    for (int level = 0; level < int(levels.size()); ++level)
        for (const auto& kv : levels[level]) {
            const OctNode& o = nodes[kv.second];
            if (!o.removed && o.is_leaf && !o.is_shadow && !o.contents.empty())
                leaf_nodes.push_back(kv.second);
        }
    // sanity: every item belongs to exactly one leaf
    std::vector<int> owner(pts.size(), 0);
    for (int n : leaf_nodes)
        for (int i : nodes[n].contents) owner[i]++;
    for (int c : owner)
        if (c != 1) throw std::runtime_error("octree: a patch is not in exactly one leaf");
}

void Octree::leaves_csr(std::vector<int>& leaf_ptr, std::vector<int>& leaf_idx,
                        std::vector<int>& nb_ptr, std::vector<int>& nb_idx) const {
    leaf_ptr.assign(1, 0);
    nb_ptr.assign(1, 0);
    leaf_idx.clear();
    nb_idx.clear();
    for (int n : leaf_nodes) {
        const OctNode& o = nodes[n];
        leaf_idx.insert(leaf_idx.end(), o.contents.begin(), o.contents.end());
        leaf_ptr.push_back(int(leaf_idx.size()));
        for (int nb = 0; nb < 27; ++nb) {
            int m = neighbour(n, nb);
            if (m < 0) continue;
            const auto& c = nodes[m].contents;
            nb_idx.insert(nb_idx.end(), c.begin(), c.end());
        }
        nb_ptr.push_back(int(nb_idx.size()));
    }
}


}  // namespace orc_tree

extern "C" int orc_octree(int n, const double* pts, const double* centre, double edge, int nbsize,
                          int* n_leaves, long long* n_nb, int** leaf_ptr, int** leaf_idx, int** nb_ptr,
                          int** nb_idx) {
    try {
        std::vector<orc_tree::Vec3> p;
        p.resize(size_t(n));
        for (int i = 0; i < n; ++i) p[i] = orc_tree::Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        orc_tree::Octree t;
        t.build(p, orc_tree::Vec3(centre[0], centre[1], centre[2]), edge, nbsize);
        std::vector<int> lp, li, np, ni;
        t.leaves_csr(lp, li, np, ni);
        auto out = [](const std::vector<int>& v) {
            int* a = static_cast<int*>(std::malloc(sizeof(int) * (v.size() + 1)));
            std::memcpy(a, v.data(), sizeof(int) * v.size());
            return a;
        };
        *n_leaves = int(lp.size()) - 1;
        *n_nb = (long long)ni.size();
        *leaf_ptr = out(lp);
        *leaf_idx = out(li);
        *nb_ptr = out(np);
        *nb_idx = out(ni);
        return 0;
    } catch (...) {
        return -1;
    }
}

extern "C" void orc_free(void* p) { std::free(p); }
