// Adaptive octree restating src/fmm/octree.h under the BEM-FMM policy of
// BEEP::init_fmm_octree (src/beep/beep.cpp:1585-1641):
//   relaxed insertion (split a childless node at > 4*nbsize, octree.h:1134-1142),
//   build_neighbourhoods (octree.h:780-805),
//   optimize_neighbourhoods (octree.h:867-986),
//   remove_empty_nodes (octree.h:1052-1082).
// The result decides which patch pairs are integrated explicitly: every leaf
// interacts explicitly with the contents of its 27 same-level neighbour boxes
// (get_neighbourhood_contents, octree.h:627-648), everything else is far field.
//
// The nodes hold no item lists: a node's items are one range of a single
// array of the items, and a child's range is part of its parent's.  Every
// decision of the build depends only on node counts.  Two ways to split a
// node's range into its children's:
//   build():        the items start in ascending order and a split is a
//                   stable in-place partition by octant (the reference's
//                   comparison of the item with the node centre), so a node
//                   that has no children keeps its items in ascending order
//                   (the reference's insertion order);
//   build_sorted(): the items are given as their descent keys (tree_key.hpp)
//                   in ascending order, and a child's range is found by binary
//                   search on its key digit (the device builds the tree so).
#include <algorithm>
#include <stdexcept>

#include "host.hpp"
#include "tree_key.hpp"

namespace beep {

namespace {
inline size_t mix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return size_t(k);
}
}  // namespace

void Octree::LevelIndex::insert(uint64_t k, int id) {
    if (2 * (count + 1) > keys.size()) {
        // grow to keep the load factor at most 1/2
        std::vector<uint64_t> ok;
        std::vector<int> oi;
        ok.swap(keys);
        oi.swap(ids);
        const size_t cap = std::max<size_t>(64, ok.size() * 2);
        keys.assign(cap, ~0ull);
        ids.assign(cap, -1);
        count = 0;
        for (size_t s = 0; s < ok.size(); ++s)
            if (ok[s] != ~0ull) insert(ok[s], oi[s]);
    }
    const size_t m = keys.size() - 1;
    size_t s = mix(k) & m;
    while (keys[s] != ~0ull && keys[s] != k) s = (s + 1) & m;
    if (keys[s] == ~0ull) ++count;
    keys[s] = k;
    ids[s] = id;
}

int Octree::LevelIndex::find(uint64_t k) const {
    if (keys.empty()) return -1;
    const size_t m = keys.size() - 1;
    size_t s = mix(k) & m;
    while (keys[s] != ~0ull) {
        if (keys[s] == k) return ids[s];
        s = (s + 1) & m;
    }
    return -1;
}

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
    while (int(levels.size()) <= level) {
        levels.emplace_back();
        index.emplace_back();
    }
    levels[level].push_back(id);
    nkey_.push_back(key(x, y, z));
    ncount_.push_back(0);
    gone_.push_back(0);
    index[level].insert(nkey_.back(), id);
    return id;
}

// the walks over a level go in key order (x, y, z bits), the order of the
// reference's per-level maps: ties among equally fat neighbourhoods and the
// leaf order depend on it
const std::vector<int>& Octree::ordered(int level) {
    if (int(sorted_.size()) < int(levels.size())) sorted_.resize(levels.size());
    std::vector<int>& v = sorted_[level];
    const std::vector<int>& all = levels[level];
    if (v.size() != all.size()) {
        // a level only grows at its end (creation order): sort the new ids
        // and merge them into the sorted ones
        const size_t old = v.size();
        v.insert(v.end(), all.begin() + old, all.end());
        auto less = [&](int a, int b) { return nkey_[a] < nkey_[b]; };
        std::sort(v.begin() + old, v.end(), less);
        std::inplace_merge(v.begin(), v.begin() + old, v.end(), less);
    }
    return v;
}

int Octree::find(int level, long x, long y, long z) const {
    if (level < 0 || level >= int(levels.size())) return -1;
    const long cells = 1L << level;
    if (x < 0 || y < 0 || z < 0 || x >= cells || y >= cells || z >= cells) return -1;
    const int id = index[level].find(key(uint32_t(x), uint32_t(y), uint32_t(z)));
    if (id < 0) return -1;
    return gone_[id] ? -1 : id;
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
    const int level = nodes[n].level + 1;
    const bool keyed = perm_.empty();
    if (keyed && level > beeptree::KEY_LEVELS)
        throw std::runtime_error("octree: deeper than 21 levels (coincident patches?)");
    nodes[n].has_children = true;
    const int lo0 = nodes[n].lo, hi = nodes[n].hi;
    int ends[8];
    if (keyed) {
        // a child's items: the part of the parent's range whose key digit at
        // the child's level is its octant
        const int shift = 3 * (beeptree::KEY_LEVELS - level);
        const uint64_t* K = skeys_.data();
        int lo = lo0;
        for (int c = 0; c < 7; ++c) {
            lo = int(std::partition_point(K + lo, K + hi, [&](uint64_t k) { return int((k >> shift) & 7) <= c; }) - K);
            ends[c] = lo;
        }
        ends[7] = hi;
    } else {
        // each child receives the parent's items that fall in its octant
        // (Node child constructor, octree.h:120-131), in the parent's order
        const Vec3 pc = nodes[n].centre;
        const int m = hi - lo0;
        oct_.resize(size_t(m));
        tmp_.resize(size_t(m));
        int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < m; ++k) {
            const Vec3& q = (*pts_)[perm_[lo0 + k]];
            oct_[k] = static_cast<unsigned char>((q.x > pc.x ? 1 : 0) + (q.y > pc.y ? 2 : 0) + (q.z > pc.z ? 4 : 0));
            ++cnt[oct_[k]];
        }
        int at[8], run = 0;
        for (int c = 0; c < 8; ++c) {
            at[c] = run;
            run += cnt[c];
            ends[c] = lo0 + run;
        }
        for (int k = 0; k < m; ++k) tmp_[at[oct_[k]]++] = perm_[lo0 + k];
        std::copy(tmp_.begin(), tmp_.end(), perm_.begin() + lo0);
    }
    int lo = lo0;
    for (int c = 0; c < 8; ++c) {
        const OctNode& p = nodes[n];
        const double h = p.edge / 2.0;
        Vec3 cc = p.centre;
        cc.x += (c & 1) ? h / 2.0 : -h / 2.0;
        cc.y += (c & 2) ? h / 2.0 : -h / 2.0;
        cc.z += (c & 4) ? h / 2.0 : -h / 2.0;
        int id = new_node(level, p.ix * 2 + (c & 1), p.iy * 2 + ((c >> 1) & 1),
                          p.iz * 2 + ((c >> 2) & 1), cc, h, n);
        nodes[id].lo = lo;
        nodes[id].hi = ends[c];
        ncount_[id] = ends[c] - lo;
        nodes[n].children[c] = id;
        lo = ends[c];
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
    long r = long(ncount_[n]);
    for (int nb = 0; nb < 27; ++nb) {
        if (nb == 13) continue;
        int m = neighbour(n, nb);
        if (m >= 0) r += long(ncount_[m]);
    }
    return r;
}

long Octree::cached_nb(int n) {
    if (nodes[n].cached_nb < 0) nodes[n].cached_nb = count_nb(n);
    return nodes[n].cached_nb;
}

double Octree::interacts_per_item(long total_items) {
    long total = 0;
    // an integer sum: the walk order does not matter
    for (int level = 0; level < int(levels.size()) && level < max_depth; ++level)
        for (int id : levels[level]) {
            const OctNode& o = nodes[id];
            if (o.is_leaf && !o.is_shadow && !o.is_deleted)
                total += cached_nb(id) * long(o.count());
        }
    return double(total) / double(total_items);
}

void Octree::build(const std::vector<Vec3>& pts, const Vec3& c, double e, int nb) {
    skeys_.clear();
    perm_.resize(pts.size());
    for (size_t i = 0; i < pts.size(); ++i) perm_[i] = int(i);
    pts_ = &pts;
    grow(int(pts.size()), c, e, nb);
    pts_ = nullptr;
}

void Octree::build_sorted(const uint64_t* sorted_keys, int n, const Vec3& c, double e, int nb) {
    perm_.clear();
    skeys_.assign(sorted_keys, sorted_keys + n);
    for (int i = 1; i < n; ++i)
        if (skeys_[i] < skeys_[i - 1]) throw std::runtime_error("octree: keys not in ascending order");
    grow(n, c, e, nb);
}

void Octree::grow(int n, const Vec3& c, double e, int nb) {
    if (n <= 0) throw std::runtime_error("octree: no items");
    nodes.clear();
    nkey_.clear();
    ncount_.clear();
    gone_.clear();
    levels.clear();
    index.clear();
    sorted_.clear();
    leaf_nodes.clear();
    centre = c;
    edge = e;
    nbsize = nb;
    nodes.reserve(64 + size_t(n) / 4);
    nkey_.reserve(64 + size_t(n) / 4);
    new_node(0, 0, 0, 0, c, e, -1);
    nodes[0].hi = n;
    ncount_[0] = n;
    // relaxed insertion: a childless node splits when an item takes it past
    // 4 * nbsize, and later items go on to its children, so a node ends up
    // split exactly when it holds more than 4 * nbsize items in all (and is
    // above max_depth)
    for (size_t q = 0; q < nodes.size(); ++q) {
        if (long(nodes[q].count()) > 4L * nbsize && nodes[q].level < max_depth) {
            subdivide(int(q));
            nodes[q].is_leaf = false;
        }
    }

    // build_neighbourhoods: every non-shadow node with children gets a
    // complete set of same-level neighbours with children
    for (int level = 0; level < int(levels.size()); ++level) {
        std::vector<int> with_children;
        for (int id : ordered(level)) {
            const OctNode& o = nodes[id];
            if (o.has_children && !o.is_shadow) with_children.push_back(id);
        }
        for (int n : with_children) subdivide_neighbourhood(n);
    }

    // optimize_neighbourhoods
    long total_items = 0;
    for (int level = 0; level < int(levels.size()); ++level)
        for (int id : ordered(level)) {
            const OctNode& o = nodes[id];
            if (!o.is_shadow && o.is_leaf && !o.is_deleted) total_items += long(o.count());
        }
    double avg = interacts_per_item(total_items);
    while (std::floor(avg) > nbsize) {
        std::vector<int> fat;
        for (int level = 0; level < int(levels.size()); ++level)
            for (int id : ordered(level)) {
                const OctNode& o = nodes[id];
                if (!o.is_shadow && o.is_leaf && o.count() > 0 && cached_nb(id) > nbsize)
                    fat.push_back(id);
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
        for (int id : ordered(level)) {
            OctNode& o = nodes[id];
            if (o.count() == 0 && id != 0 && !o.removed) {
                o.removed = true;
                gone_[id] = 1;
                if (o.parent >= 0)
                    for (int c = 0; c < 8; ++c)
                        if (nodes[o.parent].children[c] == id) nodes[o.parent].children[c] = -1;
            }
        }

    for (int level = 0; level < int(levels.size()); ++level)
        for (int id : ordered(level)) {
            const OctNode& o = nodes[id];
            if (!o.removed && o.is_leaf && !o.is_shadow && o.count() > 0)
                leaf_nodes.push_back(id);
        }
    // sanity: the leaves' ranges tile the sorted items (every item in exactly
    // one leaf)
    std::vector<std::pair<int, int>> r;
    r.reserve(leaf_nodes.size());
    for (int id : leaf_nodes) r.push_back({nodes[id].lo, nodes[id].hi});
    std::sort(r.begin(), r.end());
    int at = 0;
    for (const auto& x : r) {
        if (x.first != at) throw std::runtime_error("octree: a patch is not in exactly one leaf");
        at = x.second;
    }
    if (at != n) throw std::runtime_error("octree: a patch is not in exactly one leaf");
}

void Octree::leaves_csr(std::vector<int>& leaf_ptr, std::vector<int>& leaf_idx,
                        std::vector<int>& nb_ptr, std::vector<int>& nb_idx) const {
    if (!skeys_.empty()) throw std::runtime_error("octree: built from keys, its items are on the device");
    // a box without children holds its items in ascending order; a box with
    // children had its range partitioned, so its items are sorted once
    std::vector<int> sorted_at(nodes.size(), -1);
    std::vector<int> pool;
    auto items = [&](int m, const int*& b, const int*& e) {
        const OctNode& o = nodes[m];
        if (!o.has_children) {
            b = perm_.data() + o.lo;
            e = perm_.data() + o.hi;
            return;
        }
        if (sorted_at[m] < 0) {
            sorted_at[m] = int(pool.size());
            pool.insert(pool.end(), perm_.begin() + o.lo, perm_.begin() + o.hi);
            std::sort(pool.begin() + sorted_at[m], pool.end());
        }
        b = pool.data() + sorted_at[m];
        e = b + o.count();
    };
    size_t n_items = 0, n_nb = 0;
    for (int n : leaf_nodes) {
        n_items += size_t(nodes[n].count());
        for (int nb = 0; nb < 27; ++nb) {
            const int m = neighbour(n, nb);
            if (m >= 0) n_nb += size_t(nodes[m].count());
        }
    }
    leaf_ptr.assign(1, 0);
    nb_ptr.assign(1, 0);
    leaf_idx.clear();
    nb_idx.clear();
    leaf_ptr.reserve(leaf_nodes.size() + 1);
    nb_ptr.reserve(leaf_nodes.size() + 1);
    leaf_idx.reserve(n_items);
    nb_idx.reserve(n_nb);
    pool.reserve(n_items);
    const int *b, *e;
    for (int n : leaf_nodes) {
        items(n, b, e);
        leaf_idx.insert(leaf_idx.end(), b, e);
        leaf_ptr.push_back(int(leaf_idx.size()));
        for (int nb = 0; nb < 27; ++nb) {
            const int m = neighbour(n, nb);
            if (m < 0 || nodes[m].count() == 0) continue;
            items(m, b, e);
            nb_idx.insert(nb_idx.end(), b, e);
        }
        nb_ptr.push_back(int(nb_idx.size()));
    }
}

void Octree::leaf_segments(std::vector<int>& leaf_level, std::vector<int>& seg_ptr,
                           std::vector<int>& seg_lo, std::vector<int>& seg_hi) const {
    leaf_level.clear();
    seg_ptr.assign(1, 0);
    seg_lo.clear();
    seg_hi.clear();
    for (int n : leaf_nodes) {
        const OctNode& o = nodes[n];
        leaf_level.push_back(o.level);
        seg_lo.push_back(o.lo);
        seg_hi.push_back(o.hi);
        for (int nb = 0; nb < 27; ++nb) {
            const int m = neighbour(n, nb);
            if (m < 0 || nodes[m].count() == 0) continue;
            seg_lo.push_back(nodes[m].lo);
            seg_hi.push_back(nodes[m].hi);
        }
        seg_ptr.push_back(int(seg_lo.size()));
    }
}

}  // namespace beep
