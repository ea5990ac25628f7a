// Descent key of a point in the adaptive octree (host_octree.cpp), shared by
// the host build and the device build (octree.hip) so both place every point
// in the same boxes.  The key lists the octant chosen at each level below the
// root, three bits per level (x + 2y + 4z, the reference's child id - 1,
// octree.h:120-131), the first level in the top bits.  Each octant is decided
// as the reference's insertion decides it, by q > centre per axis, against the
// node centres computed as subdivision computes them: a child's centre is its
// parent's +- (parent edge / 2) / 2 per axis and its edge the parent's / 2.
// So the items of a node at level l are the items whose keys share the node's
// 3l-bit prefix, and sorting the keys puts every node's items in one range.
#ifndef BEEP_TREE_KEY_HPP
#define BEEP_TREE_KEY_HPP

#include <cstdint>

#if defined(__HIPCC__)
#define TREEKEY_FN __host__ __device__ inline
#else
#define TREEKEY_FN inline
#endif

namespace beeptree {

// levels a key resolves (63 bits); the octree refuses to go deeper
constexpr int KEY_LEVELS = 21;

TREEKEY_FN uint64_t descent_key(double px, double py, double pz, double cx, double cy, double cz,
                                double edge) {
    uint64_t k = 0;
    for (int l = 0; l < KEY_LEVELS; ++l) {
        const int o = (px > cx ? 1 : 0) + (py > cy ? 2 : 0) + (pz > cz ? 4 : 0);
        k = (k << 3) | uint64_t(o);
        const double h = edge / 2.0;
        cx += (o & 1) ? h / 2.0 : -h / 2.0;
        cy += (o & 2) ? h / 2.0 : -h / 2.0;
        cz += (o & 4) ? h / 2.0 : -h / 2.0;
        edge = h;
    }
    return k;
}

// the 3l-bit prefix of a key: the node at level l that holds the point
TREEKEY_FN uint64_t key_prefix(uint64_t k, int level) {
    return level == 0 ? 0 : k >> (3 * (KEY_LEVELS - level));
}

}  // namespace beeptree

#endif
