// Host-side data model: meshes, node patches, quadrature point sets and the
// adaptive octree that decides which patch pairs are integrated explicitly.
//
// These are the caller-side objects of the BEM mat-vec (the reference keeps
// them in Mesh / MeshInstance / BasicNodePatch / Octree, src/bem/mesh.h,
// src/bem/node_patch.h, src/fmm/octree.h).  They produce the flat arrays that
// cross the device boundary declared in include/beep_hip.h.
#pragma once

#include <array>
#include <cstdint>
#include <cmath>
#include <map>
#include <string>
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace beep {

struct Vec3 {
    double x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(double s) const { return {x * s, y * s, z * s}; }
    Vec3 operator/(double s) const { return {x / s, y / s, z / s}; }
    Vec3& operator+=(const Vec3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    double dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
    Vec3 cross(const Vec3& o) const {
        return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
    }
    double norm() const { return std::sqrt(dot(*this)); }
    Vec3 unit() const { double n = norm(); return {x / n, y / n, z / n}; }
};

// One quadrature / qualocation / galerkin point: position, unit normal,
// weight.  Same content as the reference QuadPointT (src/bem/quad_point.h:82).
struct QuadPt {
    Vec3 p, n;
    double w;
};

// Triangle quadrature rule in barycentric (u, v) with weights summing to 1.
struct Rule {
    int order;
    std::vector<double> u, v, w;
};
// order 0 means "single centroid point" (node_patch.cpp:210-213, 243-246);
// the others are the tables of src/bem/triangle.h:75-96.
const Rule& rule_for_order(int order);

struct Quat {  // unit quaternion (w, x, y, z)
    double w = 1, x = 0, y = 0, z = 0;
    Vec3 rotate(const Vec3& v) const;
};

// Reference mesh (the reference's Mesh + BasicNodePatch list).
// A curved triangle patch with G1 blending across its edges: the PNG1
// construction (PNG1_Triangle, src/bem/png1.h:86-121 and 135-346), a cubic
// Bezier-like surface whose edge control points lie in the tangent planes of
// the corner normals and whose centre control point blends three variants,
// each built in the averaged tangent plane of one edge.  Barycentric
// (u, v, w = 1 - u - v): w weights vertex 1, u vertex 2, v vertex 3.
struct CurvedTri {
    // the 21 control points, 63 doubles, in the order curved.hpp lists:
    // corners b300 b030 b003; corner tangent-plane points p1_210 p1_201
    // p2_120 p2_021 p3_102 p3_012; the same carried to the next corner
    // p2_201 p2_102 p3_120 p3_210 p1_012 p1_021; centre variants per edge
    // c12_1 c12_2 c23_2 c23_3 c31_3 c31_1
    double ctl[63];
    // vertices, vertex normals n1..n3, and the planar normals of the
    // triangles across edges 1-2, 2-3, 3-1
    CurvedTri(const Vec3& v1, const Vec3& v2, const Vec3& v3, const Vec3& n1, const Vec3& n2,
              const Vec3& n3, const Vec3& n12, const Vec3& n23, const Vec3& n31);
    Vec3 point(double u, double v) const;
    // d/du and d/dv with the blended control points held fixed (as the
    // construction's jacobian and normal take them)
    void tangents(double u, double v, Vec3& du, Vec3& dv) const;
};

struct Mesh {
    std::vector<Vec3> verts;
    std::vector<std::array<int, 3>> tris;          // outward CCW
    std::vector<std::array<double, 4>> charges;    // x, y, z, q
    // triangles around each vertex, in increasing triangle order (flat:
    // vertex v's are vt_tri[vt_ptr[v] .. vt_ptr[v+1]-1], with the vertex's
    // corner 0, 1, 2 in each in vt_corner)
    std::vector<int> vt_ptr, vt_tri, vt_corner;
    // node patch per vertex (node_patch.cpp:55-130)
    std::vector<Vec3> centroid, normal;
    std::vector<double> area;                      // "bezier_area" of a planar patch
    Vec3 centre;                                   // centroid of charges (mesh.cpp:1654)
    bool centre_given = false;                     // read from a file (the .mtz centre member,
                                                   // mesh.cpp:223-233): kept by finalize
    std::vector<double> init_f, init_h;            // f and h per node patch from a .mtz fh member
                                                   // (mesh.cpp:630-650), empty if none
    bool curved = false;                           // PNG1 curved patches instead of planar ones
    std::vector<CurvedTri> ctri;                   // per triangle, when curved
    double radius = 0;                             // max |patch centroid - centre|

    void finalize();                               // adjacency, patches, centre, radius
    void finalize_charges();                       // centre and radius only
    // constant-basis points of one node patch (triangle.h:388-479): for each
    // adjacent triangle, the two sub-triangles (centre, vertex, edge midpoint)
    void patch_points(int vertex, const Rule& rule, std::vector<QuadPt>& out) const;
};

// Star-shaped surface r(theta, phi) = r0 * (1 + sum_{l=2..lmax} sum_m a_lm Y_lm)
// with real orthonormal Y_lm and a_lm ~ N(0, sigma0 / l) (standard deviation).
struct StarShape {
    double r0 = 1;
    int lmax = 1;
    std::vector<double> a;  // indexed l*l + l + m
    double radius(const Vec3& dir) const;
};
StarShape star_shape(double r0, int lmax, double sigma0, uint64_t seed);

Mesh make_icosphere(int subdivisions, double radius);
Mesh make_star(int subdivisions, const StarShape& shape);

// seeded charge generators
void charges_uniform_ball(Mesh& m, int n, double rmax, double qmin, double qmax, uint64_t seed);
void charges_random_walk(Mesh& m, const StarShape& shape, int n, double step, double clearance,
                         double qsigma, uint64_t seed);

// Deterministic random numbers (splitmix64) shared by the generators.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x1234567ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    Quat quaternion() {
        // uniform random rotation (Shoemake)
        double u1 = uniform(), u2 = uniform(), u3 = uniform();
        double a = std::sqrt(1 - u1), b = std::sqrt(u1);
        const double tp = 6.283185307179586;
        return Quat{a * std::sin(tp * u2), a * std::cos(tp * u2), b * std::sin(tp * u3),
                    b * std::cos(tp * u3)};
    }
};

// ---------------------------------------------------------------------------
// worker threads for the host build: OMP_NUM_THREADS if set, else the
// hardware concurrency, at most 16
inline int host_threads() {
    int n = int(std::thread::hardware_concurrency());
    if (const char* e = std::getenv("OMP_NUM_THREADS")) n = std::atoi(e);
    return std::max(1, std::min(16, n));
}

template <typename F>
void parallel_ranges(int n, F&& f) {
    const int nt = std::min(host_threads(), std::max(1, n / 2048));
    if (nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    std::exception_ptr err;
    std::mutex mu;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            try {
                f(int(int64_t(n) * t / nt), int(int64_t(n) * (t + 1) / nt));
            } catch (...) {
                std::lock_guard<std::mutex> g(mu);
                err = std::current_exception();
            }
        });
    for (auto& x : th) x.join();
    if (err) std::rethrow_exception(err);
}

// ---------------------------------------------------------------------------
// Adaptive octree over patch centroids.  Restates src/fmm/octree.h with the
// BEM-FMM policy of BEEP::init_fmm_octree (beep.cpp:1585-1641): relaxed
// insertion, build_neighbourhoods, optimize_neighbourhoods, remove_empty_nodes.
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
    // the node's items: positions lo .. hi-1 of the key-sorted items (the
    // items whose descent keys share the node's prefix)
    int lo = 0, hi = 0;
    int count() const { return hi - lo; }
};

struct Octree {
    Vec3 centre;
    double edge = 0;
    int nbsize = 1500;
    int max_depth = 20;
    std::vector<OctNode> nodes;
    std::vector<std::vector<int>> levels;          // per level: node ids, in creation order
    // per level: key -> node id by open addressing (lookups); keys are
    // below 2^63, so all ones marks an empty slot
    struct LevelIndex {
        std::vector<uint64_t> keys;
        std::vector<int> ids;
        size_t count = 0;
        void insert(uint64_t k, int id);
        int find(uint64_t k) const;
    };
    std::vector<LevelIndex> index;

    static uint64_t key(uint32_t x, uint32_t y, uint32_t z) {
        return (uint64_t(x) << 42) | (uint64_t(y) << 21) | uint64_t(z);
    }
    int find(int level, long x, long y, long z) const;
    int neighbour(int node, int nb) const;         // -1 when absent / outside
    // the tree over points pts in the cube (c, e)
    void build(const std::vector<Vec3>& pts, const Vec3& c, double e, int nb);
    // the tree from n descent keys in ascending order (the items' own order
    // among equal keys is not needed: only counts decide the tree)
    void build_sorted(const uint64_t* sorted_keys, int n, const Vec3& c, double e, int nb);

    // flattened result: leaves (non-shadow, non-empty) with their contents and
    // neighbourhood contents (27 same-level boxes, octree.h:627-648), each
    // box's items in ascending order (the reference's insertion order);
    // after build()
    std::vector<int> leaf_nodes;
    void leaves_csr(std::vector<int>& leaf_ptr, std::vector<int>& leaf_idx,
                    std::vector<int>& nb_ptr, std::vector<int>& nb_idx) const;
    // the same as ranges of the key-sorted items: per leaf its level, and
    // segments (lo, hi) of that level's boxes: the leaf's own first, then its
    // neighbourhood's non-empty boxes in neighbour order (seg_ptr: leaves + 1,
    // in segments); after either build
    void leaf_segments(std::vector<int>& leaf_level, std::vector<int>& seg_ptr,
                       std::vector<int>& seg_lo, std::vector<int>& seg_hi) const;
    long bottom_level() const { return long(levels.size()) - 1; }

  private:
    std::vector<uint64_t> skeys_;  // build_sorted(): the descent keys, ascending
    std::vector<int> perm_;        // build(): the items, each node's a range
    std::vector<Vec3> const* pts_ = nullptr;
    std::vector<unsigned char> oct_;  // octant of each item of the node being split
    std::vector<int> tmp_;
    void grow(int n, const Vec3& c, double e, int nb);
    // per level: node ids in key order, extended when the level has grown
    std::vector<std::vector<int>> sorted_;
    // per node, compact copies of what the neighbour walks read: the (x, y, z)
    // key within its level, the item count, whether it was removed
    std::vector<uint64_t> nkey_;
    std::vector<int> ncount_;
    std::vector<unsigned char> gone_;
    const std::vector<int>& ordered(int level);
    int new_node(int level, uint32_t x, uint32_t y, uint32_t z, const Vec3& c, double e, int parent);
    void subdivide(int n);
    void subdivide_neighbourhood(int n);
    long count_nb(int n) const;
    long cached_nb(int n);
    double interacts_per_item(long total_items);
};

}  // namespace beep

// Opaque handle behind beep_mesh* of include/beep_host.h.
// A fresh value from a process-wide counter: a mesh takes one when it is
// made and again whenever its charges change, so equal stamps mean equal
// surface and charges (a copy keeps its original's stamp).
inline uint64_t next_mesh_stamp() {
    static std::atomic<uint64_t> next{1};
    return next++;
}

struct beep_mesh {
    beep::Mesh mesh;
    bool has_shape = false;
    beep::StarShape shape;
    uint64_t stamp = next_mesh_stamp();
};

// Sets the message returned by beep_host_last_error() on this thread.
void beep_host_set_error(const char* msg);

// A rigid placement of a reference mesh (MeshInstance, mesh_instance.h:23-87).
struct Instance {
    const beep_mesh* mesh;
    uint64_t stamp = 0;  // the mesh's stamp when the instance was added
    beep::Quat rot;
    beep::Vec3 offset;
    double Dprotein;
};

// Opaque handle behind beep_scene* of include/beep_host.h: solver parameters,
// instances and, after beep_scene_build, the flattened patches, point sets
// (QuadPt = 7 doubles x, y, z, nx, ny, nz, w), charges and octree lists.
struct beep_scene {
    double Dsolvent, kappa;
    int quad, qual, nbsize;
    std::vector<Instance> inst;
    bool built = false;
    bool host_points = true;  // false: point sets are generated on the device
    bool with_tree = true;    // false: no octree (the near field comes from elsewhere)
    std::vector<beep::Vec3> centroid, normal;
    std::vector<double> area, eps, gc;
    std::vector<int> patch_inst;
    std::vector<int> ptr[3];
    std::vector<beep::QuadPt> pts[3];
    std::vector<double> charges;  // xyzq
    std::vector<int> charge_inst;
    std::vector<int> inst_patch_ptr, inst_charge_ptr;
    beep::Octree tree;
    std::vector<int> leaf_ptr, leaf_idx, nb_ptr, nb_idx;
    beep::Vec3 cube_centre;
    double cube_edge = 0;
};
static_assert(sizeof(beep::Vec3) == 3 * sizeof(double), "Vec3 is 3 packed doubles");
static_assert(sizeof(beep::QuadPt) == 7 * sizeof(double), "QuadPt is 7 packed doubles");
