// Scene assembly and the host C ABI (include/beep_host.h).
//
// A scene is the reference's BEEP object seen from the hot path: solvent
// dielectric, kappa, rule orders, neighbourhood size and a list of rigid mesh
// instances.  build() places every instance's patches and points in universe
// coordinates (NodePatch::change_coordinate_frame), computes the bounding
// cube (BEEP::calculate_bounding_cube, beep.cpp:1652-1693) and the octree.
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/beep_host.h"
#include "host.hpp"

using namespace beep;

struct Instance {
    const beep_mesh* mesh;
    Quat rot;
    Vec3 offset;
    double Dprotein;
};

struct beep_scene {
    double Dsolvent, kappa;
    int quad, qual, nbsize;
    std::vector<Instance> inst;
    bool built = false;
    // flattened
    std::vector<Vec3> centroid, normal;
    std::vector<double> area, eps, gc;
    std::vector<int> patch_inst;
    std::vector<int> ptr[3];
    std::vector<QuadPt> pts[3];
    std::vector<double> charges;  // xyzq
    std::vector<int> charge_inst;
    std::vector<int> inst_patch_ptr, inst_charge_ptr;
    Octree tree;
    std::vector<int> leaf_ptr, leaf_idx, nb_ptr, nb_idx;
    Vec3 cube_centre;
    double cube_edge = 0;
};

static thread_local std::string g_err;

extern "C" const char* beep_host_last_error(void) { return g_err.c_str(); }

void beep_host_set_error(const char* msg) { g_err = msg; }

template <typename F>
static int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    } catch (...) {
        g_err = "unknown error";
        return -1;
    }
}

extern "C" int beep_mesh_icosphere(int subdiv, double radius, beep_mesh** out) {
    return guarded([&] {
        if (subdiv < 0 || subdiv > 9 || radius <= 0) throw std::runtime_error("bad icosphere args");
        auto m = std::make_unique<beep_mesh>();
        m->mesh = make_icosphere(subdiv, radius);
        m->mesh.finalize();
        *out = m.release();
    });
}

extern "C" int beep_mesh_star(int subdiv, double r0, int lmax, double sigma0, uint64_t seed,
                              beep_mesh** out) {
    return guarded([&] {
        if (subdiv < 0 || subdiv > 9 || r0 <= 0 || lmax < 2) throw std::runtime_error("bad star args");
        auto m = std::make_unique<beep_mesh>();
        m->shape = star_shape(r0, lmax, sigma0, seed);
        m->has_shape = true;
        m->mesh = make_star(subdiv, m->shape);
        m->mesh.finalize();
        *out = m.release();
    });
}

extern "C" int beep_mesh_set_charges(beep_mesh* m, int n, const double* xyzq) {
    return guarded([&] {
        if (!m || n < 0 || (n > 0 && !xyzq)) throw std::runtime_error("bad charges");
        m->mesh.charges.resize(n);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 4; ++k) m->mesh.charges[i][k] = xyzq[4 * i + k];
        m->mesh.finalize();
    });
}

extern "C" int beep_mesh_charges_ball(beep_mesh* m, int n, double rmax, double qmin, double qmax,
                                      uint64_t seed) {
    return guarded([&] {
        if (!m || n < 0 || rmax <= 0) throw std::runtime_error("bad ball charges");
        charges_uniform_ball(m->mesh, n, rmax, qmin, qmax, seed);
        m->mesh.finalize();
    });
}

extern "C" int beep_mesh_charges_walk(beep_mesh* m, int n, double step, double clearance,
                                      double qsigma, uint64_t seed) {
    return guarded([&] {
        if (!m || !m->has_shape) throw std::runtime_error("random-walk charges need a star mesh");
        if (n < 0 || step <= 0) throw std::runtime_error("bad walk charges");
        charges_random_walk(m->mesh, m->shape, n, step, clearance, qsigma, seed);
        m->mesh.finalize();
    });
}

extern "C" int beep_mesh_info(const beep_mesh* m, int* nv, int* nt, int* nc, double* centre3,
                              double* radius) {
    return guarded([&] {
        if (!m) throw std::runtime_error("null mesh");
        if (nv) *nv = int(m->mesh.verts.size());
        if (nt) *nt = int(m->mesh.tris.size());
        if (nc) *nc = int(m->mesh.charges.size());
        if (centre3) {
            centre3[0] = m->mesh.centre.x;
            centre3[1] = m->mesh.centre.y;
            centre3[2] = m->mesh.centre.z;
        }
        if (radius) *radius = m->mesh.radius;
    });
}

extern "C" int beep_mesh_vertices(const beep_mesh* m, double* xyz, int* tri3) {
    return guarded([&] {
        if (!m) throw std::runtime_error("null mesh");
        if (xyz)
            for (size_t i = 0; i < m->mesh.verts.size(); ++i) {
                xyz[3 * i] = m->mesh.verts[i].x;
                xyz[3 * i + 1] = m->mesh.verts[i].y;
                xyz[3 * i + 2] = m->mesh.verts[i].z;
            }
        if (tri3)
            for (size_t t = 0; t < m->mesh.tris.size(); ++t)
                for (int k = 0; k < 3; ++k) tri3[3 * t + k] = m->mesh.tris[t][k];
    });
}

extern "C" void beep_mesh_destroy(beep_mesh* m) { delete m; }

extern "C" int beep_scene_create(double Dsolvent, double kappa, int quad, int qual, int nbsize,
                                 beep_scene** out) {
    return guarded([&] {
        auto ok_rule = [](int r) { return r == 0 || r == 1 || r == 4 || r == 6 || r == 7 || r == 16; };
        if (!ok_rule(quad) || !ok_rule(qual)) throw std::runtime_error("rule must be 0,1,4,6,7,16");
        if (Dsolvent <= 0 || kappa < 0 || nbsize < 1) throw std::runtime_error("bad scene args");
        auto s = std::make_unique<beep_scene>();
        s->Dsolvent = Dsolvent;
        s->kappa = kappa;
        s->quad = quad;
        s->qual = qual;
        s->nbsize = nbsize;
        *out = s.release();
    });
}

extern "C" int beep_scene_add_instance(beep_scene* s, const beep_mesh* m, const double* q4,
                                       const double* off3, double Dprotein) {
    return guarded([&] {
        if (!s || !m || Dprotein <= 0) throw std::runtime_error("bad instance");
        Instance in;
        in.mesh = m;
        if (q4) {
            double n = std::sqrt(q4[0] * q4[0] + q4[1] * q4[1] + q4[2] * q4[2] + q4[3] * q4[3]);
            if (!(n > 0)) throw std::runtime_error("zero quaternion");
            in.rot = Quat{q4[0] / n, q4[1] / n, q4[2] / n, q4[3] / n};
        }
        in.offset = off3 ? Vec3(off3[0], off3[1], off3[2]) : m->mesh.centre;
        in.Dprotein = Dprotein;
        s->inst.push_back(in);
        s->built = false;
    });
}

// universe coordinates of a mesh-frame point: rotate about the mesh centre,
// then put the centre at the instance offset
static Vec3 place(const Instance& in, const Vec3& p) {
    return in.rot.rotate(p - in.mesh->mesh.centre) + in.offset;
}

extern "C" int beep_scene_build(beep_scene* s) {
    return guarded([&] {
        if (!s || s->inst.empty()) throw std::runtime_error("scene has no instances");
        s->centroid.clear(); s->normal.clear(); s->area.clear(); s->eps.clear(); s->gc.clear();
        s->patch_inst.clear(); s->charges.clear(); s->charge_inst.clear();
        s->inst_patch_ptr.assign(1, 0);
        s->inst_charge_ptr.assign(1, 0);
        for (int k = 0; k < 3; ++k) { s->ptr[k].assign(1, 0); s->pts[k].clear(); }
        const int orders[3] = {s->qual, s->quad, 4};
        std::vector<QuadPt> tmp;
        Vec3 lo(1e99, 1e99, 1e99), hi(-1e99, -1e99, -1e99);
        for (size_t ii = 0; ii < s->inst.size(); ++ii) {
            const Instance& in = s->inst[ii];
            const Mesh& m = in.mesh->mesh;
            const double eps = s->Dsolvent / in.Dprotein;
            for (size_t v = 0; v < m.verts.size(); ++v) {
                const Vec3 c = place(in, m.centroid[v]);
                const Vec3 n = in.rot.rotate(m.normal[v]);
                s->centroid.push_back(c);
                s->normal.push_back(n);
                s->area.push_back(m.area[v]);
                s->eps.push_back(eps);
                s->gc.push_back(0.5);
                s->patch_inst.push_back(int(ii));
                for (int k = 0; k < 3; ++k) {
                    tmp.clear();
                    if (orders[k] == 0) {
                        // single centroid point: weight 1 for qualocation,
                        // patch area for quadrature (node_patch.cpp:111-117)
                        tmp.push_back(QuadPt{c, n, k == 0 ? 1.0 : m.area[v]});
                    } else {
                        m.patch_points(int(v), rule_for_order(orders[k]), tmp);
                        for (auto& q : tmp) {
                            q.p = place(in, q.p);
                            q.n = in.rot.rotate(q.n);
                        }
                        // qualocation weights are renormalised to sum to 1
                        // (node_patch.cpp:221-226)
                        if (k == 0)
                            for (auto& q : tmp) q.w /= m.area[v];
                    }
                    s->pts[k].insert(s->pts[k].end(), tmp.begin(), tmp.end());
                    s->ptr[k].push_back(int(s->pts[k].size()));
                }
            }
            for (const auto& q : m.charges) {
                Vec3 p = place(in, Vec3(q[0], q[1], q[2]));
                s->charges.insert(s->charges.end(), {p.x, p.y, p.z, q[3]});
                s->charge_inst.push_back(int(ii));
            }
            s->inst_patch_ptr.push_back(int(s->centroid.size()));
            s->inst_charge_ptr.push_back(int(s->charge_inst.size()));
            const double r = m.radius;
            lo = Vec3(std::min(lo.x, in.offset.x - r), std::min(lo.y, in.offset.y - r),
                      std::min(lo.z, in.offset.z - r));
            hi = Vec3(std::max(hi.x, in.offset.x + r), std::max(hi.y, in.offset.y + r),
                      std::max(hi.z, in.offset.z + r));
        }
        Vec3 d = hi - lo;
        s->cube_edge = std::max(d.x, std::max(d.y, d.z));
        if (!(s->cube_edge > 0)) throw std::runtime_error("degenerate bounding cube");
        s->cube_centre = (hi + lo) / 2.0;
        s->tree.build(s->centroid, s->cube_centre, s->cube_edge, s->nbsize);
        s->tree.leaves_csr(s->leaf_ptr, s->leaf_idx, s->nb_ptr, s->nb_idx);
        s->built = true;
    });
}

static void need_built(const beep_scene* s) {
    if (!s || !s->built) throw std::runtime_error("scene not built");
}

extern "C" int beep_scene_get_sizes(const beep_scene* s, beep_scene_sizes* o) {
    return guarded([&] {
        need_built(s);
        std::memset(o, 0, sizeof(*o));
        o->n_patches = int(s->centroid.size());
        o->n_instances = int(s->inst.size());
        o->n_charges = int(s->charge_inst.size());
        o->n_qual = int(s->pts[0].size());
        o->n_quad = int(s->pts[1].size());
        o->n_galerkin = int(s->pts[2].size());
        o->n_leaves = int(s->leaf_ptr.size()) - 1;
        o->n_near_cols = int(s->nb_idx.size());
        long long pairs = 0;
        for (int l = 0; l < o->n_leaves; ++l)
            pairs += (long long)(s->leaf_ptr[l + 1] - s->leaf_ptr[l]) * (s->nb_ptr[l + 1] - s->nb_ptr[l]);
        o->n_near_pairs = pairs;
        o->tree_depth = int(s->tree.bottom_level());
        o->cube_centre[0] = s->cube_centre.x;
        o->cube_centre[1] = s->cube_centre.y;
        o->cube_centre[2] = s->cube_centre.z;
        o->cube_edge = s->cube_edge;
    });
}

extern "C" int beep_scene_patches(const beep_scene* s, double* c3, double* n3, double* area,
                                  double* eps, double* gc, int* inst) {
    return guarded([&] {
        need_built(s);
        for (size_t i = 0; i < s->centroid.size(); ++i) {
            if (c3) { c3[3 * i] = s->centroid[i].x; c3[3 * i + 1] = s->centroid[i].y; c3[3 * i + 2] = s->centroid[i].z; }
            if (n3) { n3[3 * i] = s->normal[i].x; n3[3 * i + 1] = s->normal[i].y; n3[3 * i + 2] = s->normal[i].z; }
            if (area) area[i] = s->area[i];
            if (eps) eps[i] = s->eps[i];
            if (gc) gc[i] = s->gc[i];
            if (inst) inst[i] = s->patch_inst[i];
        }
    });
}

extern "C" int beep_scene_points(const beep_scene* s, int kind, int* ptr, double* pts) {
    return guarded([&] {
        need_built(s);
        if (kind < 0 || kind > 2) throw std::runtime_error("point kind must be 0, 1 or 2");
        if (ptr) std::memcpy(ptr, s->ptr[kind].data(), s->ptr[kind].size() * sizeof(int));
        if (pts)
            for (size_t i = 0; i < s->pts[kind].size(); ++i) {
                const QuadPt& q = s->pts[kind][i];
                double* d = pts + 7 * i;
                d[0] = q.p.x; d[1] = q.p.y; d[2] = q.p.z;
                d[3] = q.n.x; d[4] = q.n.y; d[5] = q.n.z;
                d[6] = q.w;
            }
    });
}

extern "C" int beep_scene_tree(const beep_scene* s, int* leaf_ptr, int* leaf_idx, int* nb_ptr,
                               int* nb_idx) {
    return guarded([&] {
        need_built(s);
        auto cp = [](int* d, const std::vector<int>& v) {
            if (d) std::memcpy(d, v.data(), v.size() * sizeof(int));
        };
        cp(leaf_ptr, s->leaf_ptr);
        cp(leaf_idx, s->leaf_idx);
        cp(nb_ptr, s->nb_ptr);
        cp(nb_idx, s->nb_idx);
    });
}

extern "C" int beep_scene_charges(const beep_scene* s, double* xyzq, int* ci) {
    return guarded([&] {
        need_built(s);
        if (xyzq) std::memcpy(xyzq, s->charges.data(), s->charges.size() * sizeof(double));
        if (ci) std::memcpy(ci, s->charge_inst.data(), s->charge_inst.size() * sizeof(int));
    });
}

extern "C" int beep_scene_instances(const beep_scene* s, double* Dp, int* pp, int* cp) {
    return guarded([&] {
        need_built(s);
        for (size_t i = 0; i < s->inst.size(); ++i)
            if (Dp) Dp[i] = s->inst[i].Dprotein;
        if (pp) std::memcpy(pp, s->inst_patch_ptr.data(), s->inst_patch_ptr.size() * sizeof(int));
        if (cp) std::memcpy(cp, s->inst_charge_ptr.data(), s->inst_charge_ptr.size() * sizeof(int));
    });
}

extern "C" void beep_scene_destroy(beep_scene* s) { delete s; }
