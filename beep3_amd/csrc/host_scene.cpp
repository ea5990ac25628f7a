// Scene assembly and the host C ABI (include/beep_host.h).
//
// A scene is the reference's BEEP object seen from the hot path: solvent
// dielectric, kappa, rule orders, neighbourhood size and a list of rigid mesh
// instances.  build() places every instance's patches and points in universe
// coordinates (NodePatch::change_coordinate_frame), computes the bounding
// cube (BEEP::calculate_bounding_cube, beep.cpp:1652-1693) and the octree.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>

#include "../../include/beep_host.h"
#include "host.hpp"

using namespace beep;

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
        m->mesh.finalize_charges();
        m->stamp = next_mesh_stamp();
    });
}

extern "C" int beep_mesh_set_curved(beep_mesh* m, int on) {
    return guarded([&] {
        if (!m) throw std::runtime_error("null mesh");
        if (m->mesh.curved == (on != 0)) return;
        m->mesh.curved = on != 0;
        m->mesh.finalize();  // patches (centroid, normal, area) on the new surface
        m->stamp = next_mesh_stamp();
    });
}

extern "C" int beep_mesh_curved_point(const beep_mesh* m, int tri, double u, double v, double* p3,
                                      double* n3) {
    return guarded([&] {
        if (!m || !p3 || !n3) throw std::runtime_error("null argument");
        if (!m->mesh.curved) throw std::runtime_error("the mesh has planar patches");
        if (tri < 0 || tri >= int(m->mesh.ctri.size())) throw std::runtime_error("no such triangle");
        const CurvedTri& c = m->mesh.ctri[size_t(tri)];
        const Vec3 p = c.point(u, v);
        Vec3 du, dv;
        c.tangents(u, v, du, dv);
        const Vec3 n = du.cross(dv).unit();
        p3[0] = p.x; p3[1] = p.y; p3[2] = p.z;
        n3[0] = n.x; n3[1] = n.y; n3[2] = n.z;
    });
}

extern "C" int beep_mesh_charges_ball(beep_mesh* m, int n, double rmax, double qmin, double qmax,
                                      uint64_t seed) {
    return guarded([&] {
        if (!m || n < 0 || rmax <= 0) throw std::runtime_error("bad ball charges");
        charges_uniform_ball(m->mesh, n, rmax, qmin, qmax, seed);
        m->mesh.finalize_charges();
        m->stamp = next_mesh_stamp();
    });
}

extern "C" int beep_mesh_charges_walk(beep_mesh* m, int n, double step, double clearance,
                                      double qsigma, uint64_t seed) {
    return guarded([&] {
        if (!m || !m->has_shape) throw std::runtime_error("random-walk charges need a star mesh");
        if (n < 0 || step <= 0) throw std::runtime_error("bad walk charges");
        charges_random_walk(m->mesh, m->shape, n, step, clearance, qsigma, seed);
        m->mesh.finalize_charges();
        m->stamp = next_mesh_stamp();
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

extern "C" int beep_mesh_clone(const beep_mesh* m, beep_mesh** out) {
    return guarded([&] {
        if (!m || !out) throw std::runtime_error("null argument");
        *out = new beep_mesh(*m);
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
        in.stamp = m->stamp;
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
        const int orders[3] = {s->qual, s->quad, 4};
        // patch p = vertex v of instance ii, instances in order
        std::vector<std::pair<int, int>> plist;
        s->inst_patch_ptr.assign(1, 0);
        s->inst_charge_ptr.assign(1, 0);
        s->charges.clear();
        s->charge_inst.clear();
        Vec3 lo(1e99, 1e99, 1e99), hi(-1e99, -1e99, -1e99);
        for (size_t ii = 0; ii < s->inst.size(); ++ii) {
            const Instance& in = s->inst[ii];
            const Mesh& m = in.mesh->mesh;
            for (size_t v = 0; v < m.verts.size(); ++v) plist.push_back({int(ii), int(v)});
            for (const auto& q : m.charges) {
                Vec3 p = place(in, Vec3(q[0], q[1], q[2]));
                s->charges.insert(s->charges.end(), {p.x, p.y, p.z, q[3]});
                s->charge_inst.push_back(int(ii));
            }
            s->inst_patch_ptr.push_back(int(plist.size()));
            s->inst_charge_ptr.push_back(int(s->charge_inst.size()));
            const double r = m.radius;
            lo = Vec3(std::min(lo.x, in.offset.x - r), std::min(lo.y, in.offset.y - r),
                      std::min(lo.z, in.offset.z - r));
            hi = Vec3(std::max(hi.x, in.offset.x + r), std::max(hi.y, in.offset.y + r),
                      std::max(hi.z, in.offset.z + r));
        }
        const int N = int(plist.size());
        // point counts: one centroid point for rule 0, else k points on each of
        // the 2 x valence sub-triangles of the patch
        for (int k = 0; k < 3; ++k) {
            s->ptr[k].assign(size_t(N) + 1, 0);
            const size_t per = orders[k] == 0 ? 0 : rule_for_order(orders[k]).w.size();
            for (int p = 0; p < N; ++p) {
                const Mesh& m = s->inst[plist[p].first].mesh->mesh;
                const int v = plist[p].second;
                const size_t c = per == 0 ? 1 : 2 * size_t(m.vt_ptr[v + 1] - m.vt_ptr[v]) * per;
                if (size_t(s->ptr[k][p]) + c > size_t(0x7fffffff))
                    throw std::runtime_error("too many points for 32-bit offsets");
                s->ptr[k][p + 1] = s->ptr[k][p] + int(c);
            }
            s->pts[k].clear();
            if (s->host_points) s->pts[k].resize(size_t(s->ptr[k][N]));
        }
        s->centroid.resize(N); s->normal.resize(N); s->area.resize(N);
        s->eps.resize(N); s->gc.resize(N); s->patch_inst.resize(N);
        parallel_ranges(N, [&](int p0, int p1) {
            std::vector<QuadPt> tmp;
            for (int p = p0; p < p1; ++p) {
                const Instance& in = s->inst[plist[p].first];
                const Mesh& m = in.mesh->mesh;
                const int v = plist[p].second;
                const Vec3 c = place(in, m.centroid[v]);
                const Vec3 n = in.rot.rotate(m.normal[v]);
                s->centroid[p] = c;
                s->normal[p] = n;
                s->area[p] = m.area[v];
                s->eps[p] = s->Dsolvent / in.Dprotein;
                s->gc[p] = 0.5;
                s->patch_inst[p] = plist[p].first;
                for (int k = 0; k < 3 && s->host_points; ++k) {
                    QuadPt* out = s->pts[k].data() + s->ptr[k][p];
                    if (orders[k] == 0) {
                        // single centroid point: weight 1 for qualocation,
                        // patch area for quadrature (node_patch.cpp:111-117)
                        out[0] = QuadPt{c, n, k == 0 ? 1.0 : m.area[v]};
                        continue;
                    }
                    tmp.clear();
                    m.patch_points(v, rule_for_order(orders[k]), tmp);
                    for (size_t j = 0; j < tmp.size(); ++j) {
                        QuadPt q = tmp[j];
                        q.p = place(in, q.p);
                        q.n = in.rot.rotate(q.n);
                        // qualocation weights are renormalised to sum to 1
                        // (node_patch.cpp:221-226)
                        if (k == 0) q.w /= m.area[v];
                        out[j] = q;
                    }
                }
            }
        });
        Vec3 d = hi - lo;
        s->cube_edge = std::max(d.x, std::max(d.y, d.z));
        if (!(s->cube_edge > 0)) throw std::runtime_error("degenerate bounding cube");
        s->cube_centre = (hi + lo) / 2.0;
        if (s->with_tree) {
            s->tree.build(s->centroid, s->cube_centre, s->cube_edge, s->nbsize);
            s->tree.leaves_csr(s->leaf_ptr, s->leaf_idx, s->nb_ptr, s->nb_idx);
        } else {
            s->leaf_ptr.assign(1, 0);
            s->leaf_idx.clear();
            s->nb_ptr.assign(1, 0);
            s->nb_idx.clear();
        }
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
        o->n_galerkin = s->ptr[2].empty() ? 0 : s->ptr[2].back();
        o->n_leaves = int(s->leaf_ptr.size()) - 1;
        o->n_near_cols = int(s->nb_idx.size());
        long long pairs = 0;
        for (int l = 0; l < o->n_leaves; ++l)
            pairs += (long long)(s->leaf_ptr[l + 1] - s->leaf_ptr[l]) * (s->nb_ptr[l + 1] - s->nb_ptr[l]);
        o->n_near_pairs = pairs;
        o->tree_depth = s->with_tree ? int(s->tree.bottom_level()) : 0;
        o->cube_centre[0] = s->cube_centre.x;
        o->cube_centre[1] = s->cube_centre.y;
        o->cube_centre[2] = s->cube_centre.z;
        o->cube_edge = s->cube_edge;
        for (const Instance& in : s->inst) {
            o->n_triangles += int(in.mesh->mesh.tris.size());
            o->n_vertex_triangles += in.mesh->mesh.vt_ptr.empty() ? 0 : in.mesh->mesh.vt_ptr.back();
        }
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
        if (!s->host_points && pts)
            throw std::runtime_error("this scene leaves its point sets to the device");
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

extern "C" int beep_scene_set_host_points(beep_scene* s, int on) {
    return guarded([&] {
        if (!s) throw std::runtime_error("null scene");
        s->host_points = on != 0;
        s->built = false;
    });
}

extern "C" int beep_scene_set_tree(beep_scene* s, int on) {
    return guarded([&] {
        if (!s) throw std::runtime_error("null scene");
        s->with_tree = on != 0;
        s->built = false;
    });
}

extern "C" int beep_scene_surface(const beep_scene* s, double* verts3, int* tris3, int* vt_ptr,
                                  int* vt_tri, int* vt_corner) {
    return guarded([&] {
        need_built(s);
        size_t vbase = 0, tbase = 0, kbase = 0;
        if (vt_ptr) vt_ptr[0] = 0;
        for (const Instance& in : s->inst) {
            const Mesh& m = in.mesh->mesh;
            const int nv = int(m.verts.size()), nt = int(m.tris.size());
            // every vertex with the triangles around it, in the order its
            // points are made; the mesh's lists shifted to scene numbering
            parallel_ranges(nv, [&](int v0, int v1) {
                for (int v = v0; v < v1; ++v) {
                    if (verts3) {
                        const Vec3 q = place(in, m.verts[v]);
                        verts3[3 * (vbase + v)] = q.x;
                        verts3[3 * (vbase + v) + 1] = q.y;
                        verts3[3 * (vbase + v) + 2] = q.z;
                    }
                    if (vt_ptr) vt_ptr[vbase + v + 1] = int(kbase) + m.vt_ptr[v + 1];
                    for (int k = m.vt_ptr[v]; k < m.vt_ptr[v + 1]; ++k) {
                        if (vt_tri) vt_tri[kbase + k] = int(tbase) + m.vt_tri[k];
                        if (vt_corner) vt_corner[kbase + k] = m.vt_corner[k];
                    }
                }
            });
            if (tris3)
                parallel_ranges(nt, [&](int t0, int t1) {
                    for (int t = t0; t < t1; ++t)
                        for (int c = 0; c < 3; ++c) tris3[3 * (tbase + t) + c] = int(vbase) + m.tris[t][c];
                });
            vbase += size_t(nv);
            tbase += size_t(nt);
            kbase += size_t(m.vt_ptr[nv]);
        }
    });
}

extern "C" int beep_scene_surface_curved(const beep_scene* s, int* n_curved, int* ctl_idx, double* ctl) {
    return guarded([&] {
        need_built(s);
        // a rigid motion of the control points moves the surface with them
        // (every surface point is an affine combination of them), so placed
        // control points give the placed surface
        size_t tbase = 0;
        int nc = 0;
        for (const Instance& in : s->inst) {
            const Mesh& m = in.mesh->mesh;
            for (size_t t = 0; t < m.tris.size(); ++t) {
                if (!m.curved) {
                    if (ctl_idx) ctl_idx[tbase + t] = -1;
                    continue;
                }
                if (ctl_idx) ctl_idx[tbase + t] = nc;
                if (ctl)
                    for (int i = 0; i < 21; ++i) {
                        const double* c = m.ctri[t].ctl + 3 * i;
                        const Vec3 q = place(in, Vec3(c[0], c[1], c[2]));
                        double* d = ctl + 63 * size_t(nc) + 3 * i;
                        d[0] = q.x;
                        d[1] = q.y;
                        d[2] = q.z;
                    }
                ++nc;
            }
            tbase += m.tris.size();
        }
        if (n_curved) *n_curved = nc;
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
